"""The uneven batch at full size: 4096 circuits of 4-16 qubits and 10-60
moments of {H, X^t, Y^t, Z^t, CNOT, CZ^t, ISWAP^t, FSim}, 30% symbolic, 1-8
Pauli strings of weight <= 4 (padded to 8 with empty sums and zero
downstream weights), expectation + adjoint gradient from one persistent
launch (``expectation_and_gradient``).

Checks (semantics of ``core/ops/tfq_adj_grad_op.cc:190-278``):
* 65 rows, five of every width 4..16, against the complex128 oracle at
  1e-5 absolute, for both derivative modes (TFQ's float32 finite-difference
  gradient gates and the exact derivative);
* every row of the batched launch against the same plan launched one circuit
  per launch (1e-6: the gradients are sums of FP64 atomics, so only their
  order differs);
* every row against the float32 C++ restatement of TFQ's CPU schedule
  (``oracle/cpu_tfq.cc``) at 5e-5: both sides compute in float32 with
  different fusion and summation order, over sums of up to ~200 gate terms.
"""

import os
import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import serialize as S
from quantum_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def c4():
    _say("C4 building 4096 uneven circuits")
    circuits, names, values, psums, weights = W.config4_uneven()
    widths = np.array([len(c.all_qubits()) for c in circuits])
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(psums)
    rows = []
    for w in range(4, 17):
        idx = np.where(widths == w)[0]
        assert len(idx) >= 5, (w, len(idx))
        rows += list(idx[:5])
    return progs, names, values, ps, weights, widths, np.array(sorted(rows))


@pytest.fixture(scope="module")
def c4_gpu(c4):
    from quantum_amd import tfq_adj_grad_op as A
    progs, names, values, ps, weights, _, _ = c4
    B, O_ = len(progs), ps.shape[1]
    res = {}
    for mode in ("tfq", "analytic"):
        _say(f"C4 batched launch, derivative={mode}")
        e, g = A.expectation_and_gradient(progs, names, values, ps, weights, derivative=mode)
        res[mode] = (e.cpu().numpy(), g.cpu().numpy())
    assert res["tfq"][0].shape == (B, O_) and res["tfq"][1].shape == (B, len(names))
    return res


def test_c4_stratified_rows_match_oracle_tfq_gradient_gates(c4, c4_gpu):
    progs, names, values, ps, weights, widths, rows = c4
    e, g = c4_gpu["tfq"]
    assert sorted(set(widths[rows])) == list(range(4, 17))
    _say(f"C4 oracle on {len(rows)} rows (tfq gradient gates)")
    sub = [progs[i] for i in rows]
    want_e = O.simulate_expectation(sub, names, values[rows], ps[rows])
    want_g = O.adjoint_gradient(sub, names, values[rows], ps[rows], weights[rows])
    np.testing.assert_allclose(e[rows], want_e, atol=TOL)
    np.testing.assert_allclose(g[rows], want_g, atol=TOL)


def test_c4_stratified_rows_match_oracle_exact_derivative(c4, c4_gpu):
    progs, names, values, ps, weights, widths, rows = c4
    e, g = c4_gpu["analytic"]
    _say(f"C4 oracle on {len(rows)} rows (exact derivative)")
    sub = [progs[i] for i in rows]
    want_g = O.adjoint_gradient(sub, names, values[rows], ps[rows], weights[rows],
                                mode="analytic")
    np.testing.assert_allclose(e, c4_gpu["tfq"][0], atol=0)
    np.testing.assert_allclose(g[rows], want_g, atol=TOL)
    # The exact and TFQ finite-difference gradients differ by O(eps^2) per gate.
    np.testing.assert_allclose(g, c4_gpu["tfq"][1], atol=5e-3)


def test_c4_every_row_batched_equals_one_circuit_per_launch(c4, c4_gpu):
    import torch
    from quantum_amd import tfq_adj_grad_op as A
    from quantum_amd.tfq_simulate_ops import _device_and_stream
    progs, names, values, ps, weights, _, _ = c4
    plan, shape = A.plan_adj_grad(progs, names, values, ps, weights, with_expectation=True)
    dev, stream = _device_and_stream()
    plan.upload(dev, stream)
    batched = torch.empty(shape, dtype=torch.float32, device=dev)
    plain = torch.empty(shape, dtype=torch.float32, device=dev)
    _say("C4 batched launch (plan reused)")
    plan.launch(batched, stream)
    _say("C4 one launch per circuit")
    plan.launch(plain, stream, plain=True)
    torch.cuda.synchronize()
    B, O_ = len(progs), ps.shape[1]
    b = batched.cpu().numpy()
    np.testing.assert_allclose(plain.cpu().numpy(), b, atol=1e-6)
    e, g = c4_gpu["tfq"]
    np.testing.assert_allclose(b[:B * O_].reshape(B, O_), e, atol=1e-6)
    np.testing.assert_allclose(b[B * O_:].reshape(B, -1), g, atol=1e-6)


def test_c4_every_row_matches_cpu_restatement(c4, c4_gpu):
    from oracle import cpu_tfq
    progs, names, values, ps, weights, _, _ = c4
    e, g = c4_gpu["tfq"]
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    _say(f"C4 lowering 4096 circuits for the CPU restatement ({cores} workers)")
    chunks = cpu_tfq.lower_parallel(progs, names, values, ps, workers=cores)
    _say("C4 CPU restatement: expectation + adjoint gradient")
    ce = cpu_tfq.run_chunks(chunks, 0, threads=cores)
    cg = cpu_tfq.run_chunks(chunks, 1, downstream=weights, threads=cores)
    np.testing.assert_allclose(e, ce, atol=5e-5)
    np.testing.assert_allclose(g, cg, atol=5e-5)
