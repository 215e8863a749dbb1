"""The benchmark configurations other than the uneven batch, at their full
sizes, as parity cases.

* C1 (64 circuits, 10 qubits, 600 symbols): every expectation against the
  complex128 oracle (1e-5).
* C2 (30 qubits, 5x6 grid, 14 cycles): beyond the oracle's reach, so checked
  through size-independent properties: the 1e6-shot sampled expectation of the
  49-term sum ZZ and the ZZ estimate from 1e6 samples both agree with the exact
  expectation within their statistical error; every sample is a valid 30-bit
  string.
* C3 (26 qubits, 416 symbols): the adjoint gradient against parameter shift
  (rotation gates: [E(t + pi/2) - E(t - pi/2)] / 2) and central differences
  (h = 1e-3) computed with the exact expectation op; TFQ's finite-difference
  gradient gates agree with the exact derivative.
"""

import math
import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import serialize as S
from quantum_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_config1_full_batch_matches_oracle():
    from quantum_amd import tfq_simulate_ops as sim
    circuits, names, values, ops = W.config1_line_rotations()
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(ops)
    _say("C1 64 x 10 qubits")
    got = sim.tfq_simulate_expectation(progs, names, values, ps).cpu().numpy()
    want = O.simulate_expectation(list(progs), names, values, ps)
    np.testing.assert_allclose(got, want, atol=1e-5)


@pytest.fixture(scope="module")
def grid30():
    circuit, zz = W.config2_grid()
    return S.convert_to_tensor([circuit]), S.convert_to_tensor([[zz]]), len(zz.terms)


def test_config2_sampled_expectation_matches_exact(grid30):
    from quantum_amd import tfq_simulate_ops as sim
    progs, ps, nterms = grid30
    _say("C2 exact <sum ZZ> at 30 qubits")
    exact = float(sim.tfq_simulate_expectation(progs, [], np.zeros((1, 0)), ps).cpu()[0, 0])
    _say("C2 sampled expectation, 1e6 shots per term")
    est = float(sim.tfq_simulate_sampled_expectation(progs, [], np.zeros((1, 0)), ps,
                                                      np.full((1, 1), 10 ** 6, np.int32),
                                                      seed=1234).cpu()[0, 0])
    # Independent draws per term: std of the sum <= sqrt(49) / sqrt(1e6) = 7e-3.
    assert abs(est - exact) < 5 * math.sqrt(nterms) / 1e3, (est, exact)


def test_config2_samples_reproduce_the_zz_expectation(grid30):
    import torch
    from quantum_amd import tfq_simulate_ops as sim
    from quantum_amd import workloads as Wl
    progs, ps, nterms = grid30
    exact = float(sim.tfq_simulate_expectation(progs, [], np.zeros((1, 0)), ps).cpu()[0, 0])
    _say("C2 1e6 samples of 30 qubits")
    s = sim.tfq_simulate_samples(progs, [], np.zeros((1, 0)), [10 ** 6], seed=99)
    assert tuple(s.shape) == (1, 10 ** 6, 30)
    bits = s[0]
    assert bool(((bits == 0) | (bits == 1)).all())
    # Qubits are sorted (row, col); sample column j is qubit j.
    z = 1.0 - 2.0 * bits.to(torch.float32)
    circuit, zz = Wl.config2_grid()
    qs = sorted({q for term in zz.terms for q in term.paulis}, key=lambda q: (q.row, q.col))
    assert len(qs) == 30
    col = {q: j for j, q in enumerate(qs)}
    est = 0.0
    for term in zz.terms:
        a, b = [col[q] for q in term.paulis]
        est += float(term.coefficient.real) * float((z[:, a] * z[:, b]).mean())
    # Terms share shots; |sum| error <= sum of per-term errors (each std <= 1e-3).
    assert abs(est - exact) < 5 * nterms / 1e3, (est, exact)


@pytest.fixture(scope="module")
def hea26():
    circuit, names, values, ham = W.config3_hea()
    return S.convert_to_tensor([circuit] * 2), names, values[:2], S.convert_to_tensor([[ham]] * 2)


def test_config3_adjoint_matches_parameter_shift_and_central_difference(hea26):
    from quantum_amd import tfq_adj_grad_op as adj
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, values, ps = hea26
    ones = np.ones((2, 1), np.float32)
    _say("C3 adjoint gradient, 26 qubits, 2 x 416 symbols")
    g = adj.tfq_adj_grad(progs, names, values, ps, ones, derivative="analytic").cpu().numpy()
    g_tfq = adj.tfq_adj_grad(progs, names, values, ps, ones).cpu().numpy()
    assert g.shape == (2, 416)
    # TFQ's float32 finite-difference gradient gates track the exact derivative.
    np.testing.assert_allclose(g_tfq, g, atol=5e-3)
    cols = [0, 13, 51, 200, 415]
    for c in cols:
        vp = values.astype(np.float64).copy()
        vm = vp.copy()
        vp[:, c] += math.pi / 2
        vm[:, c] -= math.pi / 2
        _say(f"C3 parameter shift, symbol {c}")
        ep = sim.tfq_simulate_expectation(progs, names, vp.astype(np.float32), ps).cpu().numpy()
        em = sim.tfq_simulate_expectation(progs, names, vm.astype(np.float32), ps).cpu().numpy()
        np.testing.assert_allclose(g[:, c], (ep - em)[:, 0] / 2, atol=1e-4)
    for c in cols[:2]:
        h = 1e-3
        vp = values.astype(np.float64).copy()
        vm = vp.copy()
        vp[:, c] += h
        vm[:, c] -= h
        _say(f"C3 central difference, symbol {c}")
        ep = sim.tfq_simulate_expectation(progs, names, vp.astype(np.float32), ps).cpu().numpy()
        em = sim.tfq_simulate_expectation(progs, names, vm.astype(np.float32), ps).cpu().numpy()
        # float32 expectations: |dE| resolution ~1e-6 / (2h) plus the symbol's
        # float32 rounding.
        np.testing.assert_allclose(g[:, c], (ep - em)[:, 0] / (2 * h), atol=2e-2)
