"""The autotuner times every candidate setting on the device, keeps the
fastest active, caches it per key, and the chosen setting's results still
match the oracle."""

import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import circuits as C
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


@pytest.fixture(autouse=True)
def _restore():
    from quantum_amd import autotune as AT
    yield
    AT.reset()


def _batch():
    rng = np.random.default_rng(31)
    names = [f"s{i}" for i in range(4)]
    circuits, psums = [], []
    for n in (5, 9, 12, 14, 7, 13):
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, 2 * n, symbols=names, p_control=0.1)
                        + C.Circuit([C.I(x) for x in q]))
        psums.append([W.random_pauli_sum(rng, q, 3, max_weight=min(4, n))])
    vals = rng.uniform(-2, 2, size=(len(circuits), 4)).astype(np.float32)
    g = np.ones((len(circuits), 1), np.float32)
    return S.convert_to_tensor(circuits), names, vals, S.convert_to_tensor(psums), g


def test_autotune_picks_fastest_and_keeps_parity():
    import torch
    from quantum_amd import _lib
    from quantum_amd import autotune as AT
    from quantum_amd import tfq_adj_grad_op as A
    progs, names, vals, ps, g = _batch()

    def build():
        plan, shape = A.plan_adj_grad(progs, names, vals, ps, g, with_expectation=True)
        return plan, shape, torch.float32

    _say("autotune over the batch candidates")
    res = AT.autotune(build, AT.BATCH_CANDIDATES, key="mixed-batch", steps=2, warmup=1)
    assert len(res["table"]) == len(AT.BATCH_CANDIDATES)
    assert all(r["ms"] > 0 for r in res["table"])
    fastest = min(res["table"], key=lambda r: r["ms"])
    assert res["best"] == fastest["setting"]
    # Cached: the same key returns the same result without timing again.
    assert AT.autotune(build, AT.BATCH_CANDIDATES, key="mixed-batch") is res
    # The chosen setting is the active one; its results match the oracle.
    exp, grad = A.expectation_and_gradient(progs, names, vals, ps, g)
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g)
    np.testing.assert_allclose(exp.cpu().numpy(), ref_e, atol=1e-5 * max(1, np.abs(ref_e).max()))
    np.testing.assert_allclose(grad.cpu().numpy(), ref_g, atol=1e-5 * max(1, np.abs(ref_g).max()))
    _lib.set_fusion_width(0)
