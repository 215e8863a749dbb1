"""Circuits of 14-16 qubits on the HBM-streaming path (the streaming threshold
lowered to 14): expectation, adjoint gradient with expectation, and state
against the complex128 oracle at 1e-5, one circuit per plan and a mixed batch.
The default planner runs these widths inside the batched kernel; the lowered
threshold is what a lone wide circuit uses on its own."""

import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import _lib
from quantum_amd import circuits as C
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
    _lib.set_streaming_min_qubits(14)
    yield
    _lib.set_streaming_min_qubits(-1)


def _batch(seed, widths):
    rng = np.random.default_rng(seed)
    names = [f"s{i}" for i in range(4)]
    circuits, sums = [], []
    for n in widths:
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, 2 * n, symbols=names, p_symbol=0.5,
                                         p_control=0.1) + C.Circuit([C.I(x) for x in q]))
        sums.append([W.random_pauli_sum(rng, q, 3, max_weight=4) for _ in range(2)])
    B = len(widths)
    return (S.convert_to_tensor(circuits), names, rng.uniform(-3, 3, size=(B, 4)).astype(np.float32),
            S.convert_to_tensor(sums), rng.uniform(-1, 1, size=(B, 2)).astype(np.float32))


@pytest.mark.parametrize("widths", [[14], [15], [16], [16, 14, 15]])
def test_streaming_expectation_and_adjoint(widths):
    from quantum_amd import tfq_adj_grad_op as A
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, vals, ps, g = _batch(sum(widths), widths)
    plan, _ = A.plan_adj_grad(progs, names, vals, ps, g, with_expectation=True)
    assert plan.stats()["streaming"], "threshold 14 must select the streaming path"
    del plan
    _say(f"streaming expectation+adjoint widths={widths}")
    e = sim.tfq_simulate_expectation(progs, names, vals, ps).cpu().numpy()
    ee, gg = A.expectation_and_gradient(progs, names, vals, ps, g, derivative="analytic")
    eo = O.simulate_expectation(list(progs), names, vals, ps)
    go = O.adjoint_gradient(list(progs), names, vals, ps, g, mode="analytic")
    np.testing.assert_allclose(e, eo, atol=TOL * max(1, np.abs(eo).max()))
    np.testing.assert_allclose(ee.cpu().numpy(), eo, atol=TOL * max(1, np.abs(eo).max()))
    np.testing.assert_allclose(gg.cpu().numpy(), go, atol=TOL * max(1, np.abs(go).max()))


def test_streaming_state():
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, vals, _, _ = _batch(3, [15, 14])
    _say("streaming state widths=[15, 14]")
    st = sim.tfq_simulate_state(progs, names, vals).cpu().numpy()
    so = O.simulate_state(list(progs), names, vals)
    np.testing.assert_allclose(st, so, atol=TOL)
