"""Plan reuse: a plan of a heterogeneous batch rebound to new symbol values
and new downstream gradients equals a fresh plan (expectations bit for bit,
gradients to 1e-6: FP64 atomics add in varying order) and the complex128
oracle (1e-5); the op-path cache reuses plans across calls and drops them
when planner settings change."""

import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import _lib
from quantum_amd import circuits as C
from quantum_amd import plan_cache
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


def _batch(seed, widths):
    rng = np.random.default_rng(seed)
    names = [f"s{i}" for i in range(5)]
    circuits, sums = [], []
    for n in widths:
        if n == 0:
            circuits.append(C.Circuit())
            sums.append([C.PauliSum([]), C.PauliSum([])])
            continue
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, 3 * n, symbols=names, p_symbol=0.6,
                                         p_control=0.15) + C.Circuit([C.I(x) for x in q]))
        sums.append([W.random_pauli_sum(rng, q, 3, max_weight=min(4, n)) for _ in range(2)])
    return S.convert_to_tensor(circuits), names, S.convert_to_tensor(sums), rng


def _launch(plan, shape):
    import torch
    from quantum_amd.tfq_simulate_ops import _device_and_stream
    dev, stream = _device_and_stream()
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    plan.launch(out, stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


# Wave-mode (4-10 qubits, with the threshold lowered), LDS-resident (11-13),
# tiled (14-16) and empty programs in one batch.
@pytest.mark.parametrize("wave_min", [0, -1])
def test_rebound_heterogeneous_plan_equals_fresh_plan(wave_min):
    from quantum_amd import tfq_adj_grad_op as A
    progs, names, ps, rng = _batch(41 + wave_min, [5, 9, 0, 12, 14, 4, 16, 7, 15, 10])
    B = len(progs)
    v1, v2 = (rng.uniform(-3, 3, size=(B, 5)).astype(np.float32) for _ in range(2))
    g1 = rng.uniform(-1, 1, size=(B, 2)).astype(np.float32)
    g2 = rng.uniform(-1, 1, size=(B, 2)).astype(np.float32)
    g2[1, 0] = 0.0
    g2[3, 1] = 3e-6  # below TFQ's 1e-5 drop threshold once multiplied by |c| <= 1
    _lib.set_wave_min_circuits(wave_min)
    try:
        plan, shape = A.plan_adj_grad(progs, names, v1, ps, g1, with_expectation=True)
        plan.rebind(v2)
        plan.rebind_downstream(g2)
        _say(f"rebound heterogeneous plan, wave_min={wave_min}")
        got = _launch(plan, shape)
        fresh, _ = A.plan_adj_grad(progs, names, v2, ps, g2, with_expectation=True)
        want = _launch(fresh, shape)
    finally:
        _lib.set_wave_min_circuits(-1)
    np.testing.assert_array_equal(got[:B * 2], want[:B * 2])
    np.testing.assert_allclose(got[B * 2:], want[B * 2:], atol=1e-6)
    eo = O.simulate_expectation(list(progs), names, v2, ps)
    go = O.adjoint_gradient(list(progs), names, v2, ps, g2)
    np.testing.assert_allclose(got[:B * 2].reshape(B, 2), eo, atol=TOL * max(1, np.abs(eo).max()))
    np.testing.assert_allclose(got[B * 2:].reshape(B, -1), go, atol=TOL * max(1, np.abs(go).max()))


def test_rebind_is_refused_for_streaming_plans():
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, ps, rng = _batch(5, [19])
    plan, _ = sim.plan_expectation(progs, names, rng.uniform(size=(1, 5)).astype(np.float32), ps)
    with pytest.raises(_lib.InvalidArgumentError, match="cannot be rebound"):
        plan.rebind(np.zeros((1, 5), np.float32))


def test_op_calls_reuse_plans():
    from quantum_amd import tfq_adj_grad_op as A
    from quantum_amd import tfq_simulate_ops as sim
    plan_cache.clear()
    progs, names, ps, rng = _batch(77, [6, 13, 15, 3])
    B = len(progs)
    for call in range(3):
        vals = rng.uniform(-3, 3, size=(B, 5)).astype(np.float32)
        g = rng.uniform(-1, 1, size=(B, 2)).astype(np.float32)
        _say(f"cached op calls, call {call}")
        e = sim.tfq_simulate_expectation(progs, names, vals, ps).cpu().numpy()
        ee, gg = A.expectation_and_gradient(progs, names, vals, ps, g)
        grad = A.tfq_adj_grad(progs, names, vals, ps, g).cpu().numpy()
        assert len(plan_cache._CACHE) == 3
        eo = O.simulate_expectation(list(progs), names, vals, ps)
        go = O.adjoint_gradient(list(progs), names, vals, ps, g)
        np.testing.assert_allclose(e, eo, atol=TOL * max(1, np.abs(eo).max()))
        # Two different plans (the adjoint plan fuses for its reverse pass):
        # equal to float32 rounding.
        np.testing.assert_allclose(ee.cpu().numpy(), e, atol=1e-6)
        np.testing.assert_allclose(gg.cpu().numpy(), go, atol=TOL * max(1, np.abs(go).max()))
        np.testing.assert_allclose(grad, gg.cpu().numpy(), atol=1e-6)
    _lib.set_fusion_width(1)
    try:
        assert len(plan_cache._CACHE) == 0
        e1 = sim.tfq_simulate_expectation(progs, names, vals, ps).cpu().numpy()
        np.testing.assert_allclose(e1, eo, atol=TOL * max(1, np.abs(eo).max()))
    finally:
        _lib.set_fusion_width(0)
