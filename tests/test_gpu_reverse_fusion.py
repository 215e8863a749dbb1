"""Adjoint gradients with the reverse sweep fused across parameterised gates
(blocks of <= 2 qubits holding <= 2 derivatives, derivative matrices
U_a dG U_b built on the host) and with it split at every parameterised gate
(the reference's gradient circuit, core/src/adj_util.cc:155-172), against
the complex128 oracle at 1e-5: every gate family, dense symbol use with
symbols repeated inside one block and two-symbol gates (FSim, PhasedISwap,
PhasedX), controlled parameterised gates, both derivative modes; LDS-resident,
wave-mode, tiled and streaming circuits; and a rebound plan."""

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
    yield
    _lib.set_reverse_fusion(True)
    _lib.set_wave_min_circuits(-1)
    _lib.set_streaming_min_qubits(-1)


def _batch(seed, widths, p_symbol=0.7, p_control=0.15, n_symbols=3):
    rng = np.random.default_rng(seed)
    names = [f"t{i}" for i in range(n_symbols)]
    circuits, sums = [], []
    for n in widths:
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, 4 * n, symbols=names, p_symbol=p_symbol,
                                         p_control=p_control) + C.Circuit([C.I(x) for x in q]))
        sums.append([W.random_pauli_sum(rng, q, 3, max_weight=min(3, n)) for _ in range(2)])
    B = len(widths)
    return (S.convert_to_tensor(circuits), names,
            rng.uniform(-2, 2, size=(B, n_symbols)).astype(np.float32),
            S.convert_to_tensor(sums), rng.uniform(-1, 1, size=(B, 2)).astype(np.float32))


def _check(progs, names, vals, ps, g, mode):
    from quantum_amd import tfq_adj_grad_op as A
    ee, gg = A.expectation_and_gradient(progs, names, vals, ps, g, derivative=mode)
    eo = O.simulate_expectation(list(progs), names, vals, ps)
    go = O.adjoint_gradient(list(progs), names, vals, ps, g, mode=mode)
    np.testing.assert_allclose(ee.cpu().numpy(), eo, atol=TOL * max(1, np.abs(eo).max()))
    np.testing.assert_allclose(gg.cpu().numpy(), go, atol=TOL * max(1, np.abs(go).max()))
    return gg.cpu().numpy()


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_lds_and_wave_circuits(fused, mode):
    _lib.set_reverse_fusion(fused)
    for wave_min in (-1, 0):
        _lib.set_wave_min_circuits(wave_min)
        progs, names, vals, ps, g = _batch(11 + wave_min, [2, 5, 8, 10, 12, 13, 3, 9])
        _say(f"reverse fusion={fused} mode={mode} wave_min={wave_min}")
        _check(progs, names, vals, ps, g, mode)
    _lib.set_wave_min_circuits(-1)


@pytest.mark.parametrize("fused", [True, False])
def test_tiled_and_streaming_circuits(fused):
    _lib.set_reverse_fusion(fused)
    progs, names, vals, ps, g = _batch(23, [14, 16, 15], p_symbol=0.5)
    _say(f"reverse fusion={fused}: tiled 14-16 qubits")
    _check(progs, names, vals, ps, g, "analytic")
    _lib.set_streaming_min_qubits(14)
    try:
        _say(f"reverse fusion={fused}: streaming 14-16 qubits")
        _check(progs, names, vals, ps, g, "tfq")
    finally:
        _lib.set_streaming_min_qubits(-1)


def test_fused_and_split_sweeps_agree_and_fusion_applies_fewer_ops():
    from quantum_amd import tfq_adj_grad_op as A
    progs, names, vals, ps, g = _batch(5, [6, 11, 12, 7, 13], p_symbol=0.9, p_control=0.0,
                                       n_symbols=2)
    out = {}
    for fused in (True, False):
        _lib.set_reverse_fusion(fused)
        plan, _ = A.plan_adj_grad(progs, names, vals, ps, g, with_expectation=True)
        out[fused] = plan.stats()["adjoint_ops"]
        _say(f"dense symbols, reverse fusion={fused}")
        _check(progs, names, vals, ps, g, "analytic")
    _lib.set_reverse_fusion(True)
    assert out[True] < out[False]


def test_rebound_fused_plan_equals_fresh_plan():
    import torch
    from quantum_amd import tfq_adj_grad_op as A
    from quantum_amd.tfq_simulate_ops import _device_and_stream
    _lib.set_reverse_fusion(True)
    progs, names, v1, ps, g = _batch(7, [4, 9, 12, 15, 6])
    v2 = np.random.default_rng(8).uniform(-2, 2, size=v1.shape).astype(np.float32)
    dev, stream = _device_and_stream()

    def launch(plan, shape):
        out = torch.empty(shape, dtype=torch.float32, device=dev)
        plan.upload(dev, stream)
        plan.launch(out, stream)
        torch.cuda.synchronize()
        return out.cpu().numpy()

    plan, shape = A.plan_adj_grad(progs, names, v1, ps, g, with_expectation=True)
    plan.rebind(v2)
    _say("rebound plan with fused parameterised blocks")
    got = launch(plan, shape)
    fresh, _ = A.plan_adj_grad(progs, names, v2, ps, g, with_expectation=True)
    want = launch(fresh, shape)
    B = len(progs)
    np.testing.assert_array_equal(got[:B * 2], want[:B * 2])
    np.testing.assert_allclose(got[B * 2:], want[B * 2:], atol=1e-6)
