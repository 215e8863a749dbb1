"""Dense MFMA passes of the streaming path: with fusion width 5 or 6, fused
blocks of 5-6 qubits on states wider than the tiled batch kernel are applied
as (2^k x 2^k) . (2^k x 2^(n-k)) complex contractions on FP32 MFMA
(v_mfma_f32_16x16x4f32) instead of tile passes of 2-qubit ops.

Checked against the complex128 oracle: states and expectations at 19 and 20
qubits with every gate family and controls, samples at 16 qubits (draw for
draw), and the 30-qubit grid circuit's 49-term expectation against the
width-2 path.  Tolerance 1e-5 absolute (scaled by the observable's size)."""

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


@pytest.fixture(autouse=True)
def _reset_width():
    yield
    _lib.set_fusion_width(0)


def _batch(seed, n, depth, n_sym=3):
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = [f"s{i}" for i in range(n_sym)]
    circ = W.random_circuit(rng, q, depth, symbols=names, p_control=0.15) + \
        C.Circuit([C.I(x) for x in q])
    vals = rng.uniform(-2, 2, size=(1, n_sym)).astype(np.float32)
    ps = [[W.random_pauli_sum(rng, q, 4, max_weight=4), W.random_pauli_sum(rng, q, 3)]]
    return S.convert_to_tensor([circ]), names, vals, S.convert_to_tensor(ps)


def _mfma_flops(progs, names, vals, ps):
    from quantum_amd import tfq_simulate_ops as sim
    plan, _ = sim.plan_expectation(progs, names, vals, ps)
    return plan.stats()["mfma_flops"], plan.stats()["forward_passes"]


@pytest.mark.parametrize("n,width", [(19, 5), (19, 6), (20, 5), (20, 6)])
def test_dense_passes_match_oracle(n, width):
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, vals, ps = _batch(500 + n + width, n, 6 * n)
    base_mfma, _ = _mfma_flops(progs, names, vals, ps)
    _lib.set_fusion_width(width)
    mfma, passes = _mfma_flops(progs, names, vals, ps)
    # Dense passes add at least one k = 5 contraction (8 * 2^5 real flops per
    # amplitude) of MFMA work over the width-2 plan.
    assert mfma - base_mfma >= 8 * 32 * 2 ** n, (mfma, base_mfma)
    _say(f"dense passes n={n} width={width}: {passes} passes")
    st = sim.tfq_simulate_state(progs, names, vals).cpu().numpy()
    ref = O.simulate_state(list(progs), names, vals)
    np.testing.assert_allclose(st, ref, atol=TOL)
    out = sim.tfq_simulate_expectation(progs, names, vals, ps).cpu().numpy()
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref_e, atol=TOL * max(1, np.abs(ref_e).max()))


def test_dense_passes_in_sampling():
    """Samples mode has no tiled variant, so 16 qubits take the streaming path."""
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, vals, _ = _batch(77, 16, 80)
    _lib.set_fusion_width(5)
    _say("dense passes: samples n=16")
    got = sim.tfq_simulate_samples(progs, names, vals, [20000], seed=9).cpu().numpy()
    from test_gpu_rng import assert_draws_match
    assert assert_draws_match(got[0], progs[0], names, vals, 9, 0) > 0.9 * 20000


def test_dense_passes_on_the_30_qubit_grid():
    from quantum_amd import tfq_simulate_ops as sim
    circuit, zz = W.config2_grid()
    progs = S.convert_to_tensor([circuit])
    ps = S.convert_to_tensor([[zz]])
    none = np.zeros((1, 0), np.float32)
    _say("30-qubit grid, width 2")
    e2 = float(sim.tfq_simulate_expectation(progs, [], none, ps).cpu()[0, 0])
    for width in (5, 6):
        _lib.set_fusion_width(width)
        _say(f"30-qubit grid, width {width}")
        ew = float(sim.tfq_simulate_expectation(progs, [], none, ps).cpu()[0, 0])
        assert abs(ew - e2) < TOL * 49, (width, ew, e2)
