"""One-state streaming passes on 14-qubit tiles with the five-register-qubit
engine (pass_kernel's wide instantiation, 32 amplitudes per thread): states,
expectations (every gate family, controls) and samples against the
complex128 oracle at 1e-5, and the 30-qubit grid's 49-term expectation
against the 13-qubit-tile path."""

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
def _wide_tiles():
    _lib.set_tuning(2, 14, 0)
    yield
    _lib.set_tuning(2, 0, 0)


def _batch(seed, n, depth, n_sym=3):
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = [f"s{i}" for i in range(n_sym)]
    circ = W.random_circuit(rng, q, depth, symbols=names, p_control=0.15) + \
        C.Circuit([C.I(x) for x in q])
    vals = rng.uniform(-2, 2, size=(1, n_sym)).astype(np.float32)
    ps = [[W.random_pauli_sum(rng, q, 4, max_weight=4), W.random_pauli_sum(rng, q, 3)]]
    return S.convert_to_tensor([circ]), names, vals, S.convert_to_tensor(ps)


@pytest.mark.parametrize("n", [19, 20])
def test_wide_tiles_match_oracle(n):
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, vals, ps = _batch(900 + n, n, 6 * n)
    plan, _ = sim.plan_expectation(progs, names, vals, ps)
    assert plan.stats()["streaming"] and plan.stats()["forward_passes"] > 0
    _say(f"14-qubit tiles n={n}: {plan.stats()['forward_passes']} passes")
    del plan
    st = sim.tfq_simulate_state(progs, names, vals).cpu().numpy()
    ref = O.simulate_state(list(progs), names, vals)
    np.testing.assert_allclose(st, ref, atol=TOL)
    out = sim.tfq_simulate_expectation(progs, names, vals, ps).cpu().numpy()
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref_e, atol=TOL * max(1, np.abs(ref_e).max()))


def test_wide_tiles_in_sampling():
    from quantum_amd import tfq_simulate_ops as sim
    progs, names, vals, _ = _batch(78, 16, 80)
    _say("14-qubit tiles: samples n=16")
    got = sim.tfq_simulate_samples(progs, names, vals, [20000], seed=11).cpu().numpy()
    from test_gpu_rng import assert_draws_match
    assert assert_draws_match(got[0], progs[0], names, vals, 11, 0) > 0.9 * 20000


def test_wide_tiles_on_the_30_qubit_grid():
    from quantum_amd import tfq_simulate_ops as sim
    circuit, zz = W.config2_grid()
    progs = S.convert_to_tensor([circuit])
    ps = S.convert_to_tensor([[zz]])
    none = np.zeros((1, 0), np.float32)
    _say("30-qubit grid, 14-qubit tiles")
    e14 = float(sim.tfq_simulate_expectation(progs, [], none, ps).cpu()[0, 0])
    _lib.set_tuning(2, 13, 0)
    _say("30-qubit grid, 13-qubit tiles")
    e13 = float(sim.tfq_simulate_expectation(progs, [], none, ps).cpu()[0, 0])
    assert abs(e14 - e13) < 1e-4
