"""Pins the noisy-op oracle (``oracle/noise_oracle.py``) on the CPU.

* Kraus completeness of every channel (sum K^dag K = I, mixture
  probabilities sum to 1) for random strengths.
* One-qubit closed forms of each channel's action on |0>, |1> or |+>.
* Programs without channels: the density matrix is |psi><psi| of the unitary
  oracle, so expectations equal ``tfq_oracle.simulate_expectation``.
* The trajectory replica converges to the density matrix: expectations,
  sampled expectations and sample frequencies of small noisy circuits agree
  within a few standard errors (fixed seeds, so deterministic).
* Parse errors keep the reference's messages.
"""

import numpy as np
import pytest

from oracle import noise_oracle as N
from oracle import tfq_oracle as O
from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W


def _p(c):
    return S.serialize_circuit(c)


def _ps(x):
    return S.serialize_paulisum(x)


@pytest.mark.parametrize("cid", N.CHANNEL_IDS)
def test_kraus_completeness(cid):
    rng = np.random.default_rng(sum(map(ord, cid)))
    for _ in range(5):
        params = {"p": rng.uniform(0, 1), "gamma": rng.uniform(0, 1),
                  "p_x": rng.uniform(0, 0.3), "p_y": rng.uniform(0, 0.3),
                  "p_z": rng.uniform(0, 0.3)}
        kr = N.channel_kraus(cid, {k: np.float32(v) for k, v in params.items()})
        if cid in N.MIXTURES:
            assert abs(sum(p for _, p in kr) - 1.0) < 1e-12
            for U, _ in kr:
                np.testing.assert_allclose(U.conj().T @ U, np.eye(2), atol=1e-12)
        else:
            tot = sum(K.conj().T @ K for K, _ in kr)
            np.testing.assert_allclose(tot, np.eye(2), atol=1e-12)


def _one_qubit(prep, chan, obs):
    q = C.LineQubit(0)
    c = C.Circuit(prep + [chan.on(q)])
    pauli = {"X": C.pauli_x, "Z": C.pauli_z}[obs](q)
    return N.density_expectation([_p(c)], [], np.zeros((1, 0)), [[_ps(C.PauliSum([pauli]))]])[0, 0]


@pytest.mark.parametrize("p", [0.0, 0.13, 0.5])
def test_channel_closed_forms(p):
    q = C.LineQubit(0)
    one = [C.X(q)]
    plus = [C.H(q)]
    g = 0.37
    f32 = lambda x: float(np.float32(x))
    assert _one_qubit([C.I(q)], C.depolarize(p), "Z") == pytest.approx(1 - 4 * f32(p) / 3, abs=1e-7)
    assert _one_qubit([C.I(q)], C.bit_flip(p), "Z") == pytest.approx(1 - 2 * f32(p), abs=1e-7)
    assert _one_qubit(plus, C.phase_flip(p), "X") == pytest.approx(1 - 2 * f32(p), abs=1e-7)
    adp = C.asymmetric_depolarize(p / 3, p / 4, p / 5)
    px, py, pz = f32(p / 3), f32(p / 4), f32(p / 5)
    assert _one_qubit([C.I(q)], adp, "Z") == pytest.approx(1 - 2 * (px + py), abs=1e-7)
    assert _one_qubit(plus, adp, "X") == pytest.approx(1 - 2 * (py + pz), abs=1e-7)
    assert _one_qubit(one, C.amplitude_damp(p), "Z") == pytest.approx(2 * f32(p) - 1, abs=1e-7)
    assert _one_qubit(plus, C.phase_damp(p), "X") == pytest.approx(np.sqrt(1 - f32(p)), abs=1e-7)
    gad = C.generalized_amplitude_damp(p, g)
    assert _one_qubit(one, gad, "Z") == pytest.approx(2 * f32(p) * f32(g) - 1, abs=1e-7)
    assert _one_qubit([C.I(q)], gad, "Z") == pytest.approx(1 - 2 * (1 - f32(p)) * f32(g), abs=1e-7)
    assert _one_qubit(one, C.ResetChannel(), "Z") == pytest.approx(1.0, abs=1e-12)
    assert _one_qubit(plus, C.ResetChannel(), "X") == pytest.approx(0.0, abs=1e-12)


def test_channel_free_programs_match_the_unitary_oracle():
    rng = np.random.default_rng(5)
    names = ["a", "b"]
    q = C.LineQubit.range(4)
    circuits = [W.random_circuit(rng, q, 12, symbols=names, p_control=0.2) for _ in range(3)]
    sums = [[W.random_pauli_sum(rng, q, 3)] for _ in range(3)]
    vals = rng.uniform(-2, 2, size=(3, 2)).astype(np.float32)
    progs = [_p(c) for c in circuits]
    ps = [[_ps(s) for s in row] for row in sums]
    np.testing.assert_allclose(N.density_expectation(progs, names, vals, ps),
                               O.simulate_expectation(progs, names, vals, ps), atol=1e-10)


def _noisy_case(seed, n, batch, depth):
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = ["a", "b"]
    circuits = [W.random_noisy_circuit(rng, q, depth, symbols=names, p_channel=0.6)
                + C.Circuit([C.I(x) for x in q]) for _ in range(batch)]
    sums = [[W.random_pauli_sum(rng, q, 3, max_weight=2), W.random_pauli_sum(rng, q, 2)]
            for _ in range(batch)]
    vals = rng.uniform(-2, 2, size=(batch, 2)).astype(np.float32)
    return [_p(c) for c in circuits], names, vals, [[_ps(s) for s in row] for row in sums]


def test_trajectory_replica_converges_to_density_matrix():
    progs, names, vals, ps = _noisy_case(17, 3, 2, 8)
    T = 1500
    ns = np.full((2, 2), T, np.int32)
    exact = N.density_expectation(progs, names, vals, ps)
    traj = N.trajectory_expectation(progs, names, vals, ps, ns, seed=123)
    # Seeds fixed: the replica lands within 3e-3 of the exact values here,
    # while an unnormalised Kraus step (|psi| 0.9 after each damping channel)
    # moves them by 0.09-0.41.
    np.testing.assert_allclose(traj, exact, atol=0.03)
    # One +-1 shot per term per trajectory: standard errors up to ~0.05.
    samp = N.trajectory_sampled_expectation(progs, names, vals, ps, ns, seed=321)
    np.testing.assert_allclose(samp, exact, atol=0.12)


def test_trajectory_samples_match_density_probabilities():
    progs, names, vals, _ = _noisy_case(23, 3, 1, 8)
    T = 4000
    out = N.trajectory_samples(progs, names, vals, T, seed=99)
    probs, n = N.density_probabilities(progs[0], names, vals[0])
    idx = np.zeros(T, dtype=np.int64)
    for q in range(n):
        idx |= out[0, :, n - 1 - q].astype(np.int64) << q
    freq = np.bincount(idx, minlength=2 ** n) / T
    se = np.sqrt(probs * (1 - probs) / T)
    assert np.all(np.abs(freq - probs) <= 5 * se + 1e-3)


def test_replica_draws_depend_on_seed_and_row_only():
    progs, names, vals, ps = _noisy_case(29, 3, 2, 6)
    ns = np.full((2, 2), 40, np.int32)
    a = N.trajectory_expectation(progs, names, vals, ps, ns, seed=7)
    b = N.trajectory_expectation(progs, names, vals, ps, ns, seed=7)
    c = N.trajectory_expectation(progs, names, vals, ps, ns, seed=8)
    np.testing.assert_array_equal(a, b)
    assert not np.allclose(a, c)


def test_noisy_parse_errors():
    q = C.LineQubit(0)
    good = _p(C.Circuit([C.depolarize(0.1).on(q)]))
    N.NoisyProgram(good)
    with pytest.raises(O.OracleError, match="Could not parse gate id: DP"):
        O.Program(good)
    # An unknown id: rewrite DP -> QQ in the wire bytes (same length).
    bad = good.replace(b"DP", b"QQ")
    with pytest.raises(O.OracleError, match="Could not parse channel id: QQ"):
        N.NoisyProgram(bad)
