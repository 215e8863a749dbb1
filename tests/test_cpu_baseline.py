"""The C++ CPU baseline (oracle/cpu_tfq.cc, float32, TFQ's qsim schedule)
against the complex128 oracle on random circuits: expectations and TFQ-mode
adjoint gradients agree to float32 accuracy (1e-4), with and without the
basic fusion, threaded and single-threaded."""

import numpy as np

from oracle import cpu_tfq
from oracle import tfq_oracle as O
from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W


def _batch(seed, count=6):
    rng = np.random.default_rng(seed)
    names = ["a", "b", "c", "d"]
    circs, sums = [], []
    for i in range(count):
        q = C.LineQubit.range(int(rng.integers(2, 7)))
        circs.append(W.random_circuit(rng, q, 25, symbols=names, p_symbol=0.5, p_control=0.2))
        sums.append([W.random_pauli_sum(rng, q, 3), W.random_pauli_sum(rng, q, 2)])
    circs.append(C.Circuit())
    sums.append([C.PauliSum([]), C.PauliSum([])])
    vals = rng.uniform(-2, 2, (len(circs), len(names))).astype(np.float32)
    grads = rng.uniform(-1, 1, (len(circs), 2)).astype(np.float32)
    return S.convert_to_tensor(circs), names, vals, S.convert_to_tensor(sums), grads


def test_cpu_baseline_matches_oracle():
    progs, names, vals, ps, grads = _batch(3)
    low = cpu_tfq.Lowered(list(progs), names, vals, ps)
    for threads in (1, 4):
        e = cpu_tfq.run(low, 0, threads=threads)
        g = cpu_tfq.run(low, 1, downstream=grads, threads=threads)
        np.testing.assert_allclose(e, O.simulate_expectation(list(progs), names, vals, ps), atol=1e-4)
        np.testing.assert_allclose(g, O.adjoint_gradient(list(progs), names, vals, ps, grads),
                                   atol=1e-4)


def test_compute_large_schedule_matches_oracle():
    """The state-parallel (ComputeLarge) schedule gives the same results as the
    oracle, including the empty program's -2, at 1 and 4 threads."""
    progs, names, vals, ps, grads = _batch(4)
    low = cpu_tfq.Lowered(list(progs), names, vals, ps)
    want_e = O.simulate_expectation(list(progs), names, vals, ps)
    want_g = O.adjoint_gradient(list(progs), names, vals, ps, grads)
    for threads in (1, 4):
        np.testing.assert_allclose(cpu_tfq.run_large(low, 0, threads=threads), want_e, atol=1e-4)
        np.testing.assert_allclose(cpu_tfq.run_large(low, 1, downstream=grads, threads=threads),
                                   want_g, atol=1e-4)


def test_compute_large_bounded_reverse_sweep_covers_the_last_gates():
    """gate_limit = G runs the whole reverse sweep; a bound touching only the
    last gates leaves the symbols of earlier gates at zero."""
    q = C.LineQubit.range(3)
    circ = C.Circuit.from_moments([[C.rx(C.Symbol("a")).on(q[0])], [C.CNOT(q[0], q[1])],
                                   [C.ry(C.Symbol("b")).on(q[2])]])
    progs = S.convert_to_tensor([circ])
    ps = S.convert_to_tensor([[C.PauliSum([C.PauliString({q[0]: "Z", q[2]: "X"})])]])
    vals = np.array([[0.3, -0.7]], np.float32)
    low = cpu_tfq.Lowered(list(progs), ["a", "b"], vals, ps)
    ones = np.ones((1, 1), np.float32)
    full = cpu_tfq.run_large(low, 1, downstream=ones, threads=2)
    np.testing.assert_allclose(cpu_tfq.run_large(low, 1, downstream=ones, threads=2, gate_limit=3),
                               full)
    last = cpu_tfq.run_large(low, 1, downstream=ones, threads=2, gate_limit=1)
    assert last[0, 0] == 0.0 and abs(last[0, 1] - full[0, 1]) < 1e-6 and full[0, 0] != 0.0


def test_compute_large_sampler_matches_inverse_cdf_draws():
    """qsim's sampler (sorted uniforms, one scan) returns, in ascending order,
    the multiset of the oracle's inverse-CDF draws from the same Philox
    stream; only draws within float32 rounding of a CDF boundary may move."""
    rng = np.random.default_rng(9)
    q = C.LineQubit.range(8)
    circs = [W.random_circuit(rng, q, 30) for _ in range(3)]
    progs = S.convert_to_tensor(circs)
    ps = S.convert_to_tensor([[C.PauliSum([])]] * 3)
    low = cpu_tfq.Lowered(list(progs), [], np.zeros((3, 0), np.float32), ps)
    shots, seed = 20000, 31
    got = cpu_tfq.sample_large(low, shots, seed, threads=4)
    for i in range(3):
        probs = np.abs(O.simulate(O.Program(progs[i], {}))) ** 2
        want = np.sort(O.sample_indices(probs, seed, i, 0xFFFFFFFF, shots))
        assert np.all(np.diff(got[i].astype(np.int64)) >= 0)
        assert np.mean(got[i].astype(np.int64) == want) > 0.999


def test_basic_fusion_preserves_the_state():
    rng = np.random.default_rng(5)
    q = C.LineQubit.range(4)
    circ = W.random_circuit(rng, q, 40, p_control=0.3)
    prog = O.Program(S.serialize_circuit(circ), {})
    plain = [(g.unitary(), list(g.bits), sum(1 << b for b in g.cbits),
              sum(int(v) << b for b, v in zip(g.cbits, g.cvals))) for g in prog.gates]
    fused = cpu_tfq._basic_fuse(plain)
    assert len(fused) < len(plain)

    def run(gl):
        s = O.zero_state(prog.n)
        for U, bits, cm, cv in gl:
            cb = [b for b in range(prog.n) if (cm >> b) & 1]
            s = O.apply_matrix(s, U, bits, prog.n, cb, [(cv >> b) & 1 for b in cb])
        return s
    np.testing.assert_allclose(run(fused), run(plain), atol=1e-12)


def test_noisy_baseline_matches_trajectory_replica():
    """The float32 C++ trajectory loop (the noisy op's CPU baseline) draws the
    same Philox variates as the oracle's replica, so the averages agree."""
    from oracle import noise_oracle as N
    rng = np.random.default_rng(41)
    names = ["a", "b"]
    circuits, sums = [], []
    for n in (3, 5, 7):
        q = C.LineQubit.range(n)
        circuits.append(W.random_noisy_circuit(rng, q, 3 * n, symbols=names, p_channel=0.6,
                                               p_control=0.1) + C.Circuit([C.I(x) for x in q]))
        sums.append([W.random_pauli_sum(rng, q, 3), W.random_pauli_sum(rng, q, 2)])
    progs = [S.serialize_circuit(c) for c in circuits]
    ps = [[S.serialize_paulisum(x) for x in row] for row in sums]
    vals = rng.uniform(-2, 2, size=(3, 2)).astype(np.float32)
    ns = np.array([[40, 25], [30, 30], [12, 40]], np.int32)
    low = cpu_tfq.LoweredNoisy(progs, names, vals, ps)
    got = cpu_tfq.run_noisy(low, ns, seed=99, threads=4)
    ref = N.trajectory_expectation(progs, names, vals, ps, ns, seed=99)
    np.testing.assert_allclose(got, ref, atol=2e-4)
