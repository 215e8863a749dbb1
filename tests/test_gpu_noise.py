"""``tfq.noise`` trajectory ops on the GPU.

* Replica parity: with the same seed the device's trajectories make the same
  Kraus choices as the oracle's replica (``noise_oracle.trajectory_*``), so
  averages agree to float rounding; circuits of 2-16 qubits cover wave mode
  (with idle qubits added below 4 qubits) and workgroup trajectories (state
  in LDS up to 14 qubits, in a workspace buffer beyond), mixed
  in one batch with an empty program and per-slot trajectory counts.
* Statistical parity with the exact density matrix, as the reference's own
  tests check (``noisy_expectation_op_test.py:246-312``, atol 5e-2): random
  noisy circuits and each supported channel on every qubit.
* Channel-free programs equal the exact expectation.
* A plan launched twice gives the same result (accumulators are cleared).
* Input errors.
"""

import sys

import numpy as np
import pytest

from oracle import noise_oracle as N
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


def _batch(seed, widths, depth_per_qubit=3, p_channel=0.5, n_terms=2):
    rng = np.random.default_rng(seed)
    names = ["a", "b"]
    circuits, sums = [], []
    for n in widths:
        if n == 0:
            circuits.append(C.Circuit())
            sums.append([C.PauliSum([]), C.PauliSum([])])
            continue
        q = C.LineQubit.range(n)
        circuits.append(W.random_noisy_circuit(rng, q, depth_per_qubit * n, symbols=names,
                                               p_channel=p_channel, p_control=0.1)
                        + C.Circuit([C.I(x) for x in q]))
        sums.append([W.random_pauli_sum(rng, q, n_terms, max_weight=3),
                     W.random_pauli_sum(rng, q, n_terms)])
    vals = rng.uniform(-2, 2, size=(len(widths), 2)).astype(np.float32)
    return S.convert_to_tensor(circuits), names, vals, S.convert_to_tensor(sums)


@pytest.mark.parametrize("widths", [(2, 3, 0), (4, 6), (5, 8, 0, 10), (11, 12), (14, 3),
                                    (16, 0, 9)])
def test_noisy_expectation_matches_replica(widths):
    from quantum_amd import noise
    progs, names, vals, ps = _batch(100 + sum(widths), widths)
    B = len(widths)
    ns = np.full((B, 2), 48, np.int32)
    ns[:, 1] = 31
    _say(f"noisy expectation replica widths={widths}")
    out = noise.expectation(progs, names, vals, ps, ns, seed=2024).cpu().numpy()
    ref = N.trajectory_expectation(list(progs), names, vals, ps, ns, seed=2024)
    np.testing.assert_allclose(out, ref, atol=2e-4)
    for i, n in enumerate(widths):
        if n == 0:
            np.testing.assert_array_equal(out[i], -2.0)


@pytest.mark.parametrize("widths", [(3, 6), (9, 0, 11), (15, 5)])
def test_noisy_sampled_expectation_matches_replica(widths):
    from quantum_amd import noise
    progs, names, vals, ps = _batch(200 + sum(widths), widths)
    B = len(widths)
    ns = np.full((B, 2), 40, np.int32)
    ns[:, 0] = 23
    _say(f"noisy sampled expectation replica widths={widths}")
    out = noise.sampled_expectation(progs, names, vals, ps, ns, seed=77).cpu().numpy()
    ref = N.trajectory_sampled_expectation(list(progs), names, vals, ps, ns, seed=77)
    np.testing.assert_allclose(out, ref, atol=2e-4)


@pytest.mark.parametrize("widths", [(2, 5, 0), (7, 10), (12, 4), (14,), (15, 3)])
def test_noisy_samples_match_replica(widths):
    from quantum_amd import noise
    progs, names, vals, _ = _batch(300 + sum(widths), widths)
    _say(f"noisy samples replica widths={widths}")
    out = noise.samples(progs, names, vals, [64], seed=5).cpu().numpy()
    ref = N.trajectory_samples(list(progs), names, vals, 64, seed=5)
    assert out.shape == ref.shape
    # A draw that lands within float rounding of a CDF boundary may differ.
    rows_differ = np.any(out != ref, axis=2)
    assert rows_differ.mean() <= 0.01, rows_differ.sum()
    assert np.all((out == -2) == (ref == -2))


def test_noisy_expectation_statistics_match_density_matrix():
    from quantum_amd import noise
    progs, names, vals, ps = _batch(400, (4, 5, 6), depth_per_qubit=2, p_channel=0.6)
    ns = np.full((3, 2), 6000, np.int32)
    _say("noisy expectation vs density matrix (6000 trajectories)")
    out = noise.expectation(progs, names, vals, ps, ns, seed=11).cpu().numpy()
    exact = N.density_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, exact, atol=5e-2, rtol=5e-2)
    samp = noise.sampled_expectation(progs, names, vals, ps, ns, seed=12).cpu().numpy()
    np.testing.assert_allclose(samp, exact, atol=1e-1, rtol=5e-2)


@pytest.mark.parametrize("chan", C.supported_channels(), ids=lambda c: c.gate_id)
def test_single_channel_statistics(chan):
    """Each channel on every qubit after a random circuit (the reference's
    ``test_single_channel``, ``noisy_expectation_op_test.py:276-312``)."""
    from quantum_amd import noise
    rng = np.random.default_rng(sum(map(ord, chan.gate_id)))
    q = C.LineQubit.range(5)
    circuits = [W.random_circuit(rng, q, 10) + C.Circuit(chan.on_each(*q)) for _ in range(3)]
    sums = [[W.random_pauli_sum(rng, q, 3), W.random_pauli_sum(rng, q, 2)] for _ in range(3)]
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(sums)
    vals = np.zeros((3, 0), np.float32)
    ns = np.full((3, 2), 5000, np.int32)
    _say(f"single channel {chan.gate_id}")
    out = noise.expectation(progs, [], vals, ps, ns, seed=31).cpu().numpy()
    exact = N.density_expectation(list(progs), [], vals, ps)
    np.testing.assert_allclose(out, exact, atol=5e-2, rtol=5e-2)


@pytest.mark.parametrize("n,batch", [(13, 1), (14, 2), (17, 1), (6, 25)])
def test_channel_free_programs_equal_exact_expectation(n, batch):
    from quantum_amd import noise
    rng = np.random.default_rng(n * 7 + batch)
    names = ["alpha", "beta"]
    q = C.LineQubit.range(n)
    circuits = [W.random_circuit(rng, q, 2 * n, symbols=names) + C.Circuit([C.I(x) for x in q])
                for _ in range(batch)]
    sums = [[W.random_pauli_sum(rng, q, 3), W.random_pauli_sum(rng, q, 3)] for _ in range(batch)]
    vals = rng.uniform(-2, 2, size=(batch, 2)).astype(np.float32)
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(sums)
    ns = np.full((batch, 2), 3, np.int32)
    _say(f"channel-free noisy expectation n={n} batch={batch}")
    out = noise.expectation(progs, names, vals, ps, ns, seed=1).cpu().numpy()
    ref = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref, atol=5e-4, rtol=5e-4)


@pytest.mark.parametrize("n,T", [(20, 1), (18, 600)])
def test_wide_channel_free_programs_in_workspace(n, T):
    """Workspace trajectories at the widest sizes: one 20-qubit trajectory
    (a workspace of one 8 MiB state), and 600 trajectories of an 18-qubit
    program, more than the 512 workgroups of a launch, each exact."""
    from quantum_amd import noise
    rng = np.random.default_rng(n * 13 + T)
    q = C.LineQubit.range(n)
    circ = W.random_circuit(rng, q, 2 * n) + C.Circuit([C.I(x) for x in q])
    progs = S.convert_to_tensor([circ])
    ps = S.convert_to_tensor([[W.random_pauli_sum(rng, q, 3), C.PauliSum([C.pauli_z(q[0])])]])
    ns = np.full((1, 2), T, np.int32)
    plan, _ = noise.plan_expectation(progs, [], np.zeros((1, 0)), ps, ns, seed=5)
    assert plan.workspace_bytes < (min(T, 512) + 1) * 8 * 2 ** n
    _say(f"wide channel-free noisy expectation n={n} T={T}")
    out = noise.expectation(progs, [], np.zeros((1, 0)), ps, ns, seed=5).cpu().numpy()
    ref = O.simulate_expectation(list(progs), [], np.zeros((1, 0)), ps)
    np.testing.assert_allclose(out, ref, atol=5e-4, rtol=5e-4)


@pytest.mark.parametrize("n,T", [(10, 200), (7, 1000), (12, 64)])
def test_every_trajectory_counts_once(n, T):
    """Regression: many work items over many workgroups (each trajectory of a
    channel-free circuit gives the exact value, so the average must too)."""
    from quantum_amd import noise
    rng = np.random.default_rng(n + T)
    q = C.LineQubit.range(n)
    circ = W.random_circuit(rng, q, 3 * n) + C.Circuit([C.I(x) for x in q])
    progs = S.convert_to_tensor([circ, circ])
    obs = [C.PauliSum([C.pauli_z(x)]) for x in q[:4]]
    ps = S.convert_to_tensor([obs, obs])
    ns = np.full((2, 4), T, np.int32)
    ns[1] = T // 3
    _say(f"trajectory count n={n} T={T}")
    out = noise.expectation(progs, [], np.zeros((2, 0)), ps, ns, seed=3).cpu().numpy()
    ref = O.simulate_expectation(list(progs), [], np.zeros((2, 0)), ps)
    np.testing.assert_allclose(out, ref, atol=1e-4)


def test_noisy_plan_replays_identically():
    import torch
    from quantum_amd import noise
    progs, names, vals, ps = _batch(500, (3, 7, 12))
    ns = np.full((3, 2), 50, np.int32)
    plan, shape = noise.plan_expectation(progs, names, vals, ps, ns, seed=9)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    _say("noisy plan replay")
    plan.upload(dev, stream)
    plan.launch(out, stream)
    first = out.cpu().numpy().copy()
    plan.launch(out, stream)
    np.testing.assert_array_equal(out.cpu().numpy(), first)


def test_noisy_empty_inputs_and_errors():
    from quantum_amd import noise
    from quantum_amd._lib import InvalidArgumentError
    _say("noisy empty inputs and errors")
    out = noise.expectation(S.convert_to_tensor([C.Circuit()]), [], np.zeros((1, 0)),
                            np.empty((1, 0), dtype=object), np.zeros((1, 0), np.int32))
    assert tuple(out.shape) == (1, 0)
    out = noise.expectation(np.empty((0,), dtype=object), [], np.zeros((0, 0)),
                            np.empty((0, 0), dtype=object), np.zeros((0, 0), np.int32))
    assert tuple(out.shape) == (0, 0)
    q = C.LineQubit.range(2)
    prog = S.convert_to_tensor([C.Circuit([C.H(q[0]), C.depolarize(0.1).on(q[1])])])
    ps = S.convert_to_tensor([[C.PauliSum([C.pauli_z(q[0])])]])
    with pytest.raises(InvalidArgumentError, match="greater than 0"):
        noise.expectation(prog, [], np.zeros((1, 0)), ps, np.zeros((1, 1), np.int32))
    with pytest.raises(InvalidArgumentError, match="Dimension 1 of num_samples"):
        noise.expectation(prog, [], np.zeros((1, 0)), ps, np.ones((1, 2), np.int32))
    bad = np.array([prog[0].replace(b"DP", b"QQ")], dtype=object)
    with pytest.raises(InvalidArgumentError, match="Could not parse channel id: QQ"):
        noise.expectation(bad, [], np.zeros((1, 0)), ps, np.ones((1, 1), np.int32))
    from quantum_amd import tfq_simulate_ops as T
    with pytest.raises(InvalidArgumentError, match="Could not parse gate id: DP"):
        T.tfq_simulate_expectation(prog, [], np.zeros((1, 0)), ps)
    wide = S.convert_to_tensor([C.Circuit([C.bit_flip(0.1).on(x) for x in C.LineQubit.range(21)])])
    with pytest.raises(NotImplementedError, match="at most 20 qubits"):
        noise.samples(wide, [], np.zeros((1, 0)), [4])


def test_noisy_layers():
    """``backend='noisy'`` on the Expectation, SampledExpectation and Sample
    layers (``expectation.py:278-345``)."""
    import torch
    from quantum_amd import layers
    q = C.LineQubit.range(3)
    a = C.Symbol("a")
    circ = C.Circuit([C.rx(a).on(q[0]), C.CNOT(q[0], q[1]), C.amplitude_damp(0.2).on(q[1]),
                      C.depolarize(0.05).on(q[2]), C.H(q[2])])
    op = C.PauliSum([C.pauli_z(q[1])])
    _say("noisy layers")
    exp = layers.Expectation(backend="noisy", seed=5)
    with pytest.raises(RuntimeError, match="repetitions"):
        exp([circ], symbol_names=["a"], symbol_values=np.array([[0.7]], np.float32), operators=op)
    v = torch.tensor([[0.7]], device="cuda", requires_grad=True)
    out = exp([circ], symbol_names=["a"], symbol_values=v, operators=op, repetitions=20000)
    prog = [S.serialize_circuit(circ)]
    ps = [[S.serialize_paulisum(op)]]
    exact = N.density_expectation(prog, ["a"], np.array([[0.7]]), ps)[0, 0]
    assert abs(float(out[0, 0]) - exact) < 0.03
    out.sum().backward()
    h = 1e-3
    fd = (N.density_expectation(prog, ["a"], np.array([[0.7 + h]]), ps)[0, 0] -
          N.density_expectation(prog, ["a"], np.array([[0.7 - h]]), ps)[0, 0]) / (2 * h)
    assert np.isfinite(float(v.grad[0, 0]))
    assert abs(float(v.grad[0, 0]) - fd) < 0.25
    sexp = layers.SampledExpectation(backend="noisy", seed=6)
    s_out = sexp([circ], symbol_names=["a"], symbol_values=np.array([[0.7]], np.float32),
                 operators=op, repetitions=20000)
    assert abs(float(s_out[0, 0]) - exact) < 0.05
    smp = layers.Sample(backend="noisy", seed=7)
    bits = smp([circ], symbol_names=["a"], symbol_values=np.array([[0.7]], np.float32),
               repetitions=100)
    assert tuple(bits[0].shape) == (100, 3)
    with pytest.raises(RuntimeError, match="noiseless"):
        layers.Expectation()([circ], symbol_names=["a"],
                             symbol_values=np.array([[0.7]], np.float32), operators=op,
                             repetitions=10)
