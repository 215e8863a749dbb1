"""GPU parity: the HIP path (through the C-ABI) against the complex128 oracle
and the reference's golden vectors.  Tolerance: 1e-5 absolute for
expectations, states, inner products and adjoint gradients (complex64 device
arithmetic vs complex128 oracle on the same float32 inputs)."""

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
    _lib.set_fusion_width(0)
    yield
    _lib.set_fusion_width(0)


def ops():
    from quantum_amd import tfq_simulate_ops, tfq_adj_grad_op, math_ops
    return tfq_simulate_ops, tfq_adj_grad_op, math_ops


def _np(t):
    return t.detach().cpu().numpy()


def _ser(cs):
    return S.convert_to_tensor(cs)


# ----------------------------------------------------------------------------- goldens
def test_golden_state_and_expectations(goldens):
    sim, _, _ = ops()
    circ, (q0, q1) = W.two_qubit_state_circuit()
    _say("golden two-qubit state")
    st = _np(sim.tfq_simulate_state(_ser([circ]), [], np.zeros((1, 0))))[0]
    exp = np.array([complex(*a) for a in goldens["two_qubit_state"]["amplitudes"]])
    np.testing.assert_allclose(st, exp, atol=TOL)
    g = goldens["two_term_expectation"]
    keys = list(g["values"])
    psums = [[C.PauliSum([C.PauliString({q0: k[0], q1: k[1]}, g["coefficient"])]) for k in keys]]
    _say("golden expectation table")
    out = _np(sim.tfq_simulate_expectation(_ser([circ]), [], np.zeros((1, 0)), _ser(psums)))
    np.testing.assert_allclose(out[0], [g["values"][k] for k in keys], atol=TOL)
    comp = C.PauliSum([C.PauliString({q0: "Z", q1: "X"}, 0.1234), C.PauliString({q0: "X"}, -3.0),
                       C.PauliString({}, 4.0)])
    out = _np(sim.tfq_simulate_expectation(_ser([circ]), [], np.zeros((1, 0)), _ser([[comp]])))
    assert abs(out[0, 0] - goldens["compound_expectation"]["value"]) < TOL


@pytest.mark.parametrize("name", ["simple_case", "simple_case2", "simple_case_shared",
                                  "simple_case_single"])
@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_golden_adjoint(goldens, name, mode):
    _, adj, _ = ops()
    g = goldens["adjoint_gradients"]
    circ, names, q = W.adjoint_golden_cases()[name]
    vals = np.array([[g["values"][s] for s in names]], dtype=np.float32)
    psums = _ser([[C.pauli_z(q[0]), C.pauli_x(q[1])]])
    _say(f"golden adjoint {name} {mode}")
    out = _np(adj.tfq_adj_grad(_ser([circ]), names, vals, psums, np.ones((1, 2)), derivative=mode))
    np.testing.assert_allclose(out, g[name]["expected"], atol=g["tol"])
    ref = O.adjoint_gradient([S.serialize_circuit(circ)], names, vals, psums, np.ones((1, 2)),
                             mode=mode)
    np.testing.assert_allclose(out, ref, atol=TOL)


def test_golden_docstrings(goldens):
    sim, _, mo = ops()
    q = C.GridQubit(0, 0)
    circ = C.Circuit(C.H(q) ** C.Symbol("alpha"))
    ps = 3.5 * C.pauli_x(q) - 2.2 * C.pauli_y(q)
    _say("golden docstring expectation")
    out = _np(sim.tfq_simulate_expectation(_ser([circ]), ["alpha"], [[0.123]], _ser([[ps]])))
    assert abs(out[0, 0] - goldens["h_pow_expectation"]["value"]) < TOL
    g = goldens["inner_product"]
    a, b = C.Symbol("alpha"), C.Symbol("beta")
    qs = C.GridQubit.rect(1, 2)
    refs = [C.Circuit([(C.H ** a).on(x) for x in qs]), C.Circuit(C.X(qs[0]) ** a, C.Y(qs[1]) ** b)]
    others = [C.Circuit([C.X(x) for x in qs]), C.Circuit([(C.Y ** 0.125).on(x) for x in qs]),
              C.Circuit([(C.X ** 0.5).on(x) for x in qs])]
    _say("golden inner product")
    ip = _np(mo.inner_product(_ser(refs), ["alpha", "beta"], g["values"], _ser([others, others])))
    exp = np.array([[complex(*v) for v in row] for row in g["expected"]])
    np.testing.assert_allclose(ip, exp, atol=g["tol"])


# ----------------------------------------------------------------------------- random circuits
def _random_batch(seed, n, batch, depth, p_control=0.25, n_sym=3, n_ops=2, terms=4):
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = [f"s{i}" for i in range(n_sym)]
    circuits = [W.random_circuit(rng, q, depth, symbols=names, p_control=p_control)
                for _ in range(batch)]
    vals = rng.uniform(-2, 2, size=(batch, n_sym)).astype(np.float32)
    psums = [[W.random_pauli_sum(rng, q, terms, max_weight=min(4, n)) for _ in range(n_ops)]
             for _ in range(batch)]
    return circuits, names, vals, psums


@pytest.mark.parametrize("n,width", [(2, 0), (3, 1), (5, 2), (7, 3), (9, 4), (10, 0), (12, 2),
                                     (14, 0)])
def test_random_expectation_and_state(n, width):
    sim, _, _ = ops()
    _lib.set_fusion_width(width)
    circuits, names, vals, psums = _random_batch(100 + n, n, 6, 4 * n)
    progs, ps = _ser(circuits), _ser(psums)
    _say(f"random expectation n={n} width={width}")
    out = _np(sim.tfq_simulate_expectation(progs, names, vals, ps))
    ref = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))
    _say(f"random state n={n} width={width}")
    st = _np(sim.tfq_simulate_state(progs, names, vals))
    ref_st = O.simulate_state(list(progs), names, vals)
    np.testing.assert_allclose(st, ref_st, atol=TOL)


@pytest.mark.parametrize("n", [2, 4, 6, 9, 11, 13])
@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_random_adjoint(n, mode):
    _, adj, _ = ops()
    circuits, names, vals, psums = _random_batch(200 + n, n, 4, 3 * n, n_sym=4)
    progs, ps = _ser(circuits), _ser(psums)
    g = np.random.default_rng(n).uniform(-1, 1, size=(4, 2)).astype(np.float32)
    _say(f"random adjoint n={n} mode={mode}")
    out = _np(adj.tfq_adj_grad(progs, names, vals, ps, g, derivative=mode))
    ref = O.adjoint_gradient(list(progs), names, vals, ps, g, mode=mode)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))


def test_random_inner_product():
    _, _, mo = ops()
    rng = np.random.default_rng(9)
    n = 6
    q = C.LineQubit.range(n)
    names = ["a", "b"]
    refs = [W.random_circuit(rng, q, 20, symbols=names) + C.Circuit([C.I(x) for x in q])
            for _ in range(3)]
    others = [[W.random_circuit(rng, q, 15) + C.Circuit([C.I(x) for x in q]) for _ in range(4)]
              for _ in range(3)]
    vals = rng.uniform(-1, 1, size=(3, 2)).astype(np.float32)
    _say("random inner product")
    out = _np(mo.inner_product(_ser(refs), names, vals, _ser(others)))
    ref = O.inner_product(list(_ser(refs)), names, vals, _ser(others))
    np.testing.assert_allclose(out, ref, atol=TOL)


def test_empty_program_sentinels():
    sim, adj, mo = ops()
    q = C.LineQubit.range(2)
    circs = _ser([C.Circuit(), C.Circuit(C.X(q[0]), C.Y(q[1]))])
    psums = _ser([[C.pauli_z(q[0])], [C.pauli_z(q[0])]])
    _say("empty program sentinels")
    out = _np(sim.tfq_simulate_expectation(circs, [], np.zeros((2, 0)), psums))
    assert out[0, 0] == -2.0 and abs(out[1, 0] + 1.0) < TOL
    st = _np(sim.tfq_simulate_state(circs, [], np.zeros((2, 0))))
    assert st.shape == (2, 4)
    np.testing.assert_allclose(st[0], [1, -2, -2, -2])
    smp = _np(sim.tfq_simulate_samples(circs, [], np.zeros((2, 0)), [5], seed=1))
    assert smp.shape == (2, 5, 2) and (smp[0] == -2).all()
    np.testing.assert_array_equal(smp[1], np.ones((5, 2)))
    gr = _np(adj.tfq_adj_grad(circs, ["a"], np.zeros((2, 1)), psums, np.ones((2, 1))))
    assert (gr == 0).all()


def test_config1_batch():
    sim, _, _ = ops()
    circuits, names, values, psums = W.config1_line_rotations(batch=64)
    progs, ps = _ser(circuits), _ser(psums)
    _say("config1 64x10q expectation")
    out = _np(sim.tfq_simulate_expectation(progs, names, values, ps))
    ref = O.simulate_expectation(list(progs[:8]), names, values[:8], ps[:8])
    np.testing.assert_allclose(out[:8], ref, atol=TOL)


# ----------------------------------------------------------------------------- sampling
def test_samples_match_host_replica():
    sim, _, _ = ops()
    circuits, names, vals, _ = _random_batch(300, 5, 3, 20)
    progs = _ser(circuits)
    _say("samples vs host replica")
    out = _np(sim.tfq_simulate_samples(progs, names, vals, [4000], seed=12345))
    ref = O.simulate_samples(list(progs), names, vals, 4000, 12345)
    assert out.shape == ref.shape
    match = np.mean(np.all(out == ref, axis=-1))
    assert match > 0.999, match


def test_samples_padding_and_bit_order():
    sim, _, _ = ops()
    q = C.LineQubit.range(3)
    circs = _ser([C.Circuit(C.X(q[0]), C.I(q[1]), C.I(q[2])), C.Circuit(C.X(q[1]))])
    _say("samples padding")
    out = _np(sim.tfq_simulate_samples(circs, [], np.zeros((2, 0)), [3], seed=3))
    np.testing.assert_array_equal(out[0], [[1, 0, 0]] * 3)
    np.testing.assert_array_equal(out[1], [[-2, -2, 1]] * 3)


def test_sampled_expectation_statistics(goldens):
    sim, _, _ = ops()
    g = goldens["two_term_expectation"]
    circ, (q0, q1) = W.two_qubit_state_circuit()
    keys = list(g["values"])
    psums = _ser([[C.PauliSum([C.PauliString({q0: k[0], q1: k[1]}, g["coefficient"])])
                   for k in keys]])
    _say("sampled expectation 1e6 shots")
    out = _np(sim.tfq_simulate_sampled_expectation(_ser([circ]), [], np.zeros((1, 0)), psums,
                                                   np.full((1, 9), 1000000, np.int32), seed=7))
    np.testing.assert_allclose(out[0], [g["values"][k] for k in keys], atol=1e-2)
    ref = O.simulate_sampled_expectation([S.serialize_circuit(circ)], [], np.zeros((1, 0)),
                                         psums, np.full((1, 9), 20000), seed=7)
    np.testing.assert_allclose(ref[0], [g["values"][k] for k in keys], atol=2e-2)


# ----------------------------------------------------------------------------- large path
@pytest.mark.parametrize("n", [15, 16])
def test_large_state_expectation(n):
    sim, _, _ = ops()
    circuits, names, vals, psums = _random_batch(400 + n, n, 2, 3 * n, p_control=0.1)
    progs, ps = _ser(circuits), _ser(psums)
    _say(f"large expectation n={n}")
    out = _np(sim.tfq_simulate_expectation(progs, names, vals, ps))
    ref = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))
    _say(f"large state n={n}")
    st = _np(sim.tfq_simulate_state(progs, names, vals))
    np.testing.assert_allclose(st, O.simulate_state(list(progs), names, vals), atol=TOL)


@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_large_adjoint(mode):
    _, adj, _ = ops()
    n = 14
    circuits, names, vals, psums = _random_batch(500, n, 2, 2 * n, p_control=0.1, n_sym=5)
    progs, ps = _ser(circuits), _ser(psums)
    g = np.array([[1.0, -0.5], [0.25, 1.0]], np.float32)
    _say(f"large adjoint n={n} {mode}")
    out = _np(adj.tfq_adj_grad(progs, names, vals, ps, g, derivative=mode))
    ref = O.adjoint_gradient(list(progs), names, vals, ps, g, mode=mode)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))


def test_large_samples_and_inner_product():
    sim, _, mo = ops()
    n = 15
    circuits, names, vals, _ = _random_batch(600, n, 1, 2 * n, p_control=0.0)
    progs = _ser(circuits)
    _say("large samples")
    out = _np(sim.tfq_simulate_samples(progs, names, vals, [2000], seed=99))
    ref = O.simulate_samples(list(progs), names, vals, 2000, 99)
    assert np.mean(np.all(out == ref, axis=-1)) > 0.999
    rng = np.random.default_rng(1)
    q = C.LineQubit.range(n)
    others = [[W.random_circuit(rng, q, 20) + C.Circuit([C.I(x) for x in q]) for _ in range(2)]]
    _say("large inner product")
    ip = _np(mo.inner_product(progs, names, vals, _ser(others)))
    np.testing.assert_allclose(ip, O.inner_product(list(progs), names, vals, _ser(others)), atol=TOL)


def test_large_sampled_expectation():
    sim, _, _ = ops()
    n = 15
    rng = np.random.default_rng(77)
    q = C.LineQubit.range(n)
    circ = W.random_circuit(rng, q, 40) + C.Circuit([C.I(x) for x in q])
    ps = C.PauliSum([C.pauli_z(q[0]) * C.pauli_z(q[1]), C.PauliString({q[2]: "X"}, 0.5),
                     C.PauliString({q[3]: "Y", q[4]: "Z"}, -0.7), C.PauliString({}, 0.25)])
    _say("large sampled expectation")
    out = _np(sim.tfq_simulate_sampled_expectation(_ser([circ]), [], np.zeros((1, 0)), _ser([[ps]]),
                                                   np.full((1, 1), 400000, np.int32), seed=5))
    exact = O.simulate_expectation([S.serialize_circuit(circ)], [], np.zeros((1, 0)), _ser([[ps]]))
    assert abs(out[0, 0] - exact[0, 0]) < 1.5e-2


# ----------------------------------------------------------------------------- tiled batch
def _mixed_batch(seed, widths, depth_per_qubit=2, n_sym=4):
    rng = np.random.default_rng(seed)
    names = [f"s{i}" for i in range(n_sym)]
    circuits, psums = [], []
    for n in widths:
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, depth_per_qubit * n, symbols=names,
                                         p_control=0.15) + C.Circuit([C.I(x) for x in q]))
        psums.append([W.random_pauli_sum(rng, q, 3, max_weight=min(4, n)) for _ in range(2)])
    vals = rng.uniform(-2, 2, size=(len(widths), n_sym)).astype(np.float32)
    return circuits, names, vals, psums


def test_tiled_mixed_batch_expectation_state():
    sim, _, _ = ops()
    circuits, names, vals, psums = _mixed_batch(700, [3, 8, 14, 15, 16, 5, 12])
    progs, ps = _ser(circuits), _ser(psums)
    _say("tiled mixed expectation")
    out = _np(sim.tfq_simulate_expectation(progs, names, vals, ps))
    ref = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))
    _say("tiled mixed state")
    st = _np(sim.tfq_simulate_state(progs, names, vals))
    np.testing.assert_allclose(st, O.simulate_state(list(progs), names, vals), atol=TOL)


@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_tiled_mixed_batch_adjoint(mode):
    _, adj, _ = ops()
    circuits, names, vals, psums = _mixed_batch(800, [4, 13, 14, 15, 16, 9])
    progs, ps = _ser(circuits), _ser(psums)
    g = np.random.default_rng(3).uniform(-1, 1, size=(len(circuits), 2)).astype(np.float32)
    _say(f"tiled mixed adjoint {mode}")
    exp, grad = adj.expectation_and_gradient(progs, names, vals, ps, g, derivative=mode)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g, mode=mode)
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(_np(grad), ref_g, atol=TOL * max(1, np.abs(ref_g).max()))
    np.testing.assert_allclose(_np(exp), ref_e, atol=TOL * max(1, np.abs(ref_e).max()))
    grad2 = _np(adj.tfq_adj_grad(progs, names, vals, ps, g, derivative=mode))
    np.testing.assert_allclose(grad2, _np(grad), atol=1e-6)


# ----------------------------------------------------------------------------- streaming path
def test_streaming_path_19_qubits():
    sim, adj, _ = ops()
    n = 19
    circuits, names, vals, psums = _random_batch(900, n, 1, 2 * n, p_control=0.1, n_sym=4)
    progs, ps = _ser(circuits), _ser(psums)
    _say("streaming expectation n=19")
    out = _np(sim.tfq_simulate_expectation(progs, names, vals, ps))
    ref = O.simulate_expectation(list(progs), names, vals, ps)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))
    _say("streaming state n=19")
    st = _np(sim.tfq_simulate_state(progs, names, vals))
    np.testing.assert_allclose(st, O.simulate_state(list(progs), names, vals), atol=TOL)
    g = np.array([[1.0, 0.5]], np.float32)
    _say("streaming adjoint+expectation n=19")
    exp, grad = adj.expectation_and_gradient(progs, names, vals, ps, g)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g)
    np.testing.assert_allclose(_np(grad), ref_g, atol=TOL * max(1, np.abs(ref_g).max()))
    np.testing.assert_allclose(_np(exp), ref, atol=TOL * max(1, np.abs(ref).max()))


@pytest.fixture(scope="module")
def streaming_case():
    n = 19
    circuits, names, vals, psums = _random_batch(901, n, 1, 2 * n, p_control=0.1, n_sym=4)
    progs, ps = _ser(circuits), _ser(psums)
    g = np.array([[1.0, 0.5]], np.float32)
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g)
    return progs, names, vals, ps, g, ref_e, ref_g


@pytest.fixture(scope="module")
def tiled_case():
    circuits, names, vals, psums = _mixed_batch(802, [5, 12, 13, 14, 16])
    progs, ps = _ser(circuits), _ser(psums)
    g = np.random.default_rng(5).uniform(-1, 1, size=(len(circuits), 2)).astype(np.float32)
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g)
    return progs, names, vals, ps, g, ref_e, ref_g


# Register engine off (0), batched kernel only (1), batched and streaming (2),
# with streaming tiles of 12-13 qubits for one and two states, so that the
# engine runs on one-state and two-state streaming passes and the LDS sweeps
# run where it is off.
@pytest.mark.parametrize("level,tile_single,tile_two", [(0, 13, 11), (1, 12, 12), (2, 13, 13),
                                                        (2, 12, 12)])
def test_register_engine_levels_match_oracle(streaming_case, tiled_case, level, tile_single,
                                             tile_two):
    _, adj, _ = ops()
    _lib.set_tuning(level, tile_single, tile_two)
    try:
        for label, case in (("streaming n=19", streaming_case), ("tiled mixed", tiled_case)):
            progs, names, vals, ps, g, ref_e, ref_g = case
            _say(f"engine level {level}, tiles {tile_single}/{tile_two}: {label}")
            exp, grad = adj.expectation_and_gradient(progs, names, vals, ps, g)
            np.testing.assert_allclose(_np(exp), ref_e, atol=TOL * max(1, np.abs(ref_e).max()))
            np.testing.assert_allclose(_np(grad), ref_g, atol=TOL * max(1, np.abs(ref_g).max()))
    finally:
        _lib.set_tuning()


# Wave mode: circuits of 4..10 qubits run one per wavefront after the
# workgroup-sized ones; 4 qubits leaves every lane a replica of lane 0, 10 uses
# all 64 lanes.  Controls exercise the predicated register paths.
# The planner uses wave mode only for batches with many such circuits; the
# test lowers that threshold to 0 (always) and also runs the same small batch
# on whole workgroups (the default for it).
@pytest.mark.parametrize("widths", [[4, 5, 10, 14, 7, 13], [4, 6, 8, 10]])
@pytest.mark.parametrize("wave_min", [0, -1])
def test_wave_mode_batches(widths, wave_min):
    sim, adj, _ = ops()
    circuits, names, vals, psums = _mixed_batch(803 + len(widths), widths)
    progs, ps = _ser(circuits), _ser(psums)
    g = np.random.default_rng(7).uniform(-1, 1, size=(len(circuits), 2)).astype(np.float32)
    _say(f"wave mode widths {widths}, wave_min {wave_min}")
    _lib.set_wave_min_circuits(wave_min)
    try:
        exp, grad = adj.expectation_and_gradient(progs, names, vals, ps, g)
        out = _np(sim.tfq_simulate_expectation(progs, names, vals, ps))
    finally:
        _lib.set_wave_min_circuits()
    ref_e = O.simulate_expectation(list(progs), names, vals, ps)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g)
    np.testing.assert_allclose(_np(exp), ref_e, atol=TOL * max(1, np.abs(ref_e).max()))
    np.testing.assert_allclose(_np(grad), ref_g, atol=TOL * max(1, np.abs(ref_g).max()))
    np.testing.assert_allclose(out, ref_e, atol=TOL * max(1, np.abs(ref_e).max()))


def test_plain_launch_matches_batched_launch():
    """One launch per circuit (the bench's plain baseline) over a batch with
    wave-mode, LDS-resident and tiled circuits gives the batched results."""
    import torch
    from quantum_amd import tfq_adj_grad_op as A
    circuits, names, vals, psums = _mixed_batch(811, [4, 9, 12, 14, 6, 10])
    progs, ps = _ser(circuits), _ser(psums)
    g = np.random.default_rng(9).uniform(-1, 1, size=(len(circuits), 2)).astype(np.float32)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.set_wave_min_circuits(0)
    try:
        plan, shape = A.plan_adj_grad(progs, names, vals, ps, g, with_expectation=True)
    finally:
        _lib.set_wave_min_circuits()
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    out_plain = torch.empty_like(out)
    plan.upload(dev, stream)
    _say("batched launch, mixed batch")
    plan.launch(out, stream)
    torch.cuda.synchronize()
    _say("plain launches, mixed batch")
    plan.launch(out_plain, stream, plain=True)
    torch.cuda.synchronize()
    np.testing.assert_allclose(_np(out_plain), _np(out), atol=1e-6)
    ref_g = O.adjoint_gradient(list(progs), names, vals, ps, g)
    B, S = ref_g.shape
    got_g = _np(out).reshape(-1)[-B * S:].reshape(B, S)
    np.testing.assert_allclose(got_g, ref_g, atol=TOL * max(1, np.abs(ref_g).max()))


# ----------------------------------------------------------------------------- shared programs
@pytest.mark.parametrize("derivative", ["tfq", "analytic"])
def test_rebound_plan_matches_oracle_and_fresh_plan(derivative):
    """One program tiled over the batch: the plan is built once (structural
    diagonal detection, per-row matrices) and rebound to new values."""
    import torch
    _, adj, _ = ops()
    circuit, names, q = W.hea_ring(8, 3)
    circuit = circuit + C.Circuit([C.Z(q[0]) ** C.Symbol(names[0]), C.CZ(q[1], q[2])])
    ham = W.tfim(q, 0.6)
    rng = np.random.default_rng(11)
    B = 6
    progs = _ser([circuit] * B)
    ps = _ser([[ham]] * B)
    v1 = rng.uniform(-3, 3, (B, len(names))).astype(np.float32)
    v1[0, :] = 0.0  # ry(0) is numerically diagonal; the shared plan keeps it dense
    v2 = rng.uniform(-3, 3, (B, len(names))).astype(np.float32)
    dg = rng.uniform(-1, 1, (B, 1)).astype(np.float32)
    plan, shape = adj.plan_adj_grad(progs, names, v1, ps, dg, derivative, with_expectation=True)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    outs = []
    for vals in (v1, v2):
        _say(f"rebound plan derivative={derivative}")
        plan.rebind(vals)
        out = torch.empty(shape, dtype=torch.float32, device=dev)
        plan.upload(dev, stream)
        plan.launch(out, stream)
        outs.append(_np(out))
        e, g = adj.expectation_and_gradient(progs, names, vals, ps, dg, derivative)
        fresh = np.concatenate([_np(e).reshape(-1), _np(g).reshape(-1)])
        np.testing.assert_array_equal(outs[-1], fresh)
        eo = O.simulate_expectation(list(progs), names, vals, ps)
        go = O.adjoint_gradient(list(progs), names, vals, ps, dg, mode=derivative)
        np.testing.assert_allclose(outs[-1][:B], eo[:, 0], atol=TOL)
        np.testing.assert_allclose(outs[-1][B:].reshape(B, -1), go, atol=TOL)
