"""Pins the CPU oracle before it is trusted as the checker.

How each oracle function is pinned:

* gate matrices / qubit order / expectation: the reference's golden vectors
  (``util_qsim_test.cc`` expectation table, compound sum, amplitudes; the
  ``circuit_execution_ops.py`` docstrings) and brute force against dense
  Kronecker-product unitaries built independently here;
* O*psi accumulation: ``util_qsim_test.cc:437-518`` amplitudes;
* gradient gates: ``adj_util_test.cc:355-560`` FD matrices (all 4 families
  with non-trivial forms: YP, XXP, PXP x2, FSIM x2, PISP x2);
* adjoint gradient: ``tfq_adj_grad_op_test.py:265-397`` goldens, the
  ``gradient_test.py`` closed form, and agreement with parameter shift /
  central differences of the oracle's own expectation on random circuits;
* inner product: ``inner_product_op.py:85-109`` docstring matrix;
* inner-product gradient: a one-qubit closed form, and agreement with
  central differences of the oracle's own inner product on random circuits;
* Philox stream: Random123 known-answer vectors;
* sampler: chi-square against |psi|^2;
* wire decoder: the official protobuf runtime parsing the same bytes.
"""

import math

import numpy as np
import pytest

from oracle import tfq_oracle as O
from oracle import philox, wire
from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W


def _ser(c):
    return S.serialize_circuit(c)


def _psum(p):
    return S.serialize_paulisum(p)


# ----------------------------------------------------------------------------- goldens
def test_two_qubit_amplitudes(goldens):
    g = goldens["two_qubit_state"]
    circ, _ = W.two_qubit_state_circuit()
    st = O.simulate_state([_ser(circ)], [], np.zeros((1, 0)))[0]
    exp = np.array([complex(*a) for a in g["amplitudes"]])
    np.testing.assert_allclose(st, exp, atol=g["tol"])


def test_two_term_expectations(goldens):
    g = goldens["two_term_expectation"]
    circ, (q0, q1) = W.two_qubit_state_circuit()
    ops = []
    keys = list(g["values"])
    for k in keys:
        ops.append(C.PauliSum([C.PauliString({q0: k[0], q1: k[1]}, g["coefficient"])]))
    out = O.simulate_expectation([_ser(circ)], [], np.zeros((1, 0)), [[_psum(p) for p in ops]])
    np.testing.assert_allclose(out[0], [g["values"][k] for k in keys], atol=1e-5)


def test_compound_and_identity(goldens):
    circ, (q0, q1) = W.two_qubit_state_circuit()
    comp = C.PauliSum([C.PauliString({q0: "Z", q1: "X"}, 0.1234),
                       C.PauliString({q0: "X"}, -3.0), C.PauliString({}, 4.0)])
    ident = C.PauliSum([C.PauliString({}, 0.1234)])
    out = O.simulate_expectation([_ser(circ)], [], np.zeros((1, 0)), [[_psum(comp), _psum(ident)]])
    assert abs(out[0, 0] - goldens["compound_expectation"]["value"]) < 1e-5
    assert abs(out[0, 1] - goldens["identity_term"]["value"]) < 1e-5


def test_accumulate_operators(goldens):
    g = goldens["accumulate_operators"]
    circ, (q0, q1) = W.two_qubit_state_circuit()
    p1 = C.PauliSum([C.PauliString({q0: "Z", q1: "X"}, 0.1234),
                     C.PauliString({q0: "X"}, -3.0), C.PauliString({}, 4.0)])
    p2 = C.PauliSum([C.PauliString({}, -5.0)])
    prog = O.Program(_ser(circ))
    psi = O.simulate(prog)
    sums = [O.parse_paulisum(_psum(p), prog.qmap, prog.n) for p in (p1, p2)]
    lam = O.accumulate_operators(psi, sums, [0.5, 0.25], prog.n)
    np.testing.assert_allclose(lam, [complex(*a) for a in g["amplitudes"]], atol=g["tol"])


def _gate_from(gid, params):
    """Build a Gate record directly from raw float32 arguments."""
    g = O.Gate()
    g.gid = gid
    if gid in ("YP", "XXP"):
        g.params = (np.float32(params["exponent"]), np.float32(params["exponent_scalar"]),
                    np.float32(params["global_shift"]))
        g.build = lambda e, es, gs: O.eigen_unitary(gid, np.float32(e) * np.float32(es), gs)
        idx = {"exponent": 0}
    elif gid == "PXP":
        g.params = tuple(np.float32(params[k]) for k in
                         ("phase_exponent", "phase_exponent_scalar", "exponent", "exponent_scalar",
                          "global_shift"))
        g.build = lambda p, ps, e, es, gs: O.phased_x_unitary(np.float32(p) * np.float32(ps),
                                                              np.float32(e) * np.float32(es), gs)
        idx = {"phase_exponent": 0, "exponent": 2}
    elif gid == "FSIM":
        g.params = tuple(np.float32(params[k]) for k in ("theta", "theta_scalar", "phi", "phi_scalar"))
        g.build = lambda t, ts, p, ps: O.fsim_unitary(np.float32(t) * np.float32(ts),
                                                      np.float32(p) * np.float32(ps))
        idx = {"theta": 0, "phi": 2}
    else:
        g.params = tuple(np.float32(params[k]) for k in
                         ("phase_exponent", "phase_exponent_scalar", "exponent", "exponent_scalar"))
        g.build = lambda p, ps, e, es: O.phased_iswap_unitary(np.float32(p) * np.float32(ps),
                                                              np.float32(e) * np.float32(es))
        idx = {"phase_exponent": 0, "exponent": 2}
    return g, idx


@pytest.mark.parametrize("case", ["YP_t0.125", "XXP_t0.001", "PXP_phase", "PXP_exponent",
                                  "FSIM_theta", "FSIM_phi", "PISP_phase", "PISP_exponent"])
def test_fd_gradient_gates(goldens, case):
    c = goldens["fd_gradient_gates"][case]
    gid = case.split("_")[0]
    g, idx = _gate_from(gid, c["params"])
    m = np.asarray(c["matrix"]).reshape(-1, 2)
    exp = (m[:, 0] + 1j * m[:, 1]).reshape(int(math.isqrt(len(m))), -1)
    fd = O._fd_matrix(g, idx[c["wrt"]])
    np.testing.assert_allclose(fd, exp, atol=c["tol"])
    # The analytic derivative agrees with the FD gate to O(eps^2).
    an = O._analytic_matrix(g, idx[c["wrt"]])
    np.testing.assert_allclose(an, fd, atol=3e-3 * max(1.0, np.abs(an).max()))


@pytest.mark.parametrize("name", ["simple_case", "simple_case2", "simple_case_shared",
                                  "simple_case_single"])
@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_adjoint_goldens(goldens, name, mode):
    g = goldens["adjoint_gradients"]
    circ, names, q = W.adjoint_golden_cases()[name]
    vals = np.array([[g["values"][s] for s in names]], dtype=np.float32)
    ops = [[_psum(C.pauli_z(q[0])), _psum(C.pauli_x(q[1]))]]
    out = O.adjoint_gradient([_ser(circ)], names, vals, ops, np.ones((1, 2)), mode=mode)
    np.testing.assert_allclose(out, g[name]["expected"], atol=g["tol"])


def test_x_pow_alpha_closed_form(goldens):
    g = goldens["x_pow_alpha_gradient"]
    q = C.GridQubit(0, 0)
    circ = C.Circuit(C.X(q) ** C.Symbol("alpha"))
    args = ([_ser(circ)], ["alpha"], np.array([[g["alpha"]]], np.float32),
            [[_psum(C.pauli_z(q))]], np.ones((1, 1)))
    an = O.adjoint_gradient(*args, mode="analytic")[0, 0]
    fd = O.adjoint_gradient(*args, mode="tfq")[0, 0]
    # d/da cos(pi a) = -pi sin(pi a)
    assert abs(an - (-math.pi * math.sin(math.pi * g["alpha"]))) < 1e-6
    assert abs(an - g["analytic"]) < 1e-6
    assert abs(fd - g["analytic"]) < 1e-4


def test_docstring_expectation_and_state(goldens):
    q = C.GridQubit(0, 0)
    circ = C.Circuit(C.H(q) ** C.Symbol("alpha"))
    ps = 3.5 * C.pauli_x(q) - 2.2 * C.pauli_y(q)
    out = O.simulate_expectation([_ser(circ)], ["alpha"], [[0.123]], [[_psum(ps)]])
    assert abs(out[0, 0] - goldens["h_pow_expectation"]["value"]) < 1e-5
    circ = C.Circuit(C.Y(q) ** C.Symbol("alpha"))
    st = O.simulate_state([_ser(circ)], ["alpha"], [[0.5]])
    np.testing.assert_allclose(st[0], [0.5 + 0.5j, 0.5 + 0.5j], atol=1e-6)


def test_inner_product_docstring(goldens):
    g = goldens["inner_product"]
    a, b = C.Symbol("alpha"), C.Symbol("beta")
    q = C.GridQubit.rect(1, 2)
    refs = [C.Circuit([(C.H ** a).on(x) for x in q]), C.Circuit(C.X(q[0]) ** a, C.Y(q[1]) ** b)]
    others = [C.Circuit([C.X(x) for x in q]), C.Circuit([(C.Y ** 0.125).on(x) for x in q]),
              C.Circuit([(C.X ** 0.5).on(x) for x in q])]
    out = O.inner_product([_ser(c) for c in refs], ["alpha", "beta"], g["values"],
                          [[_ser(c) for c in others]] * 2)
    exp = np.array([[complex(*v) for v in row] for row in g["expected"]])
    np.testing.assert_allclose(out, exp, atol=g["tol"])


def test_inner_product_gradient_closed_form():
    """psi = X**a |0>, phi = |0>: <psi|phi> = (1 + e^{-i pi a}) / 2, so the
    gradient with weight g is g * (-i pi / 2) e^{-i pi a}."""
    q = C.LineQubit(0)
    a = C.Symbol("a")
    val, g = 0.37, 0.8
    out = O.inner_product_gradient([_ser(C.Circuit((C.X ** a).on(q)))], ["a"],
                                   np.array([[val]], np.float32),
                                   [[_ser(C.Circuit(C.I(q)))]], np.array([[g]]),
                                   mode="analytic")
    a32 = float(np.float32(val))
    want = g * (-1j * math.pi / 2) * np.exp(-1j * math.pi * a32)
    np.testing.assert_allclose(out[0, 0], want, atol=1e-9)
    tfq = O.inner_product_gradient([_ser(C.Circuit((C.X ** a).on(q)))], ["a"],
                                   np.array([[val]], np.float32),
                                   [[_ser(C.Circuit(C.I(q)))]], np.array([[g]]))
    np.testing.assert_allclose(tfq[0, 0], want, atol=1e-4)


def test_inner_product_gradient_vs_central_difference():
    rng = np.random.default_rng(12)
    q = C.LineQubit.range(4)
    names = ["a", "b", "c"]
    circ = W.random_circuit(rng, q, 25, symbols=names, p_symbol=0.6, p_control=0.2)
    others = [[_ser(W.random_circuit(rng, q, 12) + C.Circuit([C.I(x) for x in q]))
               for _ in range(2)]]
    vals = rng.uniform(-1, 1, size=(1, 3)).astype(np.float32)
    g = np.array([[0.7, -1.3]])
    an = O.inner_product_gradient([_ser(circ)], names, vals, others, g, mode="analytic")[0]
    h = 1e-3
    for k in range(3):
        vp = vals.astype(np.float64).copy()
        vm = vp.copy()
        vp[0, k] += h
        vm[0, k] -= h
        ip = (O.inner_product([_ser(circ)], names, vp, others)
              - O.inner_product([_ser(circ)], names, vm, others))[0] / (2 * h)
        # float32 symbol values limit CD accuracy to ~ eps32/h
        np.testing.assert_allclose(an[k], np.dot(g[0], ip), atol=3e-3)


# ----------------------------------------------------------------------------- brute force
def _dense(U, bits, n, cbits=(), cvals=()):
    """Full 2^n unitary of (controlled) U on ``bits`` by explicit index maps."""
    dim = 2 ** n
    M = np.zeros((dim, dim), dtype=complex)
    k = len(bits)
    for col in range(dim):
        if any(((col >> b) & 1) != v for b, v in zip(cbits, cvals)):
            M[col, col] = 1
            continue
        sub = 0
        for j, b in enumerate(bits):
            sub |= ((col >> b) & 1) << (k - 1 - j)
        for r in range(2 ** k):
            row = col
            for j, b in enumerate(bits):
                bit = (r >> (k - 1 - j)) & 1
                row = (row & ~(1 << b)) | (bit << b)
            M[row, col] += U[r, sub]
    return M


def test_brute_force_dense_unitaries():
    rng = np.random.default_rng(7)
    n = 4
    q = C.LineQubit.range(n)
    circ = W.random_circuit(rng, q, 40, symbols=("a", "b"), p_control=0.3)
    vals = {"a": np.float32(0.37), "b": np.float32(-1.1)}
    prog = O.Program(_ser(circ), vals)
    psi = O.simulate(prog)
    full = np.eye(2 ** n, dtype=complex)
    for g in prog.gates:
        full = _dense(g.unitary(), g.bits, n, g.cbits, g.cvals) @ full
    np.testing.assert_allclose(psi, full[:, 0], atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(psi), 1.0, atol=1e-12)


def test_kron_convention_matches_reference_qubit_order():
    """qubit 0 (sorted first) is the most significant state bit."""
    q = C.LineQubit.range(3)
    prog = O.Program(_ser(C.Circuit(C.X(q[0]), C.I(q[1]), C.I(q[2]))))
    psi = O.simulate(prog)
    assert abs(psi[4] - 1) < 1e-12


def test_unitarity_all_families():
    rng = np.random.default_rng(3)
    for fam in W.ALL_GATE_FAMILIES:
        circ = C.Circuit(W.random_gate_op(rng, C.LineQubit.range(3), [], fam))
        g = O.Program(_ser(circ)).gates[0]
        U = g.unitary()
        np.testing.assert_allclose(U @ U.conj().T, np.eye(len(U)), atol=1e-12)


def _two_qubit_unitary(gate):
    q = C.LineQubit.range(2)
    return O.Program(_ser(C.Circuit(gate.on(q[0], q[1])))).gates[0].unitary()


@pytest.mark.parametrize("t,gs", [(1.0, 0.0), (0.5, 0.0), (0.3125, -0.5), (-0.71, 0.25)])
def test_two_qubit_eigen_gates_closed_forms(t, gs):
    """Closed forms independent of the eigen-projector sums the oracle uses:
    ZZ^t = diag(1, w, w, 1), CZ^t = diag(1, 1, 1, w), and on the {|01>, |10>}
    block SWAP^t = [[a, b], [b, a]] with a = (1 + w)/2, b = (1 - w)/2 and
    ISWAP^t = [[c, i s], [i s, c]] with c = cos(pi t/2), s = sin(pi t/2);
    XX^t = a I + b X(x)X and YY^t = a I + b Y(x)Y (P^2 = I for a Pauli
    product P); CNOT^t = I (+) X^t with X^t = a I + b X; all times the global
    phase e^{i pi t gs} (w = e^{i pi t})."""
    w = np.exp(1j * np.pi * t)
    ph = np.exp(1j * np.pi * t * gs)
    a, b = (1 + w) / 2, (1 - w) / 2
    I4 = np.eye(4)
    X = np.array([[0, 1], [1, 0]])
    Y = np.array([[0, -1j], [1j, 0]])
    cases = {
        "ZZ": (C.ZZPowGate(exponent=t, global_shift=gs), np.diag([1, w, w, 1])),
        "CZ": (C.CZPowGate(exponent=t, global_shift=gs), np.diag([1, 1, 1, w])),
        "SWAP": (C.SwapPowGate(exponent=t, global_shift=gs),
                 np.array([[1, 0, 0, 0], [0, a, b, 0], [0, b, a, 0], [0, 0, 0, 1]])),
        "ISWAP": (C.ISwapPowGate(exponent=t, global_shift=gs),
                  np.array([[1, 0, 0, 0],
                            [0, np.cos(np.pi * t / 2), 1j * np.sin(np.pi * t / 2), 0],
                            [0, 1j * np.sin(np.pi * t / 2), np.cos(np.pi * t / 2), 0],
                            [0, 0, 0, 1]])),
        "XX": (C.XXPowGate(exponent=t, global_shift=gs), a * I4 + b * np.kron(X, X)),
        "YY": (C.YYPowGate(exponent=t, global_shift=gs), a * I4 + b * np.kron(Y, Y)),
        "CNOT": (C.CNotPowGate(exponent=t, global_shift=gs),
                 np.block([[np.eye(2), np.zeros((2, 2))],
                           [np.zeros((2, 2)), a * np.eye(2) + b * X]])),
    }
    for name, (gate, U) in cases.items():
        got = _two_qubit_unitary(gate)
        np.testing.assert_allclose(got, ph * U, atol=1e-6, err_msg=name)


def test_swap_and_iswap_special_points():
    """SWAP^1 is the permutation |01> <-> |10>; SWAP^0.5 is Cirq's sqrt-SWAP
    with entries (1 +- i)/2; ISWAP^1 maps |01> -> i|10>; ZZ^1 = Z (x) Z."""
    perm = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])
    np.testing.assert_allclose(_two_qubit_unitary(C.SwapPowGate(exponent=1.0)), perm, atol=1e-7)
    h = np.array([[1, 0, 0, 0], [0, (1 + 1j) / 2, (1 - 1j) / 2, 0],
                  [0, (1 - 1j) / 2, (1 + 1j) / 2, 0], [0, 0, 0, 1]])
    np.testing.assert_allclose(_two_qubit_unitary(C.SwapPowGate(exponent=0.5)), h, atol=1e-7)
    isw = _two_qubit_unitary(C.ISwapPowGate(exponent=1.0))
    np.testing.assert_allclose(isw @ np.array([0, 1, 0, 0]), [0, 0, 1j, 0], atol=1e-7)
    Z = np.diag([1, -1])
    np.testing.assert_allclose(_two_qubit_unitary(C.ZZPowGate(exponent=1.0)), np.kron(Z, Z),
                               atol=1e-7)


def test_adjoint_vs_parameter_shift_and_cd():
    rng = np.random.default_rng(11)
    q = C.LineQubit.range(4)
    names = ["a", "b", "c"]
    circ = W.random_circuit(rng, q, 25, symbols=names, p_symbol=0.6, p_control=0.2)
    ps = W.random_pauli_sum(rng, q, 4)
    vals = rng.uniform(-1, 1, size=(1, 3)).astype(np.float32)
    args = ([_ser(circ)], names, vals, [[_psum(ps)]])
    an = O.adjoint_gradient(*args, np.ones((1, 1)), mode="analytic")[0]
    f = O.expectation_function(*[args[0], args[1], args[3]])
    h = 1e-3
    cd = []
    for j in range(3):
        vp = vals.astype(np.float64).copy()
        vm = vp.copy()
        vp[0, j] += h
        vm[0, j] -= h
        cd.append((f(vp)[0, 0] - f(vm)[0, 0]) / (2 * h))
    # float32 symbol values limit CD accuracy to ~ eps32/h
    np.testing.assert_allclose(an, cd, atol=2e-3)


# ----------------------------------------------------------------------------- RNG / sampler
def test_philox_known_answers(goldens):
    for case in goldens["philox4x32_10_kat"]["cases"]:
        out = philox.philox4x32_10(*[[c] for c in case["ctr"]], *case["key"])
        got = [f"{int(x[0]):08x}" for x in out]
        assert got == case["out"], (case, got)


def test_sampler_chi_square():
    from scipy import stats
    rng = np.random.default_rng(5)
    q = C.LineQubit.range(4)
    circ = W.random_circuit(rng, q, 30)
    psi = O.simulate(O.Program(_ser(circ)))
    p = np.abs(psi) ** 2
    N = 200000
    idx = O.sample_indices(p, 1234, 0, 0, N)
    counts = np.bincount(idx, minlength=len(p))
    keep = p * N > 5
    chi2 = np.sum((counts[keep] - N * p[keep]) ** 2 / (N * p[keep]))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4


def test_sampled_expectation_statistics(goldens):
    g = goldens["two_term_expectation"]
    circ, (q0, q1) = W.two_qubit_state_circuit()
    ops = [C.PauliSum([C.PauliString({q0: k[0], q1: k[1]}, g["coefficient"])]) for k in g["values"]]
    out = O.simulate_sampled_expectation([_ser(circ)], [], np.zeros((1, 0)),
                                         [[_psum(p) for p in ops]], [[100000] * len(ops)], 99)
    np.testing.assert_allclose(out[0], list(g["values"].values()), atol=1e-2)


# ----------------------------------------------------------------------------- wire format
def _dynamic_messages():
    """Program / PauliSum classes built by the protobuf runtime from descriptors
    restating program.proto / pauli_sum.proto field numbers and types."""
    from google.protobuf import descriptor_pb2, descriptor_pool
    from google.protobuf import message_factory
    FD = descriptor_pb2.FieldDescriptorProto
    fdp = descriptor_pb2.FileDescriptorProto(name="tfqx_check.proto", package="chk", syntax="proto3")

    def msg(name, fields, nested=()):
        m = fdp.message_type.add(name=name)
        for fname, num, typ, label, tname in fields:
            f = m.field.add(name=fname, number=num, type=typ, label=label)
            if tname:
                f.type_name = tname
        for nm in nested:
            m.nested_type.add().CopyFrom(nm)
        return m
    O1, R = FD.LABEL_OPTIONAL, FD.LABEL_REPEATED
    msg("ArgValue", [("float_value", 1, FD.TYPE_FLOAT, O1, None),
                     ("string_value", 3, FD.TYPE_STRING, O1, None),
                     ("double_value", 4, FD.TYPE_DOUBLE, O1, None)])
    msg("Arg", [("arg_value", 1, FD.TYPE_MESSAGE, O1, ".chk.ArgValue"),
                ("symbol", 2, FD.TYPE_STRING, O1, None)])
    entry = descriptor_pb2.DescriptorProto(name="ArgsEntry")
    entry.field.add(name="key", number=1, type=FD.TYPE_STRING, label=O1)
    entry.field.add(name="value", number=2, type=FD.TYPE_MESSAGE, label=O1, type_name=".chk.Arg")
    entry.options.map_entry = True
    msg("Gate", [("id", 1, FD.TYPE_STRING, O1, None)])
    msg("Qubit", [("id", 2, FD.TYPE_STRING, O1, None)])
    msg("Operation", [("gate", 1, FD.TYPE_MESSAGE, O1, ".chk.Gate"),
                      ("args", 2, FD.TYPE_MESSAGE, R, ".chk.Operation.ArgsEntry"),
                      ("qubits", 3, FD.TYPE_MESSAGE, R, ".chk.Qubit")], [entry])
    msg("Moment", [("operations", 1, FD.TYPE_MESSAGE, R, ".chk.Operation")])
    msg("Circuit", [("scheduling_strategy", 1, FD.TYPE_INT32, O1, None),
                    ("moments", 2, FD.TYPE_MESSAGE, R, ".chk.Moment")])
    msg("Language", [("gate_set", 1, FD.TYPE_STRING, O1, None),
                     ("arg_function_language", 2, FD.TYPE_STRING, O1, None)])
    msg("Program", [("language", 1, FD.TYPE_MESSAGE, O1, ".chk.Language"),
                    ("circuit", 2, FD.TYPE_MESSAGE, O1, ".chk.Circuit")])
    msg("PauliQubitPair", [("qubit_id", 1, FD.TYPE_STRING, O1, None),
                           ("pauli_type", 2, FD.TYPE_STRING, O1, None)])
    msg("PauliTerm", [("coefficient_real", 1, FD.TYPE_FLOAT, O1, None),
                      ("coefficient_imag", 2, FD.TYPE_FLOAT, O1, None),
                      ("paulis", 3, FD.TYPE_MESSAGE, R, ".chk.PauliQubitPair")])
    msg("PauliSum", [("terms", 1, FD.TYPE_MESSAGE, R, ".chk.PauliTerm")])
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fdp)
    get = message_factory.GetMessageClass
    return get(pool.FindMessageTypeByName("chk.Program")), get(pool.FindMessageTypeByName("chk.PauliSum"))


def test_wire_format_against_protobuf_runtime():
    Program, PauliSum = _dynamic_messages()
    q = C.GridQubit.rect(2, 2)
    a = C.Symbol("alpha")
    circ = C.Circuit(C.rx(a).on(q[0]), C.CNOT(q[0], q[1]).controlled_by(q[2], control_values=[0]),
                     C.FSimGate(0.25, a * 2.0)(q[1], q[3]), C.I(q[2]))
    data = S.serialize_circuit(circ)
    p = Program()
    p.ParseFromString(data)
    assert p.language.gate_set == "tfq_gate_set"
    assert p.circuit.scheduling_strategy == 1
    op0 = p.circuit.moments[0].operations[0]
    assert op0.gate.id == "XP" and op0.qubits[0].id == "0_0"
    assert op0.args["exponent"].symbol == "alpha"
    assert abs(op0.args["exponent_scalar"].arg_value.float_value - 0.31831) < 1e-6
    assert op0.args["global_shift"].arg_value.float_value == -0.5
    ops = [o for m in p.circuit.moments for o in m.operations]
    cn = [o for o in ops if o.gate.id == "CNP"][0]
    assert cn.args["control_qubits"].arg_value.string_value == "1_0"
    assert cn.args["control_values"].arg_value.string_value == "0"
    fs = [o for o in ops if o.gate.id == "FSIM"][0]
    assert fs.args["phi"].symbol == "alpha" and fs.args["phi_scalar"].arg_value.float_value == 2.0
    # The oracle's own decoder reads the same structure.
    dec = wire.decode_program(data)
    assert [o["gate"] for m in dec for o in m] == [o.gate.id for o in ops]
    ps = 1.5 * C.pauli_z(q[1]) * C.pauli_x(q[0]) + C.PauliString({}, -0.25)
    m = PauliSum()
    m.ParseFromString(S.serialize_paulisum(ps))
    assert [(x.qubit_id, x.pauli_type) for x in m.terms[0].paulis] == [("0_0", "X"), ("0_1", "Z")]
    assert m.terms[0].coefficient_real == 1.5 and m.terms[1].coefficient_real == -0.25


def test_serializer_rounds_to_six_decimals():
    q = C.LineQubit(0)
    dec = wire.decode_program(S.serialize_circuit(C.Circuit(C.rx(0.5).on(q))))
    e = dec[0][0]["args"]["exponent"]["float"]
    assert e == np.float32(0.159155)


def test_serializer_removes_measurements_as_the_reference():
    """``serializer.py:826-860``: measurements are removed (their moment
    stays, without the ops on the qubits measured in it); a qubit measured
    twice is an error; a gate after a measurement is kept."""
    q = C.LineQubit.range(3)
    with_m = C.Circuit.from_moments([[C.H(q[0])], [C.CNOT(q[0], q[1]), C.X(q[2])],
                                     [C.measure(q[0], q[1], key="m"), C.Y(q[2])]])
    without = C.Circuit.from_moments([[C.H(q[0])], [C.CNOT(q[0], q[1]), C.X(q[2])],
                                      [C.Y(q[2])]])
    assert S.serialize_circuit(with_m) == S.serialize_circuit(without)
    later = C.Circuit.from_moments([[C.measure(q[0])], [C.X(q[0])]])
    assert S.serialize_circuit(later) == S.serialize_circuit(
        C.Circuit.from_moments([[], [C.X(q[0])]]))
    twice = C.Circuit.from_moments([[C.measure(q[0])], [C.measure(q[0], q[1])]])
    with pytest.raises(ValueError, match="non-terminal measurement"):
        S.serialize_circuit(twice)
    # The measured program simulates like the unmeasured one.
    ps = [[C.PauliSum([C.pauli_z(q[1]) * 0.5, C.pauli_x(q[2])])]]
    e1 = O.simulate_expectation([S.serialize_circuit(with_m)], [], np.zeros((1, 0)),
                                S.convert_to_tensor(ps))
    e0 = O.simulate_expectation([S.serialize_circuit(without)], [], np.zeros((1, 0)),
                                S.convert_to_tensor(ps))
    np.testing.assert_allclose(e1, e0, atol=1e-12)
