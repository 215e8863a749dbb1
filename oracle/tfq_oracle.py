"""complex128 CPU restatement of TFQ's circuit-simulation op family.

TEST INFRASTRUCTURE ONLY -- the checker, never the thing measured or shipped.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.

It restates, in numpy complex128, what TFQ 0.7.7 computes for:

* ``tfq_simulate_expectation``  (``core/ops/tfq_simulate_expectation_op.cc:45-283``,
  per-term copy/apply/inner product of ``core/src/util_qsim.h:142-188``);
* ``tfq_simulate_state``        (``core/ops/tfq_simulate_state_op.cc:43-242``);
* ``tfq_simulate_samples``      (``core/ops/tfq_simulate_samples_op.cc:48-285``);
* ``tfq_simulate_sampled_expectation`` (``core/ops/tfq_simulate_sampled_expectation_op.cc``,
  ``core/src/util_qsim.h:199-270``);
* ``tfq_inner_product``         (``core/ops/math_ops/tfq_inner_product.cc:45-323``);
* ``tfq_adj_grad``              (``core/ops/tfq_adj_grad_op.cc:46-427``,
  ``core/src/adj_util.cc:37-302``, ``core/src/util_qsim.h:363-414``);
* ``tfq_inner_product_grad``    (``core/ops/math_ops/tfq_inner_product_grad.cc:46-501``,
  ``core/src/util_qsim.h:423-442``).

Circuit decoding restates ``core/src/program_resolution.cc:49-186`` (qubit
renumbering) and ``core/src/circuit_parser_qsim.cc:53-596`` (gate ids, argument
names, ``exponent*exponent_scalar`` in float32, qubit q -> bit n-1-q, controls).

Gate matrices are Cirq's EigenGate closed forms ``sum_k e^{i pi t (l_k+s)} P_k``
(the definitions qsim's ``gates_cirq.h`` implements; qsim itself is not in the
reference tree).  How each part is pinned (golden vectors from the reference's
own tests, closed forms, brute force) is listed in ``tests/test_oracle.py``.

Sampling uses the counter-based RNG of ``oracle/philox.py`` so its stream can be
compared bit-for-bit with the device sampler; its *statistics* are what pins it
against TFQ (TFQ seeds nondeterministically, ``tfq_simulate_samples_op.cc:149-152``).
"""

import numpy as np

from . import wire
from .philox import uniform_doubles

F32 = np.float32
GRAD_EPS = F32(5e-3)  # adj_util.cc:32
INT_MAX = 2147483647

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.diag([1.0, -1.0]).astype(np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2.0)
_PAULI = {"X": _X, "Y": _Y, "Z": _Z}


class OracleError(ValueError):
    pass


# --------------------------------------------------------------------------- gate matrices
def _proj(v):
    v = np.asarray(v, dtype=np.complex128)
    return np.outer(v, v.conj())


def _eigen_components(gid):
    """Eigen-decomposition ``[(lambda, P)]`` of each EigenGate family."""
    if gid == "XP":
        return [(0.0, (_I2 + _X) / 2), (1.0, (_I2 - _X) / 2)]
    if gid == "YP":
        return [(0.0, (_I2 + _Y) / 2), (1.0, (_I2 - _Y) / 2)]
    if gid == "ZP":
        return [(0.0, np.diag([1.0, 0]).astype(complex)), (1.0, np.diag([0, 1.0]).astype(complex))]
    if gid == "HP":
        return [(0.0, (_I2 + _H) / 2), (1.0, (_I2 - _H) / 2)]
    I4 = np.eye(4, dtype=np.complex128)
    if gid in ("XXP", "YYP", "ZZP"):
        P = _PAULI[gid[0]]
        PP = np.kron(P, P)
        return [(0.0, (I4 + PP) / 2), (1.0, (I4 - PP) / 2)]
    if gid == "CZP":
        p = _proj([0, 0, 0, 1])
        return [(0.0, I4 - p), (1.0, p)]
    if gid == "CNP":
        p = np.kron(np.diag([0, 1.0]), (_I2 - _X) / 2).astype(complex)
        return [(0.0, I4 - p), (1.0, p)]
    if gid == "SP":
        p = _proj(np.array([0, 1, -1, 0]) / np.sqrt(2))
        return [(0.0, I4 - p), (1.0, p)]
    if gid == "ISP":
        return [(0.0, _proj([1, 0, 0, 0]) + _proj([0, 0, 0, 1])),
                (0.5, _proj(np.array([0, 1, 1, 0]) / np.sqrt(2))),
                (-0.5, _proj(np.array([0, 1, -1, 0]) / np.sqrt(2)))]
    raise KeyError(gid)


def eigen_unitary(gid, t, gs):
    t = float(t)
    gs = float(gs)
    return sum(np.exp(1j * np.pi * t * (lam + gs)) * P for lam, P in _eigen_components(gid))


def phased_x_unitary(p, t, gs):
    z = np.diag([1.0, np.exp(1j * np.pi * float(p))])
    return z @ eigen_unitary("XP", t, gs) @ z.conj().T


def phased_iswap_unitary(p, t):
    d = np.diag([1.0, np.exp(1j * np.pi * float(p)), np.exp(-1j * np.pi * float(p)), 1.0])
    return d @ eigen_unitary("ISP", t, 0.0) @ d.conj().T


def fsim_unitary(theta, phi):
    c, s = np.cos(float(theta)), np.sin(float(theta))
    return np.array([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0],
                     [0, 0, 0, np.exp(-1j * float(phi))]], dtype=np.complex128)


# --------------------------------------------------------------------------- parsing
def _qubit_key(qid):
    parts = qid.split("_")
    if len(parts) == 1:
        parts = [str(INT_MAX)] + parts
    if len(parts) != 2:
        raise OracleError("Unable to parse qubit: " + qid)
    try:
        return (int(parts[0]), int(parts[1])), qid
    except ValueError:
        raise OracleError("Unable to parse qubit: " + qid)


def _ctrl_str(op):
    a = op["args"].get("control_qubits")
    return a["string"] if a else ""


def resolve_qubits(moments):
    """``ResolveQubitIds`` (program_resolution.cc:90-186): id -> index map."""
    if not moments:
        return 0, {}
    ids = set()
    for ops in moments:
        for op in ops:
            for q in op["qubits"]:
                ids.add(_qubit_key(q))
            cs = _ctrl_str(op)
            if cs:
                for q in cs.split(","):
                    ids.add(_qubit_key(q))
    ordered = sorted(ids)
    return len(ordered), {qid: i for i, (_, qid) in enumerate(ordered)}


class Gate:
    """One parsed gate: matrix builder, target bits (first target = matrix MSB),
    control bits/values and the symbolic parameters it depends on."""

    __slots__ = ("gid", "bits", "cbits", "cvals", "params", "symbols", "build")

    def unitary(self, params=None):
        return self.build(*(self.params if params is None else params))


def _arg(op, name, symmap, used=None):
    a = op["args"].get(name)
    if a is None:
        raise OracleError(f"Could not find arg: {name} in op.")
    val = F32(a["float"])
    if a["kind"] == "symbol" and a["symbol"]:
        if a["symbol"] not in symmap:
            raise OracleError("Could not find symbol in parameter map: " + a["symbol"])
        val = F32(symmap[a["symbol"]])
        if used is not None:
            used.append(a["symbol"])
    return val


def parse_gate(op, symmap, n, qmap):
    """One operation -> Gate, restating the builders of circuit_parser_qsim.cc:153-562.

    ``params`` are the float32 raw arguments; ``build(*params)`` evaluates the
    float32 effective parameters (``exp*exp_s``) and returns the unitary.
    ``symbols`` lists ``(symbol, index_into_params)`` for symbolic arguments."""
    gid = op["gate"]
    g = Gate()
    g.gid = gid
    g.bits = [n - 1 - qmap[q] for q in op["qubits"]]
    cs = _ctrl_str(op)
    cv = op["args"].get("control_values")
    cv = cv["string"] if cv else ""
    g.cbits, g.cvals = [], []
    if cs or cv:
        cq = cs.split(",")
        vals = cv.split(",")
        if len(cq) != len(vals):
            raise OracleError("Mistmatched number of control qubits and control values.")
        g.cbits = [n - 1 - qmap[q] for q in cq]
        g.cvals = [int(v) for v in vals]
    g.symbols = []
    if gid in ("I", "I2"):
        g.params = ()
        k = 1 if gid == "I" else 2
        g.build = lambda: np.eye(2 ** k, dtype=np.complex128)
    elif gid in ("XP", "YP", "ZP", "HP", "XXP", "YYP", "ZZP", "CZP", "CNP", "SP", "ISP"):
        used = []
        e = _arg(op, "exponent", symmap, used)
        es = _arg(op, "exponent_scalar", symmap)
        gs = _arg(op, "global_shift", symmap)
        g.params = (e, es, gs)
        if used:
            g.symbols = [(used[0], 0)]
        g.build = lambda e, es, gs, gid=gid: eigen_unitary(gid, F32(e) * F32(es), gs)
    elif gid == "PXP":
        used_e, used_p = [], []
        e = _arg(op, "exponent", symmap, used_e)
        es = _arg(op, "exponent_scalar", symmap)
        p = _arg(op, "phase_exponent", symmap, used_p)
        ps = _arg(op, "phase_exponent_scalar", symmap)
        gs = _arg(op, "global_shift", symmap)
        g.params = (p, ps, e, es, gs)
        if used_p:
            g.symbols.append((used_p[0], 0))
        if used_e:
            g.symbols.append((used_e[0], 2))
        g.build = lambda p, ps, e, es, gs: phased_x_unitary(F32(p) * F32(ps), F32(e) * F32(es), gs)
    elif gid == "FSIM":
        used_t, used_f = [], []
        th = _arg(op, "theta", symmap, used_t)
        ts = _arg(op, "theta_scalar", symmap)
        ph = _arg(op, "phi", symmap, used_f)
        fs = _arg(op, "phi_scalar", symmap)
        g.params = (th, ts, ph, fs)
        if used_t:
            g.symbols.append((used_t[0], 0))
        if used_f:
            g.symbols.append((used_f[0], 2))
        g.build = lambda th, ts, ph, fs: fsim_unitary(F32(th) * F32(ts), F32(ph) * F32(fs))
    elif gid == "PISP":
        used_e, used_p = [], []
        e = _arg(op, "exponent", symmap, used_e)
        es = _arg(op, "exponent_scalar", symmap)
        p = _arg(op, "phase_exponent", symmap, used_p)
        ps = _arg(op, "phase_exponent_scalar", symmap)
        g.params = (p, ps, e, es)
        if used_p:
            g.symbols.append((used_p[0], 0))
        if used_e:
            g.symbols.append((used_e[0], 2))
        g.build = lambda p, ps, e, es: phased_iswap_unitary(F32(p) * F32(ps), F32(e) * F32(es))
    else:
        raise OracleError(f"Could not parse gate id: {gid}. This is likely because a "
                          "cirq.Channel was used in an op that does not support them.")
    return g


class Program:
    def __init__(self, data, symmap=None, qmap=None, n=None):
        moments = wire.decode_program(data)
        if qmap is None:
            n, qmap = resolve_qubits(moments)
        self.n = n
        self.qmap = qmap
        self.gates = []
        if n > 0:
            for ops in moments:
                for op in ops:
                    self.gates.append(parse_gate(op, symmap or {}, n, qmap))


def parse_paulisum(data, qmap, n):
    """PauliSum -> [(coef_real, [(bit, 'X'|'Y'|'Z')])] with ids resolved
    (program_resolution.cc:163-181)."""
    out = []
    for cr, ci, paulis in wire.decode_paulisum(data):
        ps = []
        for qid, pt in paulis:
            if qid not in qmap:
                raise OracleError("Found a Pauli sum operating on qubits not found in circuit.")
            ps.append((n - 1 - qmap[qid], pt))
        out.append((float(F32(cr)), ps))
    return out


# --------------------------------------------------------------------------- state ops
def zero_state(n):
    s = np.zeros(2 ** n, dtype=np.complex128)
    s[0] = 1.0
    return s


def apply_matrix(state, U, bits, n, cbits=(), cvals=()):
    """state <- (controlled) U on ``bits`` (bits[0] is the matrix-index MSB)."""
    k = len(bits)
    psi = state.reshape([2] * n)
    axes = [n - 1 - b for b in bits]
    Ut = U.reshape([2] * (2 * k))
    out = np.tensordot(Ut, psi, axes=(list(range(k, 2 * k)), axes))
    out = np.moveaxis(out, list(range(k)), axes).reshape(-1)
    if cbits:
        idx = np.arange(2 ** n, dtype=np.int64)
        ok = np.ones(2 ** n, dtype=bool)
        for b, v in zip(cbits, cvals):
            ok &= ((idx >> b) & 1) == v
        out = np.where(ok, out, state)
    return out


def apply_gate(state, g, n, dagger=False, U=None):
    U = g.unitary() if U is None else U
    if dagger:
        U = U.conj().T
    return apply_matrix(state, U, g.bits, n, g.cbits, g.cvals)


def simulate(prog):
    s = zero_state(prog.n)
    for g in prog.gates:
        s = apply_gate(s, g, prog.n)
    return s


def pauli_apply(state, paulis, n):
    s = state
    for b, pt in paulis:
        s = apply_matrix(s, _PAULI[pt], [b], n)
    return s


def expectation_of_sum(state, terms, n):
    """``ComputeExpectationQsim``: sum_t Re(c_t) Re<psi|P_t|psi>, identity terms
    add Re(c) (util_qsim.h:152-186)."""
    total = 0.0
    for c, paulis in terms:
        if not paulis:
            total += c
            continue
        total += c * np.vdot(state, pauli_apply(state, paulis, n)).real
    return total


def _symbol_maps(symbol_names, symbol_values):
    vals = np.asarray(symbol_values, dtype=np.float32)
    return [{str(name): vals[i, j] for j, name in enumerate(symbol_names)}
            for i in range(vals.shape[0])]


def _name(x):
    return x.decode() if isinstance(x, bytes) else str(x)


def simulate_expectation(programs, symbol_names, symbol_values, pauli_sums):
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    O = np.asarray(pauli_sums, dtype=object).shape[1] if B else 0
    out = np.zeros((B, O), dtype=np.float64)
    for i in range(B):
        prog = Program(programs[i], maps[i])
        sums = [parse_paulisum(pauli_sums[i][j], prog.qmap, prog.n) for j in range(O)]
        if not prog.gates:
            out[i, :] = -2.0
            continue
        psi = simulate(prog)
        for j in range(O):
            out[i, j] = expectation_of_sum(psi, sums[j], prog.n)
    return out


def simulate_state(programs, symbol_names, symbol_values):
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    progs = [Program(programs[i], maps[i]) for i in range(len(programs))]
    max_nq = max([p.n for p in progs], default=0)
    out = np.full((len(progs), 2 ** max_nq), -2.0 + 0j, dtype=np.complex128)
    for i, p in enumerate(progs):
        s = simulate(p)
        out[i, :len(s)] = s
    return out


def inner_product(programs, symbol_names, symbol_values, other_programs):
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    K = np.asarray(other_programs, dtype=object).shape[1] if B else 0
    out = np.zeros((B, K), dtype=np.complex128)
    for i in range(B):
        prog = Program(programs[i], maps[i])
        if not prog.gates:
            out[i, :] = 1.0
            continue
        psi = simulate(prog)
        for j in range(K):
            other = Program(other_programs[i][j], {}, prog.qmap, prog.n)
            out[i, j] = np.vdot(psi, simulate(other))
    return out


# --------------------------------------------------------------------------- sampling
def sample_indices(probs, seed, stream_lo, stream_hi, num_samples):
    """Inverse-CDF sampling driven by the Philox stream: shot k draws
    u = uniform(seed, k, stream) in [0,1) and returns the smallest index x with
    cdf[x] > u * cdf[-1] (FP64 inclusive prefix of |psi|^2)."""
    cdf = np.cumsum(np.asarray(probs, dtype=np.float64))
    u = uniform_doubles(seed, num_samples, stream_lo, stream_hi)
    idx = np.searchsorted(cdf, u * cdf[-1], side="right")
    return np.minimum(idx, len(cdf) - 1)


def simulate_samples(programs, symbol_names, symbol_values, num_samples, seed):
    """``tfq_simulate_samples``: bit q of each draw -> column max_nq-1-q,
    leading -2 padding (tfq_simulate_samples_op.cc:170-187)."""
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    progs = [Program(programs[i], maps[i]) for i in range(len(programs))]
    max_nq = max([p.n for p in progs], default=0)
    out = np.full((len(progs), num_samples, max_nq), -2, dtype=np.int8)
    for i, p in enumerate(progs):
        s = simulate(p)
        idx = sample_indices(np.abs(s) ** 2, seed, i, 0xFFFFFFFF, num_samples)
        for q in range(p.n):
            out[i, :, max_nq - 1 - q] = (idx >> q) & 1
    return out


def zbasis_gates(paulis):
    """``QsimZBasisCircuitFromPauliTerm``: X -> Y^-0.5, Y -> X^0.5, Z -> none."""
    out = []
    for b, pt in paulis:
        if pt == "Z":
            continue
        if pt == "X":
            out.append((eigen_unitary("YP", F32(-0.5), 0.0), b))
        else:
            out.append((eigen_unitary("XP", F32(0.5), 0.0), b))
    return out


def sampled_term_value(state, paulis, n, num_samples, seed, stream_lo, stream_hi):
    s = state
    for U, b in zbasis_gates(paulis):
        s = apply_matrix(s, U, [b], n)
    idx = sample_indices(np.abs(s) ** 2, seed, stream_lo, stream_hi, num_samples)
    mask = 0
    for b, _ in paulis:
        mask |= 1 << b
    par = np.array([bin(int(x) & mask).count("1") & 1 for x in idx])
    return float(np.sum(1 - 2 * par)) / num_samples


def simulate_sampled_expectation(programs, symbol_names, symbol_values, pauli_sums,
                                 num_samples, seed):
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    O = np.asarray(pauli_sums, dtype=object).shape[1] if B else 0
    ns = np.asarray(num_samples, dtype=np.int64)
    out = np.zeros((B, O))
    for i in range(B):
        prog = Program(programs[i], maps[i])
        sums = [parse_paulisum(pauli_sums[i][j], prog.qmap, prog.n) for j in range(O)]
        if not prog.gates:
            out[i, :] = -2.0
            continue
        psi = simulate(prog)
        for j in range(O):
            v = 0.0
            for t, (c, paulis) in enumerate(sums[j]):
                if not paulis:
                    v += c
                    continue
                v += c * sampled_term_value(psi, paulis, prog.n, int(ns[i, j]), seed,
                                            i, (j << 16) | t)
            out[i, j] = v
    return out


# --------------------------------------------------------------------------- adjoint
def _fd_matrix(g, pidx):
    """TFQ's gradient gate (adj_util.cc:175-302): float32 matrices at
    (p +- eps) * p_scalar, float32 difference, times float32(0.5/eps)."""
    plus = list(g.params)
    minus = list(g.params)
    plus[pidx] = F32(F32(g.params[pidx]) + GRAD_EPS)
    minus[pidx] = F32(F32(g.params[pidx]) - GRAD_EPS)
    up = g.unitary(plus).astype(np.complex64)
    um = g.unitary(minus).astype(np.complex64)
    d = (up - um).astype(np.complex64)
    return (d * F32(F32(0.5) / GRAD_EPS)).astype(np.complex64).astype(np.complex128)


def _analytic_matrix(g, pidx):
    """d U / d(symbol value): the derivative of the closed-form unitary in the
    raw argument (which carries the scalar), by 4th-order Richardson central
    differences in complex128 (error ~1e-13)."""
    def at(x):
        p = [float(v) for v in g.params]
        p[pidx] = x
        if g.gid in ("XP", "YP", "ZP", "HP", "XXP", "YYP", "ZZP", "CZP", "CNP", "SP", "ISP"):
            return eigen_unitary(g.gid, p[0] * p[1], p[2])
        if g.gid == "PXP":
            return phased_x_unitary(p[0] * p[1], p[2] * p[3], p[4])
        if g.gid == "FSIM":
            return fsim_unitary(p[0] * p[1], p[2] * p[3])
        if g.gid == "PISP":
            return phased_iswap_unitary(p[0] * p[1], p[2] * p[3])
        raise KeyError(g.gid)

    x0 = float(g.params[pidx])
    h = 1e-3
    d1 = (at(x0 + h) - at(x0 - h)) / (2 * h)
    d2 = (at(x0 + h / 2) - at(x0 - h / 2)) / h
    return (4 * d2 - d1) / 3


def _effective_unitary(g):
    """Unitary with float64 effective parameters (exp*exp_s in float32 as TFQ)."""
    return g.unitary()


def accumulate_operators(psi, sums, grads, n):
    """``AccumulateOperators`` (util_qsim.h:363-414): sum_j g_j O_j psi, dropping
    terms with |g_j * c| < 1e-5."""
    lam = np.zeros_like(psi)
    for terms, gj in zip(sums, grads):
        for c, paulis in terms:
            w = F32(F32(gj) * F32(c))
            if abs(w) < 1e-5:
                continue
            if not paulis:
                lam += float(w) * psi
            else:
                lam += float(w) * pauli_apply(psi, paulis, n)
    return lam


def adjoint_gradient(programs, symbol_names, symbol_values, pauli_sums, downstream_grads,
                     mode="tfq"):
    """``tfq_adj_grad``: out[i, s] = sum over gates using s of 2 Re<dG psi_k | lambda_k>.

    mode "tfq": dG is TFQ's float32 central difference of the gate matrix
    (epsilon 5e-3); mode "analytic": the exact derivative."""
    names = [_name(s) for s in symbol_names]
    col = {s: j for j, s in enumerate(names)}
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    O = np.asarray(pauli_sums, dtype=object).shape[1] if B else 0
    grads = np.asarray(downstream_grads, dtype=np.float32).reshape(B, O)
    out = np.zeros((B, len(names)))
    for i in range(B):
        prog = Program(programs[i], maps[i])
        if not prog.gates:
            continue
        n = prog.n
        sums = [parse_paulisum(pauli_sums[i][j], prog.qmap, n) for j in range(O)]
        psi = simulate(prog)
        lam = accumulate_operators(psi, sums, grads[i], n)
        for g in reversed(prog.gates):
            U = g.unitary()
            psi = apply_gate(psi, g, n, dagger=True, U=U)
            for sym, pidx in g.symbols:
                dG = _fd_matrix(g, pidx) if mode == "tfq" else _analytic_matrix(g, pidx)
                s2 = psi.copy()
                if g.cbits:
                    idx = np.arange(2 ** n, dtype=np.int64)
                    ok = np.ones(2 ** n, dtype=bool)
                    for b, v in zip(g.cbits, g.cvals):
                        ok &= ((idx >> b) & 1) == v
                    s2 = np.where(ok, s2, 0)
                s2 = apply_matrix(s2, dG, g.bits, n)
                out[i, col[sym]] += 2.0 * np.vdot(s2, lam).real
            lam = apply_gate(lam, g, n, dagger=True, U=U)
    return out


def inner_product_gradient(programs, symbol_names, symbol_values, other_programs,
                           downstream_grads, mode="tfq"):
    """``TfqInnerProductGrad`` (tfq_inner_product_grad.cc:210-315): complex
    out[i, s] = sum over gates using s of <dG psi_k | lambda_k> with
    lambda = sum_j g_ij |phi_ij> (AccumulateFusedCircuits, util_qsim.h:423-442,
    float32 weights).  Empty programs give zeros.  mode as adjoint_gradient."""
    names = [_name(s) for s in symbol_names]
    col = {s: j for j, s in enumerate(names)}
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    K = np.asarray(other_programs, dtype=object).shape[1] if B else 0
    grads = np.asarray(downstream_grads, dtype=np.float32).reshape(B, K)
    out = np.zeros((B, len(names)), dtype=np.complex128)
    for i in range(B):
        prog = Program(programs[i], maps[i])
        if not prog.gates:
            continue
        n = prog.n
        lam = np.zeros(2 ** n, dtype=np.complex128)
        for j in range(K):
            other = Program(other_programs[i][j], {}, prog.qmap, prog.n)
            lam += float(grads[i, j]) * simulate(other)
        psi = simulate(prog)
        for g in reversed(prog.gates):
            U = g.unitary()
            psi = apply_gate(psi, g, n, dagger=True, U=U)
            for sym, pidx in g.symbols:
                dG = _fd_matrix(g, pidx) if mode == "tfq" else _analytic_matrix(g, pidx)
                s2 = psi.copy()
                if g.cbits:
                    idx = np.arange(2 ** n, dtype=np.int64)
                    ok = np.ones(2 ** n, dtype=bool)
                    for b, v in zip(g.cbits, g.cvals):
                        ok &= ((idx >> b) & 1) == v
                    s2 = np.where(ok, s2, 0)
                s2 = apply_matrix(s2, dG, g.bits, n)
                out[i, col[sym]] += np.vdot(s2, lam)
            lam = apply_gate(lam, g, n, dagger=True, U=U)
    return out


def expectation_function(programs, symbol_names, pauli_sums):
    """f(values) -> expectations [B, O]; used for parameter-shift / finite
    difference checks of the adjoint gradient."""
    def f(values):
        return simulate_expectation(programs, symbol_names, values, pauli_sums)
    return f
