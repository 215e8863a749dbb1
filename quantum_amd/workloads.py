"""Seeded workload builders: the benchmark configurations and random test circuits.

Each builder returns the circuits, symbol names, symbol values and Pauli sums
of one workload, built from a seeded numpy generator.
"""

import math

import numpy as np

from quantum_amd import circuits as C


def two_qubit_state_circuit():
    """X**0.25 on q0, CNOT(q0->q1), Y**0.5 on q1 (util_qsim_test.cc:55-60)."""
    q0, q1 = C.LineQubit.range(2)
    return C.Circuit(C.X(q0) ** 0.25, C.CNOT(q0, q1), C.Y(q1) ** 0.5), (q0, q1)


def adjoint_golden_cases():
    """The four circuits of tfq_adj_grad_op_test.py:265-397."""
    a, b, g = C.Symbol("alpha"), C.Symbol("beta"), C.Symbol("gamma")
    gq = C.GridQubit.rect(1, 2)
    lq = C.LineQubit.range(2)
    cases = {}
    q = gq
    cases["simple_case"] = (C.Circuit(C.X(q[0]) ** a, C.Y(q[1]) ** b, C.CNOT(q[0], q[1])),
                            ["alpha", "beta"], q)
    cases["simple_case2"] = (C.Circuit(C.X(q[0]) ** a, C.Y(q[1]) ** b, C.CNOT(q[0], q[1]),
                                       C.FSimGate(g, 0.5)(q[0], q[1])),
                             ["alpha", "beta", "gamma"], q)
    cases["simple_case_shared"] = (C.Circuit(C.X(q[0]) ** a, C.Y(q[1]) ** b, C.CNOT(q[0], q[1]),
                                             C.FSimGate(g, g)(q[0], q[1])),
                                   ["alpha", "beta", "gamma"], q)
    q = lq
    cases["simple_case_single"] = (C.Circuit(C.X(q[0]) ** a, C.Y(q[1]) ** a, C.CNOT(q[0], q[1]),
                                             C.FSimGate(-0.56, a)(q[0], q[1])),
                                   ["alpha", "beta", "gamma"], q)
    return cases


ALL_GATE_FAMILIES = ["XP", "YP", "ZP", "HP", "XXP", "YYP", "ZZP", "CZP", "CNP", "SP", "ISP",
                     "PXP", "FSIM", "PISP", "I"]


def random_gate_op(rng, qubits, symbols, gid=None, p_symbol=0.5, p_control=0.0):
    """One random operation of family ``gid`` on distinct random qubits."""
    if gid is None:
        gid = ALL_GATE_FAMILIES[rng.integers(len(ALL_GATE_FAMILIES))]
    two = gid in ("XXP", "YYP", "ZZP", "CZP", "CNP", "SP", "ISP", "FSIM", "PISP")
    k = 2 if two else 1
    idx = rng.choice(len(qubits), size=k, replace=False)
    tq = [qubits[i] for i in idx]

    def param():
        if symbols and rng.random() < p_symbol:
            s = C.Symbol(symbols[rng.integers(len(symbols))])
            r = rng.random()
            return s if r < 0.5 else s * float(np.round(rng.uniform(-2, 2), 3))
        return float(rng.uniform(-2, 2))

    gs = float(rng.choice([0.0, -0.5, 0.25]))
    if gid == "I":
        gate = C.I
    elif gid == "PXP":
        gate = C.PhasedXPowGate(phase_exponent=param(), exponent=param(), global_shift=gs)
    elif gid == "PISP":
        gate = C.PhasedISwapPowGate(phase_exponent=param(), exponent=param())
    elif gid == "FSIM":
        gate = C.FSimGate(param(), param())
    else:
        cls = {"XP": C.XPowGate, "YP": C.YPowGate, "ZP": C.ZPowGate, "HP": C.HPowGate,
               "XXP": C.XXPowGate, "YYP": C.YYPowGate, "ZZP": C.ZZPowGate, "CZP": C.CZPowGate,
               "CNP": C.CNotPowGate, "SP": C.SwapPowGate, "ISP": C.ISwapPowGate}[gid]
        gate = cls(exponent=param(), global_shift=gs)
    op = gate.on(*tq)
    if p_control and rng.random() < p_control:
        rest = [q for q in qubits if q not in tq]
        if rest:
            nc = int(rng.integers(1, min(3, len(rest)) + 1))
            cidx = rng.choice(len(rest), size=nc, replace=False)
            op = op.controlled_by(*[rest[i] for i in cidx],
                                  control_values=[int(v) for v in rng.integers(0, 2, nc)])
    return op


def random_circuit(rng, qubits, n_ops, symbols=(), p_symbol=0.5, p_control=0.0, families=None):
    fams = families or ALL_GATE_FAMILIES
    ops = [random_gate_op(rng, qubits, list(symbols), fams[rng.integers(len(fams))],
                          p_symbol, p_control) for _ in range(n_ops)]
    return C.Circuit(ops)


def random_noisy_circuit(rng, qubits, n_ops, symbols=(), p_symbol=0.5, p_control=0.0,
                         p_channel=0.4, channels=None):
    """random_circuit with a random channel (from ``channels``, default every
    supported kind with random strengths) after each gate with probability
    p_channel, on one of that gate's qubits."""
    fams = ALL_GATE_FAMILIES
    ops = []
    for _ in range(n_ops):
        op = random_gate_op(rng, qubits, list(symbols), fams[rng.integers(len(fams))],
                            p_symbol, p_control)
        ops.append(op)
        if rng.uniform() < p_channel:
            q = op.target_qubits[rng.integers(len(op.target_qubits))]
            if channels is None:
                ops.append(random_channel(rng).on(q))
            else:
                ops.append(channels[rng.integers(len(channels))].on(q))
    return C.Circuit(ops)


def random_channel(rng):
    """A channel of a random supported kind with strengths in [0, 0.3)."""
    u = lambda: float(np.round(rng.uniform(0.0, 0.3), 4))
    kind = int(rng.integers(8))
    if kind == 0:
        return C.depolarize(u())
    if kind == 1:
        return C.asymmetric_depolarize(u() / 3, u() / 3, u() / 3)
    if kind == 2:
        return C.generalized_amplitude_damp(float(np.round(rng.uniform(), 4)), u())
    if kind == 3:
        return C.amplitude_damp(u())
    if kind == 4:
        return C.ResetChannel()
    if kind == 5:
        return C.phase_damp(u())
    if kind == 6:
        return C.phase_flip(u())
    return C.bit_flip(u())


def random_pauli_sum(rng, qubits, n_terms, max_weight=3):
    terms = []
    for _ in range(n_terms):
        w = int(rng.integers(0, min(max_weight, len(qubits)) + 1))
        idx = rng.choice(len(qubits), size=w, replace=False)
        paulis = {qubits[i]: "XYZ"[rng.integers(3)] for i in idx}
        terms.append(C.PauliString(paulis, float(np.round(rng.uniform(-2, 2), 4))))
    return C.PauliSum(terms)


# ----------------------------------------------------------------------------- workloads
def config1_line_rotations(seed=1000, batch=64, n=10, layers=20):
    """64 circuits of 10 line qubits, 20 layers of rx ry rz per qubit (600
    independent symbols) plus a CZ ladder; observables [sum Z_i Z_i+1, X_0]."""
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = []
    moments = []
    for l in range(layers):
        for kind, fn in (("x", C.rx), ("y", C.ry), ("z", C.rz)):
            ops = []
            for i in range(n):
                nm = f"{kind}{l}_{i}"
                names.append(nm)
                ops.append(fn(C.Symbol(nm)).on(q[i]))
            moments.append(ops)
        moments.append([C.CZ(q[i], q[i + 1]) for i in range(0, n - 1, 2)])
        moments.append([C.CZ(q[i], q[i + 1]) for i in range(1, n - 1, 2)])
    circuit = C.Circuit.from_moments(moments)
    values = rng.uniform(-math.pi, math.pi, size=(batch, len(names))).astype(np.float32)
    zz = C.PauliSum([C.pauli_z(q[i]) * C.pauli_z(q[i + 1]) for i in range(n - 1)])
    ops = [zz, C.PauliSum([C.pauli_x(q[0])])]
    return [circuit] * batch, names, values, [ops] * batch


def config2_grid(seed=1001, rows=5, cols=6, cycles=14):
    """30 qubits on a 5x6 grid: 14 cycles of random single-qubit gates from
    {sqrt(X), sqrt(Y), sqrt(W)} (never repeating on a qubit) followed by
    FSim(pi/2, pi/6) on one of four grid-edge patterns (horizontal even/odd
    columns, vertical even/odd rows) per cycle.  Observable: sum Z_i Z_j over
    all 49 grid edges."""
    rng = np.random.default_rng(seed)
    q = {(r, c): C.GridQubit(r, c) for r in range(rows) for c in range(cols)}
    singles = [C.X ** 0.5, C.Y ** 0.5, C.PhasedXPowGate(phase_exponent=0.25, exponent=0.5)]
    patterns = []
    for par in (0, 1):
        patterns.append([((r, c), (r, c + 1)) for r in range(rows) for c in range(cols - 1)
                         if c % 2 == par])
    for par in (0, 1):
        patterns.append([((r, c), (r + 1, c)) for r in range(rows - 1) for c in range(cols)
                         if r % 2 == par])
    order = [0, 1, 2, 3, 2, 3, 0, 1]
    fsim = C.FSimGate(math.pi / 2, math.pi / 6)
    last = {k: -1 for k in q}
    moments = []
    for cyc in range(cycles):
        ops = []
        for k in sorted(q):
            choices = [i for i in range(3) if i != last[k]]
            g = choices[rng.integers(len(choices))]
            last[k] = g
            ops.append(singles[g].on(q[k]))
        moments.append(ops)
        moments.append([fsim(q[a], q[b]) for a, b in patterns[order[cyc % 8]]])
    circuit = C.Circuit.from_moments(moments)
    edges = [((r, c), (r, c + 1)) for r in range(rows) for c in range(cols - 1)] + \
            [((r, c), (r + 1, c)) for r in range(rows - 1) for c in range(cols)]
    zz = C.PauliSum([C.pauli_z(q[a]) * C.pauli_z(q[b]) for a, b in edges])
    return circuit, zz


def hea_ring(n, layers, prefix="t"):
    """Hardware-efficient ansatz: per layer ry, rz on every qubit (own symbols)
    and a CZ ring; returns (circuit, symbol names, qubits)."""
    q = C.LineQubit.range(n)
    names, moments = [], []
    for l in range(layers):
        ry_ops, rz_ops = [], []
        for i in range(n):
            a, b = f"{prefix}{l}_y{i}", f"{prefix}{l}_z{i}"
            names += [a, b]
            ry_ops.append(C.ry(C.Symbol(a)).on(q[i]))
            rz_ops.append(C.rz(C.Symbol(b)).on(q[i]))
        moments += [ry_ops, rz_ops]
        moments.append([C.CZ(q[i], q[i + 1]) for i in range(0, n - 1, 2)])
        moments.append([C.CZ(q[i], q[(i + 1) % n]) for i in range(1, n, 2)])
    order = []
    for l in range(layers):
        order += [f"{prefix}{l}_y{i}" for i in range(n)] + [f"{prefix}{l}_z{i}" for i in range(n)]
    return C.Circuit.from_moments(moments), order, q


def tfim(q, hx, ring=False):
    n = len(q)
    m = n if ring else n - 1
    return C.PauliSum([C.pauli_z(q[i]) * C.pauli_z(q[(i + 1) % n]) for i in range(m)] +
                      [C.PauliString({q[i]: "X"}, hx) for i in range(n)])


def config3_hea(seed=1002, n=26, layers=8, batch=8):
    """26-qubit HEA (8 layers of ry, rz + CZ ring, 416 symbols), 8 parameter
    sets ~U(-pi, pi); H = sum Z_i Z_i+1 (25 terms) + 0.7 sum X_i (26 terms)."""
    rng = np.random.default_rng(seed)
    circuit, names, q = hea_ring(n, layers)
    values = rng.uniform(-math.pi, math.pi, size=(batch, len(names))).astype(np.float32)
    return circuit, names, values, tfim(q, 0.7)


def config4_uneven(seed=1003, count=4096, n_min=4, n_max=16, d_min=10, d_max=60,
                   p_symbol=0.3, n_symbols=256, max_ops=8):
    """Uneven batch: qubits ~U{4..16}, depth ~U{10..60} moments filled with
    random gates from {H, X^t, Y^t, Z^t, CNOT, CZ^t, ISWAP^t, FSim(theta, phi)},
    30% of them parameterised by symbols drawn from a shared pool of 256
    (values ~U(-pi, pi)); 1-8 Pauli strings of weight <= 4 per circuit, padded
    to 8 with empty sums (downstream weight 0)."""
    rng = np.random.default_rng(seed)
    names = [f"p{i}" for i in range(n_symbols)]
    circuits, psums, weights = [], [], []
    one_q = ["H", "XP", "YP", "ZP"]
    two_q = ["CNOT", "CZP", "ISP", "FSIM"]

    def param():
        if rng.random() < p_symbol:
            return C.Symbol(names[rng.integers(n_symbols)])
        return float(rng.uniform(-1, 1))

    for _ in range(count):
        n = int(rng.integers(n_min, n_max + 1))
        depth = int(rng.integers(d_min, d_max + 1))
        q = C.LineQubit.range(n)
        moments = []
        for _ in range(depth):
            perm = list(rng.permutation(n))
            ops = []
            while perm:
                if len(perm) >= 2 and rng.random() < 0.5:
                    a, b = perm.pop(), perm.pop()
                    kind = two_q[rng.integers(4)]
                    if kind == "CNOT":
                        ops.append(C.CNOT(q[a], q[b]))
                    elif kind == "CZP":
                        ops.append(C.CZPowGate(exponent=param()).on(q[a], q[b]))
                    elif kind == "ISP":
                        ops.append(C.ISwapPowGate(exponent=param()).on(q[a], q[b]))
                    else:
                        ops.append(C.FSimGate(param(), param()).on(q[a], q[b]))
                else:
                    a = perm.pop()
                    kind = one_q[rng.integers(4)]
                    if kind == "H":
                        ops.append(C.H(q[a]))
                    else:
                        cls = {"XP": C.XPowGate, "YP": C.YPowGate, "ZP": C.ZPowGate}[kind]
                        ops.append(cls(exponent=param()).on(q[a]))
            moments.append(ops)
        circuits.append(C.Circuit.from_moments(moments))
        k = int(rng.integers(1, max_ops + 1))
        row, w = [], []
        for _ in range(k):
            wt = int(rng.integers(1, min(4, n) + 1))
            idx = rng.choice(n, size=wt, replace=False)
            row.append(C.PauliSum([C.PauliString({q[i]: "XYZ"[rng.integers(3)] for i in idx},
                                                 float(np.round(rng.uniform(-1, 1), 4)))]))
            w.append(1.0)
        row += [C.PauliSum([]) for _ in range(max_ops - k)]
        w += [0.0] * (max_ops - k)
        psums.append(row)
        weights.append(w)
    values = rng.uniform(-math.pi, math.pi, size=(count, n_symbols)).astype(np.float32)
    return circuits, names, values, psums, np.asarray(weights, np.float32)


def config5_vqe(n=12, layers=6):
    """12-qubit TFIM VQE: 6 layers of ry, rz (144 symbols) + CZ ring;
    H = sum Z_i Z_i+1 (ring) + 1.0 sum X_i."""
    circuit, names, q = hea_ring(n, layers)
    return circuit, names, tfim(q, 1.0, ring=True)


def noisy_layers(seed=1010, batch=32, n=8, layers=8, p_dep=0.01, gamma=0.02):
    """Noisy variational circuits: ``layers`` layers of ry(symbol) on every
    qubit and a CZ ladder, each followed by depolarize(p_dep) and
    amplitude_damp(gamma) on every qubit; observables [sum Z_i Z_i+1, X_0]."""
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = [f"t{l}_{i}" for l in range(layers) for i in range(n)]
    moments = []
    for l in range(layers):
        moments.append([C.ry(C.Symbol(f"t{l}_{i}")).on(q[i]) for i in range(n)])
        moments.append([C.CZ(q[i], q[i + 1]) for i in range(0, n - 1, 2)])
        moments.append([C.CZ(q[i], q[i + 1]) for i in range(1, n - 1, 2)])
        moments.append([C.depolarize(p_dep).on(x) for x in q])
        moments.append([C.amplitude_damp(gamma).on(x) for x in q])
    circuit = C.Circuit.from_moments(moments)
    values = rng.uniform(-math.pi, math.pi, size=(batch, len(names))).astype(np.float32)
    zz = C.PauliSum([C.pauli_z(q[i]) * C.pauli_z(q[i + 1]) for i in range(n - 1)])
    ops = [zz, C.PauliSum([C.pauli_x(q[0])])]
    return [circuit] * batch, names, values, [ops] * batch
