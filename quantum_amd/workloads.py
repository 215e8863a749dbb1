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
