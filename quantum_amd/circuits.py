"""Minimal circuit model used to build the serialized programs the ops consume.

TFQ takes `cirq.Circuit` objects and serializes them (``core/serialize/
serializer.py:800-881``).  cirq is not a dependency here, so this module
provides the small subset of its object model that TFQ's gate set can express:

* qubits: `LineQubit`, `GridQubit` (``op_serializer.py:33-39``);
* symbols and scalar multiples of symbols (``serializer.py:52-118``);
* the 16 gate families of TFQ's gate set (``serializer.py:738-797`` and the C++
  ids in ``circuit_parser_qsim.cc:575-584``) with Cirq's EigenGate semantics:
  ``U = sum_k exp(i*pi*t*(lambda_k + global_shift)) P_k``;
* controlled operations (``serializer.py:861-878``);
* `Circuit` as a list of moments, with Cirq's EARLIEST insertion when built
  from a flat operation list;
* `PauliString` / `PauliSum` (``serializer.py:903-937``).
"""

from __future__ import annotations

import math
import numbers
from typing import Dict, Iterable, List, Sequence, Tuple, Union

INT_MAX = 2147483647


# --------------------------------------------------------------------------- qubits
class _Qubit:
    def _key(self):  # (row, col) sort key used by ResolveQubitIds
        raise NotImplementedError

    def proto_id(self) -> str:
        raise NotImplementedError

    def __lt__(self, other):
        return self._key() < other._key()

    def __eq__(self, other):
        return isinstance(other, _Qubit) and self._key() == other._key() and \
            type(self) is type(other)

    def __hash__(self):
        return hash((type(self).__name__, self._key()))


class LineQubit(_Qubit):
    """A qubit on a line; serialized as ``"x"`` (``op_serializer.py:37-38``)."""

    def __init__(self, x: int):
        self.x = int(x)

    def _key(self):
        return (INT_MAX, self.x)

    def proto_id(self) -> str:
        return f"{self.x}"

    @staticmethod
    def range(*args) -> List["LineQubit"]:
        return [LineQubit(i) for i in range(*args)]

    def __repr__(self):
        return f"LineQubit({self.x})"


class GridQubit(_Qubit):
    """A qubit on a grid; serialized as ``"row_col"`` (``op_serializer.py:35-36``)."""

    def __init__(self, row: int, col: int):
        self.row = int(row)
        self.col = int(col)

    def _key(self):
        return (self.row, self.col)

    def proto_id(self) -> str:
        return f"{self.row}_{self.col}"

    @staticmethod
    def rect(rows: int, cols: int, top: int = 0, left: int = 0) -> List["GridQubit"]:
        return [GridQubit(top + r, left + c) for r in range(rows) for c in range(cols)]

    def __repr__(self):
        return f"GridQubit({self.row}, {self.col})"


# --------------------------------------------------------------------------- symbols
class Symbol:
    """A named free parameter (the role of ``sympy.Symbol``)."""

    def __init__(self, name: str):
        self.name = str(name)

    def __mul__(self, other):
        if isinstance(other, numbers.Real):
            return ScaledSymbol(self, float(other))
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, numbers.Real):
            return ScaledSymbol(self, 1.0 / float(other))
        return NotImplemented

    def __neg__(self):
        return ScaledSymbol(self, -1.0)

    def __eq__(self, other):
        return isinstance(other, Symbol) and other.name == self.name

    def __hash__(self):
        return hash(("Symbol", self.name))

    def __repr__(self):
        return f"Symbol({self.name!r})"


class ScaledSymbol:
    """``scale * symbol``: the only symbolic arithmetic TFQ serializes
    (``serializer.py:52-84`` extracts the scalar, ``:87-118`` the symbol)."""

    def __init__(self, symbol: Symbol, scale: float):
        self.symbol = symbol
        self.scale = float(scale)

    def __mul__(self, other):
        if isinstance(other, numbers.Real):
            return ScaledSymbol(self.symbol, self.scale * float(other))
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, numbers.Real):
            return ScaledSymbol(self.symbol, self.scale / float(other))
        return NotImplemented

    def __neg__(self):
        return ScaledSymbol(self.symbol, -self.scale)

    def __repr__(self):
        return f"{self.scale!r}*{self.symbol!r}"


Param = Union[float, Symbol, ScaledSymbol]


def _check_param(x):
    if isinstance(x, (Symbol, ScaledSymbol)):
        return x
    if isinstance(x, numbers.Real):
        return float(x)
    raise TypeError("Invalid input argument for exponent.")


def _scale(x, s: float):
    """Multiply a parameter by a number, keeping it serializable."""
    if isinstance(x, (Symbol, ScaledSymbol)):
        return x * s
    return float(x) * s


def is_parameterized(x) -> bool:
    return isinstance(x, (Symbol, ScaledSymbol))


# --------------------------------------------------------------------------- gates
class Gate:
    """Base gate. ``gate_id`` is the serialized id (``serializer.py:738-755``)."""

    gate_id = ""
    num_qubits = 1

    def on(self, *qubits) -> "Operation":
        if len(qubits) != self.num_qubits:
            raise ValueError(f"{type(self).__name__} acts on {self.num_qubits} qubits,"
                             f" got {len(qubits)}")
        return Operation(self, tuple(qubits))

    def __call__(self, *qubits) -> "Operation":
        return self.on(*qubits)

    def on_each(self, *qubits) -> List["Operation"]:
        """One operation of this single-qubit gate on each qubit."""
        if self.num_qubits != 1:
            raise ValueError("on_each needs a single-qubit gate")
        return [self.on(q) for q in _flatten_qubits(qubits)]

    def params(self) -> Dict[str, Param]:
        return {}


class EigenGate(Gate):
    """``U = sum_k exp(i pi t (lambda_k + s)) P_k`` with exponent t, shift s."""

    def __init__(self, exponent: Param = 1.0, global_shift: float = 0.0):
        self.exponent = _check_param(exponent)
        self.global_shift = float(global_shift)

    def __pow__(self, power):
        if isinstance(self.exponent, float) and self.exponent == 1.0:
            return type(self)(exponent=power, global_shift=self.global_shift)
        if isinstance(power, numbers.Real):
            return type(self)(exponent=self.exponent * float(power),
                              global_shift=self.global_shift)
        raise TypeError("Can only raise a gate with exponent 1 to a symbolic power.")

    def params(self):
        return {"exponent": self.exponent, "global_shift": self.global_shift}

    def __repr__(self):
        return f"{type(self).__name__}(exponent={self.exponent!r}, global_shift={self.global_shift!r})"


class XPowGate(EigenGate):
    gate_id = "XP"


class YPowGate(EigenGate):
    gate_id = "YP"


class ZPowGate(EigenGate):
    gate_id = "ZP"


class HPowGate(EigenGate):
    gate_id = "HP"


class XXPowGate(EigenGate):
    gate_id = "XXP"
    num_qubits = 2


class YYPowGate(EigenGate):
    gate_id = "YYP"
    num_qubits = 2


class ZZPowGate(EigenGate):
    gate_id = "ZZP"
    num_qubits = 2


class CZPowGate(EigenGate):
    gate_id = "CZP"
    num_qubits = 2


class CNotPowGate(EigenGate):
    gate_id = "CNP"
    num_qubits = 2


class SwapPowGate(EigenGate):
    gate_id = "SP"
    num_qubits = 2


class ISwapPowGate(EigenGate):
    gate_id = "ISP"
    num_qubits = 2


class PhasedXPowGate(Gate):
    """``Z^p X^t Z^-p`` (serialized id ``PXP``, ``serializer.py:752-755``)."""

    gate_id = "PXP"

    def __init__(self, *, phase_exponent: Param, exponent: Param = 1.0,
                 global_shift: float = 0.0):
        self.phase_exponent = _check_param(phase_exponent)
        self.exponent = _check_param(exponent)
        self.global_shift = float(global_shift)

    def params(self):
        return {"phase_exponent": self.phase_exponent, "exponent": self.exponent,
                "global_shift": self.global_shift}


class PhasedISwapPowGate(Gate):
    """``(Z^-p x Z^p) ISWAP^t (Z^p x Z^-p)`` (serialized id ``PISP``)."""

    gate_id = "PISP"
    num_qubits = 2

    def __init__(self, *, phase_exponent: Param = 0.25, exponent: Param = 1.0):
        self.phase_exponent = _check_param(phase_exponent)
        self.exponent = _check_param(exponent)
        self.global_shift = 0.0

    def params(self):
        return {"phase_exponent": self.phase_exponent, "exponent": self.exponent,
                "global_shift": self.global_shift}


class FSimGate(Gate):
    """``[[1,0,0,0],[0,c,-is,0],[0,-is,c,0],[0,0,0,e^{-i phi}]]`` (id ``FSIM``)."""

    gate_id = "FSIM"
    num_qubits = 2

    def __init__(self, theta: Param, phi: Param):
        self.theta = _check_param(theta)
        self.phi = _check_param(phi)

    def params(self):
        return {"theta": self.theta, "phi": self.phi}


class IdentityGate(Gate):
    """Single-qubit identity (id ``I``; multi-qubit identities are rejected,
    ``serializer.py:582-592``)."""

    gate_id = "I"


class MeasurementGate(Gate):
    """Computational-basis measurement of ``num_qubits`` qubits.  It has no
    serialized id: ``serialize_circuit`` removes measurements (and requires
    each qubit to be measured at most once), as the reference serializer does
    (``serializer.py:826-860``)."""

    gate_id = ""

    def __init__(self, num_qubits: int = 1, key: str = ""):
        if num_qubits < 1:
            raise ValueError("a measurement acts on at least one qubit")
        self.num_qubits = int(num_qubits)
        self.key = str(key)

    def __repr__(self):
        return f"MeasurementGate({self.num_qubits}, key={self.key!r})"


def measure(*qubits, key: str = "") -> "Operation":
    """Measure the given qubits (``cirq.measure``)."""
    qs = list(_flatten_qubits(qubits))
    return MeasurementGate(len(qs), key).on(*qs)


# --------------------------------------------------------------------------- channels
class Channel(Gate):
    """A single-qubit noise channel.  Channels carry plain float arguments (no
    symbols) and serialize under the ids of ``serializer.py:174-455``; only the
    noisy ops (``tfq.noise``) accept them."""

    def _check(self, **probs):
        for name, v in probs.items():
            if not (0.0 <= v <= 1.0):
                raise ValueError(f"{name} was {v} but must be between 0 and 1.")


class DepolarizingChannel(Channel):
    """rho -> (1 - p) rho + p/3 (X rho X + Y rho Y + Z rho Z) (id ``DP``)."""

    gate_id = "DP"

    def __init__(self, p: float):
        self.p = float(p)
        self._check(p=self.p)

    def params(self):
        return {"p": self.p}


class AsymmetricDepolarizingChannel(Channel):
    """rho -> (1 - px - py - pz) rho + px X rho X + py Y rho Y + pz Z rho Z (id ``ADP``)."""

    gate_id = "ADP"

    def __init__(self, p_x: float = 0.0, p_y: float = 0.0, p_z: float = 0.0):
        self.p_x, self.p_y, self.p_z = float(p_x), float(p_y), float(p_z)
        self._check(p_x=self.p_x, p_y=self.p_y, p_z=self.p_z)
        if self.p_x + self.p_y + self.p_z > 1.0:
            raise ValueError("Probabilities must sum to at most 1.")

    def params(self):
        return {"p_x": self.p_x, "p_y": self.p_y, "p_z": self.p_z}


class GeneralizedAmplitudeDampingChannel(Channel):
    """Kraus ops sqrt(p)[[1,0],[0,sqrt(1-g)]], sqrt(p)[[0,sqrt(g)],[0,0]],
    sqrt(1-p)[[sqrt(1-g),0],[0,1]], sqrt(1-p)[[0,0],[sqrt(g),0]] (id ``GAD``)."""

    gate_id = "GAD"

    def __init__(self, p: float, gamma: float):
        self.p, self.gamma = float(p), float(gamma)
        self._check(p=self.p, gamma=self.gamma)

    def params(self):
        return {"p": self.p, "gamma": self.gamma}


class AmplitudeDampingChannel(Channel):
    """Kraus ops [[1,0],[0,sqrt(1-g)]] and [[0,sqrt(g)],[0,0]] (id ``AD``)."""

    gate_id = "AD"

    def __init__(self, gamma: float):
        self.gamma = float(gamma)
        self._check(gamma=self.gamma)

    def params(self):
        return {"gamma": self.gamma}


class ResetChannel(Channel):
    """Kraus ops |0><0| and |0><1| (id ``RST``)."""

    gate_id = "RST"


class PhaseDampingChannel(Channel):
    """Kraus ops [[1,0],[0,sqrt(1-g)]] and [[0,0],[0,sqrt(g)]] (id ``PD``)."""

    gate_id = "PD"

    def __init__(self, gamma: float):
        self.gamma = float(gamma)
        self._check(gamma=self.gamma)

    def params(self):
        return {"gamma": self.gamma}


class PhaseFlipChannel(Channel):
    """rho -> (1 - p) rho + p Z rho Z (id ``PF``)."""

    gate_id = "PF"

    def __init__(self, p: float):
        self.p = float(p)
        self._check(p=self.p)

    def params(self):
        return {"p": self.p}


class BitFlipChannel(Channel):
    """rho -> (1 - p) rho + p X rho X (id ``BF``)."""

    gate_id = "BF"

    def __init__(self, p: float):
        self.p = float(p)
        self._check(p=self.p)

    def params(self):
        return {"p": self.p}


def depolarize(p: float) -> DepolarizingChannel:
    return DepolarizingChannel(p)


def asymmetric_depolarize(p_x: float = 0.0, p_y: float = 0.0,
                          p_z: float = 0.0) -> AsymmetricDepolarizingChannel:
    return AsymmetricDepolarizingChannel(p_x, p_y, p_z)


def generalized_amplitude_damp(p: float, gamma: float) -> GeneralizedAmplitudeDampingChannel:
    return GeneralizedAmplitudeDampingChannel(p, gamma)


def amplitude_damp(gamma: float) -> AmplitudeDampingChannel:
    return AmplitudeDampingChannel(gamma)


def phase_damp(gamma: float) -> PhaseDampingChannel:
    return PhaseDampingChannel(gamma)


def phase_flip(p: float) -> PhaseFlipChannel:
    return PhaseFlipChannel(p)


def bit_flip(p: float) -> BitFlipChannel:
    return BitFlipChannel(p)


def reset(qubit) -> "Operation":
    return ResetChannel().on(qubit)


def supported_channels() -> List[Channel]:
    """One instance of every channel the noisy ops accept (the channel list of
    ``tensorflow_quantum/python/util.py`` ``get_supported_channels``)."""
    return [depolarize(0.01), asymmetric_depolarize(0.01, 0.02, 0.03),
            generalized_amplitude_damp(0.01, 0.02), amplitude_damp(0.01), ResetChannel(),
            phase_damp(0.01), phase_flip(0.01), bit_flip(0.01)]


# Cirq-style constants and constructors.
X = XPowGate()
Y = YPowGate()
Z = ZPowGate()
H = HPowGate()
S = ZPowGate(exponent=0.5)
T = ZPowGate(exponent=0.25)
CNOT = CX = CNotPowGate()
CZ = CZPowGate()
SWAP = SwapPowGate()
ISWAP = ISwapPowGate()
XX = XXPowGate()
YY = YYPowGate()
ZZ = ZZPowGate()
I = IdentityGate()


def rx(rads: Param) -> XPowGate:
    """``exp(-i X rads/2)`` = ``XPowGate(rads/pi, global_shift=-0.5)``."""
    return XPowGate(exponent=_scale(rads, 1.0 / math.pi), global_shift=-0.5)


def ry(rads: Param) -> YPowGate:
    return YPowGate(exponent=_scale(rads, 1.0 / math.pi), global_shift=-0.5)


def rz(rads: Param) -> ZPowGate:
    return ZPowGate(exponent=_scale(rads, 1.0 / math.pi), global_shift=-0.5)


# --------------------------------------------------------------------------- operations
class Operation:
    """A gate on qubits, optionally controlled (``_tfq_control_qubits``)."""

    def __init__(self, gate: Gate, qubits: Tuple[_Qubit, ...],
                 controls: Tuple[_Qubit, ...] = (), control_values: Tuple[int, ...] = ()):
        self.gate = gate
        self.target_qubits = tuple(qubits)
        self.controls = tuple(controls)
        self.control_values = tuple(int(v) for v in control_values)
        if len(set(self.target_qubits + self.controls)) != \
                len(self.target_qubits) + len(self.controls):
            raise ValueError("Operation acts on duplicate qubits.")

    @property
    def qubits(self) -> Tuple[_Qubit, ...]:
        return self.controls + self.target_qubits

    def __pow__(self, power):
        return Operation(self.gate ** power, self.target_qubits, self.controls,
                         self.control_values)

    def controlled_by(self, *control_qubits, control_values: Sequence[int] = None):
        vals = [1] * len(control_qubits) if control_values is None else list(control_values)
        if len(vals) != len(control_qubits):
            raise ValueError("control_values must match control_qubits")
        return Operation(self.gate, self.target_qubits,
                         self.controls + tuple(control_qubits),
                         self.control_values + tuple(vals))

    def __repr__(self):
        c = f" ctrl={self.controls}" if self.controls else ""
        return f"{self.gate!r}.on{self.target_qubits}{c}"


class Moment:
    def __init__(self, operations: Iterable[Operation] = ()):
        self.operations = list(operations)
        used = set()
        for op in self.operations:
            for q in op.qubits:
                if q in used:
                    raise ValueError("Overlapping operations in a moment.")
                used.add(q)

    @property
    def qubits(self):
        return {q for op in self.operations for q in op.qubits}

    def __iter__(self):
        return iter(self.operations)


class Circuit:
    """A list of moments.  ``Circuit(op, op, ...)`` places each operation in
    the earliest moment after the last moment touching its qubits (Cirq's
    EARLIEST strategy); ``Circuit.from_moments`` keeps moments as given."""

    def __init__(self, *contents):
        self.moments: List[Moment] = []
        for item in _flatten(contents):
            self.append(item)

    @staticmethod
    def from_moments(moments: Iterable[Union[Moment, Iterable[Operation]]]) -> "Circuit":
        c = Circuit()
        for m in moments:
            c.moments.append(m if isinstance(m, Moment) else Moment(m))
        return c

    def append(self, item):
        if isinstance(item, Moment):
            self.moments.append(item)
            return
        op = item
        qs = set(op.qubits)
        idx = 0
        for k in range(len(self.moments) - 1, -1, -1):
            if self.moments[k].qubits & qs:
                idx = k + 1
                break
        if idx == len(self.moments):
            self.moments.append(Moment([op]))
        else:
            self.moments[idx].operations.append(op)

    def __iter__(self):
        return iter(self.moments)

    def __len__(self):
        return len(self.moments)

    def all_qubits(self):
        return {q for m in self.moments for q in m.qubits}

    def __add__(self, other: "Circuit") -> "Circuit":
        return Circuit.from_moments(list(self.moments) + list(other.moments))


def _flatten_qubits(items):
    for it in items:
        if isinstance(it, _Qubit):
            yield it
        else:
            yield from _flatten_qubits(it)


def _flatten(items):
    for it in items:
        if isinstance(it, (Operation, Moment)):
            yield it
        else:
            yield from _flatten(it)


# --------------------------------------------------------------------------- Paulis
class PauliString:
    """``coefficient * prod_q P_q`` with ``P in {X, Y, Z}``."""

    def __init__(self, paulis: Dict[_Qubit, str] = None, coefficient: complex = 1.0):
        self.paulis = dict(paulis or {})
        for p in self.paulis.values():
            if p not in ("X", "Y", "Z"):
                raise ValueError(f"Unknown Pauli {p!r}")
        self.coefficient = complex(coefficient)

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return PauliString(self.paulis, self.coefficient * other)
        if isinstance(other, PauliString):
            if set(self.paulis) & set(other.paulis):
                raise ValueError("Products on shared qubits are not supported.")
            d = dict(self.paulis)
            d.update(other.paulis)
            return PauliString(d, self.coefficient * other.coefficient)
        return NotImplemented

    __rmul__ = __mul__

    def __add__(self, other):
        return PauliSum([self]) + other

    def __radd__(self, other):
        return PauliSum([self]) + other

    def __sub__(self, other):
        return PauliSum([self]) + (-1.0) * _as_sum(other)

    def __neg__(self):
        return self * -1.0

    def __repr__(self):
        return f"{self.coefficient}*{self.paulis}"


def _as_sum(x) -> "PauliSum":
    if isinstance(x, PauliSum):
        return x
    if isinstance(x, PauliString):
        return PauliSum([x])
    if isinstance(x, numbers.Number):
        return PauliSum([PauliString({}, x)])
    raise TypeError(f"Cannot add {type(x)} to PauliSum")


class PauliSum:
    def __init__(self, terms: Iterable[PauliString] = ()):
        self.terms: List[PauliString] = list(terms)

    @staticmethod
    def from_pauli_strings(terms) -> "PauliSum":
        if isinstance(terms, PauliString):
            terms = [terms]
        return PauliSum(terms)

    def __add__(self, other):
        return PauliSum(self.terms + _as_sum(other).terms)

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-1.0) * _as_sum(other)

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return PauliSum([t * other for t in self.terms])
        return NotImplemented

    __rmul__ = __mul__

    def __iter__(self):
        return iter(self.terms)

    @property
    def qubits(self):
        return {q for t in self.terms for q in t.paulis}


def pauli_x(q) -> PauliString:
    return PauliString({q: "X"})


def pauli_y(q) -> PauliString:
    return PauliString({q: "Y"})


def pauli_z(q) -> PauliString:
    return PauliString({q: "Z"})
