"""Wire-format encoder for TFQ's ``Program`` and ``PauliSum`` protos.

The bytes produced here are what TFQ's op kernels parse
(``core/ops/parse_context.cc:41-56``).  The schema is
``core/proto/program.proto`` and ``core/proto/pauli_sum.proto``; the gate ids,
argument names and value rules follow ``core/serialize/serializer.py``:

* exponents / angles pass through ``_symbol_extractor`` / ``_scalar_extractor``
  (``serializer.py:52-118``): numbers are rounded to 6 decimals, a symbol
  serializes as its name, ``scale*symbol`` as the symbol plus a 6-decimal
  ``*_scalar``;
* ``global_shift`` is written unrounded (``serializer.py:469-471``);
* controls are comma-separated qubit ids and values (``serializer.py:121-135``);
* the identity gate carries ``unused = 1.0`` (a Python ``True`` through
  ``_arg_to_proto``, ``op_serializer.py:64-66``);
* every circuit uses ``gate_set = "tfq_gate_set"`` and ``MOMENT_BY_MOMENT``
  (``serializable_gate_set.py:120-140,223-230``).

The encoder writes proto3 wire format directly (no generated code).
"""

from __future__ import annotations

import functools
import math
import struct
from typing import Iterable, List

import numpy as np

from . import circuits as C

GATE_SET_NAME = "tfq_gate_set"


# --------------------------------------------------------------------------- wire helpers
@functools.lru_cache(maxsize=1 << 16)
def _varint(v: int) -> bytes:
    out = bytearray()
    v &= (1 << 64) - 1
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


@functools.lru_cache(maxsize=256)
def _key(field: int, wire: int) -> bytes:
    return _varint((field << 3) | wire)


def _len_field(field: int, payload: bytes) -> bytes:
    return _key(field, 2) + _varint(len(payload)) + payload


def _str_field(field: int, s: str, always: bool = False) -> bytes:
    if not s and not always:
        return b""
    return _len_field(field, s.encode("utf-8"))


def _float_field(field: int, x: float, always: bool = False) -> bytes:
    if x == 0.0 and not always:
        return b""
    return _key(field, 5) + struct.pack("<f", float(x))


# --------------------------------------------------------------------------- values
def round6(x: float) -> float:
    """``serializer._round``: ``np.round(x, 6)``, i.e. rint(x * 1e6) / 1e6 in
    float64 (Python's ``round`` of a float rounds half to even like rint)."""
    x = float(x)
    if not math.isfinite(x):
        return float(np.round(x, 6))
    return round(x * 1e6) / 1e6


def symbol_part(x):
    """``_symbol_extractor``: number -> rounded float, symbol -> name."""
    if isinstance(x, C.Symbol):
        return x
    if isinstance(x, C.ScaledSymbol):
        return x.symbol
    return round6(x)


def scalar_part(x) -> float:
    """``_scalar_extractor``: 1.0 for numbers and bare symbols, else the
    6-decimal scalar of ``scale*symbol``."""
    if isinstance(x, C.ScaledSymbol):
        return round6(x.scale)
    return 1.0


def _arg(value) -> bytes:
    """Encode one ``Arg`` message (``op_serializer._arg_to_proto``)."""
    if isinstance(value, C.Symbol):
        return _str_field(2, value.name, always=True)
    if isinstance(value, str):
        arg_value = _str_field(3, value, always=True)
    else:
        arg_value = _float_field(1, float(value), always=True)
    return _len_field(1, arg_value)


@functools.lru_cache(maxsize=1 << 16)
def _map_entry_cached(name: str, kind: str, value, sign: float = 1.0) -> bytes:
    v = C.Symbol(value) if kind == "sym" else value
    return _len_field(2, _str_field(1, name, always=True) + _len_field(2, _arg(v)))


def _map_entry(name: str, value) -> bytes:
    if isinstance(value, C.Symbol):
        return _map_entry_cached(name, "sym", value.name)
    if isinstance(value, str):
        return _map_entry_cached(name, "str", value)
    x = float(value)
    # 0.0 and -0.0 hash alike but encode differently.
    return _map_entry_cached(name, "f", x, math.copysign(1.0, x))


@functools.lru_cache(maxsize=1 << 12)
def _qubit_field(proto_id: str) -> bytes:
    return _len_field(3, _str_field(2, proto_id, always=True))


def _op_args(op: C.Operation) -> List:
    g = op.gate
    ctrl_q = ",".join(q.proto_id() for q in op.controls)
    ctrl_v = ",".join(str(v) for v in op.control_values)
    if isinstance(g, C.Channel):
        # Channel arguments are plain floats (serializer.py:174-455).
        args = [(k, float(v)) for k, v in g.params().items()]
    elif isinstance(g, C.IdentityGate):
        args = [("unused", 1.0)]
    elif isinstance(g, C.FSimGate):
        args = [("theta", symbol_part(g.theta)), ("phi", symbol_part(g.phi)),
                ("theta_scalar", scalar_part(g.theta)), ("phi_scalar", scalar_part(g.phi))]
    elif isinstance(g, (C.PhasedXPowGate, C.PhasedISwapPowGate)):
        args = [("phase_exponent", symbol_part(g.phase_exponent)),
                ("phase_exponent_scalar", scalar_part(g.phase_exponent)),
                ("exponent", symbol_part(g.exponent)),
                ("exponent_scalar", scalar_part(g.exponent)),
                ("global_shift", float(g.global_shift))]
    elif isinstance(g, C.EigenGate):
        args = [("exponent", symbol_part(g.exponent)),
                ("exponent_scalar", scalar_part(g.exponent)),
                ("global_shift", float(g.global_shift))]
    else:
        raise ValueError(f"Cannot serialize gate {g!r}")
    args += [("control_qubits", ctrl_q), ("control_values", ctrl_v)]
    return args


def _operation(op: C.Operation) -> bytes:
    body = _len_field(1, _str_field(1, op.gate.gate_id, always=True))
    for name, value in _op_args(op):
        body += _map_entry(name, value)
    for q in op.target_qubits:
        body += _qubit_field(q.proto_id())
    return body


# --------------------------------------------------------------------------- public API
def serialize_circuit(circuit: C.Circuit) -> bytes:
    """``serializer.serialize_circuit`` -> ``Program`` bytes.

    Measurements are removed, as the reference does (``serializer.py:826-860``):
    a qubit measured a second time raises ``ValueError``; every moment that
    holds a measurement loses the operations on the qubits measured in it."""
    if not isinstance(circuit, C.Circuit):
        raise TypeError("serialize requires Circuit objects. Given: " + str(type(circuit)))
    measured_all = set()
    measured = []
    for m in circuit.moments:
        here = set()
        for op in m.operations:
            for q in op.qubits:
                if not isinstance(q, (C.GridQubit, C.LineQubit)):
                    raise ValueError("Attempted to serialize circuit that don't use "
                                     "only GridQubits or LineQubits.")
            if isinstance(op.gate, C.MeasurementGate):
                for q in op.qubits:
                    if q in measured_all:
                        raise ValueError("Serialization of circuit failed. Circuits with "
                                         "non-terminal measurement operations are not "
                                         "supported.")
                    here.add(q)
                    measured_all.add(q)
        measured.append(here)
    circ = _key(1, 0) + _varint(1)  # MOMENT_BY_MOMENT
    for m, here in zip(circuit.moments, measured):
        ops = [op for op in m.operations if not any(q in here for q in op.qubits)]
        mom = b"".join(_len_field(1, _operation(op)) for op in ops)
        circ += _len_field(2, mom)
    lang = _str_field(1, GATE_SET_NAME)
    return _len_field(1, lang) + _len_field(2, circ)


def serialize_paulisum(psum) -> bytes:
    """``serializer.serialize_paulisum`` -> ``PauliSum`` bytes (qubits of each
    term sorted, coefficient split into float32 real/imag parts)."""
    if isinstance(psum, C.PauliString):
        psum = C.PauliSum([psum])
    if not isinstance(psum, C.PauliSum):
        raise TypeError("serialize requires a PauliSum object. Given: " + str(type(psum)))
    out = b""
    for term in psum.terms:
        body = _float_field(1, term.coefficient.real) + _float_field(2, term.coefficient.imag)
        for q in sorted(term.paulis):
            pair = _str_field(1, q.proto_id(), always=True) + \
                _str_field(2, term.paulis[q], always=True)
            body += _len_field(3, pair)
        out += _len_field(1, body)
    return out


def _serialize_any(x) -> bytes:
    if isinstance(x, C.Circuit):
        return serialize_circuit(x)
    if isinstance(x, (C.PauliSum, C.PauliString)):
        return serialize_paulisum(x)
    if isinstance(x, (bytes, bytearray)):
        return bytes(x)
    raise TypeError(f"Cannot convert {type(x)} to a serialized tensor.")


def convert_to_tensor(items) -> np.ndarray:
    """``tfq.convert_to_tensor`` (``python/util.py:264-334``): a (nested) list
    of circuits or Pauli sums -> an object ndarray of serialized bytes with the
    same shape."""
    def rec(x):
        if isinstance(x, (list, tuple)):
            return [rec(y) for y in x]
        return _serialize_any(x)

    nested = rec(list(items) if not isinstance(items, list) else items)
    arr = np.empty(_shape(nested), dtype=object)
    _fill(arr, nested)
    return arr


def _shape(nested):
    shape = []
    x = nested
    while isinstance(x, list):
        shape.append(len(x))
        if not x:
            break
        x = x[0]
    return tuple(shape)


def _fill(arr, nested):
    if arr.ndim == 1:
        for i, v in enumerate(nested):
            arr[i] = v
    else:
        for i, v in enumerate(nested):
            _fill(arr[i], v)


def pack_strings(arr) -> (bytes, np.ndarray):
    """Flatten a string tensor into ``(bytes, int64 offsets[n+1])`` for the C-ABI."""
    flat = list(np.asarray(arr, dtype=object).reshape(-1))
    bs = [x.encode("utf-8") if isinstance(x, str) else bytes(x) for x in flat]
    offsets = np.zeros(len(bs) + 1, dtype=np.int64)
    if bs:
        offsets[1:] = np.cumsum([len(b) for b in bs])
    return b"".join(bs), offsets
