"""Oracle-side proto3 decoder for TFQ ``Program`` / ``PauliSum`` bytes.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product never
does.  This decoder is written independently of the product's C++ decoder so
that the two check each other.

Schema: ``core/proto/program.proto:20-184`` and ``core/proto/pauli_sum.proto:20-35``.
Decoded objects are plain dicts / lists.
"""

import struct


class DecodeError(ValueError):
    pass


def _read_varint(b, i):
    shift = 0
    v = 0
    while True:
        if i >= len(b):
            raise DecodeError("truncated varint")
        c = b[i]
        i += 1
        v |= (c & 0x7F) << shift
        if not c & 0x80:
            return v, i
        shift += 7
        if shift > 63:
            raise DecodeError("varint too long")


def _fields(b):
    """Yield (field_number, wire_type, value) for one message body."""
    i = 0
    n = len(b)
    while i < n:
        key, i = _read_varint(b, i)
        field, wire = key >> 3, key & 7
        if wire == 0:
            v, i = _read_varint(b, i)
        elif wire == 1:
            if i + 8 > n:
                raise DecodeError("truncated fixed64")
            v = b[i:i + 8]
            i += 8
        elif wire == 2:
            ln, i = _read_varint(b, i)
            if i + ln > n:
                raise DecodeError("truncated length-delimited field")
            v = b[i:i + ln]
            i += ln
        elif wire == 5:
            if i + 4 > n:
                raise DecodeError("truncated fixed32")
            v = b[i:i + 4]
            i += 4
        else:
            raise DecodeError(f"unsupported wire type {wire}")
        yield field, wire, v


def _f32(v):
    return struct.unpack("<f", v)[0]


def decode_arg(b):
    """Arg -> {'float': x} | {'symbol': s} | {'string': s} | {'bools': [...]}."""
    out = {"float": 0.0, "symbol": "", "string": "", "kind": None}
    for f, w, v in _fields(b):
        if f == 1 and w == 2:
            out["kind"] = "value"
            for f2, w2, v2 in _fields(v):
                if f2 == 1 and w2 == 5:
                    out["float"] = _f32(v2)
                elif f2 == 3 and w2 == 2:
                    out["string"] = v2.decode("utf-8")
                elif f2 == 4 and w2 == 1:
                    out["float"] = struct.unpack("<d", v2)[0]
        elif f == 2 and w == 2:
            out["kind"] = "symbol"
            out["symbol"] = v.decode("utf-8")
    return out


def decode_operation(b):
    op = {"gate": "", "args": {}, "qubits": []}
    for f, w, v in _fields(b):
        if f == 1 and w == 2:
            for f2, w2, v2 in _fields(v):
                if f2 == 1:
                    op["gate"] = v2.decode("utf-8")
        elif f == 2 and w == 2:
            key, val = "", b""
            for f2, w2, v2 in _fields(v):
                if f2 == 1:
                    key = v2.decode("utf-8")
                elif f2 == 2:
                    val = v2
            op["args"][key] = decode_arg(val)
        elif f == 3 and w == 2:
            qid = ""
            for f2, w2, v2 in _fields(v):
                if f2 == 2:
                    qid = v2.decode("utf-8")
            op["qubits"].append(qid)
    return op


def decode_program(b):
    """Program -> list of moments, each a list of operation dicts."""
    moments = []
    for f, w, v in _fields(bytes(b)):
        if f == 2 and w == 2:  # circuit
            for f2, w2, v2 in _fields(v):
                if f2 == 2 and w2 == 2:  # moment
                    ops = []
                    for f3, w3, v3 in _fields(v2):
                        if f3 == 1 and w3 == 2:
                            ops.append(decode_operation(v3))
                    moments.append(ops)
    return moments


def decode_paulisum(b):
    """PauliSum -> list of (coef_real, coef_imag, [(qubit_id, 'X'|'Y'|'Z'), ...])."""
    terms = []
    for f, w, v in _fields(bytes(b)):
        if f == 1 and w == 2:
            cr, ci, paulis = 0.0, 0.0, []
            for f2, w2, v2 in _fields(v):
                if f2 == 1 and w2 == 5:
                    cr = _f32(v2)
                elif f2 == 2 and w2 == 5:
                    ci = _f32(v2)
                elif f2 == 3 and w2 == 2:
                    qid, pt = "", ""
                    for f3, w3, v3 in _fields(v2):
                        if f3 == 1:
                            qid = v3.decode("utf-8")
                        elif f3 == 2:
                            pt = v3.decode("utf-8")
                    paulis.append((qid, pt))
            terms.append((cr, ci, paulis))
    return terms
