"""Test helper: binary Program / PauliSum protos -> protobuf text format
(as ``TextFormat::PrintToString`` lays it out), through the oracle's own
decoder.  Used to feed the text-format fallback of the parser."""

from oracle import wire

_STRING_ARGS = ("control_qubits", "control_values")


def _q(s):
    return '"' + s.replace("\\", "\\\\").replace('"', '\\"') + '"'


def program_text(b, style="{}"):
    o, c = style
    lines = [f"language {o} gate_set: \"tfq_gate_set\" {c}", f"circuit {o}",
             "  scheduling_strategy: MOMENT_BY_MOMENT"]
    for m in wire.decode_program(b):
        lines.append(f"  moments {o}")
        for op in m:
            lines.append(f"    operations {o}")
            lines.append(f"      gate {o} id: {_q(op['gate'])} {c}")
            for k, a in op["args"].items():
                if a["kind"] == "symbol":
                    val = f"symbol: {_q(a['symbol'])}"
                elif k in _STRING_ARGS or a["string"]:
                    val = f"arg_value {o} string_value: {_q(a['string'])} {c}"
                else:
                    val = f"arg_value {o} float_value: {a['float']:.9g} {c}"
                lines.append(f"      args {o} key: {_q(k)} value {o} {val} {c} {c}")
            for qid in op["qubits"]:
                lines.append(f"      qubits {o} id: {_q(qid)} {c}")
            lines.append(f"    {c}")
        lines.append(f"  {c}")
    lines.append(c)
    return "\n".join(lines).encode()


def paulisum_text(b):
    out = []
    for cr, ci, paulis in wire.decode_paulisum(b):
        body = f"coefficient_real: {cr:.9g} coefficient_imag: {ci:.9g}"
        for qid, pt in paulis:
            body += f" paulis {{ qubit_id: {_q(qid)} pauli_type: {_q(pt)} }}"
        out.append(f"terms {{ {body} }}")
    return "\n".join(out).encode()
