"""Text-format fallback of the proto parser (``parse_context.cc:41-56``):
a program or Pauli sum that is not valid binary wire format is parsed as
protobuf text format.

* Text renderings of random programs (every gate family, symbols, controls)
  parse to the same programs as their binary form: re-encoding through
  ``tfq_ps_decompose`` (identity on programs without decomposable symbolic
  gates) gives equal decoded structures.
* Syntax the text format allows: ``<>`` delimiters, comments, commas and
  semicolons, escapes, float suffixes, enum numbers, lists.
* Malformed text and unknown fields stay "Unparseable proto".
"""

import numpy as np
import pytest

from oracle import wire
from quantum_amd import circuits as C
from quantum_amd import ps_util_ops as P
from quantum_amd import serialize as S
from quantum_amd import workloads as W
from quantum_amd._lib import InvalidArgumentError
from textproto_util import paulisum_text, program_text

SIMPLE = ["XP", "YP", "ZP", "HP", "XXP", "YYP", "ZZP", "CZP", "CNP", "SP", "I"]


def _same(a_bytes, b_bytes):
    return wire.decode_program(P.tfq_ps_decompose(np.array([a_bytes], dtype=object))[0]) == \
        wire.decode_program(P.tfq_ps_decompose(np.array([b_bytes], dtype=object))[0])


@pytest.mark.parametrize("style", ["{}", "<>"])
def test_text_programs_parse_like_binary(style):
    rng = np.random.default_rng(3)
    q = C.GridQubit.rect(2, 3)
    for k in range(5):
        circ = W.random_circuit(rng, q, 20, symbols=["a", "b"], p_control=0.3,
                                families=SIMPLE if k % 2 else None)
        b = S.serialize_circuit(circ)
        t = program_text(b, style)
        assert _same(b, t)


def test_text_syntax_variants():
    b = S.serialize_circuit(C.Circuit([C.X(C.LineQubit(0)) ** 0.5]))
    text = b"""
    # a comment
    language < gate_set: 'tfq_gate_set' >
    circuit {
      scheduling_strategy: 1;
      moments { operations {
        gate: { id: "X" "P" }
        args { key: "exponent" value { arg_value { float_value: 0.5f } } },
        args { key: "exponent_scalar" value { arg_value { float_value: 1 } } }
        args { key: "global_shift" value { arg_value { float_value: 0.0 } } }
        args { key: "control_qubits" value { arg_value { string_value: "" } } }
        args { key: "control_values" value { arg_value { string_value: "\\x00"[0:0] } } }
        qubits [{ id: "0" }]
      } }
    }
    """
    # The "[0:0]" above is not text format: the parse must fail ...
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        P.tfq_ps_decompose(np.array([text], dtype=object))
    # ... and succeed without it.
    good = text.replace(b'"\\x00"[0:0]', b'""')
    assert _same(b, good)


def test_text_pauli_sums_and_errors():
    from quantum_amd import tfq_simulate_ops as T
    q = C.LineQubit.range(2)
    circ = S.convert_to_tensor([C.Circuit([C.H(q[0]), C.CNOT(q[0], q[1])])])
    ps = C.PauliSum([C.pauli_z(q[0]) * C.pauli_z(q[1]), C.PauliString({q[1]: "X"}, 0.5)])
    text = paulisum_text(S.serialize_paulisum(ps))
    # Plans build from text Pauli sums (host only).
    T.plan_expectation(circ, [], np.zeros((1, 0)), np.array([[text]], dtype=object))
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        T.plan_expectation(circ, [], np.zeros((1, 0)), np.array([[b"terms { nope: 1 }"]], dtype=object))
    with pytest.raises(InvalidArgumentError, match="Unparseable proto: junk"):
        T.plan_expectation(np.array([b"junk"], dtype=object), [], np.zeros((1, 0)),
                           np.array([[text]], dtype=object))
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        P.tfq_ps_decompose(np.array([b"circuit { moments { operations { gate { id: 'XP' } "],
                                    dtype=object))
