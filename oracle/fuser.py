"""Host replica of the planner's gate fuser (test infrastructure only).

Restates the greedy frontier fuser of ``quantum_amd/csrc/plan.cc`` (the role
of qsim's BasicGateFuser, called at ``core/src/circuit_parser_qsim.cc:857-859``)
from its specification, on qubit masks only:

* gates are visited in program order; a gate's qubits are its targets and
  controls;
* a gate on more qubits than the width is emitted alone (targets in the
  matrix, controls kept outside), after closing the open blocks it touches;
* otherwise, if the union of the gate's qubits and the open blocks it touches
  fits the width, the touched blocks merge (into the oldest) and the gate
  joins; if not, the gate opens a new block, into which the touched blocks
  that lie within the gate's own qubits move (their gates first, in order of
  the blocks' indices), while the other touched blocks close;
* closing a block emits it; at the end the still-open blocks are emitted in
  creation order.

``fuse`` returns [(bits ascending, controls_outside, gate indices)], the same
records as the C-ABI's ``tfqx_debug_fuse``.
"""


def _bits(mask):
    return tuple(b for b in range(64) if (mask >> b) & 1)


def fuse(gate_masks, width):
    """gate_masks: [(target_mask, control_mask)] in program order."""
    blocks = []  # [gates, qmask, open]
    owner = {}
    out = []

    def close(bi):
        b = blocks[bi]
        if not b[2]:
            return
        b[2] = False
        for q in _bits(b[1]):
            owner.pop(q, None)
        out.append((_bits(b[1]), False, tuple(b[0])))

    for gi, (tm, cm) in enumerate(gate_masks):
        Q = tm | cm
        touching = sorted({owner[q] for q in _bits(Q) if q in owner})
        if bin(Q).count("1") > width:
            for t in touching:
                close(t)
            out.append((_bits(tm), cm != 0, (gi,)))
            continue
        U = Q
        for t in touching:
            U |= blocks[t][1]
        if bin(U).count("1") <= width:
            if not touching:
                blocks.append([[], 0, True])
                keep = len(blocks) - 1
            else:
                keep = touching[0]
                for t in touching[1:]:
                    o = blocks[t]
                    blocks[keep][0] += o[0]
                    blocks[keep][1] |= o[1]
                    o[2] = False
                    o[0] = []
            blocks[keep][0].append(gi)
            blocks[keep][1] |= Q
            for q in _bits(blocks[keep][1]):
                owner[q] = keep
        else:
            gates = []
            for t in touching:
                if blocks[t][1] & ~Q == 0:
                    gates += blocks[t][0]
                    blocks[t][0] = []
                    blocks[t][2] = False
                else:
                    close(t)
            blocks.append([gates + [gi], Q, True])
            for q in _bits(Q):
                owner[q] = len(blocks) - 1
    for bi in range(len(blocks)):
        close(bi)
    return out


def program_masks(program_bytes, symbol_map=None):
    """(target_mask, control_mask) per gate of a serialized program, from the
    oracle's own parser."""
    from . import tfq_oracle as T
    prog = T.Program(program_bytes, symbol_map or {})
    return [(sum(1 << b for b in g.bits), sum(1 << b for b in g.cbits)) for g in prog.gates]
