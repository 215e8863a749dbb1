"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and the host side (decode, resolve, fuse, plan) accepts
valid inputs and rejects invalid ones with the reference's messages."""

import ctypes
import os
import re

import numpy as np
import pytest

from quantum_amd import _lib
from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W
from quantum_amd._lib import InvalidArgumentError, Plan, StringTensor, np_ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "tfqx.h")).read()
    declared = set(re.findall(r"\b(tfqx_[a-z_0-9]+)\s*\(", header))
    assert len(declared) >= 18
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTED_SYMBOLS) == declared
    assert b"gfx950" in _lib.lib().tfqx_version()


def _plan_expectation(circuits, names, values, psums):
    progs = S.convert_to_tensor(circuits)
    P = StringTensor(progs)
    N = StringTensor(np.array(names, dtype=object))
    vals = np.ascontiguousarray(values, dtype=np.float32)
    ps = S.convert_to_tensor(psums)
    Ps = StringTensor(ps)
    return Plan(_lib.lib().tfqx_plan_simulate_expectation, P.ptr, P.off, P.n, N.ptr, N.off, N.n,
                np_ptr(vals), Ps.ptr, Ps.off, ps.shape[1])


def test_plan_config1_on_host():
    circuits, names, values, ops = W.config1_line_rotations(batch=4)
    plan = _plan_expectation(circuits, names, values, ops)
    assert plan.max_num_qubits == 10
    assert plan.workspace_bytes > 0


def test_plan_large_state_on_host():
    q = C.LineQubit.range(20)
    circ = C.Circuit([C.H(x) for x in q] + [C.CNOT(q[i], q[i + 1]) for i in range(19)])
    plan = _plan_expectation([circ], [], np.zeros((1, 0)), [[C.pauli_z(q[0])]])
    assert plan.max_num_qubits == 20
    assert plan.workspace_bytes > 8 * 2 ** 20


def test_unparseable_proto():
    q = C.LineQubit(0)
    P = StringTensor(np.array([b"junk"], dtype=object))
    N = StringTensor(np.array([], dtype=object))
    ps = S.convert_to_tensor([[C.pauli_z(q)]])
    Ps = StringTensor(ps)
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        Plan(_lib.lib().tfqx_plan_simulate_expectation, P.ptr, P.off, 1, N.ptr, N.off, 0, None,
             Ps.ptr, Ps.off, 1)


def test_missing_symbol():
    q = C.LineQubit(0)
    circ = C.Circuit(C.X(q) ** C.Symbol("alpha"))
    with pytest.raises(InvalidArgumentError, match="Could not find symbol in parameter map: alpha"):
        _plan_expectation([circ], ["beta"], [[0.1]], [[C.pauli_z(q)]])


def test_pauli_qubit_not_in_circuit():
    q = C.LineQubit.range(2)
    circ = C.Circuit(C.X(q[0]))
    with pytest.raises(InvalidArgumentError,
                       match="Found a Pauli sum operating on qubits not found in circuit"):
        _plan_expectation([circ], [], np.zeros((1, 0)), [[C.pauli_z(q[1])]])


def test_empty_program_sentinel_plan():
    q = C.LineQubit(0)
    plan = _plan_expectation([C.Circuit()], [], np.zeros((1, 0)), [[C.pauli_z(q)]])
    assert plan.max_num_qubits == 0


def test_num_samples_must_be_positive():
    from quantum_amd import tfq_simulate_ops as ops
    q = C.LineQubit(0)
    with pytest.raises(InvalidArgumentError, match="Each element of num_samples must be greater than 0"):
        progs = S.convert_to_tensor([C.Circuit(C.X(q))])
        ps = S.convert_to_tensor([[C.pauli_z(q)]])
        P, N = StringTensor(progs), StringTensor(np.array([], dtype=object))
        Ps = StringTensor(ps)
        ns = np.zeros((1, 1), np.int32)
        Plan(_lib.lib().tfqx_plan_simulate_sampled_expectation, P.ptr, P.off, 1, N.ptr, N.off, 0,
             None, Ps.ptr, Ps.off, 1, np_ptr(ns), 1)


def test_rank_and_batch_checks():
    from quantum_amd import tfq_simulate_ops as ops
    q = C.LineQubit(0)
    progs = S.convert_to_tensor([C.Circuit(C.X(q))])
    with pytest.raises(InvalidArgumentError, match="symbol_values must be rank 2"):
        ops._common(progs, [], np.zeros(3))
    with pytest.raises(InvalidArgumentError, match="do not match"):
        ops._common(progs, [], np.zeros((2, 0)))
    with pytest.raises(InvalidArgumentError, match="programs must be rank 1"):
        ops._common(progs.reshape(1, 1), [], np.zeros((1, 0)))


def test_channel_gate_id_rejected():
    # A gate id outside the 16 the simulator parses (e.g. a noise channel).
    q = C.LineQubit(0)
    good = S.serialize_circuit(C.Circuit(C.X(q)))
    bad = good.replace(b"\n\x02XP", b"\n\x02DP")
    P = StringTensor(np.array([bad], dtype=object))
    N = StringTensor(np.array([], dtype=object))
    ps = S.convert_to_tensor([[C.pauli_z(q)]])
    Ps = StringTensor(ps)
    with pytest.raises(InvalidArgumentError, match="Could not parse gate id: DP"):
        Plan(_lib.lib().tfqx_plan_simulate_expectation, P.ptr, P.off, 1, N.ptr, N.off, 0, None,
             Ps.ptr, Ps.off, 1)


def test_rebind_for_batch_kernel_plans_but_not_streaming_plans():
    """Rebinding keeps the plan's structure (op and pass counts) for one
    program tiled over the batch, for a batch of different programs and for
    tiled circuits; streaming plans (19+ qubits) are refused."""
    circuit, names, q = W.hea_ring(6, 2)
    ham = W.tfim(q, 0.7)
    rng = np.random.default_rng(0)
    vals = rng.uniform(-3, 3, (5, len(names)))
    shared = _plan_expectation([circuit] * 5, names, vals, [[ham]] * 5)
    stats = shared.stats()
    shared.rebind(rng.uniform(-3, 3, (5, len(names))))
    assert shared.stats() == stats
    other = W.hea_ring(6, 1)[0]
    mixed = _plan_expectation([circuit, other] * 2, names, vals[:4], [[ham]] * 4)
    stats = mixed.stats()
    mixed.rebind(vals[:4] + 1.0)
    assert mixed.stats() == stats
    wide, wnames, wq = W.hea_ring(16, 1)
    tiled = _plan_expectation([wide] * 2, wnames, rng.uniform(-3, 3, (2, len(wnames))),
                              [[W.tfim(wq, 1.0)]] * 2)
    stats = tiled.stats()
    tiled.rebind(np.zeros((2, len(wnames))))
    assert tiled.stats() == stats
    huge, hnames, hq = W.hea_ring(19, 1)
    big = _plan_expectation([huge], hnames, rng.uniform(-3, 3, (1, len(hnames))),
                            [[W.tfim(hq, 1.0)]])
    with pytest.raises(InvalidArgumentError, match="cannot be rebound"):
        big.rebind(np.zeros((1, len(hnames))))


@pytest.mark.parametrize("width", [1, 2, 3, 4])
def test_fuser_decisions_match_host_replica(width):
    """The planner's fusion decisions equal the oracle's restatement exactly."""
    from oracle import fuser
    rng = np.random.default_rng(100 + width)
    for trial in range(12):
        q = C.LineQubit.range(int(rng.integers(2, 9)))
        circ = W.random_circuit(rng, q, int(rng.integers(5, 60)), p_control=0.25)
        data = S.serialize_circuit(circ)
        got = _lib.fusion_decisions(data, width=width)
        want = fuser.fuse(fuser.program_masks(data), width)
        assert got == want, (trial, width)


def test_fuser_replica_pinned_by_hand_example():
    """A hand-worked case: H(0) CZ(0,1) H(2) CNOT(1,2) X(3) at width 2."""
    from oracle import fuser
    q = C.LineQubit.range(4)
    circ = C.Circuit.from_moments([[C.H(q[0])], [C.CZ(q[0], q[1])], [C.H(q[2])],
                                   [C.CNOT(q[1], q[2])], [C.X(q[3])]])
    data = S.serialize_circuit(circ)
    # qubit q -> bit 3 - q.  H0 opens {3}; CZ joins -> {2,3}; H2 opens {1};
    # CNOT(1,2) touches {2,3} and {1}: union 3 qubits > 2 -> CNOT opens block
    # {1,2}, into which {1} (inside the gate's qubits) moves while {2,3}
    # closes; X3 opens {0}; end: {1,2} = (H2, CNOT) then {0}.
    want = [((2, 3), False, (0, 1)), ((1, 2), False, (2, 3)), ((0,), False, (4,))]
    assert fuser.fuse(fuser.program_masks(data), 2) == want
    assert _lib.fusion_decisions(data, width=2) == want


@pytest.mark.parametrize("n,T,grid", [(20, 1, 1), (16, 3, 3), (15, 700, 512)])
def test_noisy_workspace_holds_one_state_per_launched_workgroup(n, T, grid):
    """Programs over 14 qubits keep each workgroup trajectory's state in the
    workspace: one 2^n buffer per launched workgroup, min(512, trajectories)."""
    from quantum_amd import noise
    q = C.LineQubit.range(n)
    circ = C.Circuit([C.H(x) for x in q] + [C.depolarize(0.01).on(q[0])])
    progs = S.convert_to_tensor([circ])
    ps = S.convert_to_tensor([[C.PauliSum([C.pauli_z(q[0])])]])
    plan, _ = noise.plan_expectation(progs, [], np.zeros((1, 0)), ps,
                                     np.full((1, 1), T, np.int32), seed=1)
    states = grid * 8 * 2 ** n
    assert states <= plan.workspace_bytes < states + 4 * 2 ** 20


def test_stage_values_builds_device_tables_and_rejects_streaming_plans():
    """Host side of the device re-evaluation: the tables of heterogeneous and
    shared-program plans build with the sizes the workspace layout reserved
    (the library checks them), and streaming plans are refused."""
    rng = np.random.default_rng(5)
    names = ["a", "b", "c"]
    circuits, psums = [], []
    for n in (3, 9, 12, 14, 16):
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, 4 * n, symbols=names, p_symbol=0.6,
                                         p_control=0.2))
        psums.append([W.random_pauli_sum(rng, q, 2)])
    values = rng.uniform(-1, 1, size=(len(circuits), 3)).astype(np.float32)
    plan = _plan_expectation(circuits, names, values, psums)
    plan.stage_values(values)
    plan.stage_values(values * 2)
    shared = _plan_expectation([circuits[1]] * 6, names, np.repeat(values[1:2], 6, 0),
                               [psums[1]] * 6)
    shared.stage_values(np.repeat(values[1:2], 6, 0))
    q = C.LineQubit.range(20)
    wide = _plan_expectation([W.random_circuit(rng, q, 30, symbols=names)], names, values[:1],
                             [[W.random_pauli_sum(rng, q, 2)]])
    assert wide.stats()["streaming"]
    with pytest.raises(InvalidArgumentError, match="cannot be rebound"):
        wide.stage_values(values[:1])
