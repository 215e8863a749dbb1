"""Input validation of the op wrappers against the reference's error cases
(``core/ops/tfq_simulate_ops_test.py:36-204,230-300``): wrong ranks, wrong
element types, mismatched batch sizes, unparseable protos, unknown symbols,
Pauli sums on foreign qubits and channels in noiseless ops.  Host only: every
error is raised before any device work."""

import numpy as np
import pytest

from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import tfq_simulate_ops as T
from quantum_amd import workloads as W
from quantum_amd._lib import InvalidArgumentError


def _case(batch=3, n=4):
    rng = np.random.default_rng(0)
    q = C.GridQubit.rect(1, n)
    circuits = [W.random_circuit(rng, q, 8, symbols=["alpha"], p_symbol=1.0) + C.Circuit([C.I(x) for x in q])
                for _ in range(batch)]
    sums = [[W.random_pauli_sum(rng, q, 2)] for _ in range(batch)]
    vals = rng.uniform(size=(batch, 1)).astype(np.float32)
    return S.convert_to_tensor(circuits), ["alpha"], vals, S.convert_to_tensor(sums), q


def _exp(*a):
    return T.plan_expectation(*a)


def test_expectation_input_errors():
    progs, names, vals, ps, q = _case()
    B = len(progs)
    with pytest.raises(InvalidArgumentError, match="programs must be rank 1"):
        _exp(np.array([progs], dtype=object), names, vals, ps)
    with pytest.raises(InvalidArgumentError, match="symbol_names must be rank 1."):
        _exp(progs, np.array([names], dtype=object), vals, ps)
    with pytest.raises(InvalidArgumentError, match="symbol_values must be rank 2."):
        _exp(progs, names, np.array([vals]), ps)
    with pytest.raises(InvalidArgumentError, match="symbol_values must be rank 2."):
        _exp(progs, names, vals[0], ps)
    with pytest.raises(InvalidArgumentError, match="pauli_sums must be rank 2."):
        _exp(progs, names, vals, ps[:, 0])
    with pytest.raises(InvalidArgumentError, match="pauli_sums must be rank 2."):
        _exp(progs, names, vals, ps[None])
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        _exp(np.array([b"junk"] * B, dtype=object), names, vals, ps)
    with pytest.raises(InvalidArgumentError, match="Could not find symbol in parameter map"):
        _exp(progs, ["junk"], vals, ps)
    other_q = C.GridQubit.rect(1, 6)[5]
    bad_ps = S.convert_to_tensor([[C.PauliSum([C.pauli_z(other_q)])]] * B)
    with pytest.raises(InvalidArgumentError, match="qubits not found in circuit"):
        _exp(progs, names, vals, bad_ps)
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        _exp(progs, names, vals, np.array([[b"junk"]] * B, dtype=object))
    with pytest.raises(TypeError, match="Cannot convert"):
        _exp([1.0] * B, names, vals, ps)
    with pytest.raises(TypeError, match="Cannot convert"):
        _exp(progs, [0.1234], vals, ps)
    with pytest.raises(NotImplementedError, match="Cast string to float is not supported"):
        _exp(progs, names, [["junk"]] * B, ps)
    with pytest.raises(TypeError, match="Cannot convert"):
        _exp(progs, names, vals, [[1.0]] * B)
    with pytest.raises(TypeError, match="missing"):
        T.tfq_simulate_expectation(progs, names, vals)
    with pytest.raises(TypeError, match="positional arguments"):
        T.tfq_simulate_expectation(progs, names, vals, ps, [])
    with pytest.raises(InvalidArgumentError, match="do not match"):
        _exp(progs[:2], names, vals, ps)
    with pytest.raises(InvalidArgumentError, match="do not match"):
        _exp(progs, names, vals[:2], ps)
    with pytest.raises(InvalidArgumentError, match="do not match"):
        _exp(progs, names, vals, ps[:2])
    noisy = S.convert_to_tensor([C.Circuit([C.depolarize(0.1).on(q[0])])] * B)
    with pytest.raises(InvalidArgumentError, match="cirq.Channel"):
        _exp(noisy, names, vals, ps)


def test_samples_and_sampled_expectation_input_errors():
    progs, names, vals, ps, q = _case()
    B = len(progs)
    from quantum_amd import _lib
    # num_samples shapes (the plans below are host-only).
    with pytest.raises(InvalidArgumentError, match="num_samples must be rank 2"):
        T.tfq_simulate_sampled_expectation(progs, names, vals, ps, [10] * B)
    with pytest.raises(InvalidArgumentError, match="rank 1. Got rank 2"):
        T.tfq_simulate_samples(progs, names, vals, [[10]])
    with pytest.raises(InvalidArgumentError, match="num_samples must contain 1 element"):
        T.tfq_simulate_samples(progs, names, vals, [10, 20])
    with pytest.raises(TypeError, match="Cannot convert"):
        T.tfq_simulate_samples([1] * B, names, vals, [10])
    with pytest.raises(InvalidArgumentError, match="programs must be rank 1"):
        T.tfq_simulate_state(np.array([progs], dtype=object), names, vals)
    with pytest.raises(InvalidArgumentError, match="Unparseable proto"):
        T.tfq_simulate_state(np.array([b"junk"] * B, dtype=object), names, vals)
    with pytest.raises(InvalidArgumentError, match="Could not find symbol in parameter map"):
        T.tfq_simulate_state(progs, ["junk"], vals)
