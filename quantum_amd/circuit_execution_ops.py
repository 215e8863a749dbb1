"""Op getters (``core/ops/circuit_execution_ops.py:41-451``), None-backend only.

Each getter returns the HIP op with TFQ's calling convention; samples and
states are returned padded (``padded_to_ragged`` / ``padded_to_ragged2d`` make
them ragged, as the reference does, ``tfq_utility_ops.py:69-103``).  The
optional global serialisation of ``quantum_context`` (``_GLOBAL_OP_LOCK``) is
honoured.  Only ``backend=None`` / ``'noiseless'`` exists: the cirq backends are
a different engine and noisy trajectories are not part of this op family.
"""

import functools
import threading

from . import tfq_simulate_ops as sim

_GLOBAL_OP_LOCK = threading.Lock()
_quantum_concurrent = True


def set_quantum_concurrent_op_mode(mode):
    """``quantum_context.set_quantum_concurrent_op_mode`` (``quantum_context.py:82-107``)."""
    global _quantum_concurrent
    if not isinstance(mode, bool):
        raise TypeError("mode must be a python bool.")
    _quantum_concurrent = mode


def get_quantum_concurrent_op_mode():
    return _quantum_concurrent


def _check_backend(backend):
    if backend not in (None, "noiseless"):
        raise NotImplementedError("Only the native simulator backend (None) is available.")


def _locked(fn):
    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        if _quantum_concurrent:
            return fn(*args, **kwargs)
        with _GLOBAL_OP_LOCK:
            return fn(*args, **kwargs)
    return wrapper


def get_expectation_op(backend=None, *, quantum_concurrent=None):
    _check_backend(backend)

    @_locked
    def expectation_op(programs, symbol_names, symbol_values, pauli_sums):
        return sim.tfq_simulate_expectation(programs, symbol_names, symbol_values, pauli_sums)
    return expectation_op


def get_sampled_expectation_op(backend=None, *, quantum_concurrent=None, seed=None):
    _check_backend(backend)

    @_locked
    def sampled_expectation_op(programs, symbol_names, symbol_values, pauli_sums, num_samples):
        return sim.tfq_simulate_sampled_expectation(programs, symbol_names, symbol_values,
                                                    pauli_sums, num_samples, seed=seed)
    return sampled_expectation_op


def get_sampling_op(backend=None, *, quantum_concurrent=None, seed=None):
    _check_backend(backend)

    @_locked
    def sampling_op(programs, symbol_names, symbol_values, num_samples):
        return sim.tfq_simulate_samples(programs, symbol_names, symbol_values, num_samples,
                                        seed=seed)
    return sampling_op


def get_state_op(backend=None, *, quantum_concurrent=None):
    _check_backend(backend)

    @_locked
    def state_op(programs, symbol_names, symbol_values):
        return sim.tfq_simulate_state(programs, symbol_names, symbol_values)
    return state_op
