"""Op getters (``core/ops/circuit_execution_ops.py:41-451``), None-backend only.

Each getter returns the HIP op with TFQ's calling convention; samples and
states are returned padded (``padded_to_ragged`` / ``padded_to_ragged2d`` make
them ragged, as the reference does, ``tfq_utility_ops.py:69-103``).  Only
``backend=None`` / ``'noiseless'`` exists here: the cirq backends are a
different engine.  Noisy trajectories are served by ``quantum_amd.noise``
(the ``backend='noisy'`` layers), as in the reference, whose getters have no
noisy backend either.

``quantum_concurrent`` is decided per getter, as in the reference
(``circuit_execution_ops.py:41-44,133-142``): ``True`` returns the op
itself; ``False`` returns an op that runs under the process-wide
``_GLOBAL_OP_LOCK`` (``quantum_context.py:117``), so such ops never overlap
one another.  A non-bool raises ``TypeError``.  The default is the global
mode (``set_quantum_concurrent_op_mode``) current when the getter is called;
the reference evaluates it when its module is imported, which gives the same
result for the documented use (set the mode right after import).
"""

import functools
import threading

from . import tfq_simulate_ops as sim

_GLOBAL_OP_LOCK = threading.Lock()
_quantum_concurrent = True
_DEFAULT = object()


def set_quantum_concurrent_op_mode(mode):
    """``quantum_context.set_quantum_concurrent_op_mode`` (``quantum_context.py:82-107``)."""
    global _quantum_concurrent
    _quantum_concurrent = mode


def get_quantum_concurrent_op_mode():
    return _quantum_concurrent


def _check_backend(backend):
    if backend not in (None, "noiseless"):
        raise NotImplementedError("Only the native simulator backend (None) is available.")


def _check_quantum_concurrent(quantum_concurrent):
    if not isinstance(quantum_concurrent, bool):
        raise TypeError("quantum_concurrent must be type bool."
                        f" Given: {type(quantum_concurrent)}")


def _resolve(quantum_concurrent):
    if quantum_concurrent is _DEFAULT:
        quantum_concurrent = _quantum_concurrent
    _check_quantum_concurrent(quantum_concurrent)
    return quantum_concurrent


def _wrap(fn, quantum_concurrent):
    """``fn`` itself when concurrent, else ``fn`` under the global op lock."""
    if quantum_concurrent:
        return fn

    @functools.wraps(fn)
    def serialised(*args, **kwargs):
        with _GLOBAL_OP_LOCK:
            return fn(*args, **kwargs)
    return serialised


def get_expectation_op(backend=None, *, quantum_concurrent=_DEFAULT):
    _check_backend(backend)
    concurrent = _resolve(quantum_concurrent)

    def expectation_op(programs, symbol_names, symbol_values, pauli_sums):
        return sim.tfq_simulate_expectation(programs, symbol_names, symbol_values, pauli_sums)
    return _wrap(expectation_op, concurrent)


def get_sampled_expectation_op(backend=None, *, quantum_concurrent=_DEFAULT, seed=None):
    _check_backend(backend)
    concurrent = _resolve(quantum_concurrent)

    def sampled_expectation_op(programs, symbol_names, symbol_values, pauli_sums, num_samples):
        return sim.tfq_simulate_sampled_expectation(programs, symbol_names, symbol_values,
                                                    pauli_sums, num_samples, seed=seed)
    return _wrap(sampled_expectation_op, concurrent)


def get_sampling_op(backend=None, *, quantum_concurrent=_DEFAULT, seed=None):
    _check_backend(backend)
    concurrent = _resolve(quantum_concurrent)

    def sampling_op(programs, symbol_names, symbol_values, num_samples):
        return sim.tfq_simulate_samples(programs, symbol_names, symbol_values, num_samples,
                                        seed=seed)
    return _wrap(sampling_op, concurrent)


def get_state_op(backend=None, *, quantum_concurrent=_DEFAULT):
    _check_backend(backend)
    concurrent = _resolve(quantum_concurrent)

    def state_op(programs, symbol_names, symbol_values):
        return sim.tfq_simulate_state(programs, symbol_names, symbol_values)
    return _wrap(state_op, concurrent)
