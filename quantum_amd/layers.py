"""Circuit-executor layers with TFQ's call signatures, as ``torch.nn.Module``s.

* ``Expectation``        ``python/layers/circuit_executors/expectation.py:231-390``
* ``SampledExpectation`` ``sampled_expectation.py:216-342``
* ``Sample``             ``sample.py:142-201``
* ``State``              ``state.py:115-148``
* input expansion        ``input_checks.py:24-165``

``layer(inputs, symbol_names=..., symbol_values=..., operators=...,
repetitions=..., initializer=...)``: a single circuit is tiled over the batch
of symbol values, a single operator (or flat list) over the batch of circuits;
when ``symbol_values`` is omitted the layer owns learnable parameters
initialised ``U(0, 2*pi)`` (``expectation.py:372-379``).  Backends: the
native noiseless simulator (default differentiator ``Adjoint``) and
``backend="noisy"`` (Monte-Carlo trajectories, ``tfq.noise``), which needs
``repetitions`` trajectories per operator in the call (``expectation.py:278-345``).
Sample-based layers default to ``ParameterShift``, as the reference's do.
"""

import math

import numpy as np

from . import circuits as C
from . import circuit_execution_ops as ceo
from . import differentiators as diff
from . import noise
from . import serialize
from . import tfq_simulate_ops as sim


def _torch():
    import torch
    return torch


def _module_base():
    return _torch().nn.Module


def expand_circuits(inputs, symbol_names=None, symbol_values=None):
    """-> (programs ndarray[B], names list, values float32 [B, S])."""
    torch = _torch()
    if symbol_names is None:
        symbol_names = []
    names = [str(s.name if isinstance(s, C.Symbol) else s) for s in list(symbol_names)]
    if len(set(names)) != len(names):
        raise ValueError("All elements of symbol_names must be unique.")
    values_empty = symbol_values is None
    if values_empty:
        symbol_values = np.zeros((1, 0), np.float32)
    if isinstance(symbol_values, torch.Tensor):
        values = symbol_values
    else:
        values = torch.as_tensor(np.asarray(symbol_values, dtype=np.float32))
    if isinstance(inputs, C.Circuit):
        progs = serialize.convert_to_tensor([inputs])
        progs = np.repeat(progs, values.shape[0])
    elif isinstance(inputs, np.ndarray) and inputs.dtype == object:
        progs = inputs
    elif isinstance(inputs, (list, tuple)):
        progs = serialize.convert_to_tensor(list(inputs))
    else:
        raise TypeError(f"circuits cannot be parsed with given input: {inputs}")
    if values_empty:
        values = torch.zeros((len(progs), 0), dtype=torch.float32)
    return progs, names, values


def expand_operators(operators, batch):
    """Tile a single operator / flat operator list over the batch."""
    if operators is None:
        raise RuntimeError("Value for operators not provided.")
    if isinstance(operators, (C.PauliSum, C.PauliString)):
        operators = [operators]
    if isinstance(operators, np.ndarray) and operators.dtype == object:
        return operators if operators.ndim == 2 else np.tile(operators[None], (batch, 1))
    ops = list(operators)
    if ops and isinstance(ops[0], (list, tuple)):
        return serialize.convert_to_tensor(ops)
    row = serialize.convert_to_tensor(ops)
    return np.tile(row[None], (batch, 1))


class _LearnableMixin:
    def _values(self, names, values, batch, initializer, device=None):
        torch = _torch()
        if values is not None:
            return values
        if self._w is None:
            w = torch.empty(len(names), dtype=torch.float32)
            if initializer is None:
                w.uniform_(0.0, 2 * math.pi)
            else:
                initializer(w)
            self._w = torch.nn.Parameter(w.to(device) if device is not None else w)
        return self._w.unsqueeze(0).expand(batch, -1)


def _check_backend(backend):
    if backend not in (None, "noiseless", "noisy"):
        raise NotImplementedError("Only the native noiseless and noisy backends are available.")
    return backend == "noisy"


def _noisy_op(fn, seed):
    def op(programs, symbol_names, symbol_values, pauli_sums, num_samples):
        return fn(programs, symbol_names, symbol_values, pauli_sums, num_samples, seed=seed)
    return op


def _repetitions(repetitions, shape, batch):
    """Trajectories per (circuit, operator) from an int, a list per operator
    or a full [batch, n_ops] array (``expectation.py:333-352``)."""
    if isinstance(repetitions, (int, np.integer)):
        return np.full(shape, int(repetitions), np.int32)
    reps = np.asarray(repetitions, np.int32)
    if reps.ndim == 1:
        reps = np.tile(reps[None], (batch, 1))
    return reps


class Expectation(_module_base(), _LearnableMixin):
    def __init__(self, backend="noiseless", differentiator=None, seed=None):
        super().__init__()
        self.noisy = _check_backend(backend)
        if differentiator is None:
            differentiator = diff.ParameterShift() if self.noisy else diff.Adjoint()
        if not isinstance(differentiator, diff.Differentiator):
            raise TypeError("Differentiator must inherit from Differentiator")
        if self.noisy:
            self._op = differentiator.generate_differentiable_op(
                sampled_op=_noisy_op(noise.expectation, seed))
        else:
            self._op = differentiator.generate_differentiable_op(
                analytic_op=ceo.get_expectation_op())
        self._w = None

    def forward(self, inputs, *, symbol_names=None, symbol_values=None, operators=None,
                repetitions=None, initializer=None):
        if not self.noisy and repetitions is not None:
            raise RuntimeError("repetitions value provided for analytic expectation calculation "
                               "that is noiseless.")
        if self.noisy and repetitions is None:
            raise RuntimeError("Value for repetitions not provided. With backend='noisy' a number "
                               "of trajectory repetitions must be provided in the layer call "
                               "method.")
        given = symbol_values
        progs, names, values = expand_circuits(inputs, symbol_names, symbol_values)
        ops = expand_operators(operators, len(progs))
        if given is None and names:
            values = self._values(names, None, len(progs), initializer)
        if self.noisy:
            return self._op(progs, names, values, ops, _repetitions(repetitions, ops.shape, len(progs)))
        return self._op(progs, names, values, ops)


class SampledExpectation(_module_base(), _LearnableMixin):
    def __init__(self, backend="noiseless", differentiator=None, seed=None):
        super().__init__()
        noisy = _check_backend(backend)
        if differentiator is None:
            differentiator = diff.ParameterShift()
        if isinstance(differentiator, diff.Adjoint):
            raise ValueError("sample base backends are not supported by the Adjoint method.")
        op = (_noisy_op(noise.sampled_expectation, seed) if noisy
              else ceo.get_sampled_expectation_op(seed=seed))
        self._op = differentiator.generate_differentiable_op(sampled_op=op)
        self._w = None

    def forward(self, inputs, *, symbol_names=None, symbol_values=None, operators=None,
                repetitions=None, initializer=None):
        if repetitions is None:
            raise ValueError("Number of repetitions not specified.")
        given = symbol_values
        progs, names, values = expand_circuits(inputs, symbol_names, symbol_values)
        ops = expand_operators(operators, len(progs))
        if np.isscalar(repetitions):
            reps = np.full(ops.shape, int(repetitions), np.int32)
        else:
            reps = np.asarray(repetitions, np.int32)
            if reps.ndim == 1:
                reps = np.tile(reps[None], (len(progs), 1))
        if given is None and names:
            values = self._values(names, None, len(progs), initializer)
        return self._op(progs, names, values, ops, reps)


class Sample(_module_base()):
    """Bitstrings, ragged per circuit (``sample.py:185-201``)."""

    def __init__(self, backend="noiseless", seed=None):
        super().__init__()
        if _check_backend(backend):
            self._op = lambda p, n, v, r: noise.samples(p, n, v, r, seed=seed)
        else:
            self._op = ceo.get_sampling_op(seed=seed)

    def forward(self, inputs, *, symbol_names=None, symbol_values=None, repetitions=None):
        if repetitions is None:
            raise ValueError("Number of repetitions not specified.")
        progs, names, values = expand_circuits(inputs, symbol_names, symbol_values)
        padded = self._op(progs, names, values, [int(repetitions)])
        return sim.padded_to_ragged2d(padded)


class State(_module_base()):
    """Final state vectors, ragged per circuit (``state.py:115-148``)."""

    def __init__(self, backend="noiseless"):
        super().__init__()
        if backend not in (None, "noiseless"):
            raise NotImplementedError("Only the noiseless native backend is available.")
        self._op = ceo.get_state_op()

    def forward(self, inputs, *, symbol_names=None, symbol_values=None):
        progs, names, values = expand_circuits(inputs, symbol_names, symbol_values)
        return sim.padded_to_ragged(self._op(progs, names, values))
