"""``tfq.noise``: Monte-Carlo trajectory simulation of noisy circuits
(``core/ops/noise/noisy_expectation_op.py``, ``noisy_sampled_expectation_op.py``,
``noisy_samples_op.py``; kernels ``tfq_noisy_expectation.cc``,
``tfq_noisy_sampled_expectation.cc``, ``tfq_noisy_samples.cc``).

Programs may hold the channels of ``circuits.supported_channels()`` (ids DP,
ADP, GAD, AD, RST, PD, PF, BF) besides gates.  Each trajectory picks one Kraus
operator per channel at random; results average over trajectories:

* ``expectation``: term j of row i averages its exact expectation over
  ``num_samples[i][j]`` trajectories;
* ``sampled_expectation``: each trajectory adds one measured parity per Pauli
  term instead;
* ``samples``: one terminal measurement per trajectory, ``num_samples``
  trajectories per circuit (padded bitstrings; ``padded_to_ragged2d`` gives
  the reference's ragged output).

The reference draws a fresh random seed per call (``tensorflow::random::New64()``);
here ``seed`` (default: a fresh one) fixes the counter-based Philox stream,
so a call is reproducible.  All trajectories of a batch run in one persistent
GPU launch (circuits of at most 20 qubits).
"""

import numpy as np

from . import _lib
from ._lib import InvalidArgumentError, Plan, StringTensor, np_ptr
from .tfq_simulate_ops import (_common, _num_samples_scalar, _plan_args, _run, _seed, _torch,
                               as_string_tensor, as_values)


def _sums_and_counts(programs, pauli_sums, num_samples):
    pauli_sums = as_string_tensor(pauli_sums, 2, "pauli_sums")
    ns = as_values(num_samples, "num_samples", 2, np.int32)
    if pauli_sums.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"Number of circuits and PauliSums do not match. Got {programs.shape[0]} circuits "
            f"and {pauli_sums.shape[0]} paulisums.")
    if ns.shape[0] != pauli_sums.shape[0]:
        raise InvalidArgumentError(
            f"Dimension 0 of num_samples and pauli_sums do not match.Got {ns.shape[0]} lists of "
            f"sample sizes and {pauli_sums.shape[0]} lists of pauli sums.")
    if ns.shape[1] != pauli_sums.shape[1]:
        raise InvalidArgumentError(
            f"Dimension 1 of num_samples and pauli_sums do not match.Got {ns.shape[1]} lists of "
            f"sample sizes and {pauli_sums.shape[1]} lists of pauli sums.")
    return pauli_sums, ns


def plan_expectation(programs, symbol_names, symbol_values, pauli_sums, num_samples, seed=None,
                     sampled=False):
    """The plan of a noisy (sampled) expectation call and its output shape."""
    programs, names, values = _common(programs, symbol_names, symbol_values)
    pauli_sums, ns = _sums_and_counts(programs, pauli_sums, num_samples)
    P, N, args = _plan_args(programs, names, values)
    S = StringTensor(pauli_sums)
    fn = (_lib.lib().tfqx_plan_noisy_sampled_expectation if sampled
          else _lib.lib().tfqx_plan_noisy_expectation)
    plan = Plan(fn, *args, S.ptr, S.off, pauli_sums.shape[1], np_ptr(ns), _seed(seed))
    return plan, pauli_sums.shape


def expectation(programs, symbol_names, symbol_values, pauli_sums, num_samples, seed=None):
    """Noisy expectations ``[batch, n_ops]`` (float32); -2 for empty programs."""
    plan, shape = plan_expectation(programs, symbol_names, symbol_values, pauli_sums,
                                   num_samples, seed)
    return _run(plan, shape, _torch().float32)


def sampled_expectation(programs, symbol_names, symbol_values, pauli_sums, num_samples,
                        seed=None):
    """Noisy sampled expectations ``[batch, n_ops]`` (float32): one measured
    parity per Pauli term per trajectory."""
    plan, shape = plan_expectation(programs, symbol_names, symbol_values, pauli_sums,
                                   num_samples, seed, sampled=True)
    return _run(plan, shape, _torch().float32)


def samples(programs, symbol_names, symbol_values, num_samples, seed=None):
    """Bitstrings ``[batch, num_samples, max_nq]`` (int8) with leading -2
    padding, one trajectory per sample."""
    programs, names, values = _common(programs, symbol_names, symbol_values)
    n = _num_samples_scalar(num_samples)
    P, N, args = _plan_args(programs, names, values)
    plan = Plan(_lib.lib().tfqx_plan_noisy_samples, *args, n, _seed(seed))
    return _run(plan, (programs.shape[0], n, plan.max_num_qubits), _torch().int8)
