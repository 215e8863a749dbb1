"""``tfq_adj_grad`` (TFQ ``core/ops/tfq_adj_grad_op.py:22-48``).

Returns ``d(sum_j g_ij <O_ij>)/d(symbol_values[i, s])`` as float32 ``[batch,
n_symbols]`` using the adjoint method of ``core/ops/tfq_adj_grad_op.cc``.
``derivative="tfq"`` (default) reproduces TFQ's gradient gates -- the float32
central difference of each gate matrix with eps = 5e-3 (``adj_util.cc:32``);
``derivative="analytic"`` uses the exact gate derivative.
"""

import numpy as np

from . import _lib
from ._lib import InvalidArgumentError, Plan, StringTensor, np_ptr
from .tfq_simulate_ops import _common, _plan_args, _run, _torch, as_string_tensor, as_values

_MODES = {"tfq": _lib.TFQX_GRAD_TFQ_FD, "analytic": _lib.TFQX_GRAD_ANALYTIC}


def plan_adj_grad(programs, symbol_names, symbol_values, pauli_sums, prev_grad,
                  derivative="tfq", with_expectation=False):
    programs, names, values = _common(programs, symbol_names, symbol_values)
    pauli_sums = as_string_tensor(pauli_sums, 2, "pauli_sums")
    grads = as_values(prev_grad, "downstream_grads")
    if pauli_sums.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"Number of circuits and PauliSums do not match. Got {programs.shape[0]} circuits "
            f"and {pauli_sums.shape[0]} paulisums.")
    if grads.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"Number of gradients and circuits do not match. Got {grads.shape[0]} gradients "
            f"and {programs.shape[0]} circuits.")
    if grads.shape[1] != pauli_sums.shape[1]:
        raise InvalidArgumentError(
            f"Number of gradients and pauli sum dimension do not match. Got {grads.shape[1]} "
            f"gradient entries and {pauli_sums.shape[1]} paulis per circuit.")
    if derivative not in _MODES:
        raise ValueError(f"derivative must be one of {sorted(_MODES)}")
    P, N, args = _plan_args(programs, names, values)
    S = StringTensor(pauli_sums)
    fn = _lib.lib().tfqx_plan_expectation_and_gradient if with_expectation else \
        _lib.lib().tfqx_plan_adjoint_gradient
    plan = Plan(fn, *args, S.ptr, S.off, pauli_sums.shape[1], np_ptr(grads), _MODES[derivative])
    if with_expectation:
        return plan, (programs.shape[0] * (pauli_sums.shape[1] + len(names)),)
    return plan, (programs.shape[0], len(names))


def _cached(programs, symbol_names, symbol_values, pauli_sums, prev_grad, derivative,
            with_expectation):
    """Launch an adjoint plan, reusing a recent call's plan when the programs,
    names and Pauli sums are the same (see ``plan_cache``): the new values are
    staged and evaluated into the matrices on the device, the downstream
    weights rewritten, and only those are uploaded."""
    from . import plan_cache
    from .tfq_simulate_ops import _device_and_stream
    programs, names, values = _common(programs, symbol_names, symbol_values)
    pauli_sums = as_string_tensor(pauli_sums, 2, "pauli_sums")
    grads = as_values(prev_grad, "downstream_grads")
    _, stream = _device_and_stream()
    key = ("adjoint", plan_cache.digest(programs, pauli_sums), tuple(names), programs.shape,
           pauli_sums.shape, grads.shape, derivative, with_expectation, stream)
    box = {}

    def build():
        plan, box["shape"] = plan_adj_grad(programs, names, values, pauli_sums, grads,
                                           derivative, with_expectation)
        return plan

    def rebind(plan):
        # The symbolic matrices are re-evaluated on the device at upload.
        plan.stage_values(values)
        plan.rebind_downstream(grads)

    def launch(plan):
        if "shape" not in box:
            B, O, S = programs.shape[0], pauli_sums.shape[1], len(names)
            box["shape"] = (B * (O + S),) if with_expectation else (B, S)
        return _run(plan, box["shape"], _torch().float32)

    return plan_cache.run(key, build, rebind, launch)


def expectation_and_gradient(programs, symbol_names, symbol_values, pauli_sums, prev_grad,
                             derivative="tfq"):
    """Expectations ``[batch, n_ops]`` and adjoint gradients ``[batch, n_symbols]``
    from one persistent launch (the forward state is computed once)."""
    flat = _cached(programs, symbol_names, symbol_values, pauli_sums, prev_grad, derivative, True)
    B = len(programs)
    O = flat.numel() // max(B, 1) - len(list(symbol_names)) if B else 0
    exp = flat[:B * O].reshape(B, O)
    grad = flat[B * O:].reshape(B, -1)
    return exp, grad


def tfq_adj_grad(programs, symbol_names, symbol_values, pauli_sums, prev_grad,
                 derivative="tfq"):
    return _cached(programs, symbol_names, symbol_values, pauli_sums, prev_grad, derivative, False)
