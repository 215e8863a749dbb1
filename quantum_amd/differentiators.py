"""Differentiators: the custom-gradient slot of the expectation ops.

Mirrors ``python/differentiators/differentiator.py:58-187`` (the
``generate_differentiable_op`` contract and its checks), ``adjoint.py:22-110``
(Adjoint -> ``tfq_adj_grad``) and the finite-difference members of
``linear_combination.py`` (``ForwardDifference``, ``CentralDifference``).
``tf.custom_gradient`` becomes a ``torch.autograd.Function``: the forward runs
the wrapped op, the backward calls ``differentiate_analytic`` /
``differentiate_sampled`` with the upstream gradient.  ``ParameterShift``
(``parameter_shift.py``) builds its shifted circuits with the native
rewriting ops of ``ps_util_ops``.
"""

import inspect

import numpy as np

from . import ps_util_ops
from . import tfq_adj_grad_op
from .tfq_simulate_ops import as_names, as_string_tensor, as_values

PARAMETER_IMPURITY_NAME = "_impurity_for_param_shift"


def _torch():
    import torch
    return torch


def catch_empty_inputs(fn):
    """Return zero gradients when there are no programs, symbols or values
    (``differentiator.py:23-46``)."""
    def wrapper(self, programs, symbol_names, symbol_values, *args, **kwargs):
        torch = _torch()
        vals = symbol_values
        n_prog = len(programs)
        n_sym = len(list(symbol_names))
        if n_prog == 0 or n_sym == 0 or (hasattr(vals, "numel") and vals.numel() == 0):
            dev = vals.device if hasattr(vals, "device") else None
            return torch.zeros((n_prog, n_sym), dtype=torch.float32, device=dev)
        return fn(self, programs, symbol_names, symbol_values, *args, **kwargs)
    wrapper.__name__ = fn.__name__
    wrapper.__doc__ = fn.__doc__
    return wrapper


class Differentiator:
    """Base class: attach ``self`` as the gradient of an expectation op."""

    def refresh(self):
        if hasattr(self, "expectation_op"):
            del self.expectation_op
        return self

    def generate_differentiable_op(self, *, sampled_op=None, analytic_op=None):
        if hasattr(self, "expectation_op"):
            raise TypeError("This differentiator is already used for other op by calling "
                            "generate_differentiable_op before. You need to call `refresh()` "
                            "to reuse this differentiator with another op.")
        if sampled_op is None and analytic_op is None:
            raise ValueError("generate_differentiable_op requires a sample based expectation op "
                             "to be provided with arg 'sampled_op', or an analytically calculated "
                             "expectation op to be provided with arg 'analytic_op'.")
        if sampled_op is not None and analytic_op is not None:
            raise ValueError("generate_differentiable_op was given both a sampled_op and "
                             "analytic_op. Please provide analytic_op if the operation you wish "
                             "to make differentiable is analytical. Otherwise provide sampled_op "
                             "if the operation you want to make differentiable is sample based.")
        if not callable(sampled_op) and not callable(analytic_op):
            raise TypeError("Provided arguments must be callable ops.")
        if analytic_op is not None:
            sig = inspect.signature(analytic_op).parameters
            for key in ("programs", "symbol_names", "symbol_values", "pauli_sums"):
                if not any(key in s for s in sig):
                    raise ValueError(f"unexpected signature for analytic_op. Given arg: {key}.")
            if "num_samples" in sig:
                raise ValueError("found num_samples in analytic_op. Please ensure that you are "
                                 "providing an analytical expectation op in the analytic_op arg.")
        if sampled_op is not None:
            sig = inspect.signature(sampled_op).parameters
            for key in ("programs", "symbol_names", "symbol_values", "pauli_sums", "num_samples"):
                if not any(key in s for s in sig):
                    raise ValueError(f"unexpected signature for sampled_op. Given arg: {key}.")
        self.expectation_op = analytic_op if analytic_op is not None else sampled_op
        diff = self
        torch = _torch()

        class _Analytic(torch.autograd.Function):
            @staticmethod
            def forward(ctx, symbol_values, programs, symbol_names, pauli_sums):
                out = analytic_op(programs, symbol_names, symbol_values.detach(), pauli_sums)
                ctx.save_for_backward(symbol_values, out)
                ctx.inputs = (programs, symbol_names, pauli_sums)
                return out

            @staticmethod
            def backward(ctx, grad):
                values, fwd = ctx.saved_tensors
                programs, names, psums = ctx.inputs
                g = diff.differentiate_analytic(programs, names, values, psums, fwd, grad)
                return g.to(values.device, values.dtype), None, None, None

        class _Sampled(torch.autograd.Function):
            @staticmethod
            def forward(ctx, symbol_values, programs, symbol_names, pauli_sums, num_samples):
                out = sampled_op(programs, symbol_names, symbol_values.detach(), pauli_sums,
                                 num_samples)
                ctx.save_for_backward(symbol_values, out)
                ctx.inputs = (programs, symbol_names, pauli_sums, num_samples)
                return out

            @staticmethod
            def backward(ctx, grad):
                values, fwd = ctx.saved_tensors
                programs, names, psums, ns = ctx.inputs
                g = diff.differentiate_sampled(programs, names, values, psums, ns, fwd, grad)
                return g.to(values.device, values.dtype), None, None, None, None

        def _values_tensor(v):
            if isinstance(v, torch.Tensor):
                return v if v.dtype == torch.float32 else v.float()
            return torch.as_tensor(np.asarray(v, dtype=np.float32))

        if analytic_op is not None:
            def op_wrapper_analytic(programs, symbol_names, symbol_values, pauli_sums):
                return _Analytic.apply(_values_tensor(symbol_values), programs, symbol_names,
                                       pauli_sums)
            return op_wrapper_analytic

        def op_wrapper_sampled(programs, symbol_names, symbol_values, pauli_sums, num_samples):
            return _Sampled.apply(_values_tensor(symbol_values), programs, symbol_names,
                                  pauli_sums, num_samples)
        return op_wrapper_sampled

    def differentiate_analytic(self, programs, symbol_names, symbol_values, pauli_sums,
                               forward_pass_vals, grad):
        raise NotImplementedError

    def differentiate_sampled(self, programs, symbol_names, symbol_values, pauli_sums,
                              num_samples, forward_pass_vals, grad):
        raise NotImplementedError


class Adjoint(Differentiator):
    """Adjoint-method gradients (``adjoint.py:22-110``) from the HIP adjoint
    kernel.  ``derivative="tfq"`` reproduces TFQ's finite-difference gradient
    gates (eps = 5e-3); ``"analytic"`` differentiates the gate matrices exactly."""

    def __init__(self, derivative="tfq"):
        self.derivative = derivative

    def generate_differentiable_op(self, *, sampled_op=None, analytic_op=None):
        if sampled_op is not None:
            raise ValueError("sample base backends are not supported by the Adjoint method, "
                             "please use analytic expectation or choose another differentiator.")
        return super().generate_differentiable_op(analytic_op=analytic_op)

    def get_gradient_circuits(self, programs, symbol_names, symbol_values):
        raise NotImplementedError("Adjoint differentiator cannot run on a real QPU, therefore it "
                                  "has no accessible gradient circuits.")

    @catch_empty_inputs
    def differentiate_analytic(self, programs, symbol_names, symbol_values, pauli_sums,
                               forward_pass_vals, grad):
        return tfq_adj_grad_op.tfq_adj_grad(programs, symbol_names, symbol_values, pauli_sums,
                                            grad, derivative=self.derivative)

    def differentiate_sampled(self, programs, symbol_names, symbol_values, pauli_sums,
                              num_samples, forward_pass_vals, grad):
        raise NotImplementedError("Adjoint state methods are not supported in sample based "
                                  "settings. Please use analytic expectation calculation or a "
                                  "different differentiator.")


class LinearCombination(Differentiator):
    """d f/d v_s ~= sum_k w_k f(v + p_k e_s) (``linear_combination.py:25-250``),
    evaluated with one batched op call over all shifted parameter sets."""

    def __init__(self, weights, perturbations):
        if len(weights) != len(perturbations):
            raise ValueError("weights and perturbations must have the same length.")
        if len(set(perturbations)) != len(perturbations):
            raise ValueError("perturbations must be unique.")
        self.weights = [float(w) for w in weights]
        self.perturbations = [float(p) for p in perturbations]

    def _shifted(self, programs, symbol_names, symbol_values, pauli_sums, grad, evaluate):
        torch = _torch()
        progs = as_string_tensor(programs, 1, "programs")
        names = as_names(symbol_names)
        vals = as_values(symbol_values)
        psums = as_string_tensor(pauli_sums, 2, "pauli_sums")
        B, S = vals.shape
        K = len(self.perturbations)
        shifted = np.repeat(vals[None, None], K, axis=0).repeat(S, axis=1)  # [K, S, B, S]
        for k, p in enumerate(self.perturbations):
            for s in range(S):
                shifted[k, s, :, s] += p
        rows = shifted.reshape(K * S * B, S)
        out = evaluate(np.tile(progs, K * S), names, rows, np.tile(psums, (K * S, 1)),
                       K * S)
        out = out.reshape(K, S, B, -1).float()
        w = torch.tensor(self.weights, device=out.device).view(K, 1, 1, 1)
        deriv = (w * out).sum(0)  # [S, B, O]
        g = grad.to(out.device).float()  # [B, O]
        return torch.einsum("sbo,bo->bs", deriv, g)

    @catch_empty_inputs
    def differentiate_analytic(self, programs, symbol_names, symbol_values, pauli_sums,
                               forward_pass_vals, grad):
        op = self.expectation_op
        return self._shifted(programs, symbol_names, symbol_values, pauli_sums, grad,
                             lambda p, n, v, s, reps: op(p, n, v, s))

    @catch_empty_inputs
    def differentiate_sampled(self, programs, symbol_names, symbol_values, pauli_sums,
                              num_samples, forward_pass_vals, grad):
        op = self.expectation_op
        ns = as_values(num_samples, "num_samples", 2, np.int32)
        return self._shifted(programs, symbol_names, symbol_values, pauli_sums, grad,
                             lambda p, n, v, s, reps: op(p, n, v, s, np.tile(ns, (reps, 1))))


class ForwardDifference(LinearCombination):
    """First- or second-order forward difference (``linear_combination.py:156-250``)."""

    def __init__(self, error_order=1, grid_spacing=0.001):
        if error_order == 1:
            w, p = [-1.0, 1.0], [0.0, 1.0]
        elif error_order == 2:
            w, p = [-1.5, 2.0, -0.5], [0.0, 1.0, 2.0]
        else:
            raise ValueError("error_order must be 1 or 2.")
        h = float(grid_spacing)
        super().__init__([x / h for x in w], [x * h for x in p])


class CentralDifference(LinearCombination):
    """Second- or fourth-order central difference (``linear_combination.py:251-315``)."""

    def __init__(self, error_order=2, grid_spacing=0.001):
        if error_order == 2:
            w, p = [-0.5, 0.5], [-1.0, 1.0]
        elif error_order == 4:
            w, p = [1 / 12, -2 / 3, 2 / 3, -1 / 12], [-2.0, -1.0, 1.0, 2.0]
        else:
            raise ValueError("error_order must be 2 or 4.")
        h = float(grid_spacing)
        super().__init__([x / h for x in w], [x * h for x in p])


class ParameterShift(Differentiator):
    """Parameter-shift rule (``parameter_shift.py:28-150``,
    ``parameter_shift_util.py:25-110``).

    Programs are decomposed into single-symbol eigen gates with two
    eigenvalues (``tfq_ps_decompose``); each occurrence g of symbol s (with
    exponent scalar c_g) is renamed to an impurity symbol
    (``tfq_ps_symbol_replace``) and evaluated at v_s +- 0.5 / c_g.  Then
    d<O>/dv_s = sum_g (pi c_g / 2) (<O>(v_s + 0.5/c_g) - <O>(v_s - 0.5/c_g)).
    All shifted circuits of a batch go to the expectation op in one call.
    """

    def get_gradient_circuits(self, programs, symbol_names, symbol_values):
        progs = as_string_tensor(programs, 1, "programs")
        names = as_names(symbol_names)
        vals = as_values(symbol_values)
        B, S = vals.shape
        decomposed = ps_util_ops.tfq_ps_decompose(progs)
        new = ps_util_ops.tfq_ps_symbol_replace(decomposed, names,
                                                [PARAMETER_IMPURITY_NAME] * S)  # [B, S, G]
        coeff = ps_util_ops.tfq_ps_weights_from_symbols(decomposed, names)  # [B, S, G]
        if new.shape[2] != coeff.shape[2]:
            raise ValueError("parameter-shift programs and weights do not match: "
                             f"{new.shape[2]} != {coeff.shape[2]} occurrences.")
        G = new.shape[2]
        f32 = np.float32
        # weights = c * pi * 0.5 * 0.5 * delta_eig, shifts = v +- (1 / delta_eig) / c
        # (divide_no_nan), all in float32 as TensorFlow evaluates them.
        w_plus = (((coeff * f32(np.pi)) * f32(0.5)) * f32(0.5)) * f32(2.0)
        weights = np.concatenate([w_plus[..., None], -w_plus[..., None]], axis=-1)  # [B,S,G,2]
        with np.errstate(divide="ignore", invalid="ignore"):
            s_plus = np.where(coeff != 0, f32(0.5) / coeff, f32(0.0)).astype(f32)
        shifts = vals[:, :, None, None] + np.concatenate([s_plus[..., None], -s_plus[..., None]],
                                                         axis=-1)
        M = S * G * 2
        batch_programs = np.repeat(new[..., None], 2, axis=-1).reshape(B, M)
        batch_values = np.concatenate(
            [np.repeat(vals[:, None, :], M, axis=1), shifts.reshape(B, M, 1)], axis=2).astype(f32)
        batch_weights = weights.reshape(B, S, G * 2)
        batch_mapper = np.tile(np.arange(M).reshape(1, S, G * 2), (B, 1, 1))
        new_names = np.array(list(names) + [PARAMETER_IMPURITY_NAME], dtype=object)
        return batch_programs, new_names, batch_values, batch_weights, batch_mapper

    def _jacobian_contract(self, programs, symbol_names, symbol_values, pauli_sums, grad,
                           evaluate):
        torch = _torch()
        bp, nn, bv, bw, bm = self.get_gradient_circuits(programs, symbol_names, symbol_values)
        psums = as_string_tensor(pauli_sums, 2, "pauli_sums")
        B, M = bp.shape
        S = bw.shape[1]
        if M == 0:
            return torch.zeros((B, S), dtype=torch.float32, device=grad.device)
        exps = evaluate(bp.reshape(B * M), nn, bv.reshape(B * M, -1), np.repeat(psums, M, axis=0),
                        M)
        exps = exps.reshape(B, M, -1).float()
        idx = torch.as_tensor(bm, device=exps.device, dtype=torch.long)  # [B, S, K]
        w = torch.as_tensor(bw, device=exps.device)  # [B, S, K]
        picked = torch.gather(exps.unsqueeze(1).expand(B, S, M, exps.shape[-1]), 2,
                              idx.unsqueeze(-1).expand(B, S, idx.shape[2], exps.shape[-1]))
        jac = (w.unsqueeze(-1) * picked).sum(2)  # [B, S, O]
        g = grad.to(exps.device).float()
        return (jac * g.unsqueeze(1)).sum(-1)

    @catch_empty_inputs
    def differentiate_analytic(self, programs, symbol_names, symbol_values, pauli_sums,
                               forward_pass_vals, grad):
        op = self.expectation_op
        return self._jacobian_contract(programs, symbol_names, symbol_values, pauli_sums, grad,
                                       lambda p, n, v, s, reps: op(p, n, v, s))

    @catch_empty_inputs
    def differentiate_sampled(self, programs, symbol_names, symbol_values, pauli_sums,
                              num_samples, forward_pass_vals, grad):
        op = self.expectation_op
        ns = as_values(num_samples, "num_samples", 2, np.int32)
        return self._jacobian_contract(programs, symbol_names, symbol_values, pauli_sums, grad,
                                       lambda p, n, v, s, reps: op(p, n, v, s,
                                                                   np.repeat(ns, reps, axis=0)))
