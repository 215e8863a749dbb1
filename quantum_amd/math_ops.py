"""``tfq.math.inner_product`` and ``tfq.math.fidelity``
(``core/ops/math_ops/inner_product_op.py:71-151``, ``fidelity_op.py:20-96``),
and the inner product's gradient op ``inner_product_grad``
(``inner_product_op.py:23-68``; ``tfq_inner_product_grad.cc:46-501``).

``inner_product(programs, symbol_names, symbol_values, other_programs)`` returns
``<psi(programs[i], values[i]) | psi(other_programs[i][j])>`` as complex64
``[batch, n_other]``; ``other_programs`` must be free of symbols and act on
exactly the qubits of ``programs[i]``.  ``fidelity`` is ``|inner_product|^2``.

When ``symbol_values`` is a torch tensor that requires grad, ``inner_product``
(and so ``fidelity``) is differentiable with respect to it.  The backward
uses the adjoint gradient op with the real and the imaginary part of the
incoming complex gradient G: d L / d values[i, s] = sum_j Re(conj(G[i, j]) *
<d_s psi_i | phi_ij>), PyTorch's convention for a real loss.  TFQ's own
custom gradient passes only Re(G) through a float32 cast
(``inner_product_op.py:65-68,134-147``); ``inner_product_grad`` below is that
op-level function, unchanged.
"""

import numpy as np

from . import _lib
from ._lib import InvalidArgumentError, Plan, StringTensor
from .tfq_simulate_ops import _common, _plan_args, _run, _torch, as_string_tensor, as_values

_MODES = {"tfq": 0, "analytic": 1}


def _checked_others(programs, other_programs):
    other = as_string_tensor(other_programs, 2, "other_programs")
    if other.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"programs and other_programs batch dimension do not match. Foud: "
            f"{programs.shape[0]} and {other.shape[0]}")
    return other


def _inner_product_raw(programs, names, values, other):
    P, N, args = _plan_args(programs, names, values)
    Ot = StringTensor(other)
    plan = Plan(_lib.lib().tfqx_plan_inner_product, *args, Ot.ptr, Ot.off, other.shape[1])
    return _run(plan, other.shape, _torch().complex64)


def _grad_raw(programs, names, values, other, prev_grad, derivative):
    """The op: sum_j prev_grad[i, j] <d_s psi_i | phi_ij>, complex64 [B, S]."""
    if derivative not in _MODES:
        raise ValueError(f"derivative must be one of {sorted(_MODES)}")
    g = np.ascontiguousarray(prev_grad, dtype=np.float32)
    P, N, args = _plan_args(programs, names, values)
    Ot = StringTensor(other)
    plan = Plan(_lib.lib().tfqx_plan_inner_product_grad, *args, Ot.ptr, Ot.off, other.shape[1],
                g.ctypes.data if g.size else None, _MODES[derivative])
    return _run(plan, (programs.shape[0], len(names)), _torch().complex64)


class _InnerProduct:
    """torch.autograd.Function for inner products differentiable in the symbol values."""

    @staticmethod
    def make():
        torch = _torch()

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, values_t, programs, names, other):
                values = values_t.detach().to("cpu").numpy().astype(np.float32)
                ctx.saved = (programs, names, values, other)
                ctx.values_device, ctx.values_dtype = values_t.device, values_t.dtype
                return _inner_product_raw(programs, names, values, other)

            @staticmethod
            def backward(ctx, grad):
                programs, names, values, other = ctx.saved
                if len(names) == 0:
                    return torch.zeros(values.shape, dtype=ctx.values_dtype,
                                       device=ctx.values_device), None, None, None
                G = grad.detach().to("cpu").numpy()
                t_re = _grad_raw(programs, names, values, other, G.real, "tfq")
                t_im = _grad_raw(programs, names, values, other, G.imag, "tfq")
                out = (t_re.real + t_im.imag).to(ctx.values_dtype).to(ctx.values_device)
                return out, None, None, None

        return Fn


_FN = None


def inner_product(programs, symbol_names, symbol_values, other_programs):
    global _FN
    torch = _torch()
    programs, names, values = _common(programs, symbol_names, symbol_values)
    other = _checked_others(programs, other_programs)
    if isinstance(symbol_values, torch.Tensor) and symbol_values.requires_grad:
        if _FN is None:
            _FN = _InnerProduct.make()
        return _FN.apply(symbol_values, programs, names, other)
    return _inner_product_raw(programs, names, values, other)


def fidelity(programs, symbol_names, symbol_values, other_programs):
    """``|<psi|phi>|^2`` as float32 (``fidelity_op.py:84-96``)."""
    ip = inner_product(programs, symbol_names, symbol_values, other_programs)
    return (ip.real ** 2 + ip.imag ** 2).to(_torch().float32)


def inner_product_grad(programs, symbol_names, symbol_values, other_programs, prev_grad,
                       derivative="tfq"):
    """``_inner_product_grad`` (``inner_product_op.py:23-68``): complex64
    ``[batch, n_symbols]`` equal to conj(sum_j prev_grad[i, j] *
    <d psi_i / d symbol | phi_ij>), with prev_grad cast to float32 (a complex
    prev_grad keeps its real part, as TensorFlow's cast does).

    derivative "tfq": TFQ's float32 central-difference gate derivative
    (epsilon 5e-3, ``adj_util.cc:175-302``); "analytic": the exact one.
    """
    programs, names, values = _common(programs, symbol_names, symbol_values)
    other = _checked_others(programs, other_programs)
    g = as_values(prev_grad, "prev_grad")
    if g.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"Number of gradients and circuits do not match. Got {g.shape[0]} gradients and "
            f"{programs.shape[0]} circuits.")
    if g.shape[1] != other.shape[1]:
        raise InvalidArgumentError(
            f"Number of gradients and other_programs do not match. Got {g.shape[1]} gradient "
            f"entries and {other.shape[1]} other programs.")
    return _torch().conj_physical(_grad_raw(programs, names, values, other, g, derivative))
