"""``tfq.math.inner_product`` and ``tfq.math.fidelity``
(``core/ops/math_ops/inner_product_op.py:71-151``, ``fidelity_op.py:20-96``).

``inner_product(programs, symbol_names, symbol_values, other_programs)`` returns
``<psi(programs[i], values[i]) | psi(other_programs[i][j])>`` as complex64
``[batch, n_other]``; ``other_programs`` must be free of symbols and act on
exactly the qubits of ``programs[i]``.  ``fidelity`` is ``|inner_product|^2``.
The gradient of the inner product w.r.t. symbol values (TFQ's
``tfq_inner_product_grad``) is not part of this op family.
"""

from . import _lib
from ._lib import InvalidArgumentError, Plan, StringTensor
from .tfq_simulate_ops import _common, _plan_args, _run, _torch, as_string_tensor


def inner_product(programs, symbol_names, symbol_values, other_programs):
    programs, names, values = _common(programs, symbol_names, symbol_values)
    other = as_string_tensor(other_programs, 2, "other_programs")
    if other.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"programs and other_programs batch dimension do not match. Foud: "
            f"{programs.shape[0]} and {other.shape[0]}")
    P, N, args = _plan_args(programs, names, values)
    Ot = StringTensor(other)
    plan = Plan(_lib.lib().tfqx_plan_inner_product, *args, Ot.ptr, Ot.off, other.shape[1])
    return _run(plan, other.shape, _torch().complex64)


def fidelity(programs, symbol_names, symbol_values, other_programs):
    """``|<psi|phi>|^2`` as float32 (``fidelity_op.py:84-96``)."""
    ip = inner_product(programs, symbol_names, symbol_values, other_programs)
    return (ip.real ** 2 + ip.imag ** 2).to(_torch().float32)
