"""quantum_amd: TFQ's state-vector circuit-simulation op family on MI355X.

Public surface (TFQ names):
  convert_to_tensor, from circuits / Pauli sums to serialized string tensors
  get_expectation_op, get_sampled_expectation_op, get_sampling_op, get_state_op
  math.inner_product, math.fidelity (differentiable), math.inner_product_grad
  layers.Expectation, layers.SampledExpectation, layers.Sample, layers.State
  differentiators.Adjoint, differentiators.CentralDifference, ...
  tfq_simulate_ops.*, tfq_adj_grad_op.tfq_adj_grad (raw ops)

The ops run in libtfqx.so (host C++ planner + gfx950 HIP kernels); build it
with ``python -m quantum_amd.build_lib``.
"""

import importlib

from . import circuits
from .serialize import convert_to_tensor

__version__ = "0.1.0"

_LAZY = {
    "layers": "quantum_amd.layers",
    "differentiators": "quantum_amd.differentiators",
    "tfq_simulate_ops": "quantum_amd.tfq_simulate_ops",
    "tfq_adj_grad_op": "quantum_amd.tfq_adj_grad_op",
    "math": "quantum_amd.math_ops",
    "math_ops": "quantum_amd.math_ops",
    "circuit_execution_ops": "quantum_amd.circuit_execution_ops",
    "surrogate": "quantum_amd.surrogate",
    "workloads": "quantum_amd.workloads",
}
_OPS = ("get_expectation_op", "get_sampled_expectation_op", "get_sampling_op", "get_state_op",
        "set_quantum_concurrent_op_mode", "get_quantum_concurrent_op_mode")


def __getattr__(name):
    if name in _LAZY:
        return importlib.import_module(_LAZY[name])
    if name in _OPS:
        return getattr(importlib.import_module("quantum_amd.circuit_execution_ops"), name)
    raise AttributeError(name)
