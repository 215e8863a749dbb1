"""``ParameterShift`` differentiator (``parameter_shift.py``) on the GPU.

* The two-term rule is exact for the uncontrolled single-symbol eigen gates
  the decomposition leaves (two eigenvalues, one apart), so on random
  circuits with every gate family, scalars and shared symbols (controls only
  on symbol-free gates) its gradients equal the exact adjoint gradient
  (``derivative="analytic"``) to float32 accuracy.
* A controlled gate with a global shift has three eigenphases; there the
  reference's rule is not the derivative.  The test pins what it computes:
  (1/2) [f(v + pi/2) - f(v - pi/2)] for a controlled rx(v), evaluated with the
  complex128 oracle.
"""

import sys

import numpy as np
import pytest

from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _batch(seed, n, batch, depth):
    rng = np.random.default_rng(seed)
    names = ["a", "b", "c"]
    q = C.LineQubit.range(n)
    circuits = []
    for _ in range(batch):
        # Random PhasedISwaps are left out: one with a symbolic phase and a
        # numeric exponent decomposes (as in the reference) into parts with an
        # empty symbol, which the weights op rejects (test below).
        fams = [f for f in W.ALL_GATE_FAMILIES if f != "PISP"]
        c = W.random_circuit(rng, q, depth, symbols=names, p_symbol=0.6, p_control=0.0,
                             families=fams)
        c = c + C.Circuit([C.CZ(q[0], q[1]).controlled_by(q[2]), C.H(q[2]).controlled_by(q[0])])
        t = C.Symbol("a")
        extra = C.Circuit([C.PhasedISwapPowGate(phase_exponent=t * 0.7,
                                                exponent=C.Symbol("b") * 1.3).on(q[0], q[1]),
                           C.FSimGate(theta=C.Symbol("c") * 0.9, phi=t * 1.1).on(q[1], q[2]),
                           C.PhasedXPowGate(phase_exponent=C.Symbol("b") * 0.4,
                                            exponent=C.Symbol("c")).on(q[0])])
        circuits.append(c + extra)
    sums = [[W.random_pauli_sum(rng, q, 3), W.random_pauli_sum(rng, q, 2)] for _ in range(batch)]
    vals = rng.uniform(-2, 2, size=(batch, 3)).astype(np.float32)
    return S.convert_to_tensor(circuits), names, vals, S.convert_to_tensor(sums)


@pytest.mark.parametrize("n,batch", [(3, 4), (6, 3)])
def test_parameter_shift_matches_exact_gradient(n, batch):
    import torch
    from quantum_amd import circuit_execution_ops as ceo
    from quantum_amd import differentiators as diff
    from quantum_amd import tfq_adj_grad_op as A
    progs, names, vals, ps = _batch(700 + n, n, batch, 3 * n)
    op = diff.ParameterShift().generate_differentiable_op(analytic_op=ceo.get_expectation_op())
    v = torch.tensor(vals, device="cuda", requires_grad=True)
    _say(f"parameter shift n={n} batch={batch}")
    out = op(progs, names, v, ps)
    g = torch.tensor(np.random.default_rng(1).normal(size=out.shape), dtype=torch.float32,
                     device="cuda")
    (out * g).sum().backward()
    got = v.grad.cpu().numpy()
    ref = A.tfq_adj_grad(progs, names, vals, ps, g.cpu().numpy(), derivative="analytic").cpu().numpy()
    np.testing.assert_allclose(got, ref, atol=1e-4 * max(1.0, np.abs(ref).max()))


def test_parameter_shift_sampled_op_and_empty_symbols():
    import torch
    from quantum_amd import circuit_execution_ops as ceo
    from quantum_amd import differentiators as diff
    from quantum_amd import tfq_adj_grad_op as A
    progs, names, vals, ps = _batch(801, 3, 2, 6)
    op = diff.ParameterShift().generate_differentiable_op(
        sampled_op=ceo.get_sampled_expectation_op(seed=4))
    v = torch.tensor(vals, device="cuda", requires_grad=True)
    ns = np.full((2, 2), 40000, np.int32)
    _say("parameter shift on the sampled op")
    out = op(progs, names, v, ps, ns)
    out.sum().backward()
    ref = A.tfq_adj_grad(progs, names, vals, ps, np.ones((2, 2), np.float32),
                         derivative="analytic").cpu().numpy()
    # Shot noise: sums of terms with |c| <= 2, 40000 shots per term.
    np.testing.assert_allclose(v.grad.cpu().numpy(), ref, atol=0.25)
    op2 = diff.ParameterShift().generate_differentiable_op(analytic_op=ceo.get_expectation_op())
    q = C.LineQubit.range(2)
    free = S.convert_to_tensor([C.Circuit([C.H(q[0]), C.CNOT(q[0], q[1])])] * 2)
    fps = S.convert_to_tensor([[C.PauliSum([C.pauli_z(q[1])])]] * 2)
    v2 = torch.zeros((2, 0), device="cuda", requires_grad=True)
    out2 = op2(free, [], v2, fps)
    out2.sum().backward()
    assert tuple(v2.grad.shape) == (2, 0)


def test_parameter_shift_rule_on_a_controlled_rotation():
    import math
    import torch
    from oracle import tfq_oracle as O
    from quantum_amd import circuit_execution_ops as ceo
    from quantum_amd import differentiators as diff
    q = C.LineQubit.range(2)
    v = C.Symbol("v")
    circ = C.Circuit([C.H(q[0]), C.ry(0.4).on(q[1]), C.rx(v).on(q[1]).controlled_by(q[0])])
    progs = S.convert_to_tensor([circ])
    ps = S.convert_to_tensor([[C.PauliSum([C.pauli_z(q[1]), C.pauli_x(q[0])])]])
    op = diff.ParameterShift().generate_differentiable_op(analytic_op=ceo.get_expectation_op())
    vals = np.array([[0.9]], np.float32)
    t = torch.tensor(vals, device="cuda", requires_grad=True)
    _say("parameter shift on a controlled rotation")
    op(progs, ["v"], t, ps).sum().backward()
    f = O.expectation_function(list(progs), ["v"], ps)
    rule = 0.5 * (f(vals + math.pi / 2) - f(vals - math.pi / 2))[0, 0]
    np.testing.assert_allclose(float(t.grad[0, 0]), rule, atol=1e-5)


def test_parameter_shift_rejects_a_phased_iswap_with_numeric_exponent():
    """The reference's decomposition gives the ISWAP part of such a gate an
    empty symbol (tfq_ps_decompose_op.cc:84-91), which
    tfq_ps_weights_from_symbols does not find among the symbols."""
    import torch
    from quantum_amd import circuit_execution_ops as ceo
    from quantum_amd import differentiators as diff
    from quantum_amd._lib import InvalidArgumentError
    q = C.LineQubit.range(2)
    circ = C.Circuit([C.PhasedISwapPowGate(phase_exponent=C.Symbol("p"), exponent=0.5).on(q[0], q[1])])
    op = diff.ParameterShift().generate_differentiable_op(analytic_op=ceo.get_expectation_op())
    t = torch.tensor([[0.3]], device="cuda", requires_grad=True)
    out = op(S.convert_to_tensor([circ]), ["p"], t,
             S.convert_to_tensor([[C.PauliSum([C.pauli_z(q[0])])]]))
    with pytest.raises(InvalidArgumentError, match="sympy.Symbol"):
        out.sum().backward()
