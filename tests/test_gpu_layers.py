"""Layers and differentiators on the GPU: autograd through the adjoint kernel
matches the oracle; finite differences agree with the adjoint; parameter
shift (computed from the oracle's forward values) agrees with the adjoint."""

import math
import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
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


def test_expectation_layer_autograd_matches_oracle():
    import torch
    from quantum_amd import layers
    circuit, names, q = W.hea_ring(5, 2)
    ham = W.tfim(q, 0.7)
    vals = torch.tensor(np.random.default_rng(0).uniform(-3, 3, (3, len(names))), dtype=torch.float32,
                        requires_grad=True)
    layer = layers.Expectation()
    _say("Expectation layer forward/backward")
    out = layer(circuit, symbol_names=names, symbol_values=vals, operators=ham)
    out.sum().backward()
    progs = S.convert_to_tensor([circuit] * 3)
    ps = S.convert_to_tensor([[ham]] * 3)
    ref_e = O.simulate_expectation(list(progs), names, vals.detach().numpy(), ps)
    ref_g = O.adjoint_gradient(list(progs), names, vals.detach().numpy(), ps, np.ones((3, 1)))
    np.testing.assert_allclose(out.detach().cpu().numpy(), ref_e, atol=1e-5)
    np.testing.assert_allclose(vals.grad.numpy(), ref_g, atol=1e-5)


def test_learnable_weights_and_training_step():
    import torch
    from quantum_amd import layers
    circuit, names, q = W.hea_ring(4, 2)
    ham = W.tfim(q, 1.0)
    layer = layers.Expectation()
    _say("learnable weights")
    e0 = layer(circuit, symbol_names=names, operators=ham)
    assert e0.shape == (1, 1)
    opt = torch.optim.Adam(layer.parameters(), lr=0.1)
    first = None
    for _ in range(30):
        opt.zero_grad()
        e = layer(circuit, symbol_names=names, operators=ham).sum()
        e.backward()
        opt.step()
        first = float(e) if first is None else first
    assert float(e) < first - 0.5


def test_central_difference_agrees_with_adjoint():
    import torch
    from quantum_amd import circuit_execution_ops as ceo
    from quantum_amd import differentiators as diff
    rng = np.random.default_rng(2)
    q = C.LineQubit.range(4)
    names = ["a", "b", "c"]
    circ = W.random_circuit(rng, q, 20, symbols=names, p_symbol=0.7, p_control=0.2)
    progs = S.convert_to_tensor([circ, circ])
    ps = S.convert_to_tensor([[W.random_pauli_sum(rng, q, 3)]] * 2)
    vals = torch.tensor(rng.uniform(-1, 1, (2, 3)), dtype=torch.float32)
    grads = {}
    for name, d in (("adjoint", diff.Adjoint(derivative="analytic")),
                    ("cd", diff.CentralDifference(error_order=4, grid_spacing=0.01))):
        op = d.generate_differentiable_op(analytic_op=ceo.get_expectation_op())
        v = vals.clone().requires_grad_(True)
        _say(f"differentiator {name}")
        op(progs, names, v, ps).sum().backward()
        grads[name] = v.grad.numpy()
    np.testing.assert_allclose(grads["cd"], grads["adjoint"], atol=2e-3)


def test_parameter_shift_agrees_with_adjoint():
    """Rotation gates: d<O>/dv = [f(v + pi/2) - f(v - pi/2)] / 2 per symbol."""
    from quantum_amd import tfq_adj_grad_op as adj
    circuit, names, q = W.hea_ring(6, 3)
    ham = W.tfim(q, 0.7)
    vals = np.random.default_rng(4).uniform(-math.pi, math.pi, (2, len(names))).astype(np.float32)
    progs = S.convert_to_tensor([circuit] * 2)
    ps = S.convert_to_tensor([[ham]] * 2)
    _say("parameter shift check")
    g = adj.tfq_adj_grad(progs, names, vals, ps, np.ones((2, 1)), derivative="analytic")
    f = O.expectation_function(list(progs), names, ps)
    shift = np.zeros_like(vals)
    for s in (0, 7, 20, len(names) - 1):
        vp = vals.astype(np.float64).copy()
        vm = vp.copy()
        vp[:, s] += math.pi / 2
        vm[:, s] -= math.pi / 2
        ps_grad = (f(vp) - f(vm))[:, 0] / 2
        np.testing.assert_allclose(g.cpu().numpy()[:, s], ps_grad, atol=1e-5)


def test_sample_and_state_layers():
    from quantum_amd import layers
    q = C.LineQubit.range(3)
    circs = [C.Circuit(C.X(q[0]), C.I(q[1]), C.I(q[2])), C.Circuit(C.H(q[0]))]
    _say("sample/state layers")
    smp = layers.Sample(seed=3)(circs, repetitions=10)
    assert smp[0].shape == (10, 3) and (smp[0].numpy() == [1, 0, 0]).all()
    assert smp[1].shape == (10, 1)
    st = layers.State()(circs)
    assert st[0].shape[0] == 8 and st[1].shape[0] == 2
    np.testing.assert_allclose(st[1].numpy(), [2 ** -0.5, 2 ** -0.5], atol=1e-6)


def test_sampled_expectation_layer():
    import torch
    from quantum_amd import layers
    q = C.LineQubit(0)
    circ = C.Circuit(C.X(q) ** C.Symbol("a"))
    layer = layers.SampledExpectation(seed=11)
    v = torch.tensor([[0.3]], requires_grad=True)
    _say("sampled expectation layer")
    out = layer(circ, symbol_names=["a"], symbol_values=v, operators=C.pauli_z(q),
                repetitions=200000)
    assert abs(float(out) - math.cos(math.pi * 0.3)) < 1e-2
