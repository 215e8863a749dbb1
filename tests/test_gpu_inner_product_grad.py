"""``tfq.math`` inner-product gradient on the GPU against the complex128 oracle.

* The op (``inner_product_grad``, the reference's ``_inner_product_grad``):
  conj(sum_j g_ij <d psi_i / d s | phi_ij>) for random batches of 2-13 qubits,
  TFQ's finite-difference gate derivative and the exact one, 1-3 paired
  programs per row, controls included; empty programs and symbols absent from
  a circuit give zeros.
* Circuits wider than 13 qubits (the streaming passes): 14 and 16 qubits, a
  batch mixing widths, and an empty program in such a batch.
* Input checks with the reference's messages (``inner_product_grad_test.py``).
* Autograd: ``fidelity`` and ``inner_product`` losses back-propagate to the
  symbol values as the oracle's central differences say.
"""

import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _case(seed, n, batch, K, n_sym=3, depth=None, p_control=0.2):
    rng = np.random.default_rng(seed)
    q = C.LineQubit.range(n)
    names = [f"s{i}" for i in range(n_sym)]
    # Identities keep every qubit in each circuit (paired programs must act on
    # exactly the reference circuit's qubits).
    circuits = [W.random_circuit(rng, q, depth or 3 * n, symbols=names, p_control=p_control)
                + C.Circuit([C.I(x) for x in q]) for _ in range(batch)]
    # Paired programs: symbol-free, on exactly the same qubits.
    others = [[W.random_circuit(rng, q, 2 * n) + C.Circuit([C.I(x) for x in q])
               for _ in range(K)] for _ in range(batch)]
    vals = rng.uniform(-2, 2, size=(batch, n_sym)).astype(np.float32)
    g = rng.normal(size=(batch, K)).astype(np.float32)
    progs = S.convert_to_tensor(circuits)
    oth = np.array([[S.serialize_circuit(c) for c in row] for row in others], dtype=object)
    return progs, names, vals, oth, g


@pytest.mark.parametrize("n,K", [(2, 1), (5, 3), (8, 2), (10, 1), (12, 2), (13, 1)])
@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_inner_product_grad_matches_oracle(n, K, mode):
    from quantum_amd import math_ops as M
    progs, names, vals, oth, g = _case(500 + n, n, 4, K)
    _say(f"inner product grad n={n} K={K} {mode}")
    out = M.inner_product_grad(progs, names, vals, oth, g, derivative=mode).cpu().numpy()
    ref = np.conj(O.inner_product_gradient(list(progs), names, vals, oth, g, mode=mode))
    assert out.dtype == np.complex64 and out.shape == (4, 3)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))


def test_inner_product_grad_empty_program_and_unused_symbol():
    from quantum_amd import math_ops as M
    q = C.LineQubit.range(3)
    a = C.Symbol("a")
    progs = S.convert_to_tensor([C.Circuit(), C.Circuit([(C.X ** a).on(q[0]), C.I(q[1]),
                                                         C.I(q[2])])])
    # As in inner_product_grad_test.py:346-371, the empty program is paired
    # with an empty program.
    # |000> as the second row's pair: d<psi|phi>/da = (-i pi / 2) e^{-i pi a} != 0.
    oth = np.array([[S.serialize_circuit(C.Circuit())],
                    [S.serialize_circuit(C.Circuit([C.I(x) for x in q]))]], dtype=object)
    vals = np.array([[0.3, 0.1], [0.3, 0.1]], np.float32)
    g = np.ones((2, 1), np.float32)
    _say("inner product grad: empty program, unused symbol")
    out = M.inner_product_grad(progs, ["a", "b"], vals, oth, g).cpu().numpy()
    ref = np.conj(O.inner_product_gradient(list(progs), ["a", "b"], vals, oth, g))
    np.testing.assert_allclose(out[0], 0.0)
    np.testing.assert_allclose(out[1, 1], 0.0)
    np.testing.assert_allclose(out, ref, atol=TOL)
    assert abs(out[1, 0]) > 1e-3


@pytest.mark.parametrize("n,K", [(14, 2), (16, 1)])
@pytest.mark.parametrize("mode", ["tfq", "analytic"])
def test_inner_product_grad_wide_matches_oracle(n, K, mode):
    from quantum_amd import math_ops as M
    progs, names, vals, oth, g = _case(600 + n, n, 2, K, depth=2 * n)
    _say(f"inner product grad (streaming) n={n} K={K} {mode}")
    out = M.inner_product_grad(progs, names, vals, oth, g, derivative=mode).cpu().numpy()
    ref = np.conj(O.inner_product_gradient(list(progs), names, vals, oth, g, mode=mode))
    assert np.abs(ref).max() > 1e-2
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))


def test_inner_product_grad_wide_mixed_batch_with_empty_program():
    from quantum_amd import math_ops as M
    rng = np.random.default_rng(88)
    names = ["a", "b"]
    circuits, others = [], []
    for n in (4, 15, 0, 9):
        if n == 0:
            circuits.append(C.Circuit())
            others.append([C.Circuit(), C.Circuit()])
            continue
        q = C.LineQubit.range(n)
        ident = C.Circuit([C.I(x) for x in q])
        circuits.append(W.random_circuit(rng, q, 2 * n, symbols=names, p_control=0.2) + ident)
        others.append([W.random_circuit(rng, q, n) + ident for _ in range(2)])
    progs = S.convert_to_tensor(circuits)
    oth = np.array([[S.serialize_circuit(c) for c in row] for row in others], dtype=object)
    vals = rng.uniform(-2, 2, size=(4, 2)).astype(np.float32)
    g = rng.normal(size=(4, 2)).astype(np.float32)
    _say("inner product grad (streaming): mixed widths, empty program")
    out = M.inner_product_grad(progs, names, vals, oth, g).cpu().numpy()
    ref = np.conj(O.inner_product_gradient(list(progs), names, vals, oth, g))
    np.testing.assert_allclose(out[2], 0.0)
    np.testing.assert_allclose(out, ref, atol=TOL * max(1, np.abs(ref).max()))


def test_inner_product_grad_input_checks():
    from quantum_amd import math_ops as M
    from quantum_amd._lib import InvalidArgumentError
    progs, names, vals, oth, g = _case(77, 4, 3, 2)
    with pytest.raises(InvalidArgumentError, match="do not match"):
        M.inner_product_grad(progs, names, vals, oth, g[:2])
    with pytest.raises(InvalidArgumentError, match="do not match"):
        M.inner_product_grad(progs, names, vals, oth, g[:, :1])
    with pytest.raises(InvalidArgumentError, match="positive integer"):
        M.inner_product_grad(progs, [], np.zeros((3, 0), np.float32), oth, g)
    sym_others = np.array([[p, p] for p in progs], dtype=object)
    with pytest.raises(InvalidArgumentError, match="Found symbols in other_programs"):
        M.inner_product_grad(progs, names, vals, sym_others, g)


def test_fidelity_and_inner_product_autograd():
    import torch
    from quantum_amd import math_ops as M
    progs, names, vals, oth, _ = _case(91, 6, 3, 2)
    dev = torch.device("cuda", 0)
    v = torch.tensor(vals, device=dev, requires_grad=True)
    _say("fidelity autograd")
    w = torch.tensor([[0.5, -1.0]] * 3, device=dev)
    loss = (M.fidelity(progs, names, v, oth) * w).sum()
    loss.backward()
    got = v.grad.cpu().numpy()

    def f(values):
        ip = O.inner_product(list(progs), names, values, oth)
        return float(np.sum(np.abs(ip) ** 2 * np.array([[0.5, -1.0]])))

    h = 1e-3
    cd = np.zeros_like(got)
    for i in range(3):
        for k in range(3):
            vp = vals.astype(np.float64).copy()
            vm = vp.copy()
            vp[i, k] += h
            vm[i, k] -= h
            cd[i, k] = (f(vp) - f(vm)) / (2 * h)
    # TFQ's FD gate derivative (5e-3) and float32 symbol values bound the match.
    np.testing.assert_allclose(got, cd, atol=5e-3)

    _say("inner product autograd (real and imaginary parts)")
    v2 = torch.tensor(vals, device=dev, requires_grad=True)
    ip = M.inner_product(progs, names, v2, oth)
    (ip.real.sum() + 2.0 * ip.imag.sum()).backward()
    got2 = v2.grad.cpu().numpy()

    def f2(values):
        z = O.inner_product(list(progs), names, values, oth)
        return float(np.sum(z.real) + 2.0 * np.sum(z.imag))

    cd2 = np.zeros_like(got2)
    for i in range(3):
        for k in range(3):
            vp = vals.astype(np.float64).copy()
            vm = vp.copy()
            vp[i, k] += h
            vm[i, k] -= h
            cd2[i, k] = (f2(vp) - f2(vm)) / (2 * h)
    np.testing.assert_allclose(got2, cd2, atol=5e-3)
