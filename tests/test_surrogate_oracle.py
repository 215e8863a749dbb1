"""The surrogate's float64 oracle: its hand-derived Sobolev-loss gradient is
pinned by central finite differences and by torch autograd (float64 double
backward); its input gradient by finite differences of the energy; the
bfloat16 rounding helper by exact bit patterns."""

import numpy as np
import pytest

from oracle import mlp


def _net(P, H, seed):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0, 0.6, mlp.num_params(P, H))
    x = rng.uniform(-1, 1, (5, P))
    e_t = rng.normal(0, 1, 5)
    g_t = rng.normal(0, 1, (5, P))
    return theta, x, e_t, g_t


def test_num_params_of_the_144_512_network():
    assert mlp.num_params(144, 512) == 600065


def test_input_gradient_matches_finite_differences():
    P, H = 4, 6
    theta, x, _, _ = _net(P, H, 0)
    _, G, _ = mlp.forward(theta, x, P, H)
    h = 1e-6
    for p in range(P):
        xp, xm = x.copy(), x.copy()
        xp[:, p] += h
        xm[:, p] -= h
        fd = (mlp.forward(theta, xp, P, H)[0] - mlp.forward(theta, xm, P, H)[0]) / (2 * h)
        np.testing.assert_allclose(G[:, p], fd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("lam", [0.0, 0.3, 2.0])
def test_loss_gradient_matches_finite_differences(lam):
    P, H = 3, 5
    theta, x, e_t, g_t = _net(P, H, 1)
    _, _, grad = mlp.loss_grad(theta, x, e_t, g_t, lam, P, H)
    h = 1e-6
    fd = np.empty_like(theta)
    for i in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (sum(mlp.loss(tp, x, e_t, g_t, lam, P, H)) -
                 sum(mlp.loss(tm, x, e_t, g_t, lam, P, H))) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=1e-5, atol=1e-8)


def test_loss_gradient_matches_torch_double_backward():
    import torch
    P, H = 7, 16
    theta, x, e_t, g_t = _net(P, H, 2)
    lam = 0.7
    le, lg, grad = mlp.loss_grad(theta, x, e_t, g_t, lam, P, H)
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    d, o = {}, 0
    for name, shape in mlp.param_shapes(P, H):
        n = int(np.prod(shape))
        d[name] = t[o:o + n].reshape(shape)
        o += n
    hcur = xt
    for l in (1, 2, 3):
        hcur = torch.nn.functional.silu(hcur @ d["W%d" % l].T + d["b%d" % l])
    E = hcur @ d["w4"] + d["b4"][0]
    G, = torch.autograd.grad(E.sum(), xt, create_graph=True)
    L_e = ((E - torch.tensor(e_t)) ** 2).mean()
    L_g = lam * ((G - torch.tensor(g_t)) ** 2).mean()
    (L_e + L_g).backward()
    assert abs(L_e.item() - le) < 1e-12 and abs(L_g.item() - lg) < 1e-12
    np.testing.assert_allclose(grad, t.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_bf16_rounding_is_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7                      # bf16 spacing in [1, 2)
    cases = {
        one: one,
        one + ulp / 2: one,              # tie -> even (mantissa 0)
        one + 3 * ulp / 2: one + 2 * ulp,  # tie -> even (mantissa 2)
        one + ulp / 2 + 2 ** -20: one + ulp,
        -(one + ulp / 4): -one,
        3.0: 3.0,
    }
    for x, want in cases.items():
        assert mlp.bf16_round(np.array([x]))[0] == want, x


def test_adam_matches_closed_form_first_step():
    g = np.array([0.5, -2.0, 1e-3])
    th, m, v = mlp.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.1)
    # First bias-corrected step is lr * g / (|g| + eps) = lr * sign(g) (up to eps).
    np.testing.assert_allclose(th, -0.1 * np.sign(g), rtol=1e-4)
