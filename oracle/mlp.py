"""Float64 restatement of the VQE energy surrogate (test infrastructure only).

The network maps ansatz parameters x [B, P] to an energy E [B]:

    h0 = x;  a_l = h_{l-1} W_l^T + b_l;  h_l = silu(a_l)   (l = 1..3, W_l [H, fan_in])
    E  = h3 . w4 + b4

and its input gradient G = dE/dx [B, P] by the chain

    u3 = w4 * silu'(a3);  u_l = (u_{l+1} W_{l+1}) * silu'(a_l);  G = u1 W1.

Training minimises the Sobolev loss

    L = mean_b (E_b - E*_b)^2 + lam * mean_{b,p} (G_bp - G*_bp)^2.

The flat parameter vector is W1, b1, W2, b2, W3, b3, w4, b4 (row-major).

``bf16=True`` rounds every GEMM operand to bfloat16 (round-to-nearest-even of
the float32 value) at the points where the device kernels round, so the GPU
can be compared at FP32-accumulation precision.  The parameter gradient is
derived by hand (double backward through the input-gradient branch); it is
pinned against central finite differences of ``loss`` and against torch
autograd in float64 (tests/test_surrogate_oracle.py).
"""

import numpy as np


def param_shapes(P, H):
    return [("W1", (H, P)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
            ("W3", (H, H)), ("b3", (H,)), ("w4", (H,)), ("b4", (1,))]


def num_params(P, H):
    return sum(int(np.prod(s)) for _, s in param_shapes(P, H))


def unpack(theta, P, H):
    out, o = {}, 0
    for name, shape in param_shapes(P, H):
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(d, P, H):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1)
                           for name, _ in param_shapes(P, H)])


def bf16_round(x):
    """float64 -> float32 -> bfloat16 (round to nearest, ties to even) -> float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _sig(a):
    return 1.0 / (1.0 + np.exp(-a))


def silu(a):
    return a * _sig(a)


def dsilu(a):
    s = _sig(a)
    return s * (1.0 + a * (1.0 - s))


def d2silu(a):
    s = _sig(a)
    return s * (1.0 - s) * (2.0 + a * (1.0 - 2.0 * s))


def forward(theta, x, P, H, bf16=False):
    """E [B], G [B, P] and the cache used by ``loss_grad``."""
    r = bf16_round if bf16 else (lambda v: v)
    p = unpack(np.asarray(theta, dtype=np.float64), P, H)
    W = [r(p["W1"]), r(p["W2"]), r(p["W3"])]
    b = [p["b1"], p["b2"], p["b3"]]
    hin = [r(np.asarray(x, dtype=np.float64))]
    a, h = [], []
    for l in range(3):
        al = hin[l] @ W[l].T + b[l]
        hl = silu(al)
        a.append(al)
        h.append(hl)
        if l < 2:
            hin.append(r(hl))
    E = h[2] @ p["w4"] + p["b4"][0]
    u = [None, None, p["w4"][None, :] * dsilu(a[2])]
    v = [None, None]
    for l in (2, 1):
        v[l - 1] = r(u[l]) @ W[l]
        u[l - 1] = v[l - 1] * dsilu(a[l - 1])
    G = r(u[0]) @ W[0]
    cache = dict(p=p, W=W, hin=hin, a=a, h=h, u=u, v=v, E=E, G=G)
    return E, G, cache


def loss(theta, x, e_t, g_t, lam, P, H, bf16=False):
    E, G, _ = forward(theta, x, P, H, bf16)
    return float(np.mean((E - e_t) ** 2)), float(lam * np.mean((G - g_t) ** 2))


def loss_grad(theta, x, e_t, g_t, lam, P, H, bf16=False):
    """(loss_E, loss_G, dL/dtheta) with the device's rounding points when bf16."""
    r = bf16_round if bf16 else (lambda v: v)
    E, G, c = forward(theta, x, P, H, bf16)
    p, W, hin, a, h, u, v = c["p"], c["W"], c["hin"], c["a"], c["h"], c["u"], c["v"]
    B = x.shape[0]
    dE = 2.0 * (E - e_t) / B
    dG = r(2.0 * lam * (G - g_t) / (B * P))
    g = {k: np.zeros_like(val) for k, val in p.items()}
    # G = u1 W1
    g["W1"] += r(u[0]).T @ dG
    du = [None, None, None]
    du[0] = dG @ W[0].T
    dag = [None, None, None]
    dv = [None, None, None]
    # u_l = v_l * s'(a_l), v_l = u_{l+1} W_{l+1}  (l = 0, 1)
    for l in (0, 1):
        dv[l] = r(du[l] * dsilu(a[l]))
        dag[l] = du[l] * v[l] * d2silu(a[l])
        g["W%d" % (l + 2)] += r(u[l + 1]).T @ dv[l]
        du[l + 1] = dv[l] @ W[l + 1].T
    # u3 = w4 * s'(a3) and E = h3 . w4 + b4
    w4 = p["w4"]
    g["w4"] += np.sum(du[2] * dsilu(a[2]) + dE[:, None] * h[2], axis=0)
    g["b4"] += np.sum(dE)
    da = [None, None, None]
    da[2] = du[2] * w4[None, :] * d2silu(a[2]) + dE[:, None] * w4[None, :] * dsilu(a[2])
    g["b3"] += np.sum(da[2], axis=0)
    dab = r(da[2])
    g["W3"] += dab.T @ hin[2]
    dh = dab @ W[2]
    for l in (1, 0):
        da[l] = dag[l] + dh * dsilu(a[l])
        g["b%d" % (l + 1)] += np.sum(da[l], axis=0)
        dab = r(da[l])
        g["W%d" % (l + 1)] += dab.T @ hin[l]
        if l > 0:
            dh = dab @ W[l]
    le = float(np.mean((E - e_t) ** 2))
    lg = float(lam * np.mean((G - g_t) ** 2))
    return le, lg, pack(g, P, H)


def adam(theta, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), in float64."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
