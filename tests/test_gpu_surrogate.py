"""The surrogate's GPU kernels against the float64 oracle (``oracle/mlp.py``).

* the BF16 MFMA GEMM against numpy on the same bf16 values (pins the MFMA
  operand/result lane layout and every transpose case);
* forward energy and input gradient, and the Sobolev-loss parameter gradient,
  against the oracle with the device's bf16 rounding points, at the real
  144 -> 512 -> 512 -> 512 -> 1 size;
* one Adam step against the oracle's Adam on the device's own gradient;
* training on simulator-labelled data near a point fits held-out energies and
  gradients;
* the gain of the trained network over the exact-check loop without it: its
  anchored gradient is closer to the exact gradient than the stale one.
"""

import math
import sys

import numpy as np
import pytest

from oracle import mlp as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _bf(x):
    return O.bf16_round(x)


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 32, 0, 0), (100, 70, 50, 0, 0), (37, 129, 144, 0, 1),
                                         (512, 144, 33, 1, 0), (1, 512, 144, 0, 1), (70, 90, 17, 1, 1)])
def test_gemm_bf16_matches_numpy(M, N, K, ta, tb):
    import torch
    from quantum_amd import surrogate as S
    rng = np.random.default_rng(M * 7 + N)
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    c0 = rng.normal(size=(M, N)).astype(np.float32)
    at = torch.tensor(a).to(torch.bfloat16).cuda()
    bt = torch.tensor(b).to(torch.bfloat16).cuda()
    ct = torch.tensor(c0).cuda()
    _say(f"gemm {M}x{N}x{K} ta={ta} tb={tb}")
    S.gemm_bf16(at, bt, bool(ta), bool(tb), c=ct, alpha=0.5, beta=2.0)
    A = at.float().cpu().numpy().astype(np.float64)
    Bm = bt.float().cpu().numpy().astype(np.float64)
    A = A.T if ta else A
    Bm = Bm.T if tb else Bm
    ref = 0.5 * A @ Bm + 2.0 * c0
    scale = 0.5 * np.abs(A) @ np.abs(Bm) + 2.0 * np.abs(c0)
    err = np.abs(ct.cpu().numpy() - ref) / scale
    assert err.max() < 1e-5, err.max()


def _model_and_batch(B=37, seed=3, P=144, H=512):
    import torch
    from quantum_amd import surrogate as S
    model = S.MLPSurrogate(P, H, max_batch=64, seed=seed)
    rng = np.random.default_rng(seed)
    # Larger weights than the default init so silu' and silu'' vary.
    theta = (model.init_params(seed) * 2.5).astype(np.float32)
    model.set_params(theta)
    x = rng.uniform(-1, 1, (B, P)).astype(np.float32)
    e_t = rng.normal(0, 1, B).astype(np.float32)
    g_t = rng.normal(0, 0.3, (B, P)).astype(np.float32)
    return model, theta.astype(np.float64), x, e_t, g_t


def test_forward_and_input_gradient_match_oracle():
    import torch
    model, theta, x, _, _ = _model_and_batch()
    _say("mlp forward 144-512-512-512-1")
    E, G = model.predict(torch.tensor(x).cuda())
    Eo, Go, _ = O.forward(theta, x.astype(np.float64), 144, 512, bf16=True)
    E = E.cpu().numpy()
    G = G.cpu().numpy()
    # bf16 operands, FP32 accumulation: compare at 2e-3 of each quantity's scale.
    assert np.abs(E - Eo).max() <= 2e-3 * np.abs(Eo).max()
    assert np.abs(G - Go).max() <= 2e-3 * np.abs(Go).max()
    # And the float64 network without rounding is within bf16 distance.
    Ef, Gf, _ = O.forward(theta, x.astype(np.float64), 144, 512, bf16=False)
    assert np.abs(E - Ef).max() <= 5e-2 * np.abs(Ef).max()


@pytest.mark.parametrize("B", [1, 5, 32])
def test_fused_small_batch_predict(B):
    """Batches of up to 32 rows run as one fused launch; B = 37 runs layer by
    layer.  Both match the oracle at the device's rounding points, and each
    other: the products run K in the same order, so only the energy's final
    sum order differs."""
    import torch
    model, theta, x, _, _ = _model_and_batch()
    xs = torch.tensor(x).cuda()
    _say(f"fused predict B={B}")
    E, G = model.predict(xs[:B])
    E_ref, G_ref = model.predict(xs)  # 37 rows: the layer-by-layer path
    E, G = E.cpu().numpy(), G.cpu().numpy()
    E_ref, G_ref = E_ref.cpu().numpy()[:B], G_ref.cpu().numpy()[:B]
    Eo, Go, _ = O.forward(theta, x[:B].astype(np.float64), 144, 512, bf16=True)
    assert np.abs(E - Eo).max() <= 2e-3 * np.abs(Eo).max()
    assert np.abs(G - Go).max() <= 2e-3 * np.abs(Go).max()
    np.testing.assert_allclose(E, E_ref, rtol=1e-5, atol=1e-5 * np.abs(E_ref).max())
    np.testing.assert_allclose(G, G_ref, rtol=1e-5, atol=1e-6 * np.abs(G_ref).max())
    E1, G1 = model.predict(xs[:B], with_grad=False)
    np.testing.assert_array_equal(E1.cpu().numpy(), E)


@pytest.mark.parametrize("lam", [0.0, 1.0])
def test_sobolev_loss_gradient_matches_oracle(lam):
    import torch
    model, theta, x, e_t, g_t = _model_and_batch(B=48, seed=5)
    _say(f"mlp loss/grad lam={lam}")
    loss = model.loss_grad(torch.tensor(x).cuda(), torch.tensor(e_t).cuda(), torch.tensor(g_t).cuda(),
                           lam).cpu().numpy()
    grad = model.grads().cpu().numpy().astype(np.float64)
    le, lg, go = O.loss_grad(theta, x.astype(np.float64), e_t.astype(np.float64),
                             g_t.astype(np.float64), lam, 144, 512, bf16=True)
    np.testing.assert_allclose(loss, [le, lg], rtol=2e-3, atol=1e-7)
    rel = np.linalg.norm(grad - go) / np.linalg.norm(go)
    assert rel < 5e-3, rel
    # Per block (W1, b1, ..., b4): each matches to 1e-2 of its own norm.
    o = 0
    for name, shape in O.param_shapes(144, 512):
        n = int(np.prod(shape))
        blk, ref = grad[o:o + n], go[o:o + n]
        if np.linalg.norm(ref) > 0:
            assert np.linalg.norm(blk - ref) <= 1e-2 * np.linalg.norm(ref) + 1e-9, name
        o += n


def test_adam_step_matches_oracle():
    import torch
    model, theta, x, e_t, g_t = _model_and_batch(B=32, seed=7)
    xt, et, gt = torch.tensor(x).cuda(), torch.tensor(e_t).cuda(), torch.tensor(g_t).cuda()
    _say("mlp adam step")
    model.loss_grad(xt, et, gt, 0.5)
    g = model.grads().cpu().numpy().astype(np.float64)
    p0 = model.params().cpu().numpy().astype(np.float64)
    model.train_step(xt, et, gt, 0.5, 1e-3)
    p1 = model.params().cpu().numpy().astype(np.float64)
    ref, _, _ = O.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, 1e-3)
    np.testing.assert_allclose(p1, ref, atol=2e-7, rtol=1e-6)


def _c5(n_layers=6):
    from quantum_amd import workloads as W
    from quantum_amd import surrogate as S
    circuit, names, ham = W.config5_vqe(12, n_layers)
    return S.ExactEnergy(circuit, names, ham), names


def test_exact_energy_matches_oracle():
    from oracle import tfq_oracle as T
    from quantum_amd import serialize as ser
    exact, names = _c5()
    th = np.random.default_rng(0).uniform(-math.pi, math.pi, len(names)).astype(np.float32)
    _say("exact energy C5")
    e, g = exact(th)
    progs = np.array([exact.program], dtype=object)
    ps = np.array([[exact.psum]], dtype=object)
    eo = T.simulate_expectation(list(progs), names, th[None, :], ps)[0, 0]
    go = T.adjoint_gradient(list(progs), names, th[None, :], ps, np.ones((1, 1)), mode="analytic")[0]
    assert abs(e - eo) < 1e-4
    np.testing.assert_allclose(g, go, atol=1e-4)


@pytest.fixture(scope="module")
def trained():
    """A network trained on simulator labels in a neighbourhood of theta0."""
    import torch
    from quantum_amd import surrogate as S
    exact, names = _c5()
    theta0 = np.random.default_rng(7).uniform(-0.3, 0.3, len(names))
    pts = (theta0[None, :] + 0.05 * np.random.default_rng(3).normal(size=(4096, len(names))))
    pts = pts.astype(np.float32)
    _say("label 4096 C5 parameter sets near theta0")
    E, G = S.label(exact, pts)
    th = torch.tensor(pts).cuda()
    norm = S.Normalizer.fit(E)
    model = S.MLPSurrogate(len(names), 512, max_batch=512, seed=1)
    _say("train surrogate")
    hist = S.fit(model, norm, th[512:], E[512:], G[512:], epochs=40, batch=512, lr=1e-3, lam=1.0)
    return dict(model=model, norm=norm, exact=exact, hist=hist, th=th, E=E, G=G, theta0=theta0)


def test_training_fits_held_out_energies_and_gradients(trained):
    import torch
    hist = trained["hist"]
    tot = hist.sum(axis=1)
    assert np.all(np.isfinite(tot))
    assert tot[-1] < 0.3 * tot[0], tot
    model, norm = trained["model"], trained["norm"]
    th, E, G = trained["th"][:512], trained["E"][:512].double(), trained["G"][:512].double()
    e, g = model.predict(norm.x(th).float().contiguous())
    e, g = norm.energy(e.double()), norm.grad(g.double())
    rel = float(((e - E) ** 2).mean().sqrt() / E.std())
    cos = float(torch.nn.functional.cosine_similarity(g, G, dim=1).mean())
    print(f"[surrogate] held-out relative RMSE {rel:.3f}, gradient cosine {cos:.3f}",
          file=sys.stderr)
    assert rel < 0.4 and cos > 0.9


def test_network_gradient_beats_the_stale_exact_gradient(trained):
    """What the trained network adds to the exact-check loop: between checks the
    loop without it reuses the exact gradient of the last checked point; with
    it the step uses g_exact(theta0) + g_net(x) - g_net(theta0).  The network
    is trained on the residual of the exact first-order model around theta0
    (E - E0 - g0.dx, G - g0).  On held-out points near theta0 the anchored
    network gradient is the closer one."""
    import torch
    from quantum_amd import surrogate as S
    exact = trained["exact"]
    theta0 = trained["theta0"]
    e0, g0 = exact(theta0)
    th, E, G = trained["th"], trained["E"], trained["G"]
    g0t = torch.tensor(g0, dtype=torch.float32, device=th.device)
    d = th - torch.tensor(theta0, dtype=torch.float32, device=th.device)[None, :]
    Er, Gr = E - (e0 + d @ g0t), G - g0t[None, :]
    norm = S.Normalizer.fit(Er)
    model = S.MLPSurrogate(len(theta0), 512, max_batch=512, seed=1)
    _say("train residual surrogate")
    S.fit(model, norm, th[512:], Er[512:], Gr[512:], epochs=200, batch=512, lr=1e-3, lam=1.0)
    x = torch.cat([torch.tensor(theta0[None, :], dtype=torch.float32, device=th.device), th[:512]])
    _, gs = model.predict(norm.x(x).float().contiguous())
    gs = norm.grad(gs.double())
    anchored = g0t.double()[None, :] + gs[1:] - gs[0][None, :]
    Gh = G[:512].double()
    err_net = float(torch.linalg.norm(anchored - Gh, dim=1).mean())
    err_stale = float(torch.linalg.norm(g0t.double()[None, :] - Gh, dim=1).mean())
    print(f"[surrogate] mean |g_anchored - g| = {err_net:.4f} vs stale {err_stale:.4f}",
          file=sys.stderr)
    assert err_net < 0.95 * err_stale


@pytest.mark.parametrize("loop", ["surrogate", "candidate", "ablation"])
def test_surrogate_assisted_vqe_loops(loop):
    """The C5 loops run end to end (150 steps from theta0 ~ U(-0.1, 0.1)):
    the energy they report is the exact energy at the parameters they return
    (every accepted point is checked by the simulator), it lies below the
    start energy, and the exact simulator was called."""
    from quantum_amd import surrogate as S
    exact, names = _c5()
    theta0 = np.random.default_rng(1005).uniform(-0.1, 0.1, len(names))
    e0, _ = exact(theta0.astype(np.float32))
    model = None if loop == "ablation" else S.MLPSurrogate(len(names), 512, max_batch=1024, seed=0)
    vqe = S.CandidateVQE(model, exact) if loop == "candidate" else S.SurrogateVQE(model, exact)
    _say(f"C5 {loop} VQE, 150 steps")
    r = vqe.run(theta0, steps=150)
    assert np.all(np.isfinite(r["theta"]))
    e_check, _ = exact(np.asarray(r["theta"], dtype=np.float32))
    assert abs(r["energy"] - e_check) < 1e-4 * max(1.0, abs(e_check))
    assert r["energy"] < e0 - 1.0
    assert r["exact_calls"] > 0
