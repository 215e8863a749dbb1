"""Learned energy surrogate for VQE on the GPU.

A small MLP ``E(x)`` over normalised ansatz parameters (``x = wrap(theta)/pi``)
is trained on energies and adjoint gradients that the simulator produces
(``tfq_adj_grad_op.expectation_and_gradient``), with the Sobolev loss

    mean (E - E*)^2 + lam * mean ||dE/dx - G*||^2.

Forward, input gradient, the double backward of the loss, and Adam run in
``libtfqx.so`` (BF16 MFMA GEMMs with FP32 accumulation); this module only
moves batches and drives the loops.

``SurrogateVQE`` uses the network inside an optimiser the exact simulator
checks: between exact evaluations (every ``check_every`` steps) the step
direction is the surrogate's gradient, anchored to the last exact point
(``g = g_exact(theta_k) + g_s(theta) - g_s(theta_k)``, likewise for E).  At
each check the block is accepted only if the exact energy went down by at
least a quarter of the predicted decrease and the anchored prediction is
within ``tol`` (relative) of the exact energy; otherwise the parameters are
restored and the next block runs on exact gradients (fallback).

TFQ has no learned component; the reference only provides the gradient this
trains on (``core/ops/tfq_adj_grad_op.cc``).
"""

import ctypes
import math
import time

import numpy as np

from . import _lib
from . import serialize as _ser


def _torch():
    import torch
    return torch


def _err():
    return ctypes.create_string_buffer(4096)


def wrap_angles(theta):
    """Map angles into [-pi, pi) (the energy is 2*pi periodic in each)."""
    return np.mod(np.asarray(theta, dtype=np.float64) + math.pi, 2 * math.pi) - math.pi


class MLPSurrogate:
    """``in_dim -> hidden -> hidden -> hidden -> 1`` SiLU network on the GPU.

    Parameters, Adam moments and activations for up to ``max_batch`` rows live
    in one device workspace.  Weights are FP32 (BF16 copies feed the GEMMs)."""

    def __init__(self, in_dim, hidden=512, max_batch=1024, seed=0, device=None):
        torch = _torch()
        self.lib = _lib.lib()
        self.P, self.H, self.max_batch = int(in_dim), int(hidden), int(max_batch)
        out = (ctypes.c_int64 * 6)()
        rc = self.lib.tfqx_mlp_layout(self.P, self.H, self.max_batch, out, 6)
        if rc != _lib.TFQX_OK:
            raise _lib.InvalidArgumentError("invalid surrogate dimensions")
        self.ws_bytes, self.nparams, self.off_w, self.off_g, self.off_loss, _ = [int(v) for v in out]
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.ws = torch.zeros(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self.step = 0
        self.set_params(self.init_params(seed))

    # -- parameters -------------------------------------------------------
    def init_params(self, seed):
        """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, zero head bias."""
        rng = np.random.default_rng(seed)
        P, H = self.P, self.H
        parts = []
        for fan_in, shape in ((P, (H, P)), (P, (H,)), (H, (H, H)), (H, (H,)), (H, (H, H)), (H, (H,)),
                              (H, (H,))):
            b = 1.0 / math.sqrt(fan_in)
            parts.append(rng.uniform(-b, b, shape).reshape(-1))
        parts.append(np.zeros(1))
        return np.concatenate(parts).astype(np.float32)

    def _stream(self):
        return _torch().cuda.current_stream(self.device).cuda_stream

    def _ws(self):
        return ctypes.c_void_p(self.ws.data_ptr())

    def set_params(self, host_params):
        p = np.ascontiguousarray(host_params, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        e = _err()
        _lib.check(self.lib.tfqx_mlp_set_params(self._ws(), self.P, self.H, self.max_batch,
                                                p.ctypes.data, ctypes.c_void_p(self._stream()),
                                                e, 4096), e)
        _torch().cuda.current_stream(self.device).synchronize()
        self.step = 0

    def params(self):
        """Device view (float32 [nparams]) of the parameters."""
        torch = _torch()
        return self.ws[self.off_w:self.off_w + 4 * self.nparams].view(torch.float32)

    def grads(self):
        """Device view of the last parameter gradient."""
        torch = _torch()
        return self.ws[self.off_g:self.off_g + 4 * self.nparams].view(torch.float32)

    def last_loss(self):
        """Device view [2]: (energy term, lam * gradient term) of the last loss."""
        torch = _torch()
        return self.ws[self.off_loss:self.off_loss + 8].view(torch.float32)

    # -- compute ----------------------------------------------------------
    def _check_x(self, x):
        torch = _torch()
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32):
            raise TypeError("x must be a float32 device tensor")
        if x.dim() != 2 or x.shape[1] != self.P:
            raise ValueError(f"x must be [batch, {self.P}]")
        return x.contiguous()

    def predict(self, x, with_grad=True):
        """Energies [B] and input gradients [B, P] (device tensors)."""
        torch = _torch()
        x = self._check_x(x)
        B = x.shape[0]
        E = torch.empty(B, dtype=torch.float32, device=self.device)
        G = torch.empty((B, self.P), dtype=torch.float32, device=self.device) if with_grad else None
        s = ctypes.c_void_p(self._stream())
        for b0 in range(0, B, self.max_batch):
            b1 = min(B, b0 + self.max_batch)
            e = _err()
            _lib.check(self.lib.tfqx_mlp_predict(
                self._ws(), self.P, self.H, self.max_batch, ctypes.c_void_p(x[b0:b1].data_ptr()),
                b1 - b0, ctypes.c_void_p(E[b0:b1].data_ptr()),
                ctypes.c_void_p(G[b0:b1].data_ptr()) if with_grad else None, s, e, 4096), e)
        return E, G

    def _targets(self, x, e_t, g_t):
        x = self._check_x(x)
        B = x.shape[0]
        if B > self.max_batch:
            raise ValueError(f"batch {B} exceeds max_batch {self.max_batch}")
        e_t = e_t.contiguous()
        g_t = g_t.contiguous()
        if e_t.shape != (B,) or g_t.shape != (B, self.P):
            raise ValueError("target shapes do not match x")
        return x, e_t, g_t, B

    def loss_grad(self, x, e_t, g_t, lam):
        """Computes the loss (``last_loss``) and parameter gradient (``grads``)."""
        x, e_t, g_t, B = self._targets(x, e_t, g_t)
        e = _err()
        _lib.check(self.lib.tfqx_mlp_loss_grad(
            self._ws(), self.P, self.H, self.max_batch, ctypes.c_void_p(x.data_ptr()),
            ctypes.c_void_p(e_t.data_ptr()), ctypes.c_void_p(g_t.data_ptr()), B, float(lam),
            ctypes.c_void_p(self._stream()), e, 4096), e)
        return self.last_loss()

    def train_step(self, x, e_t, g_t, lam, lr):
        """One Sobolev-loss Adam step on a batch (all on the device, no sync)."""
        x, e_t, g_t, B = self._targets(x, e_t, g_t)
        self.step += 1
        e = _err()
        _lib.check(self.lib.tfqx_mlp_train_step(
            self._ws(), self.P, self.H, self.max_batch, ctypes.c_void_p(x.data_ptr()),
            ctypes.c_void_p(e_t.data_ptr()), ctypes.c_void_p(g_t.data_ptr()), B, float(lam),
            float(lr), self.step, ctypes.c_void_p(self._stream()), e, 4096), e)
        return self.last_loss()


def gemm_bf16(a, b, trans_a=False, trans_b=False, c=None, alpha=1.0, beta=0.0):
    """``alpha * op(a) @ op(b) + beta * c`` for bfloat16 device matrices (FP32 out)."""
    torch = _torch()
    lib = _lib.lib()
    if a.dtype != torch.bfloat16 or b.dtype != torch.bfloat16:
        raise TypeError("a and b must be bfloat16")
    a = a.contiguous()
    b = b.contiguous()
    M, K = (a.shape[1], a.shape[0]) if trans_a else a.shape
    K2, N = (b.shape[1], b.shape[0]) if trans_b else b.shape
    if K != K2:
        raise ValueError("inner dimensions do not match")
    if c is None:
        c = torch.zeros((M, N), dtype=torch.float32, device=a.device)
    e = _err()
    _lib.check(lib.tfqx_gemm_bf16(M, N, K, ctypes.c_void_p(a.data_ptr()), a.shape[1], int(trans_a),
                                  ctypes.c_void_p(b.data_ptr()), b.shape[1], int(trans_b),
                                  ctypes.c_void_p(c.data_ptr()), N, float(alpha), float(beta),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), e,
                                  4096), e)
    return c


# ---------------------------------------------------------------- exact side
class ExactEnergy:
    """Energy and adjoint gradient of one ansatz/Hamiltonian pair on the GPU.

    The circuit tiled over a batch is planned once per batch size (parse,
    resolve, fuse); later calls only rebind the symbol values
    (``tfqx_plan_rebind``), upload and launch the fused expectation+adjoint
    kernel."""

    def __init__(self, circuit, symbol_names, hamiltonian, derivative="analytic"):
        self.names = list(symbol_names)
        self.program = _ser.serialize_circuit(circuit)
        self.psum = _ser.serialize_paulisum(hamiltonian)
        self.derivative = derivative
        self.calls = 0
        self.rows = 0
        self._plans = {}

    def _plan(self, n):
        if n not in self._plans:
            from .tfq_adj_grad_op import plan_adj_grad
            torch = _torch()
            progs = np.empty(n, dtype=object)
            progs[:] = [self.program] * n
            ps = np.empty((n, 1), dtype=object)
            ps[:, 0] = [self.psum] * n
            plan, shape = plan_adj_grad(progs, self.names, np.zeros((n, len(self.names)), np.float32),
                                        ps, np.ones((n, 1), np.float32), self.derivative,
                                        with_expectation=True)
            dev = torch.device("cuda", torch.cuda.current_device())
            self._plans[n] = (plan, torch.empty(shape, dtype=torch.float32, device=dev), dev)
        return self._plans[n]

    def batch(self, thetas):
        """E [N] and dE/dtheta [N, P] as float32 device tensors (one launch).

        The tensors are views of a buffer the next call with the same N reuses."""
        torch = _torch()
        thetas = np.ascontiguousarray(thetas, dtype=np.float32)
        n = thetas.shape[0]
        plan, out, dev = self._plan(n)
        if n >= 64:
            plan.stage_values(thetas)  # matrices evaluated on the device at upload
        else:
            plan.rebind(thetas)
        stream = torch.cuda.current_stream(dev).cuda_stream
        plan.upload(dev, stream)
        plan.launch(out, stream)
        self.calls += 1
        self.rows += n
        return out[:n], out[n:].reshape(n, len(self.names))

    def __call__(self, theta):
        e, g = self.batch(np.asarray(theta, dtype=np.float32)[None, :])
        return float(e[0].item()), g[0].double().cpu().numpy()


def label(exact, thetas, chunk=4096):
    """Energies and gradients of many parameter sets, in chunks (device tensors)."""
    torch = _torch()
    Es, Gs = [], []
    for i in range(0, len(thetas), chunk):
        e, g = exact.batch(thetas[i:i + chunk])
        Es.append(e.clone())
        Gs.append(g.clone())
    return torch.cat(Es), torch.cat(Gs)


class Normalizer:
    """x = wrap(theta)/pi;  e' = (E - mu)/sigma;  dE'/dx = dE/dtheta * pi / sigma."""

    def __init__(self, mu, sigma):
        self.mu, self.sigma = float(mu), float(sigma)

    @classmethod
    def fit(cls, energies):
        e = energies.double()
        return cls(float(e.mean()), max(float(e.std()), 1e-6))

    def x(self, theta):
        torch = _torch()
        if isinstance(theta, torch.Tensor):
            return (torch.remainder(theta + math.pi, 2 * math.pi) - math.pi) / math.pi
        return (wrap_angles(theta) / math.pi).astype(np.float32)

    def targets(self, E, G):
        return (E - self.mu) / self.sigma, G * (math.pi / self.sigma)

    def energy(self, e_norm):
        return e_norm * self.sigma + self.mu

    def grad(self, g_norm):
        return g_norm * (self.sigma / math.pi)


def fit(model, norm, thetas, E, G, epochs=10, batch=1024, lr=1e-3, lr_final=1e-5, lam=1.0,
        seed=0, log=None):
    """Train on (theta, E, dE/dtheta) device tensors; returns the per-epoch loss."""
    torch = _torch()
    X = norm.x(thetas.float() if isinstance(thetas, torch.Tensor) else
               torch.as_tensor(thetas, device=model.device)).float().contiguous()
    Et, Gt = norm.targets(E.float(), G.float())
    N = X.shape[0]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    steps_per_epoch = max(1, N // batch)
    total = epochs * steps_per_epoch
    hist = []
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(N, generator=gen).to(model.device)
        acc = torch.zeros(2, device=model.device)
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            # cosine learning-rate schedule
            lr_k = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * k / max(1, total)))
            acc += model.train_step(X[idx], Et[idx].contiguous(), Gt[idx].contiguous(), lam, lr_k)
            k += 1
        hist.append((acc / steps_per_epoch).cpu().numpy())
        if log:
            log(f"epoch {ep}: loss E {hist[-1][0]:.4e}  G {hist[-1][1]:.4e}")
    return np.array(hist)


class _Adam:
    def __init__(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = self.v = None
        self.t = 0

    def step(self, theta, g):
        if self.m is None:
            self.m = np.zeros_like(theta)
            self.v = np.zeros_like(theta)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return theta - self.lr * mh / (np.sqrt(vh) + self.eps)

    def state(self):
        return (None if self.m is None else self.m.copy(), None if self.v is None else self.v.copy(),
                self.t)

    def restore(self, st):
        self.m, self.v, self.t = st


def vqe_exact(exact, theta0, lr=0.02, steps=1000, e_target=None, log_every=0):
    """Adam on exact adjoint gradients.  Stops at ``e_target`` if given."""
    torch = _torch()
    theta = np.array(theta0, dtype=np.float64)
    opt = _Adam(lr)
    calls0 = exact.calls
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    reached = None
    e = None
    for k in range(steps + 1):
        e, g = exact(theta)
        hist.append((k, time.perf_counter() - t0, e))
        if e_target is not None and e <= e_target:
            reached = (k, time.perf_counter() - t0)
            break
        if k < steps:
            theta = opt.step(theta, g)
    return dict(theta=theta, energy=e, history=hist, reached=reached,
                seconds=time.perf_counter() - t0, exact_calls=exact.calls - calls0)


class SurrogateVQE:
    """Surrogate-proposed Adam steps, checked by the exact simulator.

    Every ``check_every`` steps one batched launch evaluates the current point
    exactly together with a cloud of ``cloud - 1`` nearby points
    (``theta + sigma * N(0, 1)``): a batch of small circuits costs about one
    circuit's latency on the GPU.  The network is fine-tuned on the recent
    clouds (``finetune`` Adam steps on the device), and the next steps follow
    its gradient anchored to the exact point:
    ``g = g_exact(theta_k) + g_s(theta) - g_s(theta_k)`` (likewise for E).

    At each check the block is accepted when the exact energy fell by at least
    ``accept`` times the predicted decrease and the anchored prediction is
    within ``tol`` (relative) of the exact energy; otherwise the parameters
    are restored, the step size halves (down to ``lr_min``) and the next block
    runs on exact gradients (fallback).

    With ``model=None`` the same loop runs without the network (the last exact
    gradient is reused between checks): the ablation the network is measured
    against.
    """

    def __init__(self, model, exact, lr=0.02, check_every=10, tol=1e-2, accept=0.25, cloud=256,
                 sigma=0.05, finetune=10, replay=4, lam=1.0, ft_lr=3e-4, lr_min=2e-3):
        self.model, self.exact = model, exact
        self.lr, self.check_every, self.tol, self.accept = lr, check_every, tol, accept
        self.cloud, self.sigma, self.finetune, self.replay = cloud, sigma, finetune, replay
        self.lam, self.ft_lr, self.lr_min = lam, ft_lr, lr_min
        self.norm = None
        self._rng = np.random.default_rng(0)
        self._buf = []

    def _surrogate(self, theta):
        torch = _torch()
        if self.model is None:
            return 0.0, np.zeros(len(theta))
        x = torch.as_tensor(self.norm.x(theta)[None, :], device=self.model.device)
        e, g = self.model.predict(x)
        e = self.norm.energy(float(e[0].item()))
        return e, self.norm.grad(g[0].double().cpu().numpy())

    def _check(self, theta):
        """Exact E, g at theta (and a labelled cloud around it, used to fine-tune)."""
        torch = _torch()
        if self.model is None or self.cloud <= 1:
            return self.exact(theta)
        pts = np.empty((self.cloud, len(theta)), dtype=np.float32)
        pts[0] = theta
        pts[1:] = theta[None, :] + self.sigma * self._rng.standard_normal((self.cloud - 1, len(theta)))
        E, G = self.exact.batch(pts)
        E, G = E.clone(), G.clone()
        if self.norm is None:
            self.norm = Normalizer(float(E.double().mean()), max(float(E.double().std()), 1e-6))
        self._buf.append((torch.as_tensor(pts, device=E.device), E, G))
        self._buf = self._buf[-self.replay:]
        X = torch.cat([self.norm.x(b[0]).float() for b in self._buf])
        Et, Gt = self.norm.targets(torch.cat([b[1] for b in self._buf]),
                                   torch.cat([b[2] for b in self._buf]))
        n = X.shape[0]
        bs = min(n, self.model.max_batch)
        for _ in range(self.finetune):
            idx = torch.randint(0, n, (bs,), device=E.device)
            self.model.train_step(X[idx].contiguous(), Et[idx].contiguous(), Gt[idx].contiguous(),
                                  self.lam, self.ft_lr)
        return float(E[0].item()), G[0].double().cpu().numpy()

    def run(self, theta0, steps=1000, e_target=None):
        torch = _torch()
        theta = np.array(theta0, dtype=np.float64)
        opt = _Adam(self.lr)
        calls0, rows0 = self.exact.calls, self.exact.rows
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e_k, g_k = self._check(theta)
        s_k, sg_k = self._surrogate(theta)
        hist = [(0, time.perf_counter() - t0, e_k)]
        fallbacks = surrogate_steps = exact_steps = 0
        use_exact = False
        reached = None
        k = 0
        while k < steps:
            if e_target is not None and e_k <= e_target:
                reached = (k, time.perf_counter() - t0)
                break
            start, start_opt = theta.copy(), opt.state()
            pred = e_k
            for j in range(self.check_every):
                if use_exact:
                    e, g = (e_k, g_k) if j == 0 else self.exact(theta)
                    exact_steps += 1
                elif j == 0:
                    e, g = e_k, g_k
                    surrogate_steps += 1
                else:
                    s, sg = self._surrogate(theta)
                    e, g = e_k + (s - s_k), g_k + (sg - sg_k)
                    surrogate_steps += 1
                theta = opt.step(theta, g)
                k += 1
            if not use_exact:
                s_new, _ = self._surrogate(theta)
                pred = e_k + (s_new - s_k)
            e_new, g_new = self._check(theta)
            hist.append((k, time.perf_counter() - t0, e_new))
            ok = True
            if not use_exact:
                predicted_drop = e_k - pred
                actual_drop = e_k - e_new
                rel_err = abs(pred - e_new) / max(abs(e_new), 1e-12)
                ok = actual_drop > 0 and actual_drop >= self.accept * max(predicted_drop, 0.0) \
                    and rel_err <= self.tol
            if ok:
                e_k, g_k = e_new, g_new
                s_k, sg_k = self._surrogate(theta)
                use_exact = False
            else:
                # Fallback: restore, shrink the step and run the next block exactly.
                theta = start
                opt.restore(start_opt)
                opt.lr = max(self.lr_min, opt.lr * 0.5)
                fallbacks += 1
                use_exact = True
        return dict(theta=theta, energy=e_k, history=hist, reached=reached,
                    seconds=time.perf_counter() - t0, exact_calls=self.exact.calls - calls0,
                    exact_rows=self.exact.rows - rows0, surrogate_steps=surrogate_steps,
                    exact_steps=exact_steps, fallbacks=fallbacks)


class CandidateVQE:
    """Surrogate rollouts verified in parallel by the exact simulator.

    Each round starts from the current point theta (exact E, g known) and the
    optimiser state.  The network drives ``candidates`` Adam rollouts of
    ``check_every`` steps on the device, one per step size on a geometric
    grid around ``lr``, using its gradient anchored to the exact point
    (``g_exact(theta) + g_s(x) - g_s(theta)``).  The rollout end points and
    midpoints, plus a cloud of nearby points, go to the simulator in ONE
    batched expectation+adjoint launch.  The best verified point is accepted
    if it lowers the exact energy, and its rollout's optimiser state is kept.
    Otherwise the round falls back to one exact Adam step and the step-size
    grid shrinks.  Every evaluated point also becomes fine-tuning data.

    With ``model=None`` the rollouts reuse the exact gradient at theta
    (no network): the ablation the network is measured against.
    """

    def __init__(self, model, exact, lr=0.05, check_every=10, candidates=32, spread=4.0,
                 cloud=256, sigma=0.05, finetune=8, replay=4, lam=1.0, ft_lr=3e-4):
        self.model, self.exact = model, exact
        self.lr, self.check_every, self.candidates, self.spread = lr, check_every, candidates, spread
        self.cloud, self.sigma, self.finetune, self.replay = cloud, sigma, finetune, replay
        self.lam, self.ft_lr = lam, ft_lr
        self.norm = None
        self._rng = np.random.default_rng(0)
        self._buf = []

    def _fit(self, pts, E, G):
        torch = _torch()
        if self.model is None:
            return
        if self.norm is None:
            self.norm = Normalizer(float(E.double().mean()), max(float(E.double().std()), 1e-6))
        self._buf.append((pts, E, G))
        self._buf = self._buf[-self.replay:]
        X = torch.cat([self.norm.x(b[0]).float() for b in self._buf])
        Et, Gt = self.norm.targets(torch.cat([b[1] for b in self._buf]),
                                   torch.cat([b[2] for b in self._buf]))
        n = X.shape[0]
        bs = min(n, self.model.max_batch)
        for _ in range(self.finetune):
            idx = torch.randint(0, n, (bs,), device=X.device)
            self.model.train_step(X[idx].contiguous(), Et[idx].contiguous(), Gt[idx].contiguous(),
                                  self.lam, self.ft_lr)

    def _grad(self, X):
        """Surrogate dE/dtheta at rows of X (device, float32)."""
        _, g = self.model.predict(self.norm.x(X).float().contiguous())
        return self.norm.grad(g)

    def run(self, theta0, steps=1000, e_target=None, b1=0.9, b2=0.999, eps=1e-8):
        torch = _torch()
        dev = torch.device("cuda", torch.cuda.current_device())
        C, L = self.candidates, self.check_every
        P = len(theta0)
        calls0, rows0 = self.exact.calls, self.exact.rows
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # Exact start (and a first cloud for the network).
        pts = np.empty((max(1, self.cloud) if self.model is not None else 1, P), dtype=np.float32)
        pts[0] = theta0
        pts[1:] = theta0[None, :] + self.sigma * self._rng.standard_normal((len(pts) - 1, P))
        E, G = self.exact.batch(pts)
        E, G = E.clone(), G.clone()
        self._fit(torch.as_tensor(pts, device=dev), E, G)
        theta = torch.as_tensor(theta0, dtype=torch.float32, device=dev)
        e_k, g_k = float(E[0].item()), G[0].clone()
        m = torch.zeros(P, device=dev)
        v = torch.zeros(P, device=dev)
        t = 0
        scale = 1.0
        hist = [(0, time.perf_counter() - t0, e_k)]
        fallbacks = rounds = 0
        reached = None
        k = 0
        lrs = self.lr * self.spread ** torch.linspace(-1, 1, C, device=dev)
        while k < steps:
            if e_target is not None and e_k <= e_target:
                reached = (k, time.perf_counter() - t0)
                break
            rounds += 1
            X = theta.expand(C, P).clone()
            M = m.expand(C, P).clone()
            V = v.expand(C, P).clone()
            gs0 = self._grad(theta[None, :])[0] if self.model is not None else None
            lr_c = (lrs * scale)[:, None]
            mids = None
            for j in range(L):
                if self.model is None or j == 0:
                    g = g_k.expand(C, P)
                else:
                    g = g_k[None, :] + self._grad(X) - gs0[None, :]
                tt = t + j + 1
                M = b1 * M + (1 - b1) * g
                V = b2 * V + (1 - b2) * g * g
                X = X - lr_c * (M / (1 - b1 ** tt)) / (torch.sqrt(V / (1 - b2 ** tt)) + eps)
                if j == L // 2 - 1:
                    mids = (X.clone(), M.clone(), V.clone())
            props = torch.cat([X, mids[0]])
            states = (torch.cat([M, mids[1]]), torch.cat([V, mids[2]]))
            nprop = props.shape[0]
            ncloud = max(0, self.cloud - nprop) if self.model is not None else 0
            cloud = theta[None, :] + self.sigma * torch.randn(ncloud, P, device=dev)
            batch = torch.cat([props, cloud])
            E, G = self.exact.batch(batch.cpu().numpy())
            E, G = E.clone(), G.clone()
            best = int(torch.argmin(E[:nprop]).item())
            e_best = float(E[best].item())
            k += L
            if e_best < e_k:
                theta = props[best].clone()
                e_k, g_k = e_best, G[best].clone()
                m, v = states[0][best].clone(), states[1][best].clone()
                t += L if best < C else L // 2
            else:
                # Fallback: one exact Adam step, smaller step sizes next round.
                fallbacks += 1
                scale *= 0.5
                t += 1
                m = b1 * m + (1 - b1) * g_k
                v = b2 * v + (1 - b2) * g_k * g_k
                theta = theta - self.lr * scale * (m / (1 - b1 ** t)) / \
                    (torch.sqrt(v / (1 - b2 ** t)) + eps)
                e_k, g = self.exact(theta.cpu().numpy())
                g_k = torch.as_tensor(g, dtype=torch.float32, device=dev)
            hist.append((k, time.perf_counter() - t0, e_k))
            self._fit(batch, E, G)
        return dict(theta=theta.double().cpu().numpy(), energy=e_k, history=hist, reached=reached,
                    seconds=time.perf_counter() - t0, exact_calls=self.exact.calls - calls0,
                    exact_rows=self.exact.rows - rows0, rounds=rounds, fallbacks=fallbacks)
