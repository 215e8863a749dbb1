"""Benchmark: batched expectation + adjoint gradient on the uneven 4096-circuit
batch (4-16 qubits, 10-60 moments) in one persistent launch replayed from a
hipGraph, against the same kernel launched once per circuit, plus the HBM
roofline of the 30-qubit streaming passes.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  A step is one launch over the whole batch with the
device program already resident in HBM (host parse/fuse and upload happen
before timing).  `roofline` is the 30-qubit gate pass (config: 5x6 grid,
14 cycles): algorithmic bytes 16 B x 2^30 per pass / average pass time,
measured with HIP events on the launch stream.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("expectation+adjoint-grad circuits/s, batched vs plain GPU, err≤1e-5; HBM GB/s")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_uneven(count):
    from quantum_amd import serialize as S
    from quantum_amd import workloads as W
    t = time.time()
    circuits, names, values, psums, weights = W.config4_uneven(count=count)
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(psums)
    log(f"built {count} uneven circuits in {time.time() - t:.1f}s")
    return progs, names, values, ps, weights


def time_graph(plan, out, stream, steps, warmup):
    import torch
    g = plan.capture_graph(out, stream)
    s = torch.cuda.current_stream()
    for _ in range(warmup):
        g.launch(stream)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(s)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch(stream)
    end.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(end) / 1e3, wall, g


def time_launch(fn, stream, reps=1):
    import torch
    s = torch.cuda.current_stream()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(s)
    for _ in range(reps):
        fn()
    end.record(s)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3 / reps


def hbm_roofline(stream):
    """Forward passes of the 30-qubit grid circuit; returns roofline dict."""
    import torch
    from quantum_amd import serialize as S
    from quantum_amd import tfq_simulate_ops as sim
    from quantum_amd import workloads as W
    circuit, _ = W.config2_grid()
    progs = S.convert_to_tensor([circuit])
    empty = S.convert_to_tensor([[__import__("quantum_amd.circuits", fromlist=["x"]).PauliSum([])]])
    plan, shape = sim.plan_expectation(progs, [], np.zeros((1, 0), np.float32), empty)
    st = plan.stats()
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    plan.launch(out, stream)  # warm-up
    t = time_launch(lambda: plan.launch(out, stream), stream, reps=3)
    passes = st["forward_passes"]
    per_pass = t / passes
    bytes_per_pass = 16.0 * 2 ** 30
    achieved = bytes_per_pass / per_pass / 1e9
    log(f"30q grid: {passes} passes, {t * 1e3:.2f} ms forward, {per_pass * 1e3:.3f} ms/pass, "
        f"{achieved:.0f} GB/s")
    del plan
    torch.cuda.empty_cache()
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "pass_kernel (tile pass, 30-qubit state)", "passes": passes,
            "ms_per_pass": round(per_pass * 1e3, 4), "bytes_per_pass": bytes_per_pass}


def cpu_baseline(progs, names, values, ps, weights, gpu_exp, gpu_grad, budget_s):
    """The oracle (CPU restatement, one thread) on the first circuits of the
    batch until `budget_s` is spent; also the max error of the GPU rows."""
    from oracle import tfq_oracle as O
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(1)
    except Exception:
        limiter = None
    t0 = time.perf_counter()
    done = 0
    err = 0.0
    while done < len(progs) and time.perf_counter() - t0 < budget_s:
        i = done
        e = O.simulate_expectation([progs[i]], names, values[i:i + 1], ps[i:i + 1])
        g = O.adjoint_gradient([progs[i]], names, values[i:i + 1], ps[i:i + 1], weights[i:i + 1])
        err = max(err, float(np.abs(e - gpu_exp[i:i + 1]).max()),
                  float(np.abs(g - gpu_grad[i:i + 1]).max()))
        done += 1
    dt = time.perf_counter() - t0
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    return done, dt, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--circuits", type=int, default=4096)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-roofline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: this build has no multi-GPU path (each box has one "
                 "MI355X); run with --gpus 1.")
    import torch
    from quantum_amd import tfq_adj_grad_op as adj

    assert torch.cuda.is_available(), "needs a HIP device"
    dev = torch.device("cuda", 0)
    # A dedicated (default-priority) stream: graph capture is not allowed on
    # the legacy default stream.
    side = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(side)
    stream = side.cuda_stream

    progs, names, values, ps, weights = build_uneven(args.circuits)
    t = time.time()
    plan, shape = adj.plan_adj_grad(progs, names, values, ps, weights, with_expectation=True)
    log(f"host plan (parse+resolve+fuse) {time.time() - t:.2f}s; stats {plan.stats()}")
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    torch.cuda.synchronize()

    B, O, S = len(progs), ps.shape[1], len(names)
    elapsed, wall, graph = time_graph(plan, out, stream, args.steps, args.warmup)
    value = B * args.steps / elapsed
    log(f"batched: {elapsed / args.steps * 1e3:.3f} ms/step ({value:.0f} circuits/s)")
    res = out.detach().cpu().numpy()
    gpu_exp = res[:B * O].reshape(B, O)
    gpu_grad = res[B * O:].reshape(B, S)

    out_plain = torch.empty_like(out)
    t_plain = time_launch(lambda: plan.launch(out_plain, stream, plain=True), stream, reps=1)
    plain_value = B / t_plain
    same = torch.equal(out_plain, out)
    log(f"plain (one launch per circuit): {t_plain * 1e3:.1f} ms ({plain_value:.0f} circuits/s),"
        f" identical={same}")

    roof = None
    if not args.no_roofline:
        del graph
        roof = hbm_roofline(stream)

    done, cpu_dt, err = cpu_baseline(progs, names, values, ps, weights, gpu_exp, gpu_grad,
                                     args.cpu_budget)
    log(f"cpu oracle: {done} circuits in {cpu_dt:.1f}s; max |gpu - oracle| = {err:.2e}")

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "circuits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic",
        "config": {"workload": "uneven batch: 4096 circuits, 4-16 qubits, 10-60 moments of "
                               "{H,X^t,Y^t,Z^t,CNOT,CZ^t,ISWAP^t,FSim}, 30% symbolic, 1-8 "
                               "Pauli strings (weight<=4); expectation + adjoint gradient",
                   "circuits": B, "symbols": S, "ops_per_circuit": O,
                   "launch": "one persistent kernel per step, replayed from a hipGraph"},
        "plain_gpu_circuits_per_s": round(plain_value, 1),
        "batched_vs_plain": round(value / plain_value, 2),
        "max_abs_err_vs_oracle": err,
        "oracle_rows_checked": done,
        "roofline": roof,
        "cpu_baseline": {"value": round(done / cpu_dt, 3), "unit": "circuits/s", "cores": 1,
                         "kind": "port",
                         "sample": f"first {done} circuits of the same batch (expectation + "
                                   f"adjoint gradient), complex128 numpy oracle, 1 thread, "
                                   f"{cpu_dt:.1f}s"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
