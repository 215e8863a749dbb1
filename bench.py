"""Benchmark: batched expectation + adjoint gradient on the uneven 4096-circuit
batch (4-16 qubits, 10-60 moments) in one persistent launch replayed from a
hipGraph, against the same kernel launched once per circuit, plus the HBM
roofline of the 30-qubit streaming passes.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  A step is one launch over the whole batch with the
device program already resident in HBM (host parse/fuse and upload happen
before timing).

`roofline` is the batch kernel (the only kernel of a step): its algorithmic
real flops per launch (from the plan: (8*2^k - 2) per amplitude per dense
k-qubit op, 6 per diagonal op, 8 per Pauli term, twice per op in the adjoint
sweep, plus the derivative products) divided by its average launch time,
measured with HIP events on the launch stream, against the FP32 peak.
`traffic` is the kernel's HBM bytes per launch from rocprofv3 FETCH_SIZE and
WRITE_SIZE counters, read from profiles/batch_kernel_uneven4096_traffic.json
when that file matches this workload, else null.  `hbm_pass` is the 30-qubit
streaming gate pass (5x6 grid, 14 cycles): 16 B x 2^30 per pass / average
pass time, against 8 TB/s.  `inner_product_grad` times the tfq.math
inner-product gradient op on the 64 x 10-qubit rotation batch with two paired
programs per row (graph replays), spot-checked against the oracle.
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
FP32_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "batch_kernel_uneven4096_traffic.json")


def host_threads():
    """Host threads for the CPU legs: every core this process may run on,
    capped by OMP_NUM_THREADS when the machine sets it (a machine that shares
    its cores among GPUs sets it to one GPU's share, e.g. 16)."""
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(cores, int(omp))) if omp.isdigit() and int(omp) > 0 else cores


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


def hbm_roofline(stream, tune=False):
    """Forward passes of the 30-qubit grid circuit; returns roofline dict
    (with ``tune``, the autotuner first picks the streaming settings)."""
    import torch
    from quantum_amd import autotune as AT
    from quantum_amd import serialize as S
    from quantum_amd import tfq_simulate_ops as sim
    from quantum_amd import workloads as W
    circuit, _ = W.config2_grid()
    progs = S.convert_to_tensor([circuit])
    empty = S.convert_to_tensor([[__import__("quantum_amd.circuits", fromlist=["x"]).PauliSum([])]])

    def build():
        plan, shape = sim.plan_expectation(progs, [], np.zeros((1, 0), np.float32), empty)
        return plan, shape, torch.float32

    tuned = None
    if tune:
        tuned = AT.autotune(build, AT.STREAM_CANDIDATES, steps=2, warmup=1, stream=stream)
        log(f"autotune (30q streaming): {tuned['best']} from "
            f"{[(r['setting'], r['ms']) for r in tuned['table']]}")
    bytes_per_pass = 16.0 * 2 ** 30
    dev = torch.device("cuda", torch.cuda.current_device())

    def measure():
        plan, shape, _ = build()
        st = plan.stats()
        out = torch.empty(shape, dtype=torch.float32, device=dev)
        plan.upload(dev, stream)
        plan.launch(out, stream)  # warm-up
        t = time_launch(lambda: plan.launch(out, stream), stream, reps=3)
        passes = st["forward_passes"]
        del plan
        torch.cuda.empty_cache()
        return t, passes, st["mfma_flops"]

    t, passes, mfma = measure()
    per_pass = t / passes
    achieved = bytes_per_pass / per_pass / 1e9
    log(f"30q grid: {passes} passes, {t * 1e3:.2f} ms forward, {per_pass * 1e3:.3f} ms/pass, "
        f"{achieved:.0f} GB/s")
    tile13 = None
    if tune and tuned["best"].get("tile_single", 13) != 13:
        # The same passes on 13-qubit tiles (two workgroups per CU), for comparison.
        AT.apply(dict(tuned["best"], tile_single=13))
        t13, p13, _ = measure()
        tile13 = {"passes": p13, "forward_ms": round(t13 * 1e3, 3),
                  "ms_per_pass": round(t13 / p13 * 1e3, 4),
                  "achieved": round(bytes_per_pass / (t13 / p13) / 1e9, 1),
                  "frac": round(bytes_per_pass / (t13 / p13) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"30q grid on 13-qubit tiles: {p13} passes, {t13 * 1e3:.2f} ms forward, "
            f"{t13 / p13 * 1e3:.3f} ms/pass")
    if tune:
        AT.apply(AT.DEFAULT)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "kernel": "pass_kernel (tile pass, 30-qubit state)", "passes": passes,
            "ms_per_pass": round(per_pass * 1e3, 4), "bytes_per_pass": bytes_per_pass,
            "forward_ms": round(t * 1e3, 3),
            "mfma_tflops_over_forward": round(mfma / t / 1e12, 3),
            "tile13": tile13,
            "note": "per-pass figures of the autotuned setting (fewer, longer passes on 14-qubit "
                    "tiles when those finish the forward sooner); tile13: the same circuit on "
                    "13-qubit tiles",
            "autotune": tuned}


def inner_product_grad_measure(stream, steps=20, warmup=3, oracle_rows=2):
    """tfq.math inner-product gradient on the 64 x 10-qubit rotation batch
    (600 symbols), two symbol-free paired programs per row: one launch per
    step from a hipGraph."""
    import torch
    from quantum_amd import _lib
    from quantum_amd import circuits as C
    from quantum_amd import serialize as S
    from quantum_amd import workloads as W
    from quantum_amd.tfq_simulate_ops import _common, _plan_args
    circuits, names, values, _ = W.config1_line_rotations()
    rng = np.random.default_rng(2024)
    q = C.LineQubit.range(10)
    others = np.array([[S.serialize_circuit(W.random_circuit(rng, q, 40) +
                                            C.Circuit([C.I(x) for x in q]))
                        for _ in range(2)] for _ in circuits], dtype=object)
    g = rng.normal(size=(len(circuits), 2)).astype(np.float32)
    progs, pnames, pvalues = _common(S.convert_to_tensor(circuits), names, values)
    P, N, args = _plan_args(progs, pnames, pvalues)
    Ot = _lib.StringTensor(others)
    plan = _lib.Plan(_lib.lib().tfqx_plan_inner_product_grad, *args, Ot.ptr, Ot.off, 2,
                     g.ctypes.data, 0)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty((len(circuits), len(names)), dtype=torch.complex64, device=dev)
    plan.upload(dev, stream)
    elapsed, _, graph = time_graph(plan, out, stream, steps, warmup)
    per = elapsed / steps
    flops = plan.stats()["flops"]
    res = out.cpu().numpy()
    err = None
    if oracle_rows:
        from oracle import tfq_oracle as O
        ref = O.inner_product_gradient(list(progs[:oracle_rows]), names, values[:oracle_rows],
                                       others[:oracle_rows], g[:oracle_rows])
        err = float(np.abs(res[:oracle_rows] - ref).max())
    log(f"inner-product gradient: {per * 1e3:.3f} ms per 64-circuit launch, "
        f"{len(circuits) / per:.0f} circuits/s, {flops / per / 1e12:.2f} TFLOP/s, "
        f"max |gpu - oracle| = {err}")
    del graph, plan
    return {"workload": "64 circuits x 10 qubits (20 layers rx ry rz + CZ ladder, 600 symbols), "
                        "2 symbol-free paired programs per row",
            "circuits": len(circuits), "ms_per_launch": round(per * 1e3, 4),
            "circuits_per_s": round(len(circuits) / per, 1),
            "flops_per_launch": flops, "tflops": round(flops / per / 1e12, 3),
            "max_abs_err_vs_oracle": err, "oracle_rows_checked": oracle_rows}


def ops_measure(stream, steps=20, warmup=3, shots=1000):
    """Each op of the family on the 64 x 10-qubit rotation batch (600
    symbols, observables [sum Z_i Z_i+1, X_0]): one launch per step from a
    hipGraph, with 2 rows checked against the oracle (sampled ops against the
    oracle's replica of the same Philox stream)."""
    import torch
    from oracle import tfq_oracle as O
    from quantum_amd import _lib
    from quantum_amd import circuits as C
    from quantum_amd import serialize as S
    from quantum_amd import workloads as W
    from quantum_amd.tfq_simulate_ops import _common, _plan_args
    circuits, names, values, ops = W.config1_line_rotations()
    progs, pnames, pvalues = _common(S.convert_to_tensor(circuits), names, values)
    ps = S.convert_to_tensor(ops)
    B, O_ = len(circuits), ps.shape[1]
    rng = np.random.default_rng(2024)
    q = C.LineQubit.range(10)
    others = np.array([[S.serialize_circuit(W.random_circuit(rng, q, 40) +
                                            C.Circuit([C.I(x) for x in q]))
                        for _ in range(2)] for _ in circuits], dtype=object)
    ns = np.full((B, O_), shots, np.int32)
    seed = 777
    L = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    res = {}

    def run(name, make, shape, dtype, check):
        P, N, args = _plan_args(progs, pnames, pvalues)
        plan = make(args)
        out = torch.empty(shape, dtype=dtype, device=dev)
        plan.upload(dev, stream)
        elapsed, _, graph = time_graph(plan, out, stream, steps, warmup)
        per = elapsed / steps
        err = check(out.cpu().numpy())
        res[name] = {"ms_per_launch": round(per * 1e3, 4), "circuits_per_s": round(B / per, 1),
                     "max_abs_err_vs_oracle_2_rows": err}
        log(f"{name}: {per * 1e3:.3f} ms per {B}-circuit launch, {B / per:.0f} circuits/s, "
            f"err {err}")
        del graph, plan

    Ps = _lib.StringTensor(ps)
    Ot = _lib.StringTensor(others)
    two = list(progs[:2])
    run("expectation",
        lambda a: _lib.Plan(L.tfqx_plan_simulate_expectation, *a, Ps.ptr, Ps.off, O_),
        (B, O_), torch.float32,
        lambda r: float(np.abs(r[:2] - O.simulate_expectation(two, names, values[:2], ps[:2])).max()))
    from quantum_amd import tfq_adj_grad_op as A
    g = np.random.default_rng(2025).uniform(-1, 1, size=(B, O_)).astype(np.float32)
    run("adjoint_gradient",
        lambda a: A.plan_adj_grad(progs, pnames, pvalues, ps, g)[0], (B, len(names)),
        torch.float32,
        lambda r: float(np.abs(r[:2] - O.adjoint_gradient(two, names, values[:2], ps[:2],
                                                          g[:2])).max()))
    run("state", lambda a: _lib.Plan(L.tfqx_plan_simulate_state, *a), (B, 1024), torch.complex64,
        lambda r: float(np.abs(r[:2] - O.simulate_state(two, names, values[:2])).max()))
    run("inner_product",
        lambda a: _lib.Plan(L.tfqx_plan_inner_product, *a, Ot.ptr, Ot.off, 2), (B, 2),
        torch.complex64,
        lambda r: float(np.abs(r[:2] - O.inner_product(two, names, values[:2], others[:2])).max()))
    run("samples",
        lambda a: _lib.Plan(L.tfqx_plan_simulate_samples, *a, shots, seed), (B, shots, 10),
        torch.int8,
        lambda r: float(np.mean(r[:2] != O.simulate_samples(two, names, values[:2], shots, seed))))
    run("sampled_expectation",
        lambda a: _lib.Plan(L.tfqx_plan_simulate_sampled_expectation, *a, Ps.ptr, Ps.off, O_,
                            ns.ctypes.data, seed), (B, O_), torch.float32,
        lambda r: float(np.abs(r[:2] - O.simulate_sampled_expectation(
            two, names, values[:2], ps[:2], ns[:2], seed)).max()))
    res["workload"] = ("64 circuits x 10 qubits (20 layers rx ry rz + CZ ladder, 600 symbols); "
                       "observables [sum Z_i Z_i+1, X_0]; adjoint gradient: 600 symbols with "
                       "random upstream weights; inner product: 2 paired programs; "
                       f"samples / sampled expectation: {shots} shots (samples: fraction of "
                       "differing bits vs the oracle's replica)")
    return res


def noisy_measure(stream, steps=10, warmup=2, trajectories=2000, cpu_rows=8, cpu_threads=16,
                  check=True):
    """tfq.noise expectation on 32 noisy 8-qubit circuits (8 layers of ry + CZ
    ladder, depolarize + amplitude damping on every qubit after each layer),
    ``trajectories`` per (circuit, observable): one persistent launch per step
    from a hipGraph.  The CPU leg runs the float32 C++ trajectory loop
    (oracle/cpu_tfq.cc, same Philox draws) on the first ``cpu_rows`` rows."""
    import torch
    from quantum_amd import noise
    from quantum_amd import serialize as S
    from quantum_amd import workloads as W
    circuits, names, values, ops = W.noisy_layers()
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(ops)
    B = len(circuits)
    ns = np.full((B, 2), trajectories, np.int32)
    seed = 20240
    plan, shape = noise.plan_expectation(progs, names, values, ps, ns, seed=seed)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    elapsed, _, graph = time_graph(plan, out, stream, steps, warmup)
    per = elapsed / steps
    total = B * trajectories
    flops = plan.stats()["flops"]
    res = out.cpu().numpy()
    log(f"noisy expectation: {per * 1e3:.3f} ms per launch of {total} trajectories, "
        f"{total / per:.3e} trajectories/s, {flops / per / 1e12:.2f} TFLOP/s")
    result = {"workload": "32 circuits x 8 qubits, 8 layers of ry(symbol) + CZ ladder, each followed "
                          "by depolarize(0.01) and amplitude_damp(0.02) on every qubit; "
                          "observables [sum Z_i Z_i+1, X_0]",
              "trajectories_per_launch": total, "ms_per_launch": round(per * 1e3, 4),
              "trajectories_per_s": round(total / per, 1),
              "circuits_per_s": round(B / per, 2),
              "roofline": {"bound": "mfma", "achieved": round(flops / per / 1e12, 3),
                           "peak": 157.3, "unit": "TFLOP/s",
                           "frac": round(flops / per / 1e12 / 157.3, 4),
                           "flops_per_launch": flops,
                           "kernel": "noisy_kernel (wave-mode register engine, FP32)"}}
    del graph, plan
    if check:
        from oracle import cpu_tfq
        from oracle import noise_oracle as N
        low = cpu_tfq.LoweredNoisy(list(progs[:cpu_rows]), names, values[:cpu_rows],
                                   ps[:cpu_rows])
        t0 = time.perf_counter()
        cpu = cpu_tfq.run_noisy(low, ns[:cpu_rows], seed, threads=cpu_threads)
        cpu_s = time.perf_counter() - t0
        exact = N.density_expectation(list(progs[:1]), names, values[:1], ps[:1])
        result["max_abs_diff_vs_cpu_port"] = float(np.abs(res[:cpu_rows] - cpu).max())
        result["max_abs_err_vs_density_matrix_row0"] = float(np.abs(res[:1] - exact).max())
        result["cpu_baseline"] = {
            "value": round(cpu_rows * trajectories / cpu_s, 1), "unit": "trajectories/s",
            "cores": cpu_threads, "kind": "port",
            "sample": f"first {cpu_rows} circuits x {trajectories} trajectories of the same batch; "
                      f"float32 C++ trajectory loop of TFQ's noisy CPU schedule "
                      f"(oracle/cpu_tfq.cc), {cpu_threads} threads, {cpu_s:.2f}s"}
        log(f"noisy: max |gpu - cpu port| = {result['max_abs_diff_vs_cpu_port']:.2e}, "
            f"|gpu - density matrix| (row 0) = {result['max_abs_err_vs_density_matrix_row0']:.3f}, "
            f"cpu {result['cpu_baseline']['value']:.0f} trajectories/s")
    return result


def _time_plan(plan, out, stream, reps):
    """Average seconds per launch of an uploaded plan (one warm-up launch)."""
    plan.launch(out, stream)
    return time_launch(lambda: plan.launch(out, stream), stream, reps=reps)


def sampling_measure(stream, forward_ms=None, shots=10 ** 6, seed=4321, cpu=True, cpu_gates=24):
    """The 30-qubit grid circuit (5x6 GridQubits, 14 cycles of sqrt-X/Y/W +
    FSim(pi/2, pi/6)): tfq_simulate_samples at 1e6 shots and the 49-term sum
    Z_i Z_j over grid edges as a sampled expectation with 1e6 shots per term.

    Each launch re-simulates the circuit, so the sampler's own time is the
    launch minus the forward passes (``forward_ms``, measured by the HBM leg
    on the same circuit).  The sampler's algorithmic HBM bytes are one read
    of the 8 GiB state (8 B x 2^30) plus its output."""
    import torch
    from quantum_amd import _lib
    from quantum_amd import serialize as S
    from quantum_amd import workloads as W
    from quantum_amd.tfq_simulate_ops import _common, _plan_args
    circuit, zz = W.config2_grid()
    progs, names, values = _common(S.convert_to_tensor([circuit]), [], np.zeros((1, 0)))
    ps = S.convert_to_tensor([[zz]])
    L = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    res = {"workload": "30-qubit 5x6 grid circuit, 14 cycles of random sqrt-X/sqrt-Y/sqrt-W + "
                       "FSim(pi/2, pi/6); samples: 1e6 shots; sampled expectation: 49-term "
                       "sum Z_i Z_j over grid edges, 1e6 shots per term",
           "shots": shots}
    P, N, args = _plan_args(progs, names, values)
    plan = _lib.Plan(L.tfqx_plan_simulate_samples, *args, shots, seed)
    out = torch.empty((1, shots, 30), dtype=torch.int8, device=dev)
    plan.upload(dev, stream)
    t = _time_plan(plan, out, stream, reps=3)
    bits = out[0]
    valid = bool(((bits == 0) | (bits == 1)).all())
    del plan, out
    torch.cuda.empty_cache()
    samp = {"ms_per_launch": round(t * 1e3, 3), "all_bits_valid": valid}
    if forward_ms:
        st = t * 1e3 - forward_ms
        bytes_ = 8.0 * 2 ** 30 + shots * 30
        samp.update({"sampler_ms": round(st, 3),
                     "sampler_roofline": {"bound": "hbm", "achieved": round(bytes_ / st / 1e6, 1),
                                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                          "frac": round(bytes_ / st / 1e6 / HBM_PEAK_GBS, 4),
                                          "bytes": bytes_}})
    res["samples"] = samp
    log(f"C2 samples: {t * 1e3:.2f} ms per launch (1e6 shots), sampler "
        f"{samp.get('sampler_ms')} ms")
    P, N, args = _plan_args(progs, names, values)
    Ps = _lib.StringTensor(ps)
    ns = np.full((1, 1), shots, np.int32)
    plan = _lib.Plan(L.tfqx_plan_simulate_sampled_expectation, *args, Ps.ptr, Ps.off, 1,
                     ns.ctypes.data, seed)
    out = torch.empty((1, 1), dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    t2 = _time_plan(plan, out, stream, reps=2)
    est = float(out.cpu()[0, 0])
    del plan, out
    P, N, args = _plan_args(progs, names, values)
    plan = _lib.Plan(L.tfqx_plan_simulate_expectation, *args, Ps.ptr, Ps.off, 1)
    out = torch.empty((1, 1), dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    plan.launch(out, stream)
    exact = float(out.cpu()[0, 0])
    del plan, out
    torch.cuda.empty_cache()
    se = {"ms_per_launch": round(t2 * 1e3, 3), "estimate": est, "exact": exact,
          "abs_err": abs(est - exact), "five_sigma_bound": 5 * 7.0 / 1e3}
    if forward_ms:
        st = t2 * 1e3 - forward_ms
        se["sampler_ms"] = round(st, 3)
        se["shots_per_s"] = round(49 * shots / st * 1e3, 1)
    res["sampled_expectation"] = se
    log(f"C2 sampled expectation: {t2 * 1e3:.2f} ms per launch, estimate {est:.5f} vs exact "
        f"{exact:.5f}")
    if cpu:
        from oracle import cpu_tfq
        cores = host_threads()
        low = cpu_tfq.Lowered(list(progs), names, values, ps)
        fused = int(low.circ[0, 4] - low.circ[0, 3])
        k = min(cpu_gates, fused)
        def best_of(fn, reps):
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return min(ts)

        # State set-up (|0>, 8 GiB) is timed on its own, best of two, and
        # subtracted: the sampler alone is small beside its run-to-run spread.
        t_init = best_of(lambda: cpu_tfq.sample_large(low, 0, seed, threads=cores,
                                                      gate_limit=0), 2)
        t_k = best_of(lambda: cpu_tfq.sample_large(low, 0, seed, threads=cores, gate_limit=k), 1)
        t_s = max(0.0, best_of(lambda: cpu_tfq.sample_large(low, shots, seed, threads=cores,
                                                            gate_limit=0), 2) - t_init)
        per_gate = (t_k - t_init) / max(1, k)
        full = t_init + per_gate * fused + t_s
        res["cpu_baseline"] = {
            "value": round(full * 1e3, 1), "unit": "ms per samples call (extrapolated)",
            "cores": cores, "kind": "port",
            "per_gate_ms": round(per_gate * 1e3, 2), "sampler_ms": round(t_s * 1e3, 1),
            "sample": f"ComputeLarge schedule of oracle/cpu_tfq.cc (all threads on one 2^30 "
                      f"float32 state): the first {k} of {fused} basic-fused gates timed and "
                      f"extrapolated per gate, plus qsim's sampler at {shots} shots; "
                      f"{cores} threads"}
        log(f"C2 cpu: {per_gate * 1e3:.1f} ms per fused gate, sampler {t_s * 1e3:.0f} ms, "
            f"full samples call ~{full:.1f} s on {cores} threads")
    return res


def adjoint_large_measure(stream, cpu=True, cpu_layers=1):
    """C3: 26-qubit hardware-efficient ansatz (8 layers of ry, rz per qubit +
    CZ ring, 416 symbols), batch 8 parameter sets, transverse-field Ising
    observable (25 ZZ + 0.7 * 26 X = 51 terms): tfq_adj_grad in one call.

    HBM accounting per parameter set: 16 B x 2^26 per forward pass, 32 B x
    2^26 per reverse pass (psi and lambda together), and lambda's build as one
    read + one write of the state (8 + 8 B x 2^26; its per-term gathers are
    cache reuse)."""
    import torch
    from quantum_amd import serialize as S
    from quantum_amd import tfq_adj_grad_op as A
    from quantum_amd import workloads as W
    circuit, names, values, ham = W.config3_hea()
    B = values.shape[0]
    progs = S.convert_to_tensor([circuit] * B)
    ps = S.convert_to_tensor([[ham]] * B)
    ones = np.ones((B, 1), np.float32)
    plan, shape = A.plan_adj_grad(progs, names, values, ps, ones)
    st = plan.stats()
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    t = _time_plan(plan, out, stream, reps=2)
    grads = out.cpu().numpy()
    del plan, out
    torch.cuda.empty_cache()
    N = 2 ** 26
    fwd = st["forward_passes"] // B if st["forward_passes"] >= B else st["forward_passes"]
    rev = st["adjoint_passes"] // B if st["adjoint_passes"] >= B else st["adjoint_passes"]
    bytes_set = fwd * 16.0 * N + rev * 32.0 * N + 16.0 * N
    achieved = bytes_set * B / t / 1e9
    res = {"workload": "26-qubit HEA, 8 layers ry/rz + CZ ring (416 symbols), batch 8, 51-term "
                       "TFIM (sum ZZ + 0.7 sum X), TFQ gradient gates",
           "ms_per_launch": round(t * 1e3, 2), "param_sets": B,
           "grads_per_s": round(B / t, 2), "forward_passes_per_set": fwd,
           "reverse_passes_per_set": rev,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "bytes_per_param_set": bytes_set},
           "grad_abs_max": float(np.abs(grads).max())}
    log(f"C3 adjoint: {t * 1e3:.1f} ms per launch of {B} x 416 gradients ({B / t:.2f} grads/s), "
        f"{fwd} forward + {rev} reverse passes per set, {achieved:.0f} GB/s")
    if cpu:
        from oracle import cpu_tfq
        cores = host_threads()
        low = cpu_tfq.Lowered(list(progs[:1]), names, values[:1], ps[:1])
        G = int(low.circ[0, 2] - low.circ[0, 1])
        k = min(G, 78 * cpu_layers)
        t0 = time.perf_counter()
        cpu_tfq.run_large(low, 1, downstream=ones[:1], threads=cores, gate_limit=0, fwd_limit=0)
        t_lam = time.perf_counter() - t0
        t0 = time.perf_counter()
        cpu_tfq.run_large(low, 1, downstream=ones[:1], threads=cores, gate_limit=k, fwd_limit=k)
        t_k = time.perf_counter() - t0
        full = t_lam + (t_k - t_lam) * G / k
        res["cpu_baseline"] = {
            "value": round(1.0 / full, 4), "unit": "grads/s (one parameter set, extrapolated)",
            "cores": cores, "kind": "port", "seconds_per_grad": round(full, 2),
            "sample": f"ComputeLarge schedule of oracle/cpu_tfq.cc (all threads on 2^26 float32 "
                      f"states): lambda (51 terms) plus the first {k} forward and last {k} "
                      f"reverse of {G} gates (one ansatz layer), extrapolated to all gates; "
                      f"{cores} threads"}
        log(f"C3 cpu: {full:.1f} s per gradient ({cores} threads)")
    return res


def surrogate_measure(stream, train_batch=1024, steps=20):
    """C5's network: 144 -> 512 -> 512 -> 512 -> 1 SiLU, BF16 MFMA / FP32
    accumulate: one Sobolev training step at B = train_batch and one
    inference + input-gradient call at B = 1 and B = 32, against the BF16
    dense peak."""
    import torch
    from quantum_amd import surrogate as SG
    dev = torch.device("cuda", torch.cuda.current_device())
    P, H = 144, 512
    model = SG.MLPSurrogate(P, hidden=H, max_batch=train_batch, seed=0)
    rng = np.random.default_rng(5)
    X = torch.tensor(rng.uniform(-1, 1, (train_batch, P)), dtype=torch.float32, device=dev)
    E = torch.tensor(rng.normal(size=train_batch), dtype=torch.float32, device=dev)
    G = torch.tensor(rng.normal(size=(train_batch, P)), dtype=torch.float32, device=dev)
    for _ in range(3):
        model.train_step(X, E, G, 1.0, 1e-4)
    t = time_launch(lambda: model.train_step(X, E, G, 1.0, 1e-4), stream, reps=steps)
    flops = 2.0 * train_batch * (5 * P * H + 12 * H * H)
    res = {"network": "144-512-512-512-1 SiLU, BF16 weights, FP32 accumulate (600,065 params)",
           "train_step_ms": round(t * 1e3, 4), "train_batch": train_batch,
           "roofline": {"bound": "mfma", "achieved": round(flops / t / 1e12, 3), "peak": 2500.0,
                        "unit": "TFLOP/s", "frac": round(flops / t / 1e12 / 2500.0, 5),
                        "flops_per_step": flops}}
    import ctypes
    err = ctypes.create_string_buffer(4096)
    for b in (1, 32, 64):
        xb = X[:b].contiguous()
        Eb = torch.empty(b, dtype=torch.float32, device=dev)
        Gb = torch.empty((b, P), dtype=torch.float32, device=dev)
        args = (model._ws(), P, H, train_batch, ctypes.c_void_p(xb.data_ptr()), b,
                ctypes.c_void_p(Eb.data_ptr()), ctypes.c_void_p(Gb.data_ptr()),
                ctypes.c_void_p(stream), err, 4096)
        model.lib.tfqx_mlp_predict(*args)
        # Back-to-back C-ABI calls on preallocated buffers: the device time of
        # one predict (B <= 32: the fused single-launch kernel; B = 64: layer
        # by layer, 14 launches).
        td = time_launch(lambda: model.lib.tfqx_mlp_predict(*args), stream, reps=200)
        res[f"predict_with_grad_device_ms_b{b}"] = round(td * 1e3, 4)
        if b <= 32:
            model.predict(xb)
            tp = time_launch(lambda: model.predict(xb), stream, reps=steps)
            res[f"predict_with_grad_call_ms_b{b}"] = round(tp * 1e3, 4)
    res["predict_note"] = ("device_ms: back-to-back C-ABI calls with preallocated outputs; "
                           "call_ms: MLPSurrogate.predict from Python (allocates the outputs)")
    log(f"C5 training step {t * 1e3:.3f} ms ({flops / t / 1e12:.1f} TFLOP/s); predict device "
        f"b1 {res['predict_with_grad_device_ms_b1']} ms, b32 "
        f"{res['predict_with_grad_device_ms_b32']} ms, b64 (layer by layer) "
        f"{res['predict_with_grad_device_ms_b64']} ms")
    return res


def vqe_measure(stream, steps=3000, ref_steps=4000, seed=1005):
    """C5 end to end: 12-qubit TFIM VQE (6 layers ry/rz + CZ ring, 144 symbols,
    H = sum ZZ + sum X).  Wall-clock until the energy is within 1e-3
    (relative) of E_ref, the energy exact-gradient Adam converges to in
    ``ref_steps`` steps, for exact Adam and for the two surrogate-assisted
    loops (each allowed ``steps`` steps); a loop that does not get there
    reports null."""
    import torch
    from quantum_amd import surrogate as SG
    from quantum_amd import workloads as W
    circuit, names, ham = W.config5_vqe()
    exact = SG.ExactEnergy(circuit, names, ham)
    rng = np.random.default_rng(seed)
    theta0 = rng.uniform(-0.1, 0.1, len(names))
    ref = SG.vqe_exact(exact, theta0, steps=ref_steps)
    e_ref = ref["energy"]
    target = e_ref + 1e-3 * abs(e_ref)
    res = {"workload": "12-qubit TFIM VQE, 6 layers ry/rz + CZ ring (144 symbols), H = sum ZZ + "
                       "sum X; Adam lr 0.02 from theta0 ~ U(-0.1, 0.1)",
           "e_ref": e_ref, "e_target": target, "ref_steps": ref_steps}
    runs = {"exact_adam": lambda: SG.vqe_exact(exact, theta0, steps=steps, e_target=target)}
    model = SG.MLPSurrogate(len(names), hidden=512, max_batch=1024, seed=0)
    runs["surrogate_vqe"] = lambda: SG.SurrogateVQE(model, exact).run(theta0, steps=steps,
                                                                     e_target=target)
    model2 = SG.MLPSurrogate(len(names), hidden=512, max_batch=1024, seed=1)
    runs["candidate_vqe"] = lambda: SG.CandidateVQE(model2, exact).run(theta0, steps=steps,
                                                                      e_target=target)
    for name, fn in runs.items():
        r = fn()
        res[name] = {"reached_step": r["reached"][0] if r["reached"] else None,
                     "seconds_to_target": round(r["reached"][1], 3) if r["reached"] else None,
                     "final_energy": r["energy"], "seconds": round(r["seconds"], 3),
                     "exact_calls": r["exact_calls"]}
        log(f"C5 {name}: reached {res[name]['reached_step']} in {res[name]['seconds_to_target']} s,"
            f" final {r['energy']:.5f}")
    ex = res["exact_adam"]["seconds_to_target"]
    best = [res[k]["seconds_to_target"] for k in ("surrogate_vqe", "candidate_vqe")
            if res[k]["seconds_to_target"] is not None]
    res["surrogate_beats_exact"] = bool(best and ex and min(best) < ex)
    return res


def op_call_rate(progs, names, values, ps, weights, calls=5):
    """End-to-end rate of the op a training loop calls on the uneven batch:
    ``expectation_and_gradient`` (host plan + upload + launch, until the
    results are on the device).  The first call plans; later calls with new
    symbol values and downstream gradients reuse the plan (rebind + upload +
    launch)."""
    import torch
    from quantum_amd import plan_cache
    from quantum_amd import tfq_adj_grad_op as adj
    plan_cache.clear()
    rng = np.random.default_rng(7)
    B = len(progs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    adj.expectation_and_gradient(progs, names, values, ps, weights)
    torch.cuda.synchronize()
    cold = time.perf_counter() - t0
    times = []
    for _ in range(calls):
        v = (values + rng.normal(scale=0.01, size=values.shape)).astype(np.float32)
        g = (weights * rng.uniform(0.5, 1.5, size=weights.shape)).astype(np.float32)
        t0 = time.perf_counter()
        adj.expectation_and_gradient(progs, names, v, ps, g)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    warm = float(np.median(times))
    plan_cache.clear()
    log(f"op calls: first {cold * 1e3:.0f} ms ({B / cold:.0f} circuits/s), reused plan "
        f"{warm * 1e3:.1f} ms ({B / warm:.0f} circuits/s)")
    return {"op": "expectation_and_gradient (tfq_adj_grad with the forward pass)",
            "first_call_ms": round(cold * 1e3, 1), "first_call_circuits_per_s": round(B / cold, 1),
            "reused_plan_call_ms": round(warm * 1e3, 2),
            "reused_plan_circuits_per_s": round(B / warm, 1), "calls_timed": calls,
            "note": "wall clock per call until the outputs are on the device; reused-plan calls "
                    "change every symbol value and downstream gradient"}


def plain_best_measure(progs, names, values, ps, weights, stream, stride=4, chunk=256):
    """Unbatched comparator: circuits of the uneven batch each in a plan of
    their own (what a caller passing one circuit gets), launched alone on the
    faster of their single-circuit paths -- one workgroup of the batched kernel
    (LDS-resident to 13-14 qubits, tiled through scratch above), or, for 14-16
    qubits, the HBM-streaming passes, each spread over 2^(n-13) workgroups.
    Device time per launch from HIP events around it; host planning is not
    counted.  Every ``stride``-th circuit is timed (widths and depths are drawn
    independently per circuit, so the sample has the batch's mix) and the sum
    is scaled to the whole batch."""
    import torch
    from quantum_amd import _lib
    from quantum_amd import tfq_adj_grad_op as adj
    dev = torch.device("cuda", 0)
    s = torch.cuda.current_stream()
    B = len(progs)
    widths = {}

    def timed(rows):
        res = {}
        for lo in range(0, len(rows), chunk):
            pend = []
            for i in rows[lo:lo + chunk]:
                plan, shp = adj.plan_adj_grad(progs[i:i + 1], names, values[i:i + 1],
                                              ps[i:i + 1], weights[i:i + 1],
                                              with_expectation=True)
                widths[i] = plan.stats()["max_qubits"]
                out = torch.empty(shp, dtype=torch.float32, device=dev)
                plan.upload(dev, stream)
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(s)
                plan.launch(out, stream)
                b.record(s)
                pend.append((i, a, b, plan, out))
            torch.cuda.synchronize()
            for i, a, b, _, _ in pend:
                res[i] = a.elapsed_time(b)
        return res

    rows = list(range(0, B, stride))
    timed(rows[:32])  # load and warm the kernels of the batched path
    one_wg = timed(rows)
    wide_rows = [i for i in rows if widths[i] >= 14]
    _lib.set_streaming_min_qubits(14)
    try:
        timed(wide_rows[:8])
        wide = timed(wide_rows)
    finally:
        _lib.set_streaming_min_qubits(-1)
    best = {i: min(t, wide.get(i, t)) for i, t in one_wg.items()}
    chose_stream = sum(1 for i in wide if wide[i] < one_wg[i])
    scale = B / len(rows)
    total_ms = sum(best.values()) * scale
    one_ms = sum(one_wg.values()) * scale
    log(f"plain (best single-circuit path, {len(rows)} of {B} circuits timed): {total_ms:.1f} ms "
        f"({B / total_ms * 1e3:.0f} circuits/s); one-workgroup plans alone {one_ms:.1f} ms; "
        f"streaming faster for {chose_stream} of {len(wide)} circuits of 14-16 qubits")
    return {"ms": round(total_ms, 2), "circuits_per_s": round(B / total_ms * 1e3, 1),
            "one_workgroup_plans_ms": round(one_ms, 2), "circuits_timed": len(rows),
            "wide_circuits_timed": len(wide), "wide_streaming_faster": chose_stream,
            "note": f"every {stride}th circuit in its own plan and launch on its faster "
                    f"single-circuit path (batched kernel, or streaming passes for 14-16 "
                    f"qubits), summed device time scaled to the batch; host planning excluded"}


def measured_traffic(circuits, flops):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if it
    was taken on this same workload (same circuit count and flop count)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("circuits") != circuits or t.get("flops") != flops:
        return None
    return t.get("bytes_per_launch")


def cpu_baseline(progs, names, values, ps, weights, gpu_exp, gpu_grad, sample, check):
    """TFQ's CPU (qsim) schedule restated in float32 C++ (oracle/cpu_tfq.cc) on
    the host's cores (``host_threads``), on the first `sample`
    circuits of the batch: expectation (basic fusion) + adjoint gradient.
    Lowering the protos to gate lists (Python) is not timed.  Also returns the
    max |GPU - oracle| over the first `check` circuits (complex128 oracle) and
    the max |GPU - CPU baseline| over the sample."""
    from oracle import cpu_tfq
    from oracle import tfq_oracle as O
    cores = host_threads()
    k = min(sample, len(progs))
    low = cpu_tfq.Lowered(list(progs[:k]), names, values[:k], ps[:k])
    cpu_tfq.run(low, 0, threads=cores)  # warm-up (page-in, thread start)
    t0 = time.perf_counter()
    e = cpu_tfq.run(low, 0, threads=cores)
    g = cpu_tfq.run(low, 1, downstream=weights[:k], threads=cores)
    dt = time.perf_counter() - t0
    err_cpu = max(float(np.abs(e - gpu_exp[:k]).max()), float(np.abs(g - gpu_grad[:k]).max()))
    c = min(check, len(progs))
    eo = O.simulate_expectation(list(progs[:c]), names, values[:c], ps[:c])
    go = O.adjoint_gradient(list(progs[:c]), names, values[:c], ps[:c], weights[:c])
    err = max(float(np.abs(eo - gpu_exp[:c]).max()), float(np.abs(go - gpu_grad[:c]).max()))
    return k, dt, cores, err, c, err_cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--circuits", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=512,
                    help="circuits of the batch timed on the CPU baseline")
    ap.add_argument("--oracle-check", type=int, default=32,
                    help="circuits checked against the complex128 oracle")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (the --only profiling runs never tune: they use "
                         "the planner defaults)")
    ap.add_argument("--only", choices=["all", "batch", "hbm", "copy", "ipgrad", "noisy", "ops",
                                       "c2", "c3", "c5", "vqe", "opcalls",
                                       "plain"],
                    default="all",
                    help="profiling runs: 'batch' launches only the batched kernel (graph "
                         "replays, no plain launches, no CPU baseline); 'hbm' only the 30-qubit "
                         "streaming passes; 'copy' a 1 GiB device copy (counter calibration); "
                         "'ipgrad' the inner-product gradient batch; 'c2' the 30-qubit samples "
                         "and sampled expectation; 'c3' the 26-qubit adjoint; 'c5' the "
                         "surrogate network's kernels; 'vqe' the surrogate-assisted VQE; "
                         "'opcalls' the end-to-end op call rate on the uneven batch; 'plain' the "
                         "per-circuit comparator")
    ap.add_argument("--fusion-width", type=int, default=None,
                    help="planner fusion width for the run (0 = default; 5 or 6 puts dense "
                         "MFMA passes on the streaming path, e.g. with --only hbm)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines of the C2/C3 "
                    "legs")
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

    if args.fusion_width is not None:
        from quantum_amd import _lib
        _lib.set_fusion_width(args.fusion_width)
    if args.only == "copy":
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        t = time_launch(lambda: dst.copy_(src), stream, reps=3)
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                          "ms": round(t * 1e3, 4), "GB/s": round(2 * (1 << 30) / t / 1e9, 1)}),
              flush=True)
        return
    if args.only == "hbm":
        print(json.dumps({"hbm_pass": hbm_roofline(stream)}), flush=True)
        return
    if args.only == "ipgrad":
        print(json.dumps({"inner_product_grad": inner_product_grad_measure(stream)}), flush=True)
        return
    if args.only == "ops":
        print(json.dumps({"ops": ops_measure(stream)}), flush=True)
        return
    if args.only == "c2":
        hbm = hbm_roofline(stream)
        print(json.dumps({"sampling": sampling_measure(
            stream, forward_ms=hbm["forward_ms"], cpu=not args.no_cpu)}),
            flush=True)
        return
    if args.only == "c3":
        print(json.dumps({"adjoint_26q": adjoint_large_measure(stream, cpu=not args.no_cpu)}),
              flush=True)
        return
    if args.only == "c5":
        print(json.dumps({"surrogate": surrogate_measure(stream)}), flush=True)
        return
    if args.only == "vqe":
        print(json.dumps({"vqe": vqe_measure(stream)}), flush=True)
        return
    if args.only == "noisy":
        print(json.dumps({"noisy": noisy_measure(stream, steps=args.steps, warmup=args.warmup)}),
              flush=True)
        return
    progs, names, values, ps, weights = build_uneven(args.circuits)
    if args.only == "opcalls":
        print(json.dumps({"op_calls": op_call_rate(progs, names, values, ps, weights)}),
              flush=True)
        return
    if args.only == "plain":
        print(json.dumps({"plain_best": plain_best_measure(progs, names, values, ps, weights,
                                                           stream)}), flush=True)
        return
    tuned = None
    if args.only == "all" and not args.no_autotune:
        from quantum_amd import autotune as AT

        def build():
            p, shp = adj.plan_adj_grad(progs, names, values, ps, weights, with_expectation=True)
            return p, shp, torch.float32

        tuned = AT.autotune(build, AT.BATCH_CANDIDATES, steps=3, warmup=1, stream=stream)
        log(f"autotune (batch): {tuned['best']} from "
            f"{[(r['setting'], r['ms']) for r in tuned['table']]}")
    t = time.time()
    plan, shape = adj.plan_adj_grad(progs, names, values, ps, weights, with_expectation=True)
    log(f"host plan (parse+resolve+fuse) {time.time() - t:.2f}s; stats {plan.stats()}")
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    torch.cuda.synchronize()

    B, O, S = len(progs), ps.shape[1], len(names)
    st = plan.stats()
    flops, mfma_flops = st["flops"], st["mfma_flops"]
    elapsed, wall, graph = time_graph(plan, out, stream, args.steps, args.warmup)
    value = B * args.steps / elapsed
    log(f"batched: {elapsed / args.steps * 1e3:.3f} ms/step ({value:.0f} circuits/s)")
    res = out.detach().cpu().numpy()
    gpu_exp = res[:B * O].reshape(B, O)
    gpu_grad = res[B * O:].reshape(B, S)

    if args.only == "batch":
        print(json.dumps({"batch_ms_per_launch": round(elapsed / args.steps * 1e3, 4),
                          "flops_per_launch": flops, "circuits": B}), flush=True)
        return
    out_plain = torch.empty_like(out)
    t_plain = time_launch(lambda: plan.launch(out_plain, stream, plain=True), stream, reps=1)
    plain_value = B / t_plain
    # Gradients are sums of FP64 atomics whose order varies between launches,
    # so the two paths agree to rounding, not bit for bit.
    plain_diff = float((out_plain - out).abs().max())
    log(f"plain (one launch per circuit): {t_plain * 1e3:.1f} ms ({plain_value:.0f} circuits/s),"
        f" max |plain - batched| = {plain_diff:.2e}")
    plain_best = plain_best_measure(progs, names, values, ps, weights, stream)

    per_launch = elapsed / args.steps
    tflops = flops / per_launch / 1e12
    log(f"batch kernel: {flops / 1e12:.3f} TFLOP per launch, {tflops:.2f} TFLOP/s")
    roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(B, flops),
            "kernel": "batch_kernel<adjoint> (persistent queue over LDS-resident, tiled and "
                      "wave-mode circuits; FP32 MFMA and VALU)",
            "flops_per_launch": flops, "ms_per_launch": round(per_launch * 1e3, 4),
            "peak_note": "MI355X FP32 peak (vector and matrix: 157.3 TFLOP/s)",
            # The two-qubit products that run on v_mfma_f32_4x4x1f32: their share of
            # the algorithmic flops, and their rate over the whole launch time (a
            # lower bound on the MFMA pipe's rate while busy).
            "mfma": {"flops_per_launch": mfma_flops,
                     "share_of_flops": round(mfma_flops / flops, 4) if flops else None,
                     "achieved_tflops_over_launch": round(mfma_flops / per_launch / 1e12, 3),
                     "frac_of_fp32_peak": round(mfma_flops / per_launch / 1e12 /
                                                FP32_PEAK_TFLOPS, 4)}}
    hbm = None
    if not args.no_roofline:
        del graph
        batch_setting = tuned["best"] if tuned else None
        hbm = hbm_roofline(stream, tune=not args.no_autotune)
        if batch_setting:
            from quantum_amd import autotune as AT
            AT.apply(batch_setting)
    op_calls = op_call_rate(progs, names, values, ps, weights)
    ipg = inner_product_grad_measure(stream)
    noisy = noisy_measure(stream)
    other_ops = ops_measure(stream)
    # The sampling leg runs on the default 13-qubit tiles: subtract that forward.
    fwd_ms = None
    if hbm:
        fwd_ms = hbm["tile13"]["forward_ms"] if hbm.get("tile13") else hbm["forward_ms"]
    sampling = sampling_measure(stream, forward_ms=fwd_ms, cpu=not args.no_cpu)
    adj26 = adjoint_large_measure(stream, cpu=not args.no_cpu)
    surrogate = surrogate_measure(stream)
    vqe = vqe_measure(stream)

    done, cpu_dt, cores, err, checked, err_cpu = cpu_baseline(
        progs, names, values, ps, weights, gpu_exp, gpu_grad, args.cpu_sample, args.oracle_check)
    log(f"cpu baseline: {done} circuits in {cpu_dt:.2f}s on {cores} threads; "
        f"max |gpu - oracle| = {err:.2e} ({checked} circuits); max |gpu - cpu| = {err_cpu:.2e}")

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
        "autotune": tuned,
        "plain_gpu_circuits_per_s": round(plain_value, 1),
        "batched_vs_plain": round(value / plain_value, 2),
        "plain_note": "plain = the same plan launched one circuit at a time, one workgroup "
                      "each; plain_best = each circuit alone on its faster single-circuit path",
        "plain_best": plain_best,
        "batched_vs_plain_best": round(value / plain_best["circuits_per_s"], 2),
        "max_abs_diff_plain_vs_batched": plain_diff,
        "max_abs_err_vs_oracle": err,
        "oracle_rows_checked": checked,
        "max_abs_err_vs_cpu_baseline": err_cpu,
        "op_calls": op_calls,
        "roofline": roof,
        "hbm_pass": hbm,
        "inner_product_grad": ipg,
        "noisy": noisy,
        "ops": other_ops,
        "sampling_30q": sampling,
        "adjoint_26q": adj26,
        "surrogate": surrogate,
        "vqe": vqe,
        "cpu_baseline": {"value": round(done / cpu_dt, 3), "unit": "circuits/s", "cores": cores,
                         "kind": "port",
                         "sample": f"first {done} circuits of the same batch, expectation (basic "
                                   f"fusion) + adjoint gradient; float32 C++ restatement of "
                                   f"TFQ's qsim schedule (oracle/cpu_tfq.cc), {cores} threads, "
                                   f"{cpu_dt:.2f}s; proto lowering not timed"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
