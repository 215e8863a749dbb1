"""Autotuner for the planner's fusion width and tile shapes.

The planner has process-wide knobs (``include/tfqx.h``):

* ``tfqx_set_fusion_width``: widest fused block (1..4 for the batched kernel;
  2 is TFQ's basic fuser).  On the streaming path 5 or 6 also fuses blocks of
  that many qubits and applies them as dense MFMA contractions.
* ``tfqx_set_tuning``: register engine level (0 off, 1 batched kernel, 2
  batched and streaming passes) and the streaming path's tile qubits for one-
  and two-state passes.

``autotune`` builds the caller's plan under each candidate setting, times
device launches of it with HIP events on the caller's stream, keeps the
fastest setting active and returns the timings.  Results are cached per
caller-supplied key, so a workload shape is tuned once per process.

No reference counterpart: TFQ's CPU path fixes qsim's basic fuser at width 2
and has no tiles.
"""

import time

from . import _lib

_CACHE = {}

# Batched kernel (circuits of up to 18 qubits): the fusion width and whether
# the register engine runs.
BATCH_CANDIDATES = (
    {"fusion_width": 2, "register_engine": 2},
    {"fusion_width": 3, "register_engine": 2},
    {"fusion_width": 1, "register_engine": 2},
    {"fusion_width": 2, "register_engine": 0},
)

# Streaming passes (wider states): tile qubits of one-state passes and the
# register engine on or off there.
STREAM_CANDIDATES = (
    {"fusion_width": 2, "register_engine": 2, "tile_single": 13},
    {"fusion_width": 2, "register_engine": 1, "tile_single": 13},
    {"fusion_width": 2, "register_engine": 2, "tile_single": 12},
    {"fusion_width": 2, "register_engine": 2, "tile_single": 14},
    {"fusion_width": 5, "register_engine": 2, "tile_single": 13},
    {"fusion_width": 6, "register_engine": 2, "tile_single": 13},
)

DEFAULT = {"fusion_width": 0, "register_engine": 2, "tile_single": 0, "tile_two": 0}


def apply(setting):
    """Make a candidate setting the planner's current one (missing keys take
    the defaults)."""
    s = dict(DEFAULT)
    s.update(setting)
    _lib.set_fusion_width(int(s["fusion_width"]))
    _lib.set_tuning(int(s["register_engine"]), int(s["tile_single"]), int(s["tile_two"]))


def _time_plan(plan, shape, dtype, steps, warmup, stream):
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(shape, dtype=dtype, device=dev)
    plan.upload(dev, stream)
    for _ in range(warmup):
        plan.launch(out, stream)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        plan.launch(out, stream)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def autotune(build_plan, candidates=BATCH_CANDIDATES, key=None, steps=5, warmup=1,
             stream=None):
    """Time ``build_plan()`` under each candidate setting and keep the fastest.

    ``build_plan()`` returns ``(plan, output_shape, output_dtype)`` for a plan
    built under the planner's current setting.  Returns a dict with the chosen
    setting (``best``), and per candidate its launch time in ms and host plan
    time in s (``table``).  With ``key``, a cached result is re-applied
    instead of timing again.
    """
    import torch
    if key is not None and key in _CACHE:
        apply(_CACHE[key]["best"])
        return _CACHE[key]
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    table = []
    try:
        for cand in candidates:
            apply(cand)
            t0 = time.perf_counter()
            plan, shape, dtype = build_plan()
            host_s = time.perf_counter() - t0
            ms = _time_plan(plan, shape, dtype, steps, warmup, stream)
            table.append({"setting": dict(cand), "ms": round(ms, 4), "plan_s": round(host_s, 3)})
            del plan
    except BaseException:
        apply(DEFAULT)
        raise
    best = min(table, key=lambda r: r["ms"])
    apply(best["setting"])
    result = {"best": dict(best["setting"]), "table": table}
    if key is not None:
        _CACHE[key] = result
    return result


def reset():
    """Forget cached results and restore the default setting."""
    _CACHE.clear()
    apply(DEFAULT)
