"""Build ``quantum_amd/libtfqx.so`` (host C++ + gfx950 HIP kernels) with hipcc.

Usage: ``python -m quantum_amd.build_lib`` (or ``__graft_entry__.build()``).
Objects go to ``quantum_amd/_build/``; the library is written in-tree so it
travels with the repository.
"""

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtfqx.so")
ARCH = os.environ.get("TFQX_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["wire.cc", "textproto.cc", "circuit.cc", "plan.cc", "large.cc", "noisy.cc", "api.cc",
                "ps_ops.cc", "surrogate_api.cc"]
# The batch kernel's modes are separate translation units so they compile in
# parallel (each instantiates the register engine many times).
HIP_SOURCES = ["batch_mode_adjoint.hip", "batch_mode_expectation.hip", "batch_mode_other.hip",
               "batch_mode_ipgrad.hip", "batch_kernels.hip", "large_kernels.hip",
               "surrogate_kernels.hip", "noisy_kernels.hip", "noalloc.hip", "rebind_kernels.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # Packed FP32 (v_pk_*_f32) stays off for every device code object.
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          f"-I{CSRC}", f"-I{os.path.join(HERE, '..', 'include')}"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


# Device code: keep SimplifyCFG from sinking the stores of the register
# engine's per-qubit-position variants into one store through a PHI of
# addresses, which would demote the register arrays to scratch memory.
HIP_ONLY = ["-mllvm", "-simplifycfg-sink-common=false"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    lang = ["-x", "hip"] + HIP_ONLY if src.endswith(".hip") else []
    cmd = [_hipcc()] + COMMON + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    sources = HIP_SOURCES + HOST_SOURCES
    # Longest compiles first (the batch modes), so they start together.
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        results = list(ex.map(_compile, sources))
    if verbose:
        for (_, w), s in zip(results, sources):
            if w:
                print(f"[{s}]\n{w}", file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
