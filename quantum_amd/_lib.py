"""ctypes binding of ``libtfqx.so`` (the C-ABI declared in ``include/tfqx.h``).

The reference loads its native ops with ``load_op_library`` through a lazy
loader (``core/ops/load_module.py:26-85``).  Here the HIP library is loaded
lazily on first use; a missing library raises immediately -- there is no CPU
fallback.
"""

import ctypes
import os
import threading

import numpy as np

from .serialize import pack_strings

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtfqx.so")

TFQX_OK = 0
TFQX_INVALID_ARGUMENT = 3
TFQX_UNIMPLEMENTED = 12
TFQX_GRAD_TFQ_FD = 0
TFQX_GRAD_ANALYTIC = 1

_lock = threading.Lock()
_lib = None


class InvalidArgumentError(ValueError):
    """Raised for invalid op inputs (TF's ``InvalidArgumentError`` in TFQ)."""


class DeviceError(RuntimeError):
    pass


def _declare(lib):
    c = ctypes
    P = c.c_void_p
    i64 = c.c_int64
    i32 = c.c_int32
    u64 = c.c_uint64
    s = [P, P, i64, P, P, i64, P]  # programs, names, values
    lib.tfqx_plan_simulate_expectation.argtypes = s + [P, P, i64, P, c.c_char_p, i64]
    lib.tfqx_plan_simulate_sampled_expectation.argtypes = s + [P, P, i64, P, u64, P, c.c_char_p, i64]
    lib.tfqx_plan_simulate_samples.argtypes = s + [i32, u64, P, c.c_char_p, i64]
    lib.tfqx_plan_simulate_state.argtypes = s + [P, c.c_char_p, i64]
    lib.tfqx_plan_noisy_expectation.argtypes = s + [P, P, i64, P, u64, P, c.c_char_p, i64]
    lib.tfqx_plan_noisy_sampled_expectation.argtypes = s + [P, P, i64, P, u64, P, c.c_char_p, i64]
    lib.tfqx_plan_noisy_samples.argtypes = s + [i32, u64, P, c.c_char_p, i64]
    lib.tfqx_ps_decompose.argtypes = [P, P, i64, P, c.c_char_p, i64]
    lib.tfqx_ps_symbol_replace.argtypes = [P, P, i64, P, P, i64, P, P, i64, P, c.c_char_p, i64]
    lib.tfqx_ps_weights_from_symbols.argtypes = [P, P, i64, P, P, i64, P, i64, P, c.c_char_p, i64]
    for name in ("tfqx_ps_decompose", "tfqx_ps_symbol_replace", "tfqx_ps_weights_from_symbols"):
        getattr(lib, name).restype = c.c_int
    lib.tfqx_strings_count.argtypes = [P]
    lib.tfqx_strings_count.restype = i64
    lib.tfqx_strings_data.argtypes = [P]
    lib.tfqx_strings_data.restype = P
    lib.tfqx_strings_offsets.argtypes = [P]
    lib.tfqx_strings_offsets.restype = P
    lib.tfqx_strings_dim.argtypes = [P, i64]
    lib.tfqx_strings_dim.restype = i64
    lib.tfqx_strings_destroy.argtypes = [P]
    lib.tfqx_strings_destroy.restype = None
    lib.tfqx_plan_inner_product.argtypes = s + [P, P, i64, P, c.c_char_p, i64]
    lib.tfqx_plan_inner_product_grad.argtypes = s + [P, P, i64, P, i32, P, c.c_char_p, i64]
    lib.tfqx_plan_adjoint_gradient.argtypes = s + [P, P, i64, P, i32, P, c.c_char_p, i64]
    lib.tfqx_plan_expectation_and_gradient.argtypes = s + [P, P, i64, P, i32, P, c.c_char_p, i64]
    lib.tfqx_debug_fuse.argtypes = [P, i64, P, P, i64, P, i32, P, i64, P, i64, P, P,
                                    c.c_char_p, i64]
    lib.tfqx_debug_fuse.restype = c.c_int
    lib.tfqx_debug_philox_uniform.argtypes = [u64, u64, i64, c.c_uint32, c.c_uint32, P, P,
                                              c.c_char_p, i64]
    lib.tfqx_debug_philox_uniform.restype = c.c_int
    lib.tfqx_plan_rebind.argtypes = [P, P, c.c_char_p, i64]
    lib.tfqx_plan_rebind.restype = c.c_int
    lib.tfqx_plan_rebind_downstream.argtypes = [P, P, c.c_char_p, i64]
    lib.tfqx_plan_rebind_downstream.restype = c.c_int
    lib.tfqx_plan_stage_values.argtypes = [P, P, c.c_char_p, i64]
    lib.tfqx_plan_stage_values.restype = c.c_int
    lib.tfqx_plan_upload_changed.argtypes = [P, P, P, c.c_char_p, i64]
    lib.tfqx_plan_upload_changed.restype = c.c_int
    lib.tfqx_plan_stats.argtypes = [P, P, i64]
    lib.tfqx_plan_stats.restype = c.c_int
    for name in ("tfqx_plan_simulate_expectation", "tfqx_plan_simulate_sampled_expectation",
                 "tfqx_plan_simulate_samples", "tfqx_plan_simulate_state",
                 "tfqx_plan_noisy_expectation", "tfqx_plan_noisy_sampled_expectation",
                 "tfqx_plan_noisy_samples",
                 "tfqx_plan_inner_product", "tfqx_plan_inner_product_grad",
                 "tfqx_plan_adjoint_gradient", "tfqx_plan_expectation_and_gradient", "tfqx_plan_upload", "tfqx_plan_launch", "tfqx_plan_launch_plain",
                 "tfqx_graph_capture", "tfqx_graph_launch"):
        getattr(lib, name).restype = c.c_int
    for name in ("tfqx_plan_max_num_qubits", "tfqx_plan_workspace_bytes", "tfqx_plan_output_bytes"):
        getattr(lib, name).argtypes = [P]
        getattr(lib, name).restype = i64
    lib.tfqx_plan_upload.argtypes = [P, P, P, c.c_char_p, i64]
    lib.tfqx_plan_launch.argtypes = [P, P, P, P, c.c_char_p, i64]
    lib.tfqx_plan_launch_plain.argtypes = [P, P, P, P, c.c_char_p, i64]
    lib.tfqx_plan_launch_noalloc.argtypes = [P, P, P, c.c_char_p, i64]
    lib.tfqx_plan_launch_noalloc.restype = c.c_int
    lib.tfqx_graph_capture.argtypes = [P, P, P, P, P, c.c_char_p, i64]
    lib.tfqx_graph_launch.argtypes = [P, P, c.c_char_p, i64]
    lib.tfqx_graph_destroy.argtypes = [P]
    lib.tfqx_graph_destroy.restype = None
    lib.tfqx_plan_destroy.argtypes = [P]
    lib.tfqx_plan_destroy.restype = None
    lib.tfqx_set_tuning.argtypes = [i32, i32, i32]
    lib.tfqx_set_tuning.restype = None
    lib.tfqx_set_wave_min_circuits.argtypes = [i32]
    lib.tfqx_set_wave_min_circuits.restype = None
    lib.tfqx_set_streaming_min_qubits.argtypes = [i32]
    lib.tfqx_set_streaming_min_qubits.restype = None
    lib.tfqx_set_reverse_fusion.argtypes = [i32]
    lib.tfqx_set_reverse_fusion.restype = None
    lib.tfqx_set_fusion_width.argtypes = [i32]
    lib.tfqx_set_fusion_width.restype = None
    lib.tfqx_version.restype = c.c_char_p
    f32 = c.c_float
    E = [c.c_char_p, i64]
    lib.tfqx_mlp_layout.argtypes = [i32, i32, i32, P, i64]
    lib.tfqx_mlp_set_params.argtypes = [P, i32, i32, i32, P, P] + E
    lib.tfqx_mlp_predict.argtypes = [P, i32, i32, i32, P, i32, P, P, P] + E
    lib.tfqx_mlp_loss_grad.argtypes = [P, i32, i32, i32, P, P, P, i32, f32, P] + E
    lib.tfqx_mlp_train_step.argtypes = [P, i32, i32, i32, P, P, P, i32, f32, f32, i32, P] + E
    lib.tfqx_gemm_bf16.argtypes = [i32, i32, i32, P, i32, i32, P, i32, i32, P, i32, f32, f32, P] + E
    for name in ("tfqx_mlp_layout", "tfqx_mlp_set_params", "tfqx_mlp_predict", "tfqx_mlp_loss_grad",
                 "tfqx_mlp_train_step", "tfqx_gemm_bf16"):
        getattr(lib, name).restype = c.c_int
    return lib


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is not built; run `python -m quantum_amd.build_lib` "
                                  "(there is no CPU fallback).")
            _lib = _declare(ctypes.CDLL(LIB_PATH))
        return _lib


EXPORTED_SYMBOLS = (
    "tfqx_plan_simulate_expectation", "tfqx_plan_simulate_sampled_expectation",
    "tfqx_plan_simulate_samples", "tfqx_plan_simulate_state", "tfqx_plan_noisy_expectation",
    "tfqx_plan_noisy_sampled_expectation", "tfqx_plan_noisy_samples", "tfqx_ps_decompose",
    "tfqx_ps_symbol_replace", "tfqx_ps_weights_from_symbols", "tfqx_strings_count",
    "tfqx_strings_data", "tfqx_strings_offsets", "tfqx_strings_dim", "tfqx_strings_destroy",
    "tfqx_plan_inner_product",
    "tfqx_plan_inner_product_grad", "tfqx_plan_adjoint_gradient", "tfqx_plan_expectation_and_gradient", "tfqx_plan_rebind",
    "tfqx_plan_rebind_downstream", "tfqx_plan_stage_values", "tfqx_plan_upload_changed",
    "tfqx_debug_fuse", "tfqx_debug_philox_uniform",
    "tfqx_plan_stats",
    "tfqx_plan_max_num_qubits", "tfqx_plan_workspace_bytes",
    "tfqx_plan_output_bytes", "tfqx_plan_upload", "tfqx_plan_launch", "tfqx_plan_launch_plain",
    "tfqx_plan_launch_noalloc",
    "tfqx_graph_capture", "tfqx_graph_launch", "tfqx_graph_destroy", "tfqx_plan_destroy",
    "tfqx_set_fusion_width", "tfqx_set_wave_min_circuits", "tfqx_set_streaming_min_qubits", "tfqx_set_reverse_fusion", "tfqx_set_tuning", "tfqx_version", "tfqx_mlp_layout", "tfqx_mlp_set_params",
    "tfqx_mlp_predict", "tfqx_mlp_loss_grad", "tfqx_mlp_train_step", "tfqx_gemm_bf16")


def check(rc, err):
    if rc == TFQX_OK:
        return
    msg = err.value.decode("utf-8", "replace")
    if rc == TFQX_INVALID_ARGUMENT:
        raise InvalidArgumentError(msg)
    if rc == TFQX_UNIMPLEMENTED:
        raise NotImplementedError(msg)
    raise DeviceError(f"tfqx error {rc}: {msg}")


class StringTensor:
    """A flattened string tensor kept alive for the duration of a call."""

    def __init__(self, arr):
        self.bytes, self.offsets = pack_strings(arr)
        self._buf = ctypes.create_string_buffer(self.bytes, max(1, len(self.bytes)))
        self.n = len(self.offsets) - 1

    @property
    def ptr(self):
        return ctypes.addressof(self._buf)

    @property
    def off(self):
        return self.offsets.ctypes.data


def np_ptr(a):
    return a.ctypes.data if a.size else None


class Plan:
    """Owns a ``tfqx_plan*``: host parse/fuse result plus its device workspace."""

    def __init__(self, builder, *args):
        self._lib = lib()
        self.handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(4096)
        rc = builder(*args, ctypes.byref(self.handle), err, 4096)
        check(rc, err)
        self.workspace = None
        # The workspace holds this plan's last upload: later uploads copy only
        # what changed since (tfqx_plan_upload_changed).
        self._resident = False

    @property
    def max_num_qubits(self):
        return int(self._lib.tfqx_plan_max_num_qubits(self.handle))

    def stats(self):
        out = (ctypes.c_int64 * 10)()
        self._lib.tfqx_plan_stats(self.handle, out, 10)
        keys = ("circuits", "forward_ops", "forward_passes", "max_qubits", "adjoint_ops",
                "adjoint_passes", "streaming", "workspace_bytes", "flops", "mfma_flops")
        return dict(zip(keys, [int(x) for x in out]))

    def op_mix(self):
        """Op mix of a batch-kernel plan in amplitude-op units (2^n per op of
        an n-qubit circuit, / 1024); zeros for streaming and noisy plans."""
        out = (ctypes.c_int64 * 24)()
        self._lib.tfqx_plan_stats(self.handle, out, 24)
        keys = ("forward", "rev_remap", "rev_diag", "rev_dense1", "rev_dense2", "rev_param1",
                "rev_param2", "rev_param_diag", "fwd_remap", "fwd_diag", "fwd_dense1",
                "fwd_dense2", "remap_wave_local", "remap_engine")
        return dict(zip(keys, [int(x) for x in out[10:]]))

    def rebind(self, values):
        """New symbol values (float32 [batch, n_symbols]) for a batch-kernel plan
        (not streaming or noisy); upload again before launching."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        err = ctypes.create_string_buffer(4096)
        check(self._lib.tfqx_plan_rebind(self.handle, v.ctypes.data if v.size else None, err, 4096),
              err)

    def stage_values(self, values):
        """New symbol values (float32 [batch, n_symbols]) for a batch-kernel plan,
        evaluated into the plan's matrices on the device at the next upload
        (``rebind_kernel``); the host matrices are not recomputed."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        err = ctypes.create_string_buffer(4096)
        check(self._lib.tfqx_plan_stage_values(self.handle, v.ctypes.data if v.size else None,
                                               err, 4096), err)

    def rebind_downstream(self, grads):
        """New downstream gradients (float32 [batch, n_ops]) for an adjoint plan of
        the batch kernel; upload again before launching."""
        g = np.ascontiguousarray(grads, dtype=np.float32)
        err = ctypes.create_string_buffer(4096)
        check(self._lib.tfqx_plan_rebind_downstream(self.handle, g.ctypes.data if g.size else None,
                                                    err, 4096), err)

    @property
    def workspace_bytes(self):
        return int(self._lib.tfqx_plan_workspace_bytes(self.handle))

    def upload(self, device, stream):
        import torch
        nbytes = max(256, self.workspace_bytes)
        if self.workspace is None or self.workspace.numel() < nbytes:
            self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=device)
            self._resident = False
        err = ctypes.create_string_buffer(4096)
        fn = self._lib.tfqx_plan_upload_changed if self._resident else self._lib.tfqx_plan_upload
        check(fn(self.handle, ctypes.c_void_p(self.workspace.data_ptr()), ctypes.c_void_p(stream),
                 err, 4096), err)
        self._resident = True

    def launch(self, out, stream, plain=False):
        err = ctypes.create_string_buffer(4096)
        fn = self._lib.tfqx_plan_launch_plain if plain else self._lib.tfqx_plan_launch
        check(fn(self.handle, ctypes.c_void_p(self.workspace.data_ptr()),
                 ctypes.c_void_p(out.data_ptr() if out.numel() else 0),
                 ctypes.c_void_p(stream), err, 4096), err)

    def launch_noalloc(self, out, stream):
        """Upload and launch with the library's static device scratch (no
        workspace; see include/tfqx.h: tfqx_plan_launch_noalloc)."""
        err = ctypes.create_string_buffer(4096)
        check(self._lib.tfqx_plan_launch_noalloc(
            self.handle, ctypes.c_void_p(out.data_ptr() if out.numel() else 0),
            ctypes.c_void_p(stream), err, 4096), err)

    def capture_graph(self, out, stream):
        err = ctypes.create_string_buffer(4096)
        g = ctypes.c_void_p()
        check(self._lib.tfqx_graph_capture(self.handle, ctypes.c_void_p(self.workspace.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream),
                                           ctypes.byref(g), err, 4096), err)
        return Graph(self._lib, g)

    def __del__(self):
        try:
            if self.handle:
                self._lib.tfqx_plan_destroy(self.handle)
                self.handle = ctypes.c_void_p()
        except Exception:
            pass


class Graph:
    def __init__(self, lib_, handle):
        self._lib = lib_
        self.handle = handle

    def launch(self, stream):
        err = ctypes.create_string_buffer(4096)
        check(self._lib.tfqx_graph_launch(self.handle, ctypes.c_void_p(stream), err, 4096), err)

    def __del__(self):
        try:
            if self.handle:
                self._lib.tfqx_graph_destroy(self.handle)
                self.handle = ctypes.c_void_p()
        except Exception:
            pass


def fusion_decisions(program, symbol_names=(), symbol_values=(), width=2):
    """The planner's fused ops for one serialized program: a list of
    (bits tuple, controls_outside, gate index tuple)."""
    L = lib()
    names = StringTensor(np.array(list(symbol_names), dtype=object))
    vals = np.ascontiguousarray(np.asarray(symbol_values, dtype=np.float32).reshape(-1))
    buf = ctypes.create_string_buffer(program, max(1, len(program)))
    nops, ngates = ctypes.c_int64(), ctypes.c_int64()
    err = ctypes.create_string_buffer(4096)
    args = [ctypes.addressof(buf), len(program), names.ptr, names.off, names.n, np_ptr(vals),
            int(width)]
    check(L.tfqx_debug_fuse(*args, None, 0, None, 0, ctypes.byref(nops), ctypes.byref(ngates),
                            err, 4096), err)
    ops = np.zeros((max(1, nops.value), 8), np.int32)
    gates = np.zeros(max(1, ngates.value), np.int32)
    check(L.tfqx_debug_fuse(*args, ops.ctypes.data, nops.value, gates.ctypes.data, ngates.value,
                            ctypes.byref(nops), ctypes.byref(ngates), err, 4096), err)
    out = []
    for r in ops[:nops.value]:
        bits = tuple(int(b) for b in r[1:1 + r[0]])
        out.append((bits, bool(r[5]), tuple(int(g) for g in gates[r[6]:r[7]])))
    return out


def philox_uniforms(seed, count, stream_lo, stream_hi, start=0):
    """The device sampler's FP64 uniforms for shots [start, start + count) of
    one stream (a torch float64 tensor on the current device)."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(int(count), dtype=torch.float64, device=dev)
    err = ctypes.create_string_buffer(4096)
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(lib().tfqx_debug_philox_uniform(int(seed) & (2 ** 64 - 1), int(start), int(count),
                                          int(stream_lo) & 0xFFFFFFFF,
                                          int(stream_hi) & 0xFFFFFFFF,
                                          ctypes.c_void_p(out.data_ptr() if count else 0),
                                          ctypes.c_void_p(stream), err, 4096), err)
    return out


def set_wave_min_circuits(count=-1):
    """Smallest count of 4-10 qubit circuits run one per wavefront (negative:
    the default 1024; 0: always)."""
    lib().tfqx_set_wave_min_circuits(int(count))
    _settings_changed()


def set_streaming_min_qubits(num_qubits=-1):
    """Width from which expectation, adjoint and state plans take the
    HBM-streaming path instead of the batched kernel's tiles (14-19; anything
    else restores the default 19)."""
    lib().tfqx_set_streaming_min_qubits(int(num_qubits))
    _settings_changed()


def set_reverse_fusion(on=True):
    """Adjoint plans fuse the reverse sweep across parameterised gates (on,
    the default) or split it at each of them (off, the reference's gradient
    circuit)."""
    lib().tfqx_set_reverse_fusion(1 if on else 0)
    _settings_changed()


def set_tuning(register_engine=2, tile_single=0, tile_two=0):
    """Planner tuning for later plans (see include/tfqx.h: tfqx_set_tuning).
    register_engine: 0 off, 1 batched kernel, 2 batched and streaming (default)."""
    lib().tfqx_set_tuning(int(register_engine), int(tile_single), int(tile_two))
    _settings_changed()


def set_fusion_width(width):
    lib().tfqx_set_fusion_width(int(width))
    _settings_changed()


def _settings_changed():
    from . import plan_cache
    plan_cache.bump_generation()
