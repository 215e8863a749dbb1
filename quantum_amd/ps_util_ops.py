"""Parameter-shift program rewriting ops (``core/ops/tfq_ps_util_ops.py``):
``tfq_ps_decompose``, ``tfq_ps_symbol_replace`` and
``tfq_ps_weights_from_symbols``, run natively in ``libtfqx.so`` (host code;
``core/ops/tfq_ps_decompose_op.cc``, ``tfq_ps_symbol_replace_op.cc``,
``tfq_ps_weights_from_symbols_op.cc``).

Programs go in as serialized strings (or ``Circuit`` objects) and come back as
object arrays of bytes; weights as float32 arrays.
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import InvalidArgumentError, StringTensor, check
from .tfq_simulate_ops import as_string_tensor


def _strings(handle):
    L = _lib.lib()
    try:
        n = L.tfqx_strings_count(handle)
        shape = []
        i = 0
        while True:
            d = L.tfqx_strings_dim(handle, i)
            if d < 0:
                break
            shape.append(int(d))
            i += 1
        off = np.ctypeslib.as_array(ctypes.cast(L.tfqx_strings_offsets(handle),
                                                ctypes.POINTER(ctypes.c_int64)), shape=(n + 1,))
        total = int(off[-1])
        data = ctypes.string_at(L.tfqx_strings_data(handle), total) if total else b""
        out = np.empty(n, dtype=object)
        for k in range(n):
            out[k] = data[off[k]:off[k + 1]]
        return out.reshape(shape)
    finally:
        L.tfqx_strings_destroy(handle)


def _symbols(x, name):
    arr = np.asarray(x, dtype=object) if not isinstance(x, np.ndarray) else x
    if arr.ndim != 1:
        raise InvalidArgumentError(f"{name} must be rank 1. Got rank {arr.ndim}.")
    return np.array([s.encode("utf-8") if isinstance(s, str) else bytes(s) for s in arr],
                    dtype=object)


def tfq_ps_decompose(programs):
    """Symbolic PhasedISwap / ISwap / PhasedX / FSim gates decomposed into
    single-symbol eigen gates (``tfq_ps_decompose_op.cc:39-160``)."""
    progs = as_string_tensor(programs, 1, "programs")
    P = StringTensor(progs)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(4096)
    check(_lib.lib().tfqx_ps_decompose(P.ptr, P.off, P.n, ctypes.byref(h), err, 4096), err)
    return _strings(h)


def tfq_ps_symbol_replace(programs, symbols, replacement_symbols):
    """[batch, n_symbols, n_occurrences] programs, each with one occurrence of
    a symbol renamed (``tfq_ps_symbol_replace_op.cc:39-185``)."""
    progs = as_string_tensor(programs, 1, "programs")
    syms = _symbols(symbols, "symbols")
    reps = _symbols(replacement_symbols, "replacement_symbols")
    P, S, R = StringTensor(progs), StringTensor(syms), StringTensor(reps)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(4096)
    check(_lib.lib().tfqx_ps_symbol_replace(P.ptr, P.off, P.n, S.ptr, S.off, S.n, R.ptr, R.off,
                                            R.n, ctypes.byref(h), err, 4096), err)
    return _strings(h)


def tfq_ps_weights_from_symbols(programs, symbols):
    """[batch, n_symbols, n_occurrences] float32 exponent scalars of each
    symbol's gates, zero padded (``tfq_ps_weights_from_symbols_op.cc:39-170``)."""
    progs = as_string_tensor(programs, 1, "programs")
    syms = _symbols(symbols, "symbols")
    P, S = StringTensor(progs), StringTensor(syms)
    L = _lib.lib()
    err = ctypes.create_string_buffer(4096)
    pad = ctypes.c_int64()
    check(L.tfqx_ps_weights_from_symbols(P.ptr, P.off, P.n, S.ptr, S.off, S.n, None, 0,
                                         ctypes.byref(pad), err, 4096), err)
    out = np.zeros((progs.shape[0], len(syms), pad.value), np.float32)
    check(L.tfqx_ps_weights_from_symbols(P.ptr, P.off, P.n, S.ptr, S.off, S.n,
                                         out.ctypes.data if out.size else None, out.size,
                                         ctypes.byref(pad), err, 4096), err)
    return out
