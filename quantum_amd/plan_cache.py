"""Plan reuse across op calls.

The reference parses, resolves and fuses every program on every call
(``core/ops/parse_context.cc:60-92``).  Here an op call whose programs, symbol
names and Pauli sums are byte-identical to a recent call's reuses that call's
plan: only the symbolic gate matrices are re-evaluated for the new symbol
values (on the device, ``tfqx_plan_stage_values``), and for the adjoint the
lambda weights for the new downstream gradients
(``tfqx_plan_rebind_downstream``); the upload then carries only what changed
(``tfqx_plan_upload_changed``).  A training loop therefore pays the host
planner once.

Keys hold a 128-bit xxh3 digest of the string tensors' bytes and offsets, the
symbol names, the shapes, the op's other arguments, the caller's stream and
the planner's settings generation (bumped by every fusion/tuning setter), so a
changed setting never reuses a plan built under the old one.  Plans on the
HBM-streaming path and noisy plans cannot be rebound and are not cached.

Entries are least-recently-used, at most ``set_capacity(n)`` of them (default
8; 0 disables the cache).  A per-entry lock serialises rebind + upload + launch
of one plan; launches of one plan stay ordered because the stream is part of
the key.
"""

import collections
import ctypes
import threading

import numpy as np
import xxhash

from .serialize import pack_strings

_LOCK = threading.Lock()
_CACHE = collections.OrderedDict()
_CAPACITY = 8
_GENERATION = [0]


def set_capacity(n):
    """Keep at most ``n`` plans (0 disables reuse and empties the cache)."""
    global _CAPACITY
    with _LOCK:
        _CAPACITY = max(0, int(n))
        while len(_CACHE) > _CAPACITY:
            _CACHE.popitem(last=False)


def clear():
    with _LOCK:
        _CACHE.clear()
        _MEMO.clear()


def bump_generation():
    """Planner settings changed: plans built before must not be reused."""
    with _LOCK:
        _GENERATION[0] += 1
        _CACHE.clear()


_MEMO_SIZE = 16
_MEMO = collections.OrderedDict()  # element ids -> (elements, digest, id array or None)


def _element_ids(arr):
    """The element ids of a C-contiguous object ndarray as uint64 (a copy of
    the array's PyObject pointer buffer, read in one step), or None."""
    if not (isinstance(arr, np.ndarray) and arr.dtype == object and arr.flags.c_contiguous
            and arr.size > 0 and ctypes.sizeof(ctypes.c_void_p) == 8):
        return None
    buf = (ctypes.c_uint64 * arr.size).from_address(arr.ctypes.data)
    return np.frombuffer(buf, dtype=np.uint64).copy()


def _tensor_digest(arr):
    """xxh3-128 of one string tensor (offsets, then bytes).  A training loop
    passes the same array of ``bytes`` objects on every call (139 MB for a
    4096-circuit batch); its digest is memoised on the identity of the
    elements.  The memo holds the elements themselves, so their ids cannot be
    reused by other objects while the entry lives, and ``bytes`` are
    immutable: equal ids mean equal contents.  For object ndarrays the ids
    are the array's pointer buffer, compared in bulk."""
    ids = _element_ids(arr)
    if ids is not None:
        ident = (arr.shape, xxhash.xxh3_128_digest(ids.tobytes()))
        elems = None
    else:
        elems = tuple(arr.reshape(-1)) if hasattr(arr, "reshape") else tuple(arr)
        ident = (getattr(arr, "shape", None), tuple(map(id, elems)))
    with _LOCK:
        hit = _MEMO.get(ident)
        if hit is not None and (ids is None or np.array_equal(hit[2], ids)):
            _MEMO.move_to_end(ident)
            return hit[1]
    if elems is None:
        elems = tuple(arr.reshape(-1))
    if not all(type(x) is bytes for x in elems):
        data, offsets = pack_strings(arr)
        h = xxhash.xxh3_128()
        h.update(offsets.tobytes())
        h.update(data)
        return h.digest()
    h = xxhash.xxh3_128()
    for x in elems:
        h.update(len(x).to_bytes(8, "little"))
        h.update(x)
    d = h.digest()
    with _LOCK:
        _MEMO[ident] = (elems, d, ids)
        while len(_MEMO) > _MEMO_SIZE:
            _MEMO.popitem(last=False)
    return d


def digest(*string_tensors):
    h = xxhash.xxh3_128()
    for arr in string_tensors:
        if arr is None:
            h.update(b"\x00none")
            continue
        h.update(_tensor_digest(arr))
    return h.digest()


class _Entry:
    __slots__ = ("plan", "lock")

    def __init__(self, plan):
        self.plan = plan
        self.lock = threading.Lock()


def run(key, build, rebind, launch):
    """``launch(plan)`` with a cached plan for ``key`` (rebound by
    ``rebind(plan)``) or a fresh one from ``build()``; returns launch's result."""
    if _CAPACITY == 0:
        return launch(build())
    key = key + (_GENERATION[0],)
    with _LOCK:
        entry = _CACHE.get(key)
        if entry is not None:
            _CACHE.move_to_end(key)
    if entry is not None:
        with entry.lock:
            rebind(entry.plan)
            return launch(entry.plan)
    plan = build()
    st = plan.stats()
    if st["streaming"] or getattr(plan, "noisy", False):
        return launch(plan)
    entry = _Entry(plan)
    with entry.lock:
        out = launch(plan)
    with _LOCK:
        _CACHE[key] = entry
        _CACHE.move_to_end(key)
        while len(_CACHE) > _CAPACITY:
            _CACHE.popitem(last=False)
    return out
