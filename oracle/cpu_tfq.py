"""Driver of ``oracle/cpu_tfq.cc``: TFQ's CPU (qsim) schedule restated in
float32 C++ with std::threads -- the reported CPU baseline of ``bench.py``.

TEST AND MEASUREMENT INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
Programs are lowered with the oracle's parser (``tfq_oracle.Program``): gate
matrices in complex64, TFQ's float32 finite-difference gradient gates
(``adj_util.cc:175-302``), Pauli sums resolved to (bit, type).  The
expectation path applies qsim-style basic fusion (single-qubit gates
absorbed into the next two-qubit gate on their qubit); the adjoint path
applies gates one by one, as ``tfq_adj_grad_op.cc:214-276`` does between
parameterised gates.
"""

import ctypes
import os
import subprocess

import numpy as np

from . import tfq_oracle as T

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_tfq.cc")
LIB = os.path.join(HERE, "libcpu_tfq.so")
_lib = None
_PT = {"X": 0, "Y": 1, "Z": 2}


def build():
    """Compile the restatement (g++, x86-64-v3, -O3)."""
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread",
           "-fopenmp", SRC,
           "-o", LIB + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.cpu_tfq_run.argtypes = [ctypes.c_int, ctypes.c_int64, P, P, P, P, P, P, P,
                                  ctypes.c_int64, ctypes.c_int64, P, ctypes.c_int, P, P]
        L.cpu_tfq_run.restype = ctypes.c_int
        L.cpu_tfq_run_large.argtypes = [ctypes.c_int, ctypes.c_int64, P, P, P, P, P, P, P,
                                        ctypes.c_int64, ctypes.c_int64, P, ctypes.c_int,
                                        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_uint64, P, P, P]
        L.cpu_tfq_run_large.restype = ctypes.c_int
        L.cpu_tfq_noisy.argtypes = [ctypes.c_int64, P, P, P, P, P, P, P, ctypes.c_int64, P,
                                    ctypes.c_uint64, ctypes.c_int, P]
        L.cpu_tfq_noisy.restype = ctypes.c_int
        _lib = L
    return _lib


def _basic_fuse(gates):
    """[(U, bits, cmask, cval)] -> fused list: uncontrolled single-qubit gates
    are absorbed into the next uncontrolled two-qubit gate on their qubit."""
    pending, out = {}, []

    def flush(b):
        if b in pending:
            out.append((pending.pop(b), [b], 0, 0))

    for U, bits, cmask, cval in gates:
        if len(bits) == 1 and cmask == 0:
            b = bits[0]
            pending[b] = U @ pending.get(b, np.eye(2))
        elif len(bits) == 2 and cmask == 0:
            A = pending.pop(bits[0], np.eye(2))
            Bm = pending.pop(bits[1], np.eye(2))
            out.append((U @ np.kron(A, Bm), list(bits), 0, 0))
        else:
            for b in list(bits) + [q for q in range(64) if (cmask >> q) & 1]:
                flush(b)
            out.append((U, list(bits), cmask, cval))
    for b in sorted(pending):
        flush(b)
    return out


class Lowered:
    """Flat arrays of a batch for ``cpu_tfq_run``."""

    def __init__(self, programs, symbol_names, symbol_values, pauli_sums):
        names = [T._name(s) for s in symbol_names]
        col = {s: j for j, s in enumerate(names)}
        maps = T._symbol_maps(names, np.asarray(symbol_values, dtype=np.float32).reshape(len(programs), -1))
        self.B, self.S = len(programs), len(names)
        ps = np.asarray(pauli_sums, dtype=object)
        self.O = ps.shape[1] if self.B else 0
        circ, gi, gm, mats, ti, tc, pl = [], [], [], [], [], [], []
        moff = [0]
        empty = []

        def push_mat(U):
            off = moff[0]
            m = np.asarray(U, dtype=np.complex64).reshape(-1)
            mats.append(m)
            moff[0] += m.size
            return off

        def push_gate(U, bits, cmask, cval, derivs=()):
            mo = push_mat(U)
            d = [(push_mat(dU), col[s]) for dU, s in derivs]
            d += [(0, 0)] * (2 - len(d))
            gi.append([len(bits), bits[0], bits[1] if len(bits) > 1 else 0, mo, len(derivs),
                       d[0][0], d[1][0], d[0][1], d[1][1]])
            gm.append([cmask, cval])

        for i in range(self.B):
            prog = T.Program(programs[i], maps[i])
            n = prog.n
            empty.append(not prog.gates)
            gb = len(gi)
            plain = []
            for g in prog.gates:
                U = g.unitary()
                cmask = sum(1 << b for b in g.cbits)
                cval = sum(int(v) << b for b, v in zip(g.cbits, g.cvals))
                derivs = [(T._fd_matrix(g, pidx), sym) for sym, pidx in g.symbols]
                push_gate(U, list(g.bits), cmask, cval, derivs)
                plain.append((U, list(g.bits), cmask, cval))
            ge = len(gi)
            for U, bits, cmask, cval in _basic_fuse(plain):
                push_gate(U, bits, cmask, cval)
            fe = len(gi)
            tb = len(ti)
            if prog.gates:
                for j in range(self.O):
                    for c, paulis in T.parse_paulisum(ps[i][j], prog.qmap, n):
                        ti.append([j, len(pl), len(pl) + len(paulis)])
                        tc.append(c)
                        pl.extend([b, _PT[t]] for b, t in paulis)
            circ.append([n, gb, ge, ge, fe, tb, len(ti), i])
        self.empty = np.array(empty, dtype=bool)
        self.circ = np.array(circ, np.int32).reshape(-1, 8)
        self.gi = np.array(gi, np.int32).reshape(-1, 9)
        self.gm = np.array(gm, np.uint64).reshape(-1, 2)
        self.mats = np.concatenate(mats) if mats else np.zeros(1, np.complex64)
        self.ti = np.array(ti, np.int32).reshape(-1, 3)
        self.tc = np.array(tc, np.float32)
        self.pl = np.array(pl, np.int32).reshape(-1, 2)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data if a.size else 0)


def large_schedule(low, mode):
    """TFQ's choice of ComputeLarge: widest circuit >= 26 qubits (>= 25 for the
    adjoint) or a batch of one (tfq_simulate_expectation_op.cc:120-126,
    tfq_adj_grad_op.cc:161-169)."""
    max_nq = int(low.circ[:, 0].max()) if low.B else 0
    return max_nq >= (25 if mode == 1 else 26) or low.B == 1


def run_large(low, mode, downstream=None, threads=1, gate_limit=-1, fwd_limit=-1):
    """The ComputeLarge schedule (every thread on one state).  gate_limit >= 0
    bounds the work: mode 0 applies only the first gate_limit fused gates;
    mode 1 runs the last gate_limit reverse gates (and, with fwd_limit >= 0,
    only the first fwd_limit forward gates; lambda is always built)."""
    exp = np.zeros((low.B, max(low.O, 1)), np.float32)
    grad = np.zeros((low.B, max(low.S, 1)), np.float32)
    ds = np.ascontiguousarray(downstream if downstream is not None else np.zeros((low.B, low.O)),
                              dtype=np.float32)
    mats = np.ascontiguousarray(low.mats.view(np.float32))
    rc = lib().cpu_tfq_run_large(mode, low.B, _p(low.circ), _p(low.gi), _p(low.gm), _p(mats),
                                 _p(low.ti), _p(low.tc), _p(low.pl), low.O, low.S, _p(ds),
                                 int(threads), int(gate_limit), int(fwd_limit), 0, 0, _p(exp),
                                 _p(grad), None)
    if rc != 0:
        raise RuntimeError("cpu_tfq_run_large failed")
    exp = exp[:, :low.O]
    exp[low.empty] = -2.0
    return exp if mode == 0 else grad[:, :low.S]


def sample_large(low, num_samples, seed, threads=1, gate_limit=-1):
    """ComputeLarge simulation of the fused gates (the first gate_limit when
    >= 0), then qsim's sampler: indices [B, num_samples] in ascending order,
    drawn from the Philox stream (row, 0xFFFFFFFF)."""
    out = np.zeros((low.B, max(1, int(num_samples))), np.uint64)
    mats = np.ascontiguousarray(low.mats.view(np.float32))
    ds = np.zeros((low.B, max(low.O, 1)), np.float32)
    exp = np.zeros((low.B, max(low.O, 1)), np.float32)
    rc = lib().cpu_tfq_run_large(2, low.B, _p(low.circ), _p(low.gi), _p(low.gm), _p(mats),
                                 _p(low.ti), _p(low.tc), _p(low.pl), low.O, low.S, _p(ds),
                                 int(threads), int(gate_limit), -1, int(num_samples),
                                 int(seed) & (2 ** 64 - 1), _p(exp), None, _p(out))
    if rc != 0:
        raise RuntimeError("cpu_tfq_run_large failed")
    return out[:, :int(num_samples)]


def run(low, mode, downstream=None, threads=1):
    """mode 0: expectations [B, O]; mode 1: adjoint gradients [B, S]
    (ComputeSmall schedule: one circuit per thread)."""
    exp = np.zeros((low.B, max(low.O, 1)), np.float32)
    grad = np.zeros((low.B, max(low.S, 1)), np.float32)
    ds = np.ascontiguousarray(downstream if downstream is not None else np.zeros((low.B, low.O)),
                              dtype=np.float32)
    mats = np.ascontiguousarray(low.mats.view(np.float32))
    rc = lib().cpu_tfq_run(mode, low.B, _p(low.circ), _p(low.gi), _p(low.gm), _p(mats), _p(low.ti),
                           _p(low.tc), _p(low.pl), low.O, low.S, _p(ds), int(threads), _p(exp),
                           _p(grad))
    if rc != 0:
        raise RuntimeError("cpu_tfq_run failed")
    exp = exp[:, :low.O]
    exp[low.empty] = -2.0
    return exp if mode == 0 else grad[:, :low.S]


def _lower_chunk(args):
    return Lowered(*args)


def lower_parallel(programs, symbol_names, symbol_values, pauli_sums, workers=8, chunks=None):
    """``Lowered`` over contiguous row chunks of a batch, built in ``workers``
    fresh (spawned) interpreters: the Python proto decode dominates the cost
    of lowering thousands of circuits.  Returns the chunks in row order."""
    import multiprocessing as mp
    B = len(programs)
    k = max(1, chunks or 2 * workers)
    edges = [B * c // k for c in range(k + 1)]
    vals = np.asarray(symbol_values, dtype=np.float32).reshape(B, -1)
    ps = np.asarray(pauli_sums, dtype=object)
    jobs = [(list(programs[a:b]), list(symbol_names), vals[a:b], ps[a:b])
            for a, b in zip(edges[:-1], edges[1:]) if b > a]
    if workers <= 1 or len(jobs) == 1:
        return [_lower_chunk(j) for j in jobs]
    with mp.get_context("spawn").Pool(min(workers, len(jobs))) as pool:
        return pool.map(_lower_chunk, jobs)


def run_chunks(chunks, mode, downstream=None, threads=1):
    """``run`` over the chunks of ``lower_parallel``; rows concatenated."""
    out, r0 = [], 0
    for low in chunks:
        ds = None if downstream is None else np.asarray(downstream)[r0:r0 + low.B]
        out.append(run(low, mode, ds, threads))
        r0 += low.B
    return np.concatenate(out, axis=0)


class LoweredNoisy:
    """Flat arrays of a noisy batch for ``cpu_tfq_noisy``: gates one by one and
    channels as records with nt = -1 (programs parsed by ``noise_oracle``)."""

    def __init__(self, programs, symbol_names, symbol_values, pauli_sums):
        from . import noise_oracle as N
        names = [T._name(s) for s in symbol_names]
        maps = T._symbol_maps(names, np.asarray(symbol_values, dtype=np.float32).reshape(len(programs), -1))
        self.B = len(programs)
        ps = np.asarray(pauli_sums, dtype=object)
        self.O = ps.shape[1] if self.B else 0
        circ, gi, gm, mats, ti, tc, pl = [], [], [], [], [], [], []
        moff = [0]

        def push(arr):
            off = moff[0]
            m = np.asarray(arr, dtype=np.complex64).reshape(-1)
            mats.append(m)
            moff[0] += m.size
            return off

        for i in range(self.B):
            prog = N.NoisyProgram(programs[i], maps[i])
            gb = len(gi)
            ch = 0
            for kind, it in prog.items:
                if kind == "g":
                    cmask = sum(1 << b for b in it.cbits)
                    cval = sum(int(v) << b for b, v in zip(it.cbits, it.cvals))
                    bits = list(it.bits)
                    gi.append([len(bits), bits[0], bits[1] if len(bits) > 1 else 0,
                               push(it.unitary()), 0, 0, 0, 0, 0])
                    gm.append([cmask, cval])
                    continue
                mo = push(np.concatenate([K.reshape(-1) for K, _ in it.kraus]))
                if it.cid in N.MIXTURES:
                    aux = push(N._mixture_cum(it.kraus).astype(np.complex64))
                    mix = 1
                else:
                    aux = push(np.concatenate([(K.conj().T @ K).reshape(-1) for K, _ in it.kraus]))
                    mix = 0
                gi.append([-1, it.bit, 0, mo, len(it.kraus), aux, 0, ch, mix])
                gm.append([0, 0])
                ch += 1
            ge = len(gi)
            tb = len(ti)
            for j in range(self.O):
                for c, paulis in T.parse_paulisum(ps[i][j], prog.qmap, prog.n):
                    ti.append([j, len(pl), len(pl) + len(paulis)])
                    tc.append(c)
                    pl.extend([b, _PT[t]] for b, t in paulis)
            circ.append([prog.n, gb, ge, 0, 0, tb, len(ti), i])
        self.circ = np.array(circ, np.int32).reshape(-1, 8)
        self.gi = np.array(gi, np.int32).reshape(-1, 9)
        self.gm = np.array(gm, np.uint64).reshape(-1, 2)
        self.mats = np.concatenate(mats) if mats else np.zeros(1, np.complex64)
        self.ti = np.array(ti, np.int32).reshape(-1, 3)
        self.tc = np.array(tc, np.float32)
        self.pl = np.array(pl, np.int32).reshape(-1, 2)


def run_noisy(low, num_samples, seed, threads=1):
    """Noisy expectations [B, O]: term j of row i averaged over
    num_samples[i][j] trajectories (same Philox draws as the device)."""
    out = np.zeros((low.B, max(low.O, 1)), np.float32)
    ns = np.ascontiguousarray(num_samples, dtype=np.int32)
    mats = np.ascontiguousarray(low.mats.view(np.float32))
    rc = lib().cpu_tfq_noisy(low.B, _p(low.circ), _p(low.gi), _p(low.gm), _p(mats), _p(low.ti),
                             _p(low.tc), _p(low.pl), low.O, _p(ns), int(seed) & (2 ** 64 - 1),
                             int(threads), _p(out))
    if rc != 0:
        raise RuntimeError("cpu_tfq_noisy failed")
    return out[:, :low.O]
