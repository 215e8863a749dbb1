"""complex128 CPU restatement of TFQ's noisy (Monte-Carlo trajectory) ops.

TEST INFRASTRUCTURE ONLY -- the checker, never the thing measured or shipped.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.

What it restates:

* Program -> noisy circuit: ``NoisyQsimCircuitFromProgram``
  (``core/src/circuit_parser_qsim.cc:828-870``): each operation is parsed as a
  gate (``parse_gate`` of ``tfq_oracle``); an id that is not a gate id is
  parsed as a channel (``ParseAppendChannel``, ``:598-826``: ids DP, ADP, GAD,
  AD, RST, PD, PF, BF with float arguments p, p_x, p_y, p_z, gamma), and an id
  that is neither fails with "Could not parse channel id: <id>".
* The channels: qsim v0.21.0 ``lib/channels_cirq.h`` (a Bazel dependency
  pinned in ``WORKSPACE:75-81``, absent from the reference tree), which
  implements Cirq's channel definitions.  Restated here from Cirq's published
  Kraus operators (``channel_kraus``).
* The ops: ``tfq_noisy_expectation`` (``core/ops/noise/tfq_noisy_expectation.cc:57-283``),
  ``tfq_noisy_sampled_expectation`` (``tfq_noisy_sampled_expectation.cc``, one
  shot per Pauli term per trajectory through ``ComputeSampledExpectationQsim``,
  ``core/src/util_qsim.h:199-270``) and ``tfq_noisy_samples``
  (``tfq_noisy_samples.cc:150-205``, one terminal measurement per trajectory).

Two restatements:

* ``density_*``: the exact mixed-state result (rho through every gate and
  channel, <P> = Tr(rho P)), which the reference's trajectory averages estimate.
  The reference seeds its trajectories from ``tensorflow::random::New64()``, so
  its outputs are random; parity with it is statistical against these values.
* ``trajectory_*``: a replica of the device's trajectory sampler.  It draws the
  same counter-based Philox variates (``oracle/philox.py``) and applies the
  same selection rule, so its averages match the device's to float rounding.

The selection rule (qsim's ``QuantumTrajectorySimulator`` picks Kraus
operators the same way, up to the order of its random draws): channel c of
trajectory t of row i draws u = uniform(seed, shot=t, stream=(i, c)).
* Unitary mixtures (DP, ADP, PF, BF): operator k is the first with
  u < float32(p_0 + ... + p_k) (the last when none); its unitary is applied.
* Other channels (AD, GAD, PD, RST): p_k = <psi|K_k^dag K_k|psi> from the
  qubit's reduced density matrix; operator k is the first with
  u < p_0 + ... + p_k (the last with p_k > 0 when none), and psi becomes
  K_k psi / sqrt(p_k).
"""

import numpy as np

from . import wire
from .philox import uniform_doubles
from .tfq_oracle import (F32, OracleError, _PAULI, _name, _symbol_maps, apply_matrix,
                         parse_gate, parse_paulisum, resolve_qubits, zero_state)

_I2 = np.eye(2, dtype=np.complex128)
_X = _PAULI["X"]
_Y = _PAULI["Y"]
_Z = _PAULI["Z"]

CHANNEL_IDS = ("DP", "ADP", "GAD", "AD", "RST", "PD", "PF", "BF")
MIXTURES = ("DP", "ADP", "PF", "BF")
_CHANNEL_ARGS = {"DP": ("p",), "ADP": ("p_x", "p_y", "p_z"), "GAD": ("p", "gamma"),
                 "AD": ("gamma",), "RST": (), "PD": ("gamma",), "PF": ("p",), "BF": ("p",)}
STREAM_SAMPLE = 0xFFFFFFFF
STREAM_SHOT = 0x80000000


def channel_kraus(cid, params):
    """Kraus operators of a channel from its float32 arguments (evaluated in
    double).  Unitary mixtures return ``[(U_k, p_k)]``, other channels
    ``[(K_k, None)]``."""
    f = {k: float(v) for k, v in params.items()}
    if cid == "DP":
        p = f["p"]
        return [(_I2, 1.0 - p), (_X, p / 3.0), (_Y, p / 3.0), (_Z, p / 3.0)]
    if cid == "ADP":
        px, py, pz = f["p_x"], f["p_y"], f["p_z"]
        return [(_I2, 1.0 - px - py - pz), (_X, px), (_Y, py), (_Z, pz)]
    if cid == "PF":
        return [(_I2, 1.0 - f["p"]), (_Z, f["p"])]
    if cid == "BF":
        return [(_I2, 1.0 - f["p"]), (_X, f["p"])]
    if cid == "AD":
        g = f["gamma"]
        return [(np.array([[1, 0], [0, np.sqrt(1 - g)]], complex), None),
                (np.array([[0, np.sqrt(g)], [0, 0]], complex), None)]
    if cid == "PD":
        g = f["gamma"]
        return [(np.array([[1, 0], [0, np.sqrt(1 - g)]], complex), None),
                (np.array([[0, 0], [0, np.sqrt(g)]], complex), None)]
    if cid == "GAD":
        p, g = f["p"], f["gamma"]
        a, b = np.sqrt(p), np.sqrt(1 - p)
        return [(a * np.array([[1, 0], [0, np.sqrt(1 - g)]], complex), None),
                (a * np.array([[0, np.sqrt(g)], [0, 0]], complex), None),
                (b * np.array([[np.sqrt(1 - g), 0], [0, 1]], complex), None),
                (b * np.array([[0, 0], [np.sqrt(g), 0]], complex), None)]
    if cid == "RST":
        return [(np.array([[1, 0], [0, 0]], complex), None),
                (np.array([[0, 1], [0, 0]], complex), None)]
    raise OracleError(f"Could not parse channel id: {cid}")


class Channel:
    __slots__ = ("cid", "bit", "params", "kraus")


def _channel_arg(op, name):
    a = op["args"].get(name)
    if a is None:
        raise OracleError(f"Could not find arg: {name} in op.")
    if a["kind"] == "symbol" and a["symbol"]:
        raise OracleError("Could not find symbol in parameter map: " + a["symbol"])
    return F32(a["float"])


def parse_channel(op, n, qmap):
    cid = op["gate"]
    if cid not in CHANNEL_IDS:
        raise OracleError(f"Could not parse channel id: {cid}")
    c = Channel()
    c.cid = cid
    c.bit = n - 1 - qmap[op["qubits"][0]]
    c.params = {k: _channel_arg(op, k) for k in _CHANNEL_ARGS[cid]}
    c.kraus = channel_kraus(cid, c.params)
    return c


class NoisyProgram:
    """A program as an ordered list of gates and channels."""

    def __init__(self, data, symmap=None):
        moments = wire.decode_program(data)
        self.n, self.qmap = resolve_qubits(moments)
        self.items = []
        if self.n > 0:
            for ops in moments:
                for op in ops:
                    try:
                        self.items.append(("g", parse_gate(op, symmap or {}, self.n, self.qmap)))
                    except OracleError as e:
                        if not str(e).startswith("Could not parse gate id"):
                            raise
                        self.items.append(("c", parse_channel(op, self.n, self.qmap)))

    @property
    def num_channels(self):
        return sum(1 for kind, _ in self.items if kind == "c")


# --------------------------------------------------------------------------- exact (density matrix)
def _rho_apply(rho, U, bits, n, cbits=(), cvals=()):
    """rho <- U rho U^dag (rho flattened: row index in the high n bits)."""
    v = apply_matrix(rho, U, [b + n for b in bits], 2 * n, [b + n for b in cbits], cvals)
    return apply_matrix(v, U.conj(), list(bits), 2 * n, list(cbits), cvals)


def density_matrix(prog):
    """Final density matrix of a noisy program, flattened [row * 2^n + col]."""
    n = prog.n
    rho = np.zeros(4 ** n, dtype=np.complex128)
    rho[0] = 1.0
    for kind, it in prog.items:
        if kind == "g":
            rho = _rho_apply(rho, it.unitary(), it.bits, n, it.cbits, it.cvals)
        else:
            out = np.zeros_like(rho)
            for K, p in it.kraus:
                M = K if p is None else np.sqrt(max(p, 0.0)) * K
                out += _rho_apply(rho, M, [it.bit], n)
            rho = out
    return rho


def _rho_expectation(rho, terms, n):
    """sum_t c_t Re Tr(rho P_t) (identity terms add c)."""
    N = 2 ** n
    R = rho.reshape(N, N)
    total = 0.0
    for c, paulis in terms:
        if not paulis:
            total += c
            continue
        # Tr(rho P) = Tr(P rho): P applied to the row index, then the trace.
        PR = R.copy()
        for b, pt in paulis:
            PR = apply_matrix(PR.reshape(-1), _PAULI[pt], [b + n], 2 * n).reshape(N, N)
        total += c * np.trace(PR).real
    return total


def density_expectation(programs, symbol_names, symbol_values, pauli_sums):
    """Exact noisy expectations [B, O]; empty programs give -2."""
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    O = np.asarray(pauli_sums, dtype=object).shape[1] if B else 0
    out = np.zeros((B, O))
    for i in range(B):
        prog = NoisyProgram(programs[i], maps[i])
        if not prog.items:
            out[i, :] = -2.0
            continue
        rho = density_matrix(prog)
        for j in range(O):
            out[i, j] = _rho_expectation(rho, parse_paulisum(pauli_sums[i][j], prog.qmap, prog.n),
                                         prog.n)
    return out


def density_probabilities(program, symbol_names, symbol_values_row):
    """Outcome probabilities of a noisy program (diagonal of rho)."""
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, np.asarray(symbol_values_row, np.float32).reshape(1, -1))
    prog = NoisyProgram(program, maps[0])
    N = 2 ** prog.n
    return density_matrix(prog).reshape(N, N).diagonal().real.copy(), prog.n


# --------------------------------------------------------------------------- trajectory replica
def _mixture_cum(kraus):
    """float32 cumulative probabilities of a unitary mixture (sums in double)."""
    return np.array([np.float32(sum(p for _, p in kraus[:k + 1])) for k in range(len(kraus))],
                    dtype=np.float32)


def _apply_channel(psi, ch, u, n):
    if ch.cid in MIXTURES:
        cum = _mixture_cum(ch.kraus)
        pick = len(ch.kraus) - 1
        for k in range(len(ch.kraus) - 1):
            if u < float(cum[k]):
                pick = k
                break
        return apply_matrix(psi, ch.kraus[pick][0], [ch.bit], n)
    # Reduced density matrix of the channel's qubit.
    idx = np.arange(2 ** n)
    on = ((idx >> ch.bit) & 1).astype(bool)
    a0, a1 = psi[~on], psi[on]
    s00 = float(np.sum(np.abs(a0) ** 2))
    s11 = float(np.sum(np.abs(a1) ** 2))
    s01 = complex(np.sum(np.conj(a0) * a1))
    probs = []
    for K, _ in ch.kraus:
        E = K.conj().T @ K
        probs.append(E[0, 0].real * s00 + E[1, 1].real * s11 + 2.0 * (E[0, 1] * s01).real)
    pick = -1
    cum = 0.0
    for k, p in enumerate(probs):
        cum += p
        if u < cum:
            pick = k
            break
    if pick < 0:
        pick = max(k for k, p in enumerate(probs) if p > 0)
    out = apply_matrix(psi, ch.kraus[pick][0], [ch.bit], n)
    return out / np.sqrt(probs[pick])


def run_trajectories(prog, row, seed, count):
    """Final states of trajectories 0..count-1 of a program (row = output row)."""
    n = prog.n
    nch = prog.num_channels
    draws = [uniform_doubles(seed, count, row, c) for c in range(nch)]
    states = []
    for t in range(count):
        psi = zero_state(n)
        c = 0
        for kind, it in prog.items:
            if kind == "g":
                psi = apply_matrix(psi, it.unitary(), it.bits, n, it.cbits, it.cvals)
            else:
                psi = _apply_channel(psi, it, float(draws[c][t]), n)
                c += 1
        states.append(psi)
    return states


def _term_value(psi, paulis, n):
    s = psi
    for b, pt in paulis:
        s = apply_matrix(s, _PAULI[pt], [b], n)
    return np.vdot(psi, s).real


def trajectory_expectation(programs, symbol_names, symbol_values, pauli_sums, num_samples,
                           seed):
    """Replica of the device's noisy expectation: term j of row i averages its
    exact expectation over trajectories 0..num_samples[i][j]-1."""
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    O = np.asarray(pauli_sums, dtype=object).shape[1] if B else 0
    ns = np.asarray(num_samples, dtype=np.int64).reshape(B, O)
    out = np.zeros((B, O))
    for i in range(B):
        prog = NoisyProgram(programs[i], maps[i])
        if not prog.items:
            out[i, :] = -2.0
            continue
        sums = [parse_paulisum(pauli_sums[i][j], prog.qmap, prog.n) for j in range(O)]
        T = int(ns[i].max()) if O else 0
        states = run_trajectories(prog, i, seed, T)
        for j in range(O):
            ident = sum(c for c, p in sums[j] if not p)
            acc = 0.0
            for t in range(int(ns[i, j])):
                acc += sum(c * _term_value(states[t], p, prog.n) for c, p in sums[j] if p)
            out[i, j] = acc / ns[i, j] + ident
    return out


def trajectory_sampled_expectation(programs, symbol_names, symbol_values, pauli_sums,
                                   num_samples, seed):
    """Replica of the device's noisy sampled expectation: per trajectory, each
    non-identity term k (numbered over the row's terms, slot by slot) yields
    one shot +-1 with P(+1) = (1 + <P_k>/<psi|psi>) / 2 from
    u = uniform(seed, shot=t, stream=(i, 0x80000000 | k)), scaled by its
    coefficient.  (One measurement of the term's rotated state gives parity +1
    with exactly that probability, ``util_qsim.h:199-270``.)"""
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    B = len(programs)
    O = np.asarray(pauli_sums, dtype=object).shape[1] if B else 0
    ns = np.asarray(num_samples, dtype=np.int64).reshape(B, O)
    out = np.zeros((B, O))
    for i in range(B):
        prog = NoisyProgram(programs[i], maps[i])
        if not prog.items:
            out[i, :] = -2.0
            continue
        sums = [parse_paulisum(pauli_sums[i][j], prog.qmap, prog.n) for j in range(O)]
        T = int(ns[i].max()) if O else 0
        states = run_trajectories(prog, i, seed, T)
        k = 0
        for j in range(O):
            ident = sum(c for c, p in sums[j] if not p)
            acc = 0.0
            for c, p in sums[j]:
                if not p:
                    continue
                u = uniform_doubles(seed, int(ns[i, j]), i, STREAM_SHOT | k)
                for t in range(int(ns[i, j])):
                    psi = states[t]
                    e = _term_value(psi, p, prog.n) / np.vdot(psi, psi).real
                    acc += c * (1.0 if u[t] < 0.5 * (1.0 + e) else -1.0)
                k += 1
            out[i, j] = acc / ns[i, j] + ident
    return out


def trajectory_samples(programs, symbol_names, symbol_values, num_samples, seed):
    """Replica of the device's noisy samples: trajectory t of row i measures
    all qubits once, u = uniform(seed, shot=t, stream=(i, 0xFFFFFFFF)) picking
    the smallest index x with cdf[x] > u * cdf[-1]; bit q of x goes to column
    max_nq-1-q, with -2 padding (``tfq_noisy_samples.cc:186-200``)."""
    names = [_name(s) for s in symbol_names]
    maps = _symbol_maps(names, symbol_values)
    progs = [NoisyProgram(programs[i], maps[i]) for i in range(len(programs))]
    max_nq = max([p.n for p in progs], default=0)
    out = np.full((len(progs), num_samples, max_nq), -2, dtype=np.int8)
    for i, prog in enumerate(progs):
        if prog.n == 0:
            continue
        states = run_trajectories(prog, i, seed, num_samples)
        u = uniform_doubles(seed, num_samples, i, STREAM_SAMPLE)
        for t, psi in enumerate(states):
            cdf = np.cumsum(np.abs(psi) ** 2)
            x = min(int(np.searchsorted(cdf, u[t] * cdf[-1], side="right")), len(cdf) - 1)
            for q in range(prog.n):
                out[i, t, max_nq - 1 - q] = (x >> q) & 1
    return out
