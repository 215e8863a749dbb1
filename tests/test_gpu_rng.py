"""The sampler's RNG stream on the device, bit for bit against the host
replica (``oracle/philox.py``, itself pinned by the Random123 known-answer
vectors), and samples that equal the oracle's draw for draw wherever the
draw is not decided by float32 rounding.

Streams checked are the ones the kernels use: samples (row, 0xFFFFFFFF),
sampled expectation (row, slot << 16 | term) and noisy channels (row,
channel), with shot counters past 2^32 and seeds using all 64 bits."""

import sys

import numpy as np
import pytest

from oracle import philox
from oracle import tfq_oracle as O
from quantum_amd import _lib
from quantum_amd import circuits as C
from quantum_amd import serialize as S
from quantum_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def assert_draws_match(bits, program, names, vals, seed, row, min_gap=1e-6):
    """Samples [shots, max_nq] of one row equal the oracle's inverse-CDF draws
    from the Philox stream (row, 0xFFFFFFFF), except draws whose uniform lies
    within `min_gap` of a CDF boundary (float32 state rounding decides those);
    returns the number of draws checked."""
    p = O.Program(program, O._symbol_maps(names, vals)[0])
    probs = np.abs(O.simulate(p)) ** 2
    cdf = np.cumsum(probs)
    shots = bits.shape[0]
    u = philox.uniform_doubles(seed, shots, row, 0xFFFFFFFF) * cdf[-1]
    want = np.minimum(np.searchsorted(cdf, u, side="right"), len(cdf) - 1)
    pos = np.searchsorted(cdf, u)
    gap = np.minimum(np.abs(u - cdf[np.minimum(pos, len(cdf) - 1)]),
                     np.abs(u - cdf[np.maximum(pos - 1, 0)]))
    safe = gap > min_gap
    got = (bits[:, ::-1][:, :p.n].astype(np.int64) << np.arange(p.n)).sum(axis=1)
    np.testing.assert_array_equal(got[safe], want[safe])
    assert np.all(bits[:, :bits.shape[1] - p.n] == -2)
    return int(safe.sum())


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("seed,start,count,s_lo,s_hi", [
    (12345, 0, 100000, 0, 0xFFFFFFFF),
    (0, 0, 4096, 7, 0xFFFFFFFF),
    (2 ** 64 - 1, 2 ** 32 - 1000, 5000, 3, (2 << 16) | 5),
    (0x9E3779B97F4A7C15, 2 ** 40 + 17, 3000, 4095, 0),
    (777, 1_000_000 - 50, 50, 63, 131),
])
def test_device_uniforms_equal_host_replica_bit_for_bit(seed, start, count, s_lo, s_hi):
    _say(f"philox seed={seed} start={start} count={count} stream=({s_lo},{s_hi})")
    got = _lib.philox_uniforms(seed, count, s_lo, s_hi, start=start).cpu().numpy()
    want = philox.uniform_doubles(seed, count, s_lo, s_hi, start=start)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    assert got.min() >= 0.0 and got.max() < 1.0


def test_samples_equal_oracle_except_at_rounding_boundaries():
    """Every shot whose uniform lies farther than 1e-5 (in units of the total
    mass) from a boundary of the oracle's CDF must land on the oracle's
    index: the device state is float32 (amplitude errors ~1e-7), so only
    shots that close to a boundary may differ."""
    from quantum_amd import tfq_simulate_ops as sim
    rng = np.random.default_rng(300)
    n, batch, shots, seed = 5, 40, 4000, 12345
    q = C.LineQubit.range(n)
    names = ["s0", "s1", "s2"]
    circuits = [W.random_circuit(rng, q, 20, symbols=names, p_control=0.25) for _ in range(batch)]
    vals = rng.uniform(-2, 2, size=(batch, 3)).astype(np.float32)
    progs = S.convert_to_tensor(circuits)
    _say("samples vs oracle draw for draw")
    out = sim.tfq_simulate_samples(progs, names, vals, [shots], seed=seed).cpu().numpy()
    maps = O._symbol_maps(names, vals)
    decided = 0
    for i in range(batch):
        p = O.Program(progs[i], maps[i])
        probs = np.abs(O.simulate(p)) ** 2
        cdf = np.cumsum(probs)
        u = philox.uniform_doubles(seed, shots, i, 0xFFFFFFFF) * cdf[-1]
        want = np.minimum(np.searchsorted(cdf, u, side="right"), len(cdf) - 1)
        gap = np.min(np.abs(u[:, None] - cdf[None, :]), axis=1)
        safe = gap > 1e-5
        bits = out[i, :, ::-1][:, :p.n]
        got = (bits.astype(np.int64) << np.arange(p.n)).sum(axis=1)
        np.testing.assert_array_equal(got[safe], want[safe], err_msg=f"row {i}")
        decided += int(safe.sum())
    assert decided > 0.99 * batch * shots


def test_streaming_sampler_matches_oracle_draws():
    """The HBM-streaming sampler (block sums, sub-block prefixes, one
    sub-block scan per shot) on a batch of 19-, 3- and 0-qubit programs:
    samples equal the oracle's inverse-CDF draws from the same Philox stream
    (a draw may move only if its uniform lies within 1e-6 of a CDF boundary,
    where float32 rounding of the state decides), and the sampled expectation
    of X/Y/Z terms (per-term basis change, per-term streams) equals the
    oracle's to two shots' worth (2 * 2 / shots)."""
    from quantum_amd import tfq_simulate_ops as sim
    rng = np.random.default_rng(19)
    names = ["s0", "s1"]
    circuits, sums = [], []
    for n in (19, 3, 0, 19):
        if n == 0:
            circuits.append(C.Circuit())
            sums.append([C.PauliSum([]), C.PauliSum([])])
            continue
        q = C.LineQubit.range(n)
        circuits.append(W.random_circuit(rng, q, 2 * n, symbols=names, p_control=0.1) +
                        C.Circuit([C.I(x) for x in q]))
        sums.append([W.random_pauli_sum(rng, q, 3, max_weight=3),
                     W.random_pauli_sum(rng, q, 2, max_weight=2)])
    vals = rng.uniform(-2, 2, size=(4, 2)).astype(np.float32)
    progs, ps = S.convert_to_tensor(circuits), S.convert_to_tensor(sums)
    shots, seed = 20000, 4242
    _say("streaming samples n=19 vs oracle")
    got = sim.tfq_simulate_samples(progs, names, vals, [shots], seed=seed).cpu().numpy()
    assert got.shape == (4, shots, 19)
    checked = sum(assert_draws_match(got[i], progs[i], names, vals[i:i + 1], seed, i)
                  for i in (0, 1, 3))
    assert checked > 0.9 * 3 * shots
    np.testing.assert_array_equal(got[2], -2)
    ns = np.full((4, 2), shots, np.int32)
    ns[:, 1] = 7000
    _say("streaming sampled expectation n=19 vs oracle")
    got_e = sim.tfq_simulate_sampled_expectation(progs, names, vals, ps, ns, seed=seed).cpu().numpy()
    want_e = O.simulate_sampled_expectation(list(progs), names, vals, ps, ns, seed)
    np.testing.assert_allclose(got_e, want_e, atol=4.0 / 7000)
