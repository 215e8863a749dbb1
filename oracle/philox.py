"""Host replica of the device sampler's counter-based RNG (Philox4x32-10).

TEST INFRASTRUCTURE ONLY (see ``oracle/tfq_oracle.py``).

Philox4x32-10 is the published generator of Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3" (SC'11): a 10-round
multiply/xor bijection of a 128-bit counter under a 64-bit key with
multipliers 0xD2511F53, 0xCD9E8D57 and Weyl key increments 0x9E3779B9,
0xBB67AE85.  TFQ itself seeds ``tensorflow::random::PhiloxRandom`` from
``New64()`` (``core/ops/tfq_simulate_samples_op.cc:149-152``), so its streams
are not reproducible; the device sampler fixes the layout below so that
its stream is:

    key     = (seed & 0xffffffff, seed >> 32)
    counter = (shot & 0xffffffff, shot >> 32, stream_lo, stream_hi)
    u       = ((x0 << 32 | x1) >> 11) * 2**-53        in [0, 1)
"""

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 over uint32 arrays; returns four uint32 arrays."""
    c0 = np.asarray(c0, dtype=np.uint32)
    c1 = np.asarray(c1, dtype=np.uint32)
    c2 = np.asarray(c2, dtype=np.uint32)
    c3 = np.asarray(c3, dtype=np.uint32)
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for r in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            if r < 9:
                k0 = np.uint32(k0 + W0)
                k1 = np.uint32(k1 + W1)
    return c0, c1, c2, c3


def uniform_doubles(seed, count, stream_lo, stream_hi, start=0):
    """``count`` FP64 uniforms in [0,1) for shots ``start..start+count-1``."""
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    shots = np.arange(start, start + count, dtype=np.uint64)
    x0, x1, _, _ = philox4x32_10((shots & MASK32).astype(np.uint32),
                                 (shots >> np.uint64(32)).astype(np.uint32),
                                 np.full(count, stream_lo & 0xFFFFFFFF, dtype=np.uint32),
                                 np.full(count, stream_hi & 0xFFFFFFFF, dtype=np.uint32),
                                 seed & 0xFFFFFFFF, seed >> 32)
    bits = (x0.astype(np.uint64) << np.uint64(32)) | x1.astype(np.uint64)
    return (bits >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
