"""Philox4x32-10 counter-based RNG in numpy (test infrastructure only).

Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11.  Round: (hi0, lo0) = mulhilo(0xD2511F53, c0), (hi1, lo1) =
mulhilo(0xCD9E8D57, c2); c = (hi1^c1^k0, lo1, hi0^c3^k1, lo0); key bump by
(0x9E3779B9, 0xBB67AE85) between rounds; 10 rounds.

The engine's uniform for an attempt (chain seed s, global attempt index q) is
philox(ctr = (q_lo, q_hi, 0, 0x51D3), key = (s_lo, s_hi)) -> words x0, x1 ->
((x0 >> 5) * 2^26 + (x1 >> 6)) * 2^-53, a double in [0, 1).
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)
STREAM_TAG = 0x51D3


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32-valued, key: (..., 2). Returns (..., 4) uint64 holding 32-bit words."""
    c = [np.asarray(ctr[..., j], dtype=np.uint64) & MASK for j in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64) & MASK
    k1 = np.asarray(key[..., 1], dtype=np.uint64) & MASK
    for r in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [(hi1 ^ c[1] ^ k0) & MASK, lo1, (hi0 ^ c[3] ^ k1) & MASK, lo0]
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return np.stack(c, axis=-1)


def uniforms(seed, first, count):
    """Uniforms for global attempt indices first..first+count-1 of the chain with this seed."""
    q = np.arange(first, first + count, dtype=np.uint64)
    ctr = np.stack([q & MASK, q >> np.uint64(32),
                    np.zeros_like(q), np.full_like(q, np.uint64(STREAM_TAG))], axis=-1)
    key = np.stack([np.full_like(q, np.uint64(seed) & MASK),
                    np.full_like(q, np.uint64(seed) >> np.uint64(32))], axis=-1)
    x = philox4x32_10(ctr, key)
    hi = (x[..., 0] >> np.uint64(5)).astype(np.float64)
    lo = (x[..., 1] >> np.uint64(6)).astype(np.float64)
    return (hi * 67108864.0 + lo) * (1.0 / 9007199254740992.0)
