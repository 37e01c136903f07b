"""Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC'11) in numpy, vectorised over counters: the checker of the
device generator behind bf_hip_synthesize."""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(counters, key):
    """counters: uint32[..., 4]; key: (k0, k1) ints. Returns uint32[..., 4]."""
    c = [np.asarray(counters[..., i], np.uint64) for i in range(4)]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
    return np.stack([v.astype(np.uint32) for v in c], axis=-1)


def unit(words):
    """Uniform in (0, 1] from the high 24 bits, offset by half a step (float32)."""
    return ((words >> np.uint32(8)).astype(np.float32) + np.float32(0.5)) * np.float32(2.0 ** -24)


def synth_parameters(n, lower, upper, seed, fit_offset=0, layout=0, sep=(2.0, 6.0)):
    """Parameters [n, 8] the generator draws (counters (fit, fit>>32, ~0, 0..2))."""
    fit = np.arange(fit_offset, fit_offset + n, dtype=np.uint64)
    words = []
    for j in range(3):
        ctr = np.stack([(fit & MASK).astype(np.uint32), (fit >> np.uint64(32)).astype(np.uint32),
                        np.full(n, 0xFFFFFFFF, np.uint32), np.full(n, j, np.uint32)], axis=-1)
        words.append(philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32)))
    w = np.concatenate(words, axis=-1)
    lo, hi = np.asarray(lower, np.float32), np.asarray(upper, np.float32)
    p = lo + (hi - lo) * unit(w[:, :8])
    if layout == 1:
        p[:, 5] = p[:, 2]
        s = np.float32(sep[0]) + (np.float32(sep[1]) - np.float32(sep[0])) * unit(w[:, 8])
        p[:, 4] = p[:, 1] + s * p[:, 2]
    return p.astype(np.float32)


def unit53(hi, lo):
    """Uniform in (0, 1) with 53 bits from two words (float64)."""
    return (float(((int(hi) >> 5) << 26) | (int(lo) >> 6)) + 0.5) * 2.0 ** -53


def _words(fit, j, d, seed):
    ctr = np.array([[fit & 0xFFFFFFFF, fit >> 32, j, d]], np.uint32)
    return philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))[0]


def poisson_count(mu, fit, j, seed):
    """The device's Poisson count for the stream (seed, fit, point j): inversion
    by sequential search below a mean of 10, PTRS (W. Hörmann, Insurance:
    Mathematics and Economics 12 (1993) 39-45) from 10; try t of PTRS reads
    counter (fit, fit>>32, j, 8 + t)."""
    import math
    if not mu > 0:
        return 0.0
    if mu < 10:
        w = _words(fit, j, 8, seed)
        u = unit53(w[0], w[1])
        p = math.exp(-mu)
        f, k = p, 0.0
        while u > f and k < 1000:
            k += 1.0
            p *= mu / k
            f += p
        return k
    smu, lmu = math.sqrt(mu), math.log(mu)
    b = 0.931 + 2.53 * smu
    a = -0.059 + 0.02483 * b
    inv_alpha = 1.1239 + 1.1328 / (b - 3.4)
    vr = 0.9277 - 3.6224 / (b - 2.0)
    for t in range(256):
        w = _words(fit, j, 8 + t, seed)
        u = unit53(w[0], w[1]) - 0.5
        v = unit53(w[2], w[3])
        us = 0.5 - abs(u)
        k = math.floor((2.0 * a / us + b) * u + mu + 0.43)
        if us >= 0.07 and v <= vr:
            return float(k)
        if k < 0 or (us < 0.013 and v > us):
            continue
        if math.log(v) + math.log(inv_alpha) - math.log(a / (us * us) + b) <= \
                -mu + k * lmu - math.lgamma(k + 1.0):
            return float(k)
    return mu
