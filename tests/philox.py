"""numpy restatement of the seeded perturbed-Maxwellian sampler of the device
library (pic_sample_maxwellian): the Philox4x32-10 counter-based generator
(Salmon, Moraes, Dror and Shaw, SC 2011) and the draw of each particle from
it.  Test infrastructure: the GPU sample is compared with this one."""
import math

import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10: counters (c0..c3) uint32 arrays, key
    (k0, k1) ints; returns the four output words as uint64 arrays."""
    c = [np.asarray(v, dtype=np.uint64) & MASK for v in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & MASK, (p0 >> np.uint64(32)) ^ c[3] ^ k1,
             p0 & MASK]
        k0 = (k0 + np.uint64(W0)) & MASK
        k1 = (k1 + np.uint64(W1)) & MASK
    return c


def u53(hi, lo):
    return ((hi >> np.uint64(5)) << np.uint64(26) | (lo >> np.uint64(6))).astype(np.float64) \
        * 2.0 ** -53


def maxwellian(n, alpha, kx, ky, vth, seed, xmin, dimx, ymin, dimy):
    """x, y, vx, vy of the device sampler's n particles."""
    k = np.arange(n, dtype=np.uint64)
    lo, hi = k & MASK, k >> np.uint64(32)
    z = np.zeros(n, np.uint64)
    r = philox4x32_10(lo, hi, z, z, seed & 0xFFFFFFFF, seed >> 32)
    s = philox4x32_10(lo, hi, z + np.uint64(1), z, seed & 0xFFFFFFFF, seed >> 32)
    xr = u53(r[0], r[1]) * dimx
    a = alpha * np.cos(kx * xr)
    target = u53(r[2], r[3]) * dimy
    y = target.copy()
    for _ in range(60):
        step = (y + a * np.sin(ky * y) / ky - target) / (1 + a * np.cos(ky * y))
        y = y - step
    y = np.clip(y, 0.0, dimy)
    y[y >= dimy] = 0.0
    rad = vth * np.sqrt(-2 * np.log(1.0 - u53(s[0], s[1])))
    th = 2 * math.pi * u53(s[2], s[3])
    return xmin + xr, ymin + y, rad * np.cos(th), rad * np.sin(th)
