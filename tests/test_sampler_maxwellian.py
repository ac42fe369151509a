"""The seeded perturbed-Maxwellian sampler (pic_sample_maxwellian) on the
CPU: its Philox4x32-10 restatement against the generator's published
known-answer vectors (Random123's kat_vectors), and the sample's statistics
against the density it draws from."""
import math

import numpy as np
import pytest

from philox import maxwellian, philox4x32_10


@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, out):
    got = philox4x32_10(*[np.array([v], np.uint64) for v in ctr], *key)
    assert [int(g[0]) for g in got] == list(out)


def test_maxwellian_statistics():
    """Density 1 + alpha cos(kx x) cos(ky y) on [0, 4pi]^2 (kx = ky = 1/2):
    E[cos(kx x) cos(ky y)] = alpha/4, E[cos(kx x)] = E[cos(ky y)] = 0; the
    velocities have mean 0 and variance vth^2; all within 4 standard
    errors of 400,000 draws."""
    n, alpha, L = 400_000, 0.3, 4 * math.pi
    x, y, vx, vy = maxwellian(n, alpha, 0.5, 0.5, 1.3, 1234, 0.0, L, 0.0, L)
    assert x.min() >= 0 and x.max() < L and y.min() >= 0 and y.max() < L
    se = 4 / math.sqrt(n)
    cc = np.cos(0.5 * x) * np.cos(0.5 * y)
    assert abs(cc.mean() - alpha / 4) < se
    assert abs(np.cos(0.5 * x).mean()) < se and abs(np.cos(0.5 * y).mean()) < se
    assert abs(vx.mean()) < 1.3 * se and abs(vy.mean()) < 1.3 * se
    assert abs(vx.var() / 1.69 - 1) < 2 * se and abs(vy.var() / 1.69 - 1) < 2 * se


def test_maxwellian_inversion_solves_cdf():
    """y is the root of its conditional distribution function:
    y + a sin(ky y)/ky = u dimy, a = alpha cos(kx x)."""
    from philox import u53
    n, alpha, L = 10_000, 0.6, 4 * math.pi
    x, y, _, _ = maxwellian(n, alpha, 0.5, 0.5, 1.0, 99, 0.0, L, 0.0, L)
    k = np.arange(n, dtype=np.uint64)
    z = np.zeros(n, np.uint64)
    r = philox4x32_10(k, z, z, z, 99, 0)
    u = u53(r[2], r[3])
    a = alpha * np.cos(0.5 * x)
    assert np.max(np.abs(y + a * np.sin(0.5 * y) / 0.5 - u * L)) < 1e-12
