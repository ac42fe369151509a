"""Pins of the oracle's FP32 state for every shape (oracle_f32.c) and of the
FP64 M6 push's gather alignment (pic_oracle.c or_push_v_m6), by what the
mathematics fixes rather than by retyping the formulas:

* the FP32 weights equal the FP64 reference weights (f_m6, f_tsc, f_m4 of
  pushers_m6.jl / pushers_tsc.jl / pushers_m4.jl, themselves pinned in
  test_oracle.py and test_oracle_shapes.py) to FP32 rounding, with the
  B-spline moments (unity, zero first moment);
* NGP's node is numpy's round-half-to-even of the position;
* B-splines of order >= 2 reproduce linear fields exactly: a gather of
  e = a + b x + c y returns e at the particle, so the kick of push_v! with
  bz = 0 is dt e(x_p) (test/test_uapic.jl:43-74 checks the same invariant
  on the MeshFields gather).  A stencil shifted by one node, or ex and ey
  (or the two grid indices) swapped, changes the result by b dx or more;
* the FP32 step of every shape tracks the FP64 reference sequence
  push_v!; push_x!(dt/2); push_x!(dt/2); compute_rho to FP32 resolution.
"""
import math

import numpy as np
import pytest

from conftest import random_particles

SEP = ["m6", "ngp", "tsc", "m4"]


def _ref_weights(O, kernel, t, gather):
    """FP64 reference weights of the S nodes (offsets -HL .. S-1-HL)."""
    t = float(t)
    if kernel == "m6":
        return np.array([O.f_m6(abs(t - k)) for k in range(-2, 4)])
    f = O.f_m4 if (kernel == "m4" or (kernel == "tsc" and gather)) else O.f_tsc
    return np.array([f(1 + t), f(t), f(1 - t), f(2 - t)])


@pytest.mark.parametrize("gather", [True, False])
@pytest.mark.parametrize("kernel", ["m6", "tsc", "m4"])
def test_f32_shape_weights_equal_reference(O, kernel, gather):
    hl = 2 if kernel == "m6" else 1
    for t in np.linspace(0, 1, 257, endpoint=False).astype(np.float32):
        w = O.f32_shape_weights(kernel, t, gather=gather).astype(np.float64)
        ref = _ref_weights(O, kernel, t, gather)
        assert np.allclose(w, ref, rtol=0, atol=3e-7), (kernel, t)
        assert abs(w.sum() - 1) < 1e-6
        offs = np.arange(w.size) - hl
        assert abs(np.dot(w, offs - float(t))) < 2e-6


def test_f32_ngp_ties_to_even(O):
    """The NGP node is round(Int, cell + t) (ties to even): the upper node
    for t > 1/2, and for t == 1/2 in an odd cell."""
    for cell in range(6):
        for t in (0.0, 0.25, 0.4999999, 0.5, 0.5000001, 0.75, 0.9999999):
            t = np.float32(t)
            w = O.f32_shape_weights("ngp", t, cell=cell)
            up = int(np.rint(np.float64(cell) + np.float64(t))) - cell
            assert w.tolist() == ([0.0, 1.0] if up == 1 else [1.0, 0.0]), (cell, t)


def test_f32_unknown_shape_raises(O):
    st = tuple(np.zeros(1, d) for d in (np.int32, np.float32, np.float32, np.float32,
                                        np.float32))
    with pytest.raises(ValueError, match="unknown shape"):
        O.f32_deposit(st, 8, 8, "quadratic", np.float32(1.0))
    with pytest.raises(ValueError, match="unknown shape"):
        O.f32_step(st, 8, 8, "m5", np.zeros(128, np.float32), 0.1, 0.1, 0.1, 1.0)


def _linear_setup(rng, nx, ny, dx, dy, n, margin):
    """Particles whose stencils stay `margin` nodes inside the periodic seam,
    and the linear fields ex = 0.3 + 1.1 x - 0.4 y, ey = -0.2 + 0.7 y + 0.05 x
    on the nodes (i dx, j dy)."""
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(margin * dx, (nx - margin - 1) * dx, n)
    p[1] = rng.uniform(margin * dy, (ny - margin - 1) * dy, n)
    p[2] = rng.normal(size=n)
    p[3] = rng.normal(size=n)
    p[4] = 0.01
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    ex = np.asfortranarray(0.3 + 1.1 * X - 0.4 * Y)
    ey = np.asfortranarray(-0.2 + 0.7 * Y + 0.05 * X)
    return p, ex, ey


def linear_e(x, y):
    return 0.3 + 1.1 * x - 0.4 * y, -0.2 + 0.7 * y + 0.05 * x


def test_m6_push_reproduces_linear_field(O):
    """push_v!(p, M6(), ...) with bz = 0 kicks every particle by dt e(x_p)
    for a linear e: the 7-node stencil of pushers_m6.jl:24-207 is centred
    on the particle, ex and ey are not swapped, and the grid's first index
    is x."""
    rng = np.random.default_rng(71)
    nx, ny, dx, dy = 32, 24, 0.3, 0.45
    p, ex, ey = _linear_setup(rng, nx, ny, dx, dy, 4000, 4)
    q = p.copy(order="F")
    dt = 0.1
    O.push_v(q, "m6", nx, ny, dx, dy, ex, ey, None, dt)
    e1, e2 = linear_e(p[0], p[1])
    assert np.max(np.abs((q[2] - p[2]) / dt - e1)) < 1e-12
    assert np.max(np.abs((q[3] - p[3]) / dt - e2)) < 1e-12
    assert np.array_equal(q[:2], p[:2])


def test_m6_push_reproduces_linear_bz(O):
    """With a linear bz as well, the velocity is the Boris update of
    pushers_m6.jl:209-220 with e and b taken at the particle."""
    rng = np.random.default_rng(72)
    nx, ny, dx, dy = 32, 32, 0.25, 0.25
    p, ex, ey = _linear_setup(rng, nx, ny, dx, dy, 4000, 4)
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    bz = np.asfortranarray(1.5 - 0.2 * X + 0.3 * Y)
    q = p.copy(order="F")
    dt = 0.2
    O.push_v(q, "m6", nx, ny, dx, dy, ex, ey, bz, dt)
    e1, e2 = linear_e(p[0], p[1])
    b3 = 1.5 - 0.2 * p[0] + 0.3 * p[1]
    v1 = p[2] + 0.5 * dt * e1
    v2 = p[3] + 0.5 * dt * e2
    tn = 0.5 * dt * b3
    sn = 2 * tn / (1 + tn * tn)
    v1 = v1 + v2 * tn
    v2 = v2 - v1 * sn
    v1 = v1 + v2 * tn
    assert np.max(np.abs(q[2] - (v1 + 0.5 * dt * e1))) < 1e-12
    assert np.max(np.abs(q[3] - (v2 + 0.5 * dt * e2))) < 1e-12


@pytest.mark.parametrize("kernel", ["m4", "tsc", "cic"])
def test_shape_push_reproduces_linear_field(O, kernel):
    rng = np.random.default_rng(73)
    nx, ny, dx, dy = 32, 24, 0.3, 0.45
    p, ex, ey = _linear_setup(rng, nx, ny, dx, dy, 3000, 3)
    q = p.copy(order="F")
    O.push_v(q, kernel, nx, ny, dx, dy, ex, ey, None, 0.1)
    e1, e2 = linear_e(p[0], p[1])
    assert np.max(np.abs((q[2] - p[2]) / 0.1 - e1)) < 1e-12
    assert np.max(np.abs((q[3] - p[3]) / 0.1 - e2)) < 1e-12


def _e32(ex, ey):
    nx, ny = ex.shape
    e32 = np.empty(2 * nx * ny, np.float32)
    e32[0::2] = ex.ravel(order="F")
    e32[1::2] = ey.ravel(order="F")
    return e32


@pytest.mark.parametrize("kernel", ["cic", "m6", "tsc", "m4"])
def test_f32_step_reproduces_linear_field(O, kernel):
    """The FP32 step's gather of a linear field is the field at the particle
    to FP32 rounding: the kick is dt e(x_p) within 1e-5 (a one-node shift
    would be off by 1.1 dx = 0.33)."""
    rng = np.random.default_rng(74)
    nx, ny, dx, dy = 32, 24, 0.3, 0.45
    p, ex, ey = _linear_setup(rng, nx, ny, dx, dy, 3000, 4)
    st = O.f32_from_f64(p, nx, ny, dx, dy)
    x = O.f32_to_f64(st, nx, dx, dy, 0.01)  # the FP32 positions
    v0x, v0y = st[3].astype(np.float64), st[4].astype(np.float64)
    dt = 0.1
    O.f32_step(st, nx, ny, kernel, _e32(ex, ey), np.float32(dt / 2), np.float32(0.0),
               np.float32(0.0), np.float32(0.01), deposit=False)
    e1, e2 = linear_e(x[0], x[1])
    assert np.max(np.abs((st[3] - v0x) / dt - e1)) < 1e-4
    assert np.max(np.abs((st[4] - v0y) / dt - e2)) < 1e-4


@pytest.mark.parametrize("kernel", ["ngp", "tsc", "m4"])
def test_f32_step_tracks_f64_shapes(O, kernel):
    """One FP32 step (kick, drift dt, deposit) of the NGP, TSC and M4 shapes
    against the FP64 reference sequence push_v!; push_x!(dt/2);
    push_x!(dt/2); compute_rho from the same start: positions within FP32
    resolution of a cell, velocities within 1e-6, charge within 1e-4
    normwise.  NGP's node can differ where FP32 and FP64 round a position
    on either side of a half cell: at most a few particles of 20,000."""
    rng = np.random.default_rng(22)
    nx, ny, dimx, dimy = 32, 32, 4 * np.pi, 4 * np.pi
    dx, dy = dimx / nx, dimy / ny
    n = 20000
    p = random_particles(rng, n, dimx, dimy)
    ex = np.asfortranarray(rng.normal(size=(nx, ny)) * 0.1)
    ey = np.asfortranarray(rng.normal(size=(nx, ny)) * 0.1)
    dt = 0.1
    st = O.f32_from_f64(p, nx, ny, dx, dy)
    p32 = O.f32_to_f64(st, nx, dx, dy, p[4, 0])
    acc = O.f32_step(st, nx, ny, kernel, _e32(ex, ey), np.float32(dt / 2), np.float32(dt / dx),
                     np.float32(dt / dy), np.float32(p[4, 0]))
    O.push_v(p32, kernel, nx, ny, dx, dy, ex, ey, None, dt)
    O.push_x(p32, dimx, dimy, dt / 2)
    O.push_x(p32, dimx, dimy, dt / 2)
    q = O.f32_to_f64(st, nx, dx, dy, p[4, 0])
    dv = np.abs(q[2] - p32[2])
    off = dv > 1e-6
    assert off.sum() <= (3 if kernel == "ngp" else 0)
    dxw = np.abs(q[0] - p32[0])[~off]
    dxw = np.minimum(dxw, dimx - dxw)
    assert dxw.max() < 1e-5 * dx
    rho64 = O.compute_rho(p32, kernel, nx, ny, dx, dy, dimx, dimy)
    rho = acc / (dx * dy)
    rho -= rho.sum() * dx * dy / dimx / dimy
    assert np.linalg.norm(rho - rho64) / np.linalg.norm(rho64) < 1e-4
    assert math.isclose(acc.sum(), n * p[4, 0], rel_tol=1e-6)


@pytest.mark.parametrize("kernel", SEP)
def test_f32_deposit_single_particle_moments(O, kernel):
    """One particle's FP32 deposit carries its charge w, with its first
    moment at the particle (M6, TSC, M4) or at its rounded node (NGP)."""
    rng = np.random.default_rng(75)
    nx = ny = 16
    for _ in range(25):
        c = (int(rng.integers(4, 11)), int(rng.integers(4, 11)))
        o = rng.uniform(0, 1, 2).astype(np.float32)
        st = (np.array([c[0] + nx * c[1]], np.int32), o[:1].copy(), o[1:].copy(),
              np.zeros(1, np.float32), np.zeros(1, np.float32))
        w = np.float32(0.7)
        acc = O.f32_deposit(st, nx, ny, kernel, w)
        assert acc.sum() == pytest.approx(float(w), rel=1e-6)
        gi, gj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        mx, my = (acc * gi).sum() / acc.sum(), (acc * gj).sum() / acc.sum()
        if kernel == "ngp":
            ex_ = c[0] + (1 if o[0] > 0.5 else 0)
            ey_ = c[1] + (1 if o[1] > 0.5 else 0)
            assert (mx, my) == pytest.approx((ex_, ey_), abs=1e-6)
        else:
            assert (mx, my) == pytest.approx((c[0] + o[0], c[1] + o[1]), abs=2e-6)
