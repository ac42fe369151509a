"""The oracle's NGP, TSC and M4 shapes (pushers_ngp.jl, pushers_tsc.jl,
pushers_m4.jl) pinned by what the mathematics fixes: B-spline moments and
symmetry, values at the nodes, charge conservation, first-moment
reproduction by the deposits, exact gathers of linear fields, and
round-half-to-even for NGP (against numpy's rint)."""
import math

import numpy as np
import pytest

from conftest import random_particles


def _weights(f, d):
    # the reference's order: offsets -1, 0, +1, +2 around the cell's node
    return np.array([f(1 + d), f(d), f(1 - d), f(2 - d)])


@pytest.mark.parametrize("name,var", [("tsc", 0.25), ("m4", 1.0 / 3.0)])
def test_shape_weight_moments(O, name, var):
    """Quadratic (TSC) and cubic (M4) B-splines: partition of unity, zero
    first moment about the particle, variance (order + 1)/12."""
    f = O.f_tsc if name == "tsc" else O.f_m4
    offs = np.array([-1.0, 0.0, 1.0, 2.0])
    for d in np.linspace(0.0, 0.999, 37):
        w = _weights(f, d)
        assert w.sum() == pytest.approx(1.0, abs=1e-15)
        assert np.dot(w, offs - d) == pytest.approx(0.0, abs=1e-15)
        assert np.dot(w, (offs - d) ** 2) == pytest.approx(var, abs=1e-15)
        # symmetry: w_k(d) = w_{3-k}(1-d)
        assert np.allclose(w, _weights(f, 1.0 - d)[::-1], atol=1e-15)


def test_shape_node_values(O):
    assert O.f_tsc(0.0) == 0.75 and O.f_tsc(0.5) == 0.5 and O.f_tsc(1.0) == 0.125
    assert O.f_tsc(1.5) == 0.0 and O.f_tsc(2.0) == 0.0
    assert O.f_m4(0.0) == pytest.approx(2 / 3, abs=1e-16)
    assert O.f_m4(1.0) == pytest.approx(1 / 6, abs=1e-16)
    assert O.f_m4(2.0) == 0.0 and O.f_m4(3.0) == 0.0


@pytest.mark.parametrize("kernel", ["ngp", "tsc", "m4"])
def test_shape_deposit_charge_and_first_moment(O, kernel):
    """One particle away from the boundary: the raw deposit (rho dx dy + the
    removed mean) carries its charge w, and its first moment sits at the
    particle (TSC, M4) or at the rounded node (NGP)."""
    nx = ny = 16
    dimx, dimy = 4.0, 2.0
    dx, dy = dimx / nx, dimy / ny
    rng = np.random.default_rng(3)
    for _ in range(20):
        x, y = rng.uniform(5.2, 9.7) * dx, rng.uniform(5.2, 9.7) * dy
        w = 0.7
        p = np.asfortranarray(np.array([[x], [y], [0.0], [0.0], [w]]))
        rho = O.compute_rho(p, kernel, nx, ny, dx, dy, dimx, dimy)
        raw = (rho + w / (dimx * dimy)) * dx * dy  # undo the mean removal
        assert raw.sum() == pytest.approx(w, rel=1e-14)
        gx, gy = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
        mx, my = (raw * gx).sum() / w, (raw * gy).sum() / w
        if kernel == "ngp":
            assert mx == pytest.approx(np.rint(x / dx) * dx, abs=1e-13)
            assert my == pytest.approx(np.rint(y / dy) * dy, abs=1e-13)
        else:
            assert mx == pytest.approx(x, abs=1e-13)
            assert my == pytest.approx(y, abs=1e-13)


def test_ngp_rounds_half_to_even(O):
    """Particles exactly half-way between nodes go to the even node, as
    Julia's round(Int, x) and numpy's rint do."""
    nx = ny = 8
    dx = dy = 1.0
    xs = np.array([0.5, 1.5, 2.5, 3.5, 4.5, 5.25, 5.75])
    p = np.zeros((5, xs.size), order="F")
    p[0], p[1], p[4] = xs, 2.0, 1.0
    rho = O.compute_rho(p, "ngp", nx, ny, dx, dy, 8.0, 8.0)
    raw = (rho + xs.size / 64.0) * dx * dy
    want = np.zeros(nx)
    for x in xs:
        want[int(np.rint(x)) % nx] += 1.0
    assert np.allclose(raw[:, 2], want, atol=1e-13)


@pytest.mark.parametrize("kernel", ["ngp", "tsc", "m4"])
def test_shape_push_uniform_and_linear_fields(O, kernel):
    """Uniform fields: every shape's kick equals CIC's (the weights sum to
    one).  A field linear in x and y (periodic pieces away from the seam):
    TSC and M4, which gather with the cubic spline, reproduce it exactly like
    CIC; NGP takes the rounded node's value."""
    nx, ny = 32, 32
    dimx, dimy = 8.0, 8.0
    dx, dy = dimx / nx, dimy / ny
    rng = np.random.default_rng(5)
    p0 = random_particles(rng, 500, dimx, dimy)
    p0[0] = rng.uniform(3 * dx, dimx - 4 * dx, 500)
    p0[1] = rng.uniform(3 * dy, dimy - 4 * dy, 500)
    ex = np.full((nx, ny), 0.3, order="F")
    ey = np.full((nx, ny), -0.2, order="F")
    bz = np.full((nx, ny), 1.5, order="F")
    a, b = p0.copy(order="F"), p0.copy(order="F")
    O.push_v(a, "cic", nx, ny, dx, dy, ex, ey, bz, 0.1)
    O.push_v(b, kernel, nx, ny, dx, dy, ex, ey, bz, 0.1)
    assert np.allclose(a, b, rtol=0, atol=1e-14)
    gx, gy = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    ex = np.asfortranarray(0.1 + 0.05 * gx + 0.02 * gy)
    ey = np.asfortranarray(-0.3 + 0.01 * gx - 0.04 * gy)
    a, b = p0.copy(order="F"), p0.copy(order="F")
    O.push_v(a, "cic", nx, ny, dx, dy, ex, ey, None, 0.1)
    O.push_v(b, kernel, nx, ny, dx, dy, ex, ey, None, 0.1)
    if kernel == "ngp":
        xr, yr = np.rint(p0[0] / dx) * dx, np.rint(p0[1] / dy) * dy
        e1 = 0.1 + 0.05 * xr + 0.02 * yr
        assert np.allclose(b[2], p0[2] + 0.1 * e1, atol=1e-14)
    else:
        assert np.allclose(a[2:4], b[2:4], rtol=0, atol=1e-14)


@pytest.mark.parametrize("kernel", ["ngp", "tsc", "m4"])
def test_shape_landau_loop_runs(O, kernel):
    """The loop with each shape (examples/landau_damping_poisson_cic_m6.jl
    :88-103): finite energies, and every shape sees the same initial field
    mode to a few percent."""
    import __graft_entry__ as entry
    pic = entry.load_package()
    nx = ny = 32
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, nx, 0.0, 4 * math.pi, ny)
    p = np.asfortranarray(pic.sampling.landau_damping(20000, mesh, 0.05, 0.5))
    e = O.landau_loop(p.copy(order="F"), kernel, nx, ny, mesh.dimx, mesh.dimy, 0.1, 3)
    ec = O.landau_loop(p.copy(order="F"), "cic", nx, ny, mesh.dimx, mesh.dimy, 0.1, 3)
    assert np.all(np.isfinite(e))
    assert e[0] == pytest.approx(ec[0], rel=0.1)
