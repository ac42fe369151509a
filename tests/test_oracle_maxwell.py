"""The oracle's Vlasov-Maxwell pieces (src/fdtd.jl, src/cloud_in_cell.jl)
pinned by the reference's own tests: the analytic Maxwell mode of
test/test_fdtd.jl, the known total current of test/test_deposition.jl, the
Fortran twin's deposition (src/pic.f90) and charge-free current sums."""
import math

import numpy as np
import pytest


def test_fdtd_analytic_mode(O):
    """test/test_fdtd.jl:5-60 restated: a TE mode of the periodic box, 8
    steps of ampere_maxwell!, update_fields!, faraday!, against the exact
    solution to the reference's bounds (1e-6, and 1e-3 for the node bz)."""
    dimx = dimy = 2 * math.pi
    nx = ny = 128
    dt, nstep = 1e-4, 8
    dx, dy = dimx / nx, dimy / ny
    f = O.FDTD(nx, ny)
    w = math.sqrt(2)
    xn = np.linspace(0, dimx, nx + 1)[:, None]
    yn = np.linspace(0, dimy, ny + 1)[None, :]
    xc = ((xn[:-1] + xn[1:]) / 2)
    yc = ((yn[:, :-1] + yn[:, 1:]) / 2)
    t = 0.5 * dt
    f.bz[...] = -np.cos(xc) * np.cos(yc) * math.cos(w * t)
    jx = np.zeros((nx + 1, ny + 1), order="F")
    jy = np.zeros((nx + 1, ny + 1), order="F")
    for _ in range(nstep):
        O.ampere_maxwell(f, nx, ny, dx, dy, jx, jy, dt)
        t += 0.5 * dt
        assert np.max(np.abs(f.ex - np.cos(xc) * np.sin(yn) / w * math.sin(w * t))) < 1e-6
        assert np.max(np.abs(f.ey + np.sin(xn) * np.cos(yc) / w * math.sin(w * t))) < 1e-6
        ex, ey, bz = O.update_fields(f, nx, ny)
        assert np.max(np.abs(ex - np.cos(xn) * np.sin(yn) / w * math.sin(w * t))) < 1e-6
        assert np.max(np.abs(ey + np.sin(xn) * np.cos(yn) / w * math.sin(w * t))) < 1e-6
        O.faraday(f, nx, ny, dx, dy, dt)
        t += 0.5 * dt
        assert np.max(np.abs(f.bz + np.cos(xc) * np.cos(yc) * math.cos(w * t))) < 1e-6
        ex, ey, bz = O.update_fields(f, nx, ny)
        assert np.max(np.abs(bz + np.cos(xn) * np.cos(yn) * math.cos(w * t))) < 1e-3


def test_compute_current_known_total(O):
    """test/test_deposition.jl: 128^2 on [0, 4pi]^2, 1e6 particles around
    2pi, v = w = (4pi)^2/1e6: sum(jx) = sum(jy) = 2.5872575761191685 (the
    particles come from numpy's generator, so the sums agree to sampling,
    1.5e-8); the interior sums equal the particle currents exactly."""
    rng = np.random.default_rng(1234)
    n, nx = 1_000_000, 128
    dim = 4 * math.pi
    p = np.zeros((5, n), order="F")
    p[0:2] = np.clip(rng.normal(size=(2, n)) + 2 * math.pi, 0, dim - 1e-9)
    p[2:5] = dim * dim / n
    jx, jy = O.compute_current(p, nx, nx, dim / nx, dim / nx)
    assert math.isclose(jx.sum(), 2.5872575761191685, rel_tol=1.5e-8)
    assert math.isclose(jy.sum(), 2.5872575761191685, rel_tol=1.5e-8)
    dxdy = (dim / nx) ** 2
    assert math.isclose(jx[:nx, :nx].sum(), np.sum(p[2] * p[4]) / dxdy, rel_tol=1e-12)


def test_compute_current_matches_fortran_twin(O):
    """The Julia method against src/pic.f90's deposition (whose scale factor
    nx*ny/n is computed in float32, pic.f90:116)."""
    if not O.ref_available():
        pytest.skip("reference Fortran twin not built")
    rng = np.random.default_rng(5)
    n, nx, ny = 20000, 32, 16
    dimx, dimy = 4 * math.pi, 1.0
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(0, dimx, n)
    p[1] = rng.uniform(0, dimy, n)
    p[2:4] = rng.normal(size=(2, n))
    p[4] = dimx * dimy / n
    jx, jy = O.compute_current(p, nx, ny, dimx / nx, dimy / ny)
    rx, ry = O.ref_deposition(p, nx, ny, dimx / nx, dimy / ny)
    # the twin leaves the ghost row and column unfolded; fold them as
    # compute_current! does (cloud_in_cell.jl:66-79)
    for a in (rx, ry):
        a[:, 0] += a[:, ny]
        a[:, ny] = a[:, 0]
        a[0, :] += a[nx, :]
        a[nx, :] = a[0, :]
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(jx, rx) < 1e-7 and rel(jy, ry) < 1e-7


def test_maxwell_loop_landau(O):
    """The loop of examples/landau_damping_maxwell_fdtd.jl on a small case:
    the field energy log decreases over the first steps (the Landau mode is
    damped) and stays finite."""
    import __graft_entry__ as entry
    pic = entry.load_package()
    nx, ny = 32, 8
    kx, alpha = 0.5, 0.1
    mesh = pic.TwoDGrid(0.0, 2 * math.pi / kx, nx, 0.0, 1.0, ny)
    p = np.asfortranarray(pic.sampling.landau_damping(100 * nx * ny, mesh, alpha, kx))
    f = O.FDTD(nx, ny)
    x = np.linspace(0, mesh.dimx, nx + 1)[:nx]
    f.ex[...] = (alpha / kx * np.sin(kx * x))[:, None]
    e = O.maxwell_loop(p, f, nx, ny, mesh.dimx, mesh.dimy, 0.01, 40)
    assert np.all(np.isfinite(e))
    assert e[-1] < e[0]
