"""The GPU UA step (src/ua_steps.jl, the UA deposit and gather of
src/compute_rho_m6.jl:29-179 and src/interpolation_m6.jl:3-123, the loop of
test/bupdate.jl) against the oracle.  The GPU's sin, cos and tau-FFTs round
differently from the oracle's libm and radix-2 transform (and FFTW's), so
the bars are relative tolerances written in each test."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMX, DIMY = 4 * math.pi, 2 * math.pi


def _particles(pic, rng, n):
    p = pic.Particles(n, DIMX * DIMY / n)
    p.x[0] = rng.uniform(0, DIMX, n)
    p.x[1] = rng.uniform(0, DIMY, n)
    p.v[...] = rng.normal(size=(2, n))
    p.e[...] = 0.3 * rng.normal(size=(2, n))
    return p


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("ntau", [4, 16, 64])
def test_preparation_matches_oracle(gpu, O, ntau):
    pic = gpu
    rng = np.random.default_rng(ntau)
    n, eps, dt = 3000, 0.1, math.pi / 16
    p = _particles(pic, rng, n)
    ua = pic.UA(ntau, eps, n)
    xt, yt = (np.zeros((ntau, 2, n), np.complex128, order="F") for _ in range(2))
    pic.preparation_(ua, dt, p, xt, yt)
    b, t, pl, ql, oxt, oyt = O.ua_preparation(ntau, eps, dt, p.x, p.v, p.e)
    assert _rel(p.b, b) < 1e-15 and _rel(p.t, t) < 1e-15
    for a, o in ((ua.pl, pl), (ua.ql, ql), (xt, oxt), (yt, oyt)):
        assert _rel(a, o) < 1e-12


def _fields(pic, rng, mesh):
    f = pic.MeshFields2D(mesh)
    e = rng.normal(size=(2, mesh.nx, mesh.ny))
    f.e[:, :mesh.nx, :mesh.ny] = e
    f.e[:, mesh.nx, :] = f.e[:, 0, :]
    f.e[:, :, mesh.ny] = f.e[:, :, 0]
    return f


def test_interpol_matches_oracle(gpu, O):
    """update_particles_e!: the M6 gather at every tau-position."""
    pic = gpu
    rng = np.random.default_rng(2)
    ntau, n = 16, 2000
    mesh = pic.TwoDGrid(0.0, DIMX, 64, 0.0, DIMY, 32)
    f = _fields(pic, rng, mesh)
    p = _particles(pic, rng, n)
    ua = pic.UA(ntau, 0.1, n)
    xt = np.asfortranarray(rng.uniform(-1.0, DIMX + 1.0, size=(ntau, 2, n)) + 0j)
    et = np.zeros((ntau, 2, n), order="F")
    pic.update_particles_e_(p, et, f, ua, xt)
    oet = O.ua_interpol(xt, f.e, mesh.nx, mesh.ny, 0.0, DIMX, 0.0, DIMY)
    assert _rel(et, oet) < 1e-13


def test_compute_f_and_ua_step_match_oracle(gpu, O):
    """compute_f! (with its tau-FFTs) and both ua_step! forms, from the
    oracle's own preparation! outputs."""
    pic = gpu
    rng = np.random.default_rng(3)
    ntau, n, eps, dt = 16, 2000, 0.1, math.pi / 16
    p = _particles(pic, rng, n)
    b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, p.x, p.v, p.e)
    p.b[...], p.t[...] = b, t
    ua = pic.UA(ntau, eps, n)
    et = np.asfortranarray(rng.normal(size=(ntau, 2, n)))
    fx, fy = (np.zeros((ntau, 2, n), np.complex128, order="F") for _ in range(2))
    pic.compute_f_(fx, fy, ua, p, xt, yt, et)
    ofx, ofy = O.ua_compute_f(eps, b, xt, yt, et)
    assert _rel(fx, ofx) < 1e-12 and _rel(fy, ofy) < 1e-12
    xtil = np.asfortranarray(np.fft.fft(xt, axis=0))
    gx = np.asfortranarray(rng.normal(size=(ntau, 2, n)) + 1j * rng.normal(size=(ntau, 2, n)))
    out = np.zeros((ntau, 2, n), np.complex128, order="F")
    pic.ua_step_(out, xtil, ua, p, ofx)
    assert _rel(out, O.ua_step(eps, t, pl, xtil, ofx)) < 1e-12
    pic.ua_step_(out, xtil, ua, p, ofx, gx)
    assert _rel(out, O.ua_step(eps, t, pl, xtil, ofx, ql, gx)) < 1e-12


def test_compute_rho_and_compute_v_match_oracle(gpu, O):
    """update_particles_x! (reconstruction, Particles.x, centred M6 deposit,
    MeshFields rho) and compute_v!."""
    pic = gpu
    rng = np.random.default_rng(4)
    ntau, n, eps, dt = 16, 5000, 0.1, math.pi / 16
    mesh = pic.TwoDGrid(0.0, DIMX, 64, 0.0, DIMY, 32)
    p = _particles(pic, rng, n)
    b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, p.x, p.v, p.e)
    p.b[...], p.t[...] = b, t
    ua = pic.UA(ntau, eps, n)
    f = pic.MeshFields2D(mesh)
    pic.update_particles_x_(p, f, ua, xt)
    ox, orho = O.ua_compute_rho(eps, t, xt, p.w, mesh.nx, mesh.ny, 0.0, DIMX, 0.0, DIMY)
    assert _rel(p.x, ox) < 1e-13
    assert _rel(f.rho, orho) < 1e-11
    ytil = np.asfortranarray(np.fft.fft(yt, axis=0))
    pic.compute_v_(ytil, p, ua)
    assert _rel(p.v, O.ua_compute_v(eps, t, ytil)) < 1e-12


def _bupdate_start(pic, O, n, nx, ny, seed):
    """Particles of test/bupdate.jl's problem (plasma(mesh, n)'s densities,
    sampled here with numpy) and the prologue on the oracle side."""
    rng = np.random.default_rng(seed)
    kx, alpha = 0.5, 0.05
    x, y = [], []
    while len(x) < n:
        xi, yi = rng.uniform(0, DIMX, n), rng.uniform(0, DIMY, n)
        zi = (2.0 + alpha) * rng.uniform(size=n)
        keep = 1.0 + np.sin(yi) + alpha * np.cos(kx * xi) >= zi
        x.extend(xi[keep])
        y.extend(yi[keep])
    p = pic.Particles(n, DIMX * DIMY / n)
    p.x[0], p.x[1] = x[:n], y[:n]
    p.v[...] = rng.normal(size=(2, n)) + np.array([[2.0], [0.0]]) * rng.choice([-1, 1], n)
    xo = np.asfortranarray(p.x.copy())
    rho = O.compute_rho_m6_fields(xo, p.w, nx, ny, 0.0, DIMX, 0.0, DIMY, compat=True)
    e_f, _ = O.solve_fields(rho, 0.0, DIMX, 0.0, DIMY)
    eo = O.interpol_eb_m6(xo, e_f, nx, ny, 0.0, DIMX, 0.0, DIMY)
    return p, xo, e_f, eo


@pytest.mark.parametrize("ntau,nx,ny,n,steps", [(16, 128, 64, 20000, 4), (8, 32, 16, 3000, 6)])
def test_ua_loop_matches_oracle(gpu, O, ntau, nx, ny, n, steps):
    """pic_ua_run against the oracle's loop of test/bupdate.jl from the same
    prologue (eps 0.1, dt = pi/16): positions, velocities, fields and the
    solve! energies within 1e-9.  The loop queues its energy sums on the
    device and reads them once; pic_solve_fields on the final charge gives
    the last one bit for bit."""
    pic = gpu
    eps, dt = 0.1, math.pi / 16
    mesh = pic.TwoDGrid(0.0, DIMX, nx, 0.0, DIMY, ny)
    p, xo, e_f, eo = _bupdate_start(pic, O, n, nx, ny, ntau + n)
    sim = pic.UASimulation(mesh, p, ntau, eps)
    en = sim.run(dt, steps)
    x, v = sim.particles()
    ef = sim.fields_e()
    again = ctypes.c_double()
    pic._lib.call("pic_solve_fields", sim.ctx.h, ctypes.byref(again))
    assert again.value == en[1, -1]
    sim.close()
    vo = np.asfortranarray(p.v.copy())
    oen = O.ua_run(ntau, eps, dt, steps, xo, vo, eo, p.w, e_f, 0.0, DIMX, 0.0, DIMY)
    assert _rel(en, oen) < 1e-9
    assert _rel(x, xo) < 1e-9 and _rel(v, vo) < 1e-9
    assert _rel(ef, e_f) < 1e-9


def test_ua_rejects_bad_input(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(0.0, DIMX, 16, 0.0, DIMY, 16)
    p = _particles(pic, np.random.default_rng(5), 100)
    with pytest.raises(pic._lib.PicError):
        pic.UASimulation(mesh, p, 12, 0.1)  # ntau not a power of two
    with pytest.raises(pic._lib.PicError):
        pic.UASimulation(mesh, p, 16, 0.0)  # eps must be positive
