"""The oracle's UA pieces (src/ua_type.jl, src/ua_steps.jl, the UA deposit of
src/compute_rho_m6.jl:29-179 and gather of src/interpolation_m6.jl:3-123,
the loop of test/bupdate.jl) pinned by what the mathematics fixes: numpy's
FFT, the exact solution of the linear tau-ODE each ua_step! advances
(scipy's integrator), trigonometric interpolation of band-limited samples,
the zero-velocity closed form of preparation!, and the MeshFields M6 gather
and deposit the UA versions reduce to."""
import math

import numpy as np
import pytest


def _particles(rng, n, dimx, dimy):
    x = np.asfortranarray(np.stack([rng.uniform(0, dimx, n), rng.uniform(0, dimy, n)]))
    v = np.asfortranarray(rng.normal(size=(2, n)))
    e = np.asfortranarray(0.3 * rng.normal(size=(2, n)))
    return x, v, e


def _ltau(ntau):
    return np.array([m if m < ntau // 2 else m - ntau for m in range(ntau)], dtype=float)


@pytest.mark.parametrize("ntau", [2, 4, 8, 16, 32, 64])
def test_fft_matches_numpy(O, ntau):
    rng = np.random.default_rng(ntau)
    a = rng.normal(size=ntau) + 1j * rng.normal(size=ntau)
    assert np.max(np.abs(O.fft_c(a, -1) - np.fft.fft(a))) < 1e-14 * ntau
    assert np.max(np.abs(O.fft_c(a, +1) - np.fft.ifft(a) * ntau)) < 1e-14 * ntau


def test_preparation_b_t_and_tau_circle(O):
    """b = 1 + sin(x1) sin(x2)/2, t = dt b; xt(tau) runs on the circle of
    radius eps |v| / b around x + eps (vy, -vx) / b and starts at x."""
    rng = np.random.default_rng(3)
    ntau, eps, dt, n = 16, 0.1, math.pi / 16, 50
    x, v, e = _particles(rng, n, 4 * math.pi, 2 * math.pi)
    b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, x, v, e)
    assert np.allclose(b, 1 + 0.5 * np.sin(x[0]) * np.sin(x[1]), rtol=0, atol=1e-15)
    assert np.allclose(t, dt * b, rtol=1e-15, atol=0)
    assert np.all(xt.imag == 0)
    centre = x + eps * np.stack([v[1], -v[0]]) / b
    rad = np.abs((xt.real[:, 0, :] - centre[0]) + 1j * (xt.real[:, 1, :] - centre[1]))
    assert np.allclose(rad, eps * np.hypot(v[0], v[1]) / b, rtol=1e-12, atol=1e-15)
    assert np.allclose(xt.real[0], x, rtol=0, atol=1e-14)


def test_preparation_pl_ql_are_the_step_integrals(O):
    """pl = int_0^t exp(-i l (t-s)/eps) ds and ql = int_0^t exp(-i l (t-s)/eps) s ds
    (the weights of the exact step of u' = -i l u / eps + f(s)), by quadrature."""
    from scipy.integrate import quad
    rng = np.random.default_rng(4)
    ntau, eps, dt, n = 8, 0.1, math.pi / 16, 3
    x, v, e = _particles(rng, n, 4 * math.pi, 2 * math.pi)
    b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, x, v, e)
    lt = _ltau(ntau)
    for k in range(n):
        for m in range(ntau):
            def g(s, part, pw):
                z = np.exp(-1j * lt[m] * (t[k] - s) / eps) * s ** pw
                return z.real if part == 0 else z.imag
            for arr, pw in ((pl, 0), (ql, 1)):
                re = quad(g, 0, t[k], args=(0, pw), epsabs=1e-14, epsrel=1e-13, limit=200)[0]
                im = quad(g, 0, t[k], args=(1, pw), epsabs=1e-14, epsrel=1e-13, limit=200)[0]
                assert abs(arr[m, k] - (re + 1j * im)) < 1e-12


def test_preparation_yt_zero_velocity_closed_form(O):
    """v = 0: xt = x, the bracket of interv vanishes and r(tau) is the field
    rotated by tau over b, whose zero-mean antiderivative is exact on the tau
    grid: yt1 = eps (sin tau exb + (cos tau - 1) eyb),
    yt2 = eps ((1 - cos tau) exb + sin tau eyb)."""
    rng = np.random.default_rng(5)
    ntau, eps, dt, n = 16, 0.1, 0.2, 40
    x, v, e = _particles(rng, n, 4 * math.pi, 2 * math.pi)
    v[...] = 0.0
    b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, x, v, e)
    tau = 2 * math.pi / ntau * np.arange(ntau)[:, None]
    exb, eyb = e[0] / b, e[1] / b
    y1 = eps * (np.sin(tau) * exb + (np.cos(tau) - 1) * eyb)
    y2 = eps * ((1 - np.cos(tau)) * exb + np.sin(tau) * eyb)
    assert np.max(np.abs(yt[:, 0, :] - y1)) < 1e-15
    assert np.max(np.abs(yt[:, 1, :] - y2)) < 1e-15
    assert np.max(np.abs(xt.real - x[None, :, :])) == 0.0


def test_preparation_yt_affine_in_e(O):
    """yt depends affinely on the particle's field: yt(e1 + e2) - yt(e1) -
    yt(e2) + yt(0) = 0 to rounding (a dropped or doubled field term fails)."""
    rng = np.random.default_rng(6)
    ntau, eps, dt, n = 16, 0.1, 0.2, 30
    x, v, e1 = _particles(rng, n, 4 * math.pi, 2 * math.pi)
    e2 = np.asfortranarray(rng.normal(size=(2, n)))
    y = [O.ua_preparation(ntau, eps, dt, x, v, ee)[5]
         for ee in (e1 + e2, e1, e2, np.zeros((2, n), order="F"))]
    assert np.max(np.abs(y[0] - y[1] - y[2] + y[3])) < 1e-13


@pytest.mark.parametrize("corrected", [False, True])
def test_ua_step_is_the_exact_linear_step(O, corrected):
    """ua_step! advances every tau-mode of u' = -i l u / eps + f(s) exactly
    from 0 to t, with f = fx (prediction) or f(s) = fx + (gx - fx) s / t
    (correction): against scipy's integrator at tight tolerances."""
    from scipy.integrate import solve_ivp
    rng = np.random.default_rng(7)
    ntau, eps, dt, n = 8, 0.1, 0.3, 2
    x, v, e = _particles(rng, n, 4 * math.pi, 2 * math.pi)
    b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, x, v, e)
    shape = (ntau, 2, n)
    u0 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    fx = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    gx = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    out = O.ua_step(eps, t, pl, u0, fx, ql if corrected else None, gx if corrected else None)
    lt = _ltau(ntau)
    for k in range(n):
        for c in range(2):
            for m in range(ntau):
                a = -1j * lt[m] / eps

                def rhs(s, y, m=m, c=c, k=k):
                    u = y[0] + 1j * y[1]
                    f = fx[m, c, k] + ((gx[m, c, k] - fx[m, c, k]) * s / t[k] if corrected else 0)
                    d = a * u + f
                    return [d.real, d.imag]
                sol = solve_ivp(rhs, (0, t[k]), [u0[m, c, k].real, u0[m, c, k].imag],
                                method="DOP853", rtol=1e-12, atol=1e-13)
                ref = sol.y[0, -1] + 1j * sol.y[1, -1]
                assert abs(out[m, c, k] - ref) < 1e-9


def _bandlimited(rng, ntau, n):
    """Samples on the tau grid of g(tau) = a + b e^{i tau} + c e^{-2 i tau},
    per particle and component, and g itself."""
    coef = rng.normal(size=(3, 2, n)) + 1j * rng.normal(size=(3, 2, n))
    tau = 2 * math.pi / ntau * np.arange(ntau)
    g = lambda s: coef[0] + coef[1] * np.exp(1j * s) + coef[2] * np.exp(-2j * s)
    samples = np.stack([g(s) for s in tau])
    return samples, g


def test_compute_v_evaluates_the_tau_series(O):
    """compute_v!: yt holds the FFT of tau-samples of a band-limited g; the
    series at tau = t/eps is g(t/eps), rotated by t/eps."""
    rng = np.random.default_rng(8)
    ntau, eps, n = 16, 0.1, 20
    samples, g = _bandlimited(rng, ntau, n)
    yt = np.asfortranarray(np.fft.fft(samples, axis=0))
    t = rng.uniform(0.05, 0.3, n)
    v = O.ua_compute_v(eps, t, yt)
    s = t / eps
    gv = g(s)
    assert np.allclose(v[0], (np.cos(s) * gv[0] + np.sin(s) * gv[1]).real, atol=1e-12)
    assert np.allclose(v[1], (np.cos(s) * gv[1] - np.sin(s) * gv[0]).real, atol=1e-12)


def test_compute_rho_reconstruction_and_deposit(O):
    """update_particles_x!: x(t) = real g(t/eps) brought into the domain for
    tau-samples of a band-limited g, and rho equals the centred MeshFields
    M6 deposit (compat = 0) at those positions."""
    rng = np.random.default_rng(9)
    ntau, eps, n, nx, ny = 16, 0.1, 300, 32, 16
    dimx, dimy = 4 * math.pi, 2 * math.pi
    samples, g = _bandlimited(rng, ntau, n)
    samples += np.array([dimx / 2, dimy / 2])[None, :, None]
    t = rng.uniform(0.05, 0.3, n)
    w = dimx * dimy / n
    x, rho = O.ua_compute_rho(eps, t, np.asfortranarray(samples), w, nx, ny, 0.0, dimx, 0.0, dimy)
    gv = g(t / eps).real + np.array([dimx / 2, dimy / 2])[:, None]
    want = np.stack([np.mod(gv[0], dimx), np.mod(gv[1], dimy)])
    assert np.allclose(x, want, atol=1e-12)
    x2 = np.asfortranarray(x.copy())
    rho2 = O.compute_rho_m6_fields(x2, w, nx, ny, 0.0, dimx, 0.0, dimy, compat=False)
    assert np.max(np.abs(rho - rho2)) <= 1e-13 * np.max(np.abs(rho2))


def test_ua_interpol_is_the_meshfields_gather(O):
    """update_particles_e!: every (tau, particle) position gathers as
    interpol_eb_m6!(particles, fields) does; a linear field is reproduced
    inside the domain (test/test_uapic.jl's bound, 1e-6)."""
    rng = np.random.default_rng(10)
    ntau, n, nx, ny = 8, 40, 20, 20
    xmin, xmax, ymin, ymax = 0.0, 20.0, 0.0, 20.0
    dx, dy = 1.0, 1.0
    e_f = np.zeros((2, nx + 1, ny + 1), order="F")
    e_f[0] = (np.arange(nx + 1) * dx)[:, None]
    e_f[1] = (np.arange(ny + 1) * dy)[None, :]
    xt = np.asfortranarray(rng.uniform(6.0, 14.0, size=(ntau, 2, n)) + 0j)
    et = O.ua_interpol(xt, e_f, nx, ny, xmin, xmax, ymin, ymax)
    assert np.max(np.abs(et - xt.real)) < 1e-6
    e_r = np.asfortranarray(rng.normal(size=(2, nx + 1, ny + 1)))
    et = O.ua_interpol(xt, e_r, nx, ny, xmin, xmax, ymin, ymax)
    pts = np.asfortranarray(xt.real.transpose(1, 0, 2).reshape(2, -1))
    ep = O.interpol_eb_m6(pts, e_r, nx, ny, xmin, xmax, ymin, ymax)
    assert np.array_equal(et.transpose(1, 0, 2).reshape(2, -1), ep)


def ua_loop_from_pieces(O, ntau, eps, dt, nsteps, x, v, e, w, e_f, bounds):
    """test/bupdate.jl:62-116 written with the oracle's per-function pieces."""
    xmin, xmax, ymin, ymax = bounds
    nx, ny = e_f.shape[1] - 1, e_f.shape[2] - 1

    def fft(a, sign):
        out = np.empty_like(a)
        for c in range(a.shape[1]):
            for k in range(a.shape[2]):
                out[:, c, k] = O.fft_c(a[:, c, k], sign)
        return out if sign < 0 else out / ntau

    en = []
    for _ in range(nsteps):
        b, t, pl, ql, xt, yt = O.ua_preparation(ntau, eps, dt, x, v, e)
        et = O.ua_interpol(xt, e_f, nx, ny, *bounds)
        fx, fy = O.ua_compute_f(eps, b, xt, yt, et)
        xtt, ytt = fft(xt, -1), fft(yt, -1)
        xt = fft(O.ua_step(eps, t, pl, xtt, fx), +1)
        yt = fft(O.ua_step(eps, t, pl, ytt, fy), +1)
        x, rho = O.ua_compute_rho(eps, t, xt, w, nx, ny, *bounds)
        e_f, en1 = O.solve_fields(rho, *bounds)
        et = O.ua_interpol(xt, e_f, nx, ny, *bounds)
        gx, gy = O.ua_compute_f(eps, b, xt, yt, et)
        xt = fft(O.ua_step(eps, t, pl, xtt, fx, ql, gx), +1)
        yt = O.ua_step(eps, t, pl, ytt, fy, ql, gy)
        x, rho = O.ua_compute_rho(eps, t, xt, w, nx, ny, *bounds)
        e_f, en2 = O.solve_fields(rho, *bounds)
        v = O.ua_compute_v(eps, t, yt)
        en.append((en1, en2))
    return x, v, e_f, np.array(en).T


def test_ua_run_is_the_loop_of_its_pieces(O):
    """or_ua_run against the loop composed in Python from the per-function
    pieces, on the same bits: the wiring of bupdate.jl's loop."""
    rng = np.random.default_rng(11)
    ntau, eps, dt, nsteps, n, nx, ny = 8, 0.1, math.pi / 16, 2, 64, 16, 8
    bounds = (0.0, 4 * math.pi, 0.0, 2 * math.pi)
    x, v, e = _particles(rng, n, 4 * math.pi, 2 * math.pi)
    w = 8 * math.pi ** 2 / n
    e_f = np.asfortranarray(0.1 * rng.normal(size=(2, nx + 1, ny + 1)))
    xa, va, ea = x.copy(order="F"), v.copy(order="F"), e_f.copy(order="F")
    en_a = O.ua_run(ntau, eps, dt, nsteps, xa, va, e, w, ea, *bounds)
    xb, vb, eb, en_b = ua_loop_from_pieces(O, ntau, eps, dt, nsteps, x.copy(order="F"),
                                           v.copy(order="F"), e, w, e_f.copy(order="F"), bounds)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(ea, eb)
    assert np.array_equal(en_a, en_b)
