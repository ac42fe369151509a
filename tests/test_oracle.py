"""Pins of the CPU oracle against the reference's own data and against the
mathematics, so that a wrong index, sign or dropped term in the oracle fails
here (CPU only)."""
import math

import numpy as np
import pytest

from conftest import golden, random_particles


# ------------------------------------------------------------- Fortran twin

def test_push_matches_f90_twin_fixture(O):
    """push_v!/push_x! (pushers_cic.jl:3-72) == src/pic.f90 push_v/push_x,
    bit for bit, as test/test_push.jl:41 asserts for Julia."""
    g = golden("f90_push.npz")
    nx, ny = int(g["nx"]), int(g["ny"])
    dx, dy = float(g["dimx"]) / nx, float(g["dimy"]) / ny
    p = np.asfortranarray(g["p_in"].copy())
    O.push_v(p, "cic", nx, ny, dx, dy, g["ex"], g["ey"], g["bz"], float(g["dt"]))
    assert np.array_equal(p, g["p_push_v"])
    O.push_x(p, float(g["dimx"]), float(g["dimy"]), float(g["dt"]))
    assert np.array_equal(p, g["p_push_x"])
    # the Julia method on nx x ny fields (mod1 neighbours) gives the same bits
    p = np.asfortranarray(g["p_in"].copy())
    O.push_v(p, "cic", nx, ny, dx, dy, g["ex"][:nx, :ny], g["ey"][:nx, :ny], g["bz"][:nx, :ny],
             float(g["dt"]))
    assert np.array_equal(p, g["p_push_v"])


def test_push_known_answer(O):
    """test_push.jl setup: 1 particle at (0.1, 0.1), ex=1, ey=0, bz=1,
    dt=0.1, 100x100 on [0,1]^2, 50 x (push_v, push_x)."""
    traj = golden("f90_push.npz")["single_traj"]
    p = np.zeros((5, 1), order="F")
    p[0, 0] = p[1, 0] = 0.1
    p[4, 0] = 1.0
    g1 = np.ones((100, 100), order="F")
    g0 = np.zeros((100, 100), order="F")
    for s in range(50):
        O.push_v(p, "cic", 100, 100, 0.01, 0.01, g1, g0, g1, 0.1)
        O.push_x(p, 1.0, 1.0, 0.1)
        assert np.array_equal(p[:, 0], traj[s])
    assert p[0, 0] == 0.7723249868055071
    assert p[1, 0] == 0.10388738211141334
    assert p[2, 0] == -0.9600961282277396
    assert p[3, 0] == -0.72032979321689439


@pytest.mark.skipif("not __import__('oracle').ref_available()")
def test_push_matches_f90_twin_live(O):
    rng = np.random.default_rng(7)
    nx, ny = 16, 64
    dimx, dimy = 3.0, 5.0
    dx, dy = dimx / nx, dimy / ny
    p = random_particles(rng, 5000, dimx, dimy)
    g = [np.asfortranarray(rng.normal(size=(nx + 1, ny + 1))) for _ in range(3)]
    for a in g:
        a[nx, :] = a[0, :]
        a[:, ny] = a[:, 0]
    pj, pf = p.copy(order="F"), p.copy(order="F")
    for _ in range(3):
        O.push_v(pj, "cic", nx, ny, dx, dy, *g, 0.05)
        O.push_x(pj, dimx, dimy, 0.05)
        O.ref_push_v(pf, nx, ny, dx, dy, *g, 0.05)
        O.ref_push_x(pf, dimx, dimy, 0.05)
        assert np.array_equal(pj, pf)


def test_rotation_preserves_speed(O):
    """With E = 0 the kick is a rotation by 2 atan(dt b/2): |v| is kept."""
    rng = np.random.default_rng(3)
    p = random_particles(rng, 1000, 1.0, 1.0)
    v0 = np.hypot(p[2], p[3])
    z = np.zeros((8, 8), order="F")
    b = np.full((8, 8), 3.7, order="F")
    O.push_v(p, "m6", 8, 8, 0.125, 0.125, z, z, b, 0.2)
    assert np.allclose(np.hypot(p[2], p[3]), v0, rtol=1e-14, atol=0)
    p2 = random_particles(rng, 1000, 1.0, 1.0)
    v0 = np.hypot(p2[2], p2[3])
    O.push_v(p2, "cic", 8, 8, 0.125, 0.125, z, z, b, 0.2)
    assert np.allclose(np.hypot(p2[2], p2[3]), v0, rtol=1e-14, atol=0)


def test_m6_push_uniform_field_equals_cic(O):
    """Both stencils are partitions of unity: a uniform field gives the same
    kick to rounding."""
    rng = np.random.default_rng(4)
    p = random_particles(rng, 2000, 2.0, 2.0)
    q = p.copy(order="F")
    ex = np.full((16, 16), 0.3, order="F")
    ey = np.full((16, 16), -1.1, order="F")
    bz = np.full((16, 16), 0.8, order="F")
    O.push_v(p, "cic", 16, 16, 0.125, 0.125, ex, ey, bz, 0.1)
    O.push_v(q, "m6", 16, 16, 0.125, 0.125, ex, ey, bz, 0.1)
    assert np.allclose(p, q, rtol=0, atol=1e-14)


# ------------------------------------------------------------ f_m6 shape

def test_f_m6_closed_form(O):
    assert O.f_m6(0.0) * 120 == 66.0
    assert O.f_m6(1.0) * 120 == 26.0
    assert O.f_m6(2.0) * 120 == 1.0
    assert O.f_m6(3.0) == 0.0 and O.f_m6(3.5) == 0.0
    for t in np.linspace(0, 1, 37, endpoint=False):
        s = sum(O.f_m6(abs(t - k)) for k in range(-3, 4))
        assert abs(s - 1) < 2e-15
        assert abs(O.f_m6(1 + t) - O.f_m6(abs(-1 - t))) == 0
    # the pieces join smoothly at the knots: one-sided slopes agree to O(h)
    h = 1e-5
    for q in (1.0, 2.0):
        left = (O.f_m6(q) - O.f_m6(q - h)) / h
        right = (O.f_m6(q + h) - O.f_m6(q)) / h
        assert abs(left - right) < 1e-4


# ------------------------------------------------------------ deposits

def _cic_numpy(p, nx, ny, dx, dy):
    """compute_rho(p, CloudInCell(), m) written with numpy scatter-adds (an
    independent spelling of pushers_cic.jl:82-111)."""
    xp, yp, w = p[0] / dx, p[1] / dy, p[4]
    i = np.trunc(xp).astype(np.int64)
    j = np.trunc(yp).astype(np.int64)
    fx, fy = xp - i, yp - j
    rho = np.zeros((nx, ny))
    for di, dj, a in ((0, 0, (1 - fx) * (1 - fy)), (1, 0, fx * (1 - fy)), (1, 1, fx * fy),
                      (0, 1, (1 - fx) * fy)):
        np.add.at(rho, ((i + di) % nx, (j + dj) % ny), a * w)
    return rho


def test_cic_deposit(O):
    rng = np.random.default_rng(11)
    nx, ny, dimx, dimy = 32, 16, 4 * np.pi, 2.0
    dx, dy = dimx / nx, dimy / ny
    p = random_particles(rng, 20000, dimx, dimy, w=rng.uniform(0.5, 1.5, 20000))
    raw = _cic_numpy(p, nx, ny, dx, dy)
    rho = O.compute_rho(p, "cic", nx, ny, dx, dy)
    want = raw / (dx * dy)
    want -= want.mean()
    assert np.allclose(rho, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert abs(rho.sum()) < 1e-9 * np.abs(rho).max()
    # total charge before the mean removal is sum(w) (partition of unity)
    assert math.isclose(raw.sum(), p[4].sum(), rel_tol=1e-13)


def test_cic_deposit_matches_f90_deposition(O):
    """pic.f90's deposition with v = w and n = nx*ny (factor 1) scatters the
    same CIC weights; folding its ghosts gives the raw charge."""
    g = golden("f90_push.npz")
    nx, ny = int(g["nx"]), int(g["ny"])
    dimx, dimy = float(g["dimx"]), float(g["dimy"])
    dx, dy = dimx / nx, dimy / ny
    q = np.asfortranarray(g["dep_in"])
    jx = g["dep_jx"].copy()
    jx[0, :] += jx[nx, :]
    jx[:, 0] += jx[:, ny]
    raw = jx[:nx, :ny]
    rho = O.compute_rho(q, "cic", nx, ny, dx, dy)
    want = raw / (dx * dy)
    want -= want.mean()
    assert np.allclose(rho, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_cic_deposit_landau_statistical(O, pic):
    """test/test_particles.jl: 64x16 cells, dimx = 4pi, dimy = 1, 1000 ppc,
    quiet start; compute_rho ~ alpha cos(kx x) with atol 1e-2."""
    nx, ny, alpha, kx = 64, 16, 0.1, 0.5
    mesh = pic.TwoDGrid(2 * np.pi / kx, nx, 1.0, ny)
    p = pic.sampling.landau_damping(1000 * nx * ny, mesh, alpha, kx)
    rho = O.compute_rho(p, "cic", nx, ny, mesh.dx, mesh.dy)
    want = np.repeat((alpha * np.cos(kx * mesh.x[:nx]))[:, None], ny, axis=1)
    assert np.max(np.abs(rho - want)) < 1e-2


def _m6_numpy(p, nx, ny, dx, dy, O):
    xp, yp, w = p[0] / dx, p[1] / dy, p[4]
    i = np.floor(xp).astype(np.int64)
    j = np.floor(yp).astype(np.int64)
    fx, fy = xp - i, yp - j
    f = np.vectorize(O.f_m6)
    rho = np.zeros((nx, ny))
    for a in range(-2, 4):
        for b in range(-2, 4):
            np.add.at(rho, ((i + a) % nx, (j + b) % ny), f(np.abs(fx - a)) * f(np.abs(fy - b)) * w)
    return rho


def test_m6_deposit(O):
    rng = np.random.default_rng(12)
    nx, ny, dimx, dimy = 16, 32, 2.0, 4 * np.pi
    dx, dy = dimx / nx, dimy / ny
    p = random_particles(rng, 3000, dimx, dimy)
    raw = _m6_numpy(p, nx, ny, dx, dy, O)
    rho = O.compute_rho(p, "m6", nx, ny, dx, dy, dimx, dimy)
    want = raw / (dx * dy)
    want -= want.sum() * dx * dy / dimx / dimy
    assert np.allclose(rho, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert math.isclose(raw.sum(), p[4].sum(), rel_tol=1e-13)


def test_m6_deposit_first_moment(O):
    """The centred B-spline reproduces linear functions: for particles at
    least 3 cells from the edges, sum_nodes x_node rho_raw = sum w x_p."""
    rng = np.random.default_rng(13)
    nx = ny = 32
    dx = dy = 0.5
    n = 4000
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(4 * dx, (nx - 5) * dx, n)
    p[1] = rng.uniform(4 * dy, (ny - 5) * dy, n)
    p[4] = 1.0 / n
    rho = O.compute_rho(p, "m6", nx, ny, dx, dy, nx * dx, ny * dy)
    raw = (rho + (1.0 / (nx * dx * ny * dy))) * dx * dy  # undo mean removal and scaling
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    assert math.isclose((raw * X).sum(), (p[4] * p[0]).sum(), rel_tol=1e-12)
    assert math.isclose((raw * Y).sum(), (p[4] * p[1]).sum(), rel_tol=1e-12)


def test_mf_m6_compat_pins_reference_indexing(O):
    """compute_rho_m6!(fields, particles) (compute_rho_m6.jl:181-316) puts
    the off-centre weights at mod1(i+k, nx), one node below pushers_m6.jl's
    stencil; compat=False is the corrected stencil."""
    nx, ny = 16, 16
    L = 4.0
    dx = L / nx
    x = np.array([[5.3 * dx], [7.6 * dx]], order="F")
    rho_c = O.compute_rho_m6_fields(x.copy(order="F"), 1.0, nx, ny, 0.0, L, 0.0, L, compat=True)
    rho_f = O.compute_rho_m6_fields(x.copy(order="F"), 1.0, nx, ny, 0.0, L, 0.0, L, compat=False)
    shift = 1.0 / (L * L)
    raw_c = (rho_c[:nx, :ny] + shift) * dx * dx
    raw_f = (rho_f[:nx, :ny] + shift) * dx * dx
    f = O.f_m6
    i, j, tx, ty = 5, 7, 0.3, 0.6
    wx = {k: f(abs(tx - k)) for k in range(-2, 4)}
    wy = {k: f(abs(ty - k)) for k in range(-2, 4)}
    node_c = {k: (i + k - (0 if k == 0 else 1)) % nx for k in range(-2, 4)}
    exp_c = np.zeros((nx, ny))
    exp_f = np.zeros((nx, ny))
    for a in range(-2, 4):
        for b in range(-2, 4):
            exp_c[node_c[a], (j + b - (0 if b == 0 else 1)) % ny] += wx[a] * wy[b]
            exp_f[(i + a) % nx, (j + b) % ny] += wx[a] * wy[b]
    assert np.allclose(raw_c, exp_c, atol=1e-13)
    assert np.allclose(raw_f, exp_f, atol=1e-13)
    assert raw_c[i - 1, j] == pytest.approx(0.0, abs=1e-15)  # node i-1 gets nothing
    # the swapped bounds move the rewritten positions by +dimx (compat only)
    xc = x.copy(order="F")
    O.compute_rho_m6_fields(xc, 1.0, nx, ny, 0.0, L, 0.0, L, compat=True)
    assert np.allclose(xc, x + L, atol=1e-14)


# ------------------------------------------------------------ gathers

def test_cic_gather_reproduces_bilinear(O):
    rng = np.random.default_rng(5)
    nx, ny, dx, dy = 16, 16, 0.25, 0.5
    n = 3000
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(0, (nx - 1) * dx, n)
    p[1] = rng.uniform(0, (ny - 1) * dy, n)
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    ex = np.asfortranarray(2 + 3 * X - Y + 0.5 * X * Y)
    ey = np.asfortranarray(np.full((nx, ny), -1.25))
    epx, epy = O.interpolate_cic(p, nx, ny, dx, dy, ex, ey)
    assert np.allclose(epx, 2 + 3 * p[0] - p[1] + 0.5 * p[0] * p[1], atol=1e-13)
    assert np.allclose(epy, -1.25, atol=1e-15)


def test_m6_gather_reproduces_linear(O):
    """test/test_uapic.jl:47-74: the M6 gather of e = (x, y) returns the
    particle positions (particles 3 cells inside the edges)."""
    rng = np.random.default_rng(6)
    nx, ny = 32, 32
    L = 8.0
    dx = L / nx
    n = 2000
    x = np.asfortranarray(np.vstack([rng.uniform(3 * dx, L - 4 * dx, n),
                                     rng.uniform(3 * dx, L - 4 * dx, n)]))
    e = np.zeros((2, nx + 1, ny + 1), order="F")
    for i in range(nx + 1):
        e[0, i, :] = i * dx
    for j in range(ny + 1):
        e[1, :, j] = j * dx
    ep = O.interpol_eb_m6(x.copy(order="F"), e, nx, ny, 0.0, L, 0.0, L)
    assert np.max(np.abs(ep - x)) < 1e-12


# ------------------------------------------------------------ Poisson

def test_poisson_2d_analytic(O):
    """test/test_poisson_2d.jl:26-43: mode 2 on 128^2, [0,2pi]^2."""
    n, L, mode = 128, 2 * np.pi, 2
    x = np.arange(n) * L / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    rho = np.asfortranarray(-2 * mode ** 3 * np.sin(mode * X) * np.cos(mode * Y))
    phi = O.solve_phi(rho, L, L)
    ex, ey = O.solve_e(rho, L, L)
    assert np.allclose(phi, -mode * np.sin(mode * X) * np.cos(mode * Y), rtol=0, atol=1e-13)
    assert np.allclose(ex, mode ** 2 * np.cos(mode * X) * np.cos(mode * Y), rtol=0, atol=1e-13)
    assert np.allclose(ey, -mode ** 2 * np.sin(mode * X) * np.sin(mode * Y), rtol=0, atol=1e-13)


def test_poisson_meshfields_analytic_and_fixture(O):
    """test/test_poisson.jl: both cases, max error atol 1e-14, and the FFTW
    output test/test2.dat of the second case."""
    g = golden("poisson_test2.npz")
    nx, ny, xmax, ymax = int(g["nx"]), int(g["ny"]), float(g["xmax"]), float(g["ymax"])
    xs = np.linspace(0, xmax, nx + 1)
    ys = np.linspace(0, ymax, ny + 1)
    rho1 = np.asfortranarray(-8 * np.sin(2 * xs)[:, None] * np.cos(2 * ys)[None, :])
    e, _ = O.solve_fields(rho1, 0.0, xmax, 0.0, ymax)
    assert np.max(np.abs(e[0] - 2 * np.cos(2 * xs)[:, None] * np.cos(2 * ys)[None, :])) < 1e-14
    assert np.max(np.abs(e[1] + 2 * np.sin(2 * xs)[:, None] * np.sin(2 * ys)[None, :])) < 1e-14
    e, _ = O.solve_fields(np.asfortranarray(g["rho"]), 0.0, xmax, 0.0, ymax)
    assert np.max(np.abs(e[0] - 2 * np.cos(2 * xs)[:, None])) < 1e-14
    assert np.max(np.abs(e[1] + 2 * np.sin(2 * ys)[None, :])) < 1e-14
    assert np.max(np.abs(e[0] - g["ex"])) < 2e-15
    assert np.max(np.abs(e[1] - g["ey"])) < 2e-15


def test_poisson_matches_pocketfft(O):
    """irfft with the self-conjugate columns projected = numpy's irfft2 on the
    same non-Hermitian spectrum (what FFTW's c2r returns)."""
    rng = np.random.default_rng(9)
    nx, ny, dimx, dimy = 32, 64, 4 * np.pi, 2.0
    r = np.asfortranarray(rng.normal(size=(nx, ny)))
    ex, ey = O.solve_e(r, dimx, dimy)
    kx, ky, k2 = O.poisson_tables(nx, ny, dimx, dimy)
    R = np.fft.rfft2(r.T).T
    R[0, 0] = 0
    exn = np.fft.irfft2((-1j * kx * R).T, s=(ny, nx)).T
    eyn = np.fft.irfft2((-1j * ky * R).T, s=(ny, nx)).T
    assert np.allclose(ex, exn, rtol=0, atol=1e-14)
    assert np.allclose(ey, eyn, rtol=0, atol=1e-14)
    phi = O.solve_phi(r, dimx, dimy)
    phin = np.fft.irfft2((np.fft.rfft2(r.T).T / k2).T, s=(ny, nx)).T
    assert np.allclose(phi, phin, rtol=0, atol=1e-13)


def test_poisson_tables(O):
    """poisson_2d.jl:40-57: kx[1,1] = 1, negative ky from ny/2+1 on
    (Nyquist included)."""
    kx, ky, k2 = O.poisson_tables(8, 8, 2 * np.pi, 4 * np.pi)
    assert k2[0, 0] == 1.0 and kx[0, 0] == 1.0
    assert ky[1, 4] == pytest.approx(-4 * 0.5 / (1 + 4.0), rel=1e-15)  # jk = 5 (Nyquist)
    assert ky[0, 3] == pytest.approx(1 / 1.5, rel=1e-15)


def test_energy(O):
    nx, ny = 16, 8
    dx, dy = 0.5, 0.25
    x = np.arange(nx) * dx
    ex = np.asfortranarray(np.repeat(np.cos(2 * np.pi * x / (nx * dx))[:, None], ny, axis=1))
    assert O.energy(ex, dx, dy) == pytest.approx(math.sqrt(0.5 * nx * dx * ny * dy), rel=1e-14)


# ------------------------------------------------------------ sort

def test_counting_sort_is_stable_sort(O):
    rng = np.random.default_rng(2)
    key = rng.integers(0, 300, 50000).astype(np.int32)
    perm, off = O.counting_sort(key, 300)
    assert np.array_equal(perm, np.argsort(key, kind="stable"))
    assert off[0] == 0 and off[-1] == key.size
    assert np.array_equal(np.diff(off), np.bincount(key, minlength=300))


def test_lane_spread_definition(O):
    """Worked example: cells [0,0,1,0,1,2] -> ranks in cell 0: particles
    0,1,3; cell 1: 2,4; cell 2: 5 -> round 0 (0,2,5), round 1 (1,4),
    round 2 (3); a second segment (cells 64..) follows the first."""
    assert O.lane_spread([0, 0, 1, 0, 1, 2], 128).tolist() == [0, 2, 5, 1, 4, 3]
    assert O.lane_spread([64, 1, 64, 1, 0], 128).tolist() == [4, 1, 3, 0, 2]
    rng = np.random.default_rng(5)
    cell = rng.integers(0, 300, 5000)
    perm = O.lane_spread(cell, 300)
    assert np.array_equal(np.sort(perm), np.arange(5000))
    sc = cell[perm]
    assert np.all(np.diff(sc // 64) >= 0)  # segments stay in order


# ------------------------------------------------------------ FP32 state

def test_f32_weights(O):
    for t in np.linspace(0, 1, 101, endpoint=False):
        w = O.f32_m6_weights(np.float32(t)).astype(np.float64)
        exact = [O.f_m6(abs(float(np.float32(t)) - k)) for k in range(-2, 4)]
        assert np.allclose(w, exact, rtol=0, atol=2e-7)
        assert abs(w.sum() - 1) < 1e-6


def test_f32_roundtrip(O):
    rng = np.random.default_rng(21)
    nx, ny, dimx, dimy = 64, 32, 4 * np.pi, 4 * np.pi
    dx, dy = dimx / nx, dimy / ny
    p = random_particles(rng, 10000, dimx, dimy)
    st = O.f32_from_f64(p, nx, ny, dx, dy)
    cell, ox, oy = st[:3]
    assert cell.min() >= 0 and cell.max() < nx * ny
    assert ox.min() >= 0 and ox.max() < 1 and oy.min() >= 0 and oy.max() < 1
    q = O.f32_to_f64(st, nx, dx, dy, p[4, 0])
    assert np.max(np.abs(q[0] - p[0])) < 1e-7 * dx
    assert np.max(np.abs(q[2] - p[2])) < 1e-6 * np.abs(p[2]).max()


@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_f32_step_tracks_f64(O, kernel):
    """One FP32 step (kick, drift dt, deposit) against the FP64 reference
    sequence push_v; push_x(dt/2); push_x(dt/2); compute_rho on the same
    input: positions within FP32 resolution of a cell, charge within 1e-6."""
    rng = np.random.default_rng(22)
    nx, ny, dimx, dimy = 32, 32, 4 * np.pi, 4 * np.pi
    dx, dy = dimx / nx, dimy / ny
    n = 20000
    p = random_particles(rng, n, dimx, dimy)
    ex = np.asfortranarray(rng.normal(size=(nx, ny)) * 0.1)
    ey = np.asfortranarray(rng.normal(size=(nx, ny)) * 0.1)
    dt = 0.1
    st = O.f32_from_f64(p, nx, ny, dx, dy)
    p32 = O.f32_to_f64(st, nx, dx, dy, p[4, 0])  # the FP64 run starts where FP32 does
    e32 = np.empty(2 * nx * ny, np.float32)
    e32[0::2] = ex.ravel(order="F")
    e32[1::2] = ey.ravel(order="F")
    acc = O.f32_step(st, nx, ny, kernel, e32, np.float32(dt / 2), np.float32(dt / dx),
                     np.float32(dt / dy), np.float32(p[4, 0]))
    O.push_v(p32, kernel, nx, ny, dx, dy, ex, ey, None, dt)
    O.push_x(p32, dimx, dimy, dt / 2)
    O.push_x(p32, dimx, dimy, dt / 2)
    q = O.f32_to_f64(st, nx, dx, dy, p[4, 0])
    dxw = np.abs(q[0] - p32[0])
    dxw = np.minimum(dxw, dimx - dxw)
    assert dxw.max() < 1e-5 * dx
    assert np.max(np.abs(q[2] - p32[2])) < 1e-6
    if kernel == "cic":
        rho64 = O.compute_rho(p32, "cic", nx, ny, dx, dy)
        rho = acc / (dx * dy)
        rho -= rho.mean()
    else:
        rho64 = O.compute_rho(p32, "m6", nx, ny, dx, dy, dimx, dimy)
        rho = acc / (dx * dy)
        rho -= rho.sum() * dx * dy / dimx / dimy
    assert np.linalg.norm(rho - rho64) / np.linalg.norm(rho64) < 1e-4
    assert math.isclose(acc.sum(), n * p[4, 0], rel_tol=1e-6)


# ------------------------------------------------------------ time loop

def test_landau_loop_cic_damps(O, pic):
    """The reference loop (examples/landau_damping_poisson_cic_m6.jl:70-83)
    on the CPU: the electric energy of the k=0.5 mode damps at about the
    Landau rate 0.153 (the CPU end-to-end check of the oracle)."""
    from conftest import landau_particles
    mesh, p = landau_particles(pic, 200 * 128 * 8, 128, 8, alpha=0.1, dimy=1.0)
    e = O.landau_loop(p, "cic", mesh.nx, mesh.ny, mesh.dimx, mesh.dimy, 0.1, 200)
    t = np.arange(1, 201) * 0.1
    g, w, nmax = pic.diagnostics.damping_rate(t, e, tmin=0.5, tmax=15.0)
    assert nmax >= 5
    assert abs(g - (-0.153359)) < 0.02, g
    assert abs(w - 1.415662) < 0.08, w
