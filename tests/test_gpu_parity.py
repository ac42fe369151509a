"""Parity of the HIP library with the CPU oracle, per entry point, on the same
seeded inputs.  Bars: bit-exact for cell indices, sort order and the CIC
push (the reference's Fortran twin is bit-identical to Julia); 1e-12
normwise relative for rho and E in FP64."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import golden, landau_particles, random_particles

pytestmark = pytest.mark.gpu

RTOL = 1e-12  # FP64 rho and E, normwise relative


def slow_count(pic, ctx):
    """Particles the segment passes completed on their global-memory path."""
    v = C.c_uint64()
    pic._lib.call("pic_slow_chunks", ctx.h, C.byref(v))
    return v.value


def nrel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# ------------------------------------------------------------------ push

def test_push_v_push_x_cic_bitwise_f90(gpu, O):
    pic = gpu
    g = golden("f90_push.npz")
    nx, ny = int(g["nx"]), int(g["ny"])
    mesh = pic.TwoDGrid(float(g["dimx"]), nx, float(g["dimy"]), ny)
    p = pic.ParticleGroup(g["p_in"].shape[1])
    p.array[...] = g["p_in"]
    pic.push_v_(p, pic.CloudInCell(), mesh, g["ex"][:nx, :ny], g["ey"][:nx, :ny],
                g["bz"][:nx, :ny], float(g["dt"]))
    assert np.array_equal(p.array, g["p_push_v"])
    pic.push_x_(p, mesh, float(g["dt"]))
    assert np.array_equal(p.array, g["p_push_x"])


def test_f90_dropins_bitwise(gpu):
    """ParticleInCell.F90's entry points (pic.jl:60-114) through pic_f90_*:
    the fixture of pic.f90 and the test_push.jl 50-step trajectory."""
    pic = gpu
    g = golden("f90_push.npz")
    nx, ny = int(g["nx"]), int(g["ny"])
    mesh = pic.TwoDGrid(float(g["dimx"]), nx, float(g["dimy"]), ny)
    p = pic.ParticleGroup(g["p_in"].shape[1])
    p.array[...] = g["p_in"]
    pic.F90.push_v_(p, pic.CloudInCell(), mesh, g["ex"], g["ey"], g["bz"], float(g["dt"]))
    assert np.array_equal(p.array, g["p_push_v"])
    pic.F90.push_x_(p, mesh, float(g["dt"]))
    assert np.array_equal(p.array, g["p_push_x"])
    traj = g["single_traj"]
    m = pic.TwoDGrid(1.0, 100, 1.0, 100)
    q = pic.ParticleGroup(1)
    q.array[0, 0] = q.array[1, 0] = 0.1
    q.array[4, 0] = 1.0
    ones = np.ones((101, 101), order="F")
    zeros = np.zeros((101, 101), order="F")
    for s in range(50):
        pic.F90.push_v_(q, pic.CloudInCell(), m, ones, zeros, ones, 0.1)
        pic.F90.push_x_(q, m, 0.1)
        assert np.array_equal(q.array[:, 0], traj[s]), s
    assert q.array[0, 0] == 0.7723249868055071


@pytest.mark.parametrize("bz", ["none", "uniform", "grid"])
def test_push_v_m6(gpu, O, bz):
    pic = gpu
    rng = np.random.default_rng(31)
    nx, ny = 32, 64
    mesh = pic.TwoDGrid(4 * np.pi, nx, 2 * np.pi, ny)
    p = pic.ParticleGroup(50000)
    p.array[...] = random_particles(rng, 50000, mesh.dimx, mesh.dimy)
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(2))
    bzg = np.asfortranarray(rng.normal(size=(nx, ny)))
    bz_arg = {"none": None, "uniform": 1.7, "grid": bzg}[bz]
    bz_o = {"none": None, "uniform": np.full((nx, ny), 1.7, order="F"), "grid": bzg}[bz]
    q = p.array.copy(order="F")
    pic.push_v_(p, pic.M6(), mesh, ex, ey, bz_arg, 0.05)
    O.push_v(q, "m6", nx, ny, mesh.dx, mesh.dy, ex, ey, bz_o, 0.05)
    assert np.array_equal(p.array[:2], q[:2])
    assert np.max(np.abs(p.array[2:4] - q[2:4])) < 1e-14


@pytest.mark.parametrize("bz", ["none", "uniform", "grid"])
def test_push_v_cic_bitwise(gpu, O, bz):
    pic = gpu
    rng = np.random.default_rng(32)
    nx, ny = 64, 32
    mesh = pic.TwoDGrid(4 * np.pi, nx, 2.0, ny)
    p = pic.ParticleGroup(50000)
    p.array[...] = random_particles(rng, 50000, mesh.dimx, mesh.dimy)
    ex, ey, bzg = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(3))
    bz_arg = {"none": None, "uniform": -2.5, "grid": bzg}[bz]
    bz_o = {"none": None, "uniform": np.full((nx, ny), -2.5, order="F"), "grid": bzg}[bz]
    q = p.array.copy(order="F")
    pic.push_v_(p, pic.CloudInCell(), mesh, ex, ey, bz_arg, 0.1)
    O.push_v(q, "cic", nx, ny, mesh.dx, mesh.dy, ex, ey, bz_o, 0.1)
    if bz == "uniform":
        # a scalar field is used as is; the reference interpolates a constant
        # grid, a1*b + a2*b + a3*b + a4*b, which differs from b by rounding
        assert np.array_equal(p.array[:2], q[:2])
        assert np.max(np.abs(p.array[2:4] - q[2:4])) < 1e-15
    else:
        assert np.array_equal(p.array, q)


def test_push_edge_positions(gpu, O):
    """Particles on x = 0, x = dimx (kept by push_x!'s strict >) and just
    below the upper bound: the cell wraps periodically on both sides."""
    pic = gpu
    nx, ny = 16, 16
    mesh = pic.TwoDGrid(1.0, nx, 1.0, ny)
    p = pic.ParticleGroup(4)
    p.array[0] = [0.0, 1.0, np.nextafter(1.0, 0), 0.5]
    p.array[1] = [1.0, 0.0, 0.5, np.nextafter(1.0, 0)]
    p.array[2:4] = 0.3
    p.array[4] = 0.25
    ex = np.asfortranarray(np.arange(nx * ny, dtype=np.float64).reshape(nx, ny, order="F"))
    ey = np.asfortranarray(-ex)
    q = p.array.copy(order="F")
    pic.push_v_(p, pic.CloudInCell(), mesh, ex, ey, None, 0.1)
    O.push_v(q, "cic", nx, ny, mesh.dx, mesh.dy, ex, ey, None, 0.1)
    assert np.array_equal(p.array, q)
    rho = pic.compute_rho(p, pic.CloudInCell(), mesh)
    rho_o = O.compute_rho(q, "cic", nx, ny, mesh.dx, mesh.dy)
    assert nrel(rho, rho_o) < RTOL


# ------------------------------------------------------------ deposits

@pytest.mark.parametrize("kernel", ["cic", "m6"])
@pytest.mark.parametrize("order", ["random", "sorted"])
def test_compute_rho(gpu, O, kernel, order):
    pic = gpu
    rng = np.random.default_rng(33)
    nx, ny, n = 128, 64, 200000
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = pic.ParticleGroup(n)
    p.array[...] = random_particles(rng, n, mesh.dimx, mesh.dimy,
                                    w=rng.uniform(0.5, 1.5, n) * mesh.dimx * mesh.dimy / n)
    if order == "sorted":
        cell = (np.floor(p.array[0] / mesh.dx) + nx * np.floor(p.array[1] / mesh.dy)).astype(int)
        p.array[...] = p.array[:, np.argsort(cell, kind="stable")]
    k = pic.CloudInCell() if kernel == "cic" else pic.M6()
    rho = pic.compute_rho(p, k, mesh)
    rho_o = O.compute_rho(p.array, kernel, nx, ny, mesh.dx, mesh.dy, mesh.dimx, mesh.dimy)
    assert nrel(rho, rho_o) < RTOL


def test_compute_rho_uniform_and_empty(gpu, O):
    pic = gpu
    mesh, p = landau_particles(pic, 65536, 64, 64)
    pg = pic.ParticleGroup(p.shape[1])
    pg.array[...] = p
    rho = pic.compute_rho(pg, pic.M6(), mesh)
    rho_o = O.compute_rho(p, "m6", 64, 64, mesh.dx, mesh.dy, mesh.dimx, mesh.dimy)
    assert nrel(rho, rho_o) < RTOL
    empty = pic.ParticleGroup(0)
    rho0 = pic.compute_rho(empty, pic.CloudInCell(), mesh)
    assert np.all(rho0 == 0)


def test_f90_deposition_total_current(gpu, O):
    """test/test_deposition.jl: 128^2 on [0,4pi]^2, 1e6 particles around 2pi,
    v = w = (4pi)^2/1e6: sum(jx) == sum(jy) ~ 2.5872575761191685."""
    pic = gpu
    rng = np.random.default_rng(1234)
    n, nx = 1_000_000, 128
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, nx)
    p = pic.ParticleGroup(n)
    p.array[0:2] = np.clip(rng.normal(size=(2, n)) + 2 * np.pi, 0, 4 * np.pi - 1e-9)
    p.array[2:5] = (mesh.dimx * mesh.dimy) / n
    jx = np.zeros((nx + 1, nx + 1), order="F")
    jy = np.zeros((nx + 1, nx + 1), order="F")
    pic.F90.compute_current_(jx, jy, mesh, pic.CloudInCell(), p)
    assert math.isclose(jx.sum(), 2.5872575761191685, rel_tol=1.5e-8)
    assert math.isclose(jy.sum(), 2.5872575761191685, rel_tol=1.5e-8)
    if O.ref_available():
        rx, ry = O.ref_deposition(p.array, nx, nx, mesh.dx, mesh.dy)
        assert nrel(jy, ry) < 1e-13


# ------------------------------------------------------------ gathers

def test_interpolate_cic_bitwise(gpu, O):
    pic = gpu
    rng = np.random.default_rng(34)
    nx, ny, n = 32, 32, 20000
    mesh = pic.TwoDGrid(2.0, nx, 3.0, ny)
    p = pic.ParticleGroup(n)
    p.array[...] = random_particles(rng, n, mesh.dimx, mesh.dimy)
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(2))
    epx, epy = np.zeros(n), np.zeros(n)
    pic.interpolate_cic_(epx, epy, mesh, p, ex, ey)
    ox, oy = O.interpolate_cic(p.array, nx, ny, mesh.dx, mesh.dy, ex, ey)
    assert np.array_equal(epx, ox) and np.array_equal(epy, oy)


@pytest.mark.parametrize("compat", [True, False])
def test_meshfields_m6(gpu, O, compat):
    """compute_rho_m6!(fields, particles) and interpol_eb_m6!(particles,
    fields), MeshFields API (compute_rho_m6.jl:181, interpolation_m6.jl:125)."""
    pic = gpu
    rng = np.random.default_rng(35)
    nx, ny, n = 64, 32, 30000
    mesh = pic.TwoDGrid(0.0, 4 * np.pi, nx, 0.0, 2 * np.pi, ny)
    w = mesh.dimx * mesh.dimy / n
    parts = pic.Particles(n, w)
    parts.x[0] = rng.uniform(-mesh.dimx, 2 * mesh.dimx, n)  # outside the box: wrapped
    parts.x[1] = rng.uniform(0, mesh.dimy, n)
    fields = pic.MeshFields2D(mesh)
    xo = parts.x.copy(order="F")
    pic.compute_rho_m6_(fields, parts, compat=compat)
    rho_o = O.compute_rho_m6_fields(xo, w, nx, ny, 0.0, mesh.xmax, 0.0, mesh.ymax, compat=compat)
    assert nrel(fields.rho, rho_o) < RTOL
    assert np.array_equal(parts.x, xo)
    fields.e[0] = rng.normal(size=(nx + 1, ny + 1))
    fields.e[1] = rng.normal(size=(nx + 1, ny + 1))
    for c in (0, 1):  # periodic ghosts as solve! leaves them
        fields.e[c, nx, :] = fields.e[c, 0, :]
        fields.e[c, :, ny] = fields.e[c, :, 0]
    pic.interpol_eb_m6_(parts, fields)
    ep = O.interpol_eb_m6(xo, fields.e, nx, ny, 0.0, mesh.xmax, 0.0, mesh.ymax)
    assert np.max(np.abs(parts.e - ep)) < 1e-13 * np.abs(ep).max()
    assert np.array_equal(parts.x, xo)


# ------------------------------------------------------------ Poisson

def test_solve_e_phi(gpu, O):
    pic = gpu
    rng = np.random.default_rng(36)
    for nx, ny, dimx, dimy in ((64, 64, 4 * np.pi, 4 * np.pi), (128, 32, 1.0, 3.0),
                               (16, 256, 2 * np.pi, 4 * np.pi), (256, 128, 3.0, 1.5)):
        grid = pic.TwoDGrid(0.0, dimx, nx, 0.0, dimy, ny)
        ps = pic.TwoDPoissonPeriodic(grid)
        rho = np.asfortranarray(rng.normal(size=(nx, ny)))
        ex, ey = np.zeros((nx, ny), order="F"), np.zeros((nx, ny), order="F")
        pic.solve_(ex, ey, ps, rho)
        eo, fo = O.solve_e(rho, dimx, dimy)
        assert nrel(ex, eo) < RTOL and nrel(ey, fo) < RTOL
        phi = np.zeros((nx, ny), order="F")
        pic.solve_(phi, ps, rho)
        assert nrel(phi, O.solve_phi(rho, dimx, dimy)) < RTOL


@pytest.mark.parametrize("nx,ny", [(8, 8), (32, 16), (128, 128), (512, 64), (64, 512),
                                   (256, 256), (128, 256), (256, 64), (512, 128)])
def test_solve_e_fused_matches_rocfft(gpu, nx, ny):
    """The three-kernel E solve (power-of-two grids up to 512: transforms
    across the lanes of a wave from 64 nodes a side, 1, 2, 4 or 8 points per
    lane; in LDS below) against the rocFFT transforms around the same
    spectral kernel: fields to 1e-13 normwise, and the fused path's field
    energy as the reduction of those fields (one pic_run step from a random
    deposit)."""
    pic = gpu
    rng = np.random.default_rng(nx + 7 * ny)
    grid = pic.TwoDGrid(0.0, 2.0, nx, 0.0, 3.0, ny)
    rho = np.asfortranarray(rng.normal(size=(nx, ny)))
    out = []
    for fused in (True, False):
        ctx = pic.Context(grid)
        pic._lib.call("pic_set_fused_solve", ctx.h, 1 if fused else 0)
        ctx.set_grid(pic._lib.PIC_GRID_RHO, rho)
        pic._lib.call("pic_solve_e", ctx.h)
        out.append((ctx.get_grid(pic._lib.PIC_GRID_EX), ctx.get_grid(pic._lib.PIC_GRID_EY)))
        ctx.close()
    (ef, ff), (er, fr) = out
    assert nrel(ef, er) < 1e-13 and nrel(ff, fr) < 1e-13
    # the loop's energy history with either solve
    n = 4096
    p = random_particles(rng, n, grid.dimx, grid.dimy)
    en = []
    for fused in (True, False):
        sim = pic.Simulation(grid, p, pic.CloudInCell())
        sim.ctx.sync()
        pic._lib.call("pic_set_fused_solve", sim.ctx.h, 1 if fused else 0)
        en.append(sim.run(0.05, 3))
        sim.close()
    assert np.allclose(en[0], en[1], rtol=1e-12, atol=0)


def test_solve_fields_fixture(gpu, O):
    """solve!(fields, Poisson2D) on the input of test/test2.dat: FFTW's output
    and the energy of meshfields.jl:105-106."""
    pic = gpu
    g = golden("poisson_test2.npz")
    nx, ny = int(g["nx"]), int(g["ny"])
    mesh = pic.TwoDGrid(0.0, float(g["xmax"]), nx, 0.0, float(g["ymax"]), ny)
    fields = pic.MeshFields2D(mesh)
    fields.rho[...] = g["rho"]
    en = pic.solve_(fields, pic.Poisson2D(mesh))
    assert np.max(np.abs(fields.e[0] - g["ex"])) < 4e-15
    assert np.max(np.abs(fields.e[1] - g["ey"])) < 4e-15
    eo, en_o = O.solve_fields(np.asfortranarray(g["rho"]), 0.0, float(g["xmax"]), 0.0,
                              float(g["ymax"]))
    assert en == pytest.approx(en_o, rel=1e-13)


def test_field_energy(gpu, O):
    pic = gpu
    rng = np.random.default_rng(37)
    nx, ny = 64, 32
    mesh = pic.TwoDGrid(3.0, nx, 2.0, ny)
    ctx = pic.Context(mesh)
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(2))
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    out = (C.c_double * 2)()
    pic._lib.call("pic_field_energy", ctx.h, out)
    assert out[0] == pytest.approx(np.sum(ex ** 2) * mesh.dx * mesh.dy, rel=1e-13)
    assert out[1] == pytest.approx(np.sum(ey ** 2) * mesh.dx * mesh.dy, rel=1e-13)
    assert math.sqrt(out[0]) == pytest.approx(O.energy(ex, mesh.dx, mesh.dy), rel=1e-13)
    ctx.close()


# ------------------------------------------------------------ sort

@pytest.mark.parametrize("state", ["f64", "f32"])
@pytest.mark.parametrize("nx,ny,n", [(64, 64, 100000), (128, 128, 300001), (512, 512, 70000),
                                     (8, 8, 5), (8, 8, 150000)])
def test_sort_bitwise(gpu, O, state, nx, ny, n):
    pic = gpu
    rng = np.random.default_rng(nx + n)
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    sim = pic.Simulation(mesh, p, pic.CloudInCell(),
                         state=pic._lib.PIC_STATE_F32 if state == "f32" else 0)
    if state == "f32":
        cell = sim.state_f32()[0]
    else:
        cell = (np.floor(p[0] / mesh.dx).astype(int) % nx
                + nx * (np.floor(p[1] / mesh.dy).astype(int) % ny)).astype(np.int32)
    perm = sim.sort()
    want, off = O.counting_sort(cell, nx * ny)
    assert np.array_equal(perm, want)
    assert np.array_equal(sim.cell_offsets(), off)
    if state == "f32":
        assert np.array_equal(sim.state_f32()[0], cell[want])
    else:
        assert np.array_equal(sim.particles(), p[:, want])
    sim.close()


@pytest.mark.parametrize("nx,ny,n", [(64, 64, 200000), (128, 128, 300001), (32, 16, 9000),
                                     (512, 512, 50000), (8, 8, 3), (8, 8, 150000)])
def test_sort_lanes_bitwise(gpu, O, nx, ny, n):
    """pic_sort_lanes: the stable cell sort followed by the lane-spread order
    within 64-cell segments, against the oracle's lexsort definition.  On the
    8 x 8 grid with 150,000 particles every cell holds ~2,300, so the
    segment's rank tables span more than one round of their chunked fill
    (2,048 ranks) and every sort tile holds long runs of equal keys."""
    pic = gpu
    rng = np.random.default_rng(nx * 7 + n)
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    sim = pic.Simulation(mesh, p, pic.M6(), state=pic._lib.PIC_STATE_F32)
    cell = sim.state_f32()[0]
    perm = sim.sort(lanes=True)
    want = O.lane_spread(cell, nx * ny)
    assert np.array_equal(perm, want)
    got = sim.state_f32()[0]
    assert np.array_equal(got, cell[want])
    if n >= 64 * 64:
        # a wave of 64 consecutive particles covers many distinct cells
        w = got[: (n // 64) * 64].reshape(-1, 64)
        distinct = np.array([len(np.unique(r)) for r in w[:200]])
        assert distinct.mean() > 32
    sim.close()


def test_sort_after_context_growth(gpu, O):
    """A context whose particle capacity grows between sorts (same number of
    sort tiles) reallocates its sort scratch: the second sort is still the
    stable sort."""
    pic = gpu
    rng = np.random.default_rng(17)
    nx, ny = 32, 32
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    ctx = pic.Context(mesh, capacity=1000)
    for n in (1000, 3000):
        p = random_particles(rng, n, mesh.dimx, mesh.dimy)
        ctx.set_particles(p)
        perm = np.zeros(n, np.int32)
        pic._lib.call("pic_sort", ctx.h, perm.ctypes.data)
        cell = (np.floor(p[0] / mesh.dx).astype(int) % nx
                + nx * (np.floor(p[1] / mesh.dy).astype(int) % ny)).astype(np.int32)
        want, _ = O.counting_sort(cell, nx * ny)
        assert np.array_equal(perm, want), n
    ctx.close()


@pytest.mark.parametrize("state", [0, 1])
def test_run_large_grid_few_particles(gpu, O, state):
    """pic_run on a 512 x 512 grid with 300 particles (the per-cell scans
    are far longer than the particle arrays): energies of the FP64 state
    match the oracle loop, the FP32 state tracks them."""
    pic = gpu
    rng = np.random.default_rng(23)
    nx = ny = 512
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, 300, mesh.dimx, mesh.dimy)
    sim = pic.Simulation(mesh, p, pic.CloudInCell(), state=state)
    e = sim.run(0.1, 6, sort_every=2)
    sim.close()
    eo = O.landau_loop(p.copy(order="F"), "cic", nx, ny, mesh.dimx, mesh.dimy, 0.1, 6)
    tol = 1e-10 if state == 0 else 1e-3
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < tol


# ------------------------------------------------------------ FP32 state

@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_f32_step_bitwise(gpu, O, kernel):
    """One fused FP32 pass (kick, drift, deposit) from the same state and the
    same FP32 field: cells, offsets and velocities bit-identical to the
    oracle; the compensated LDS tile sums within 1e-12 of the oracle's FP64
    accumulation of the same FP32 contributions."""
    pic = gpu
    rng = np.random.default_rng(38)
    nx, ny, n = 64, 64, 150000
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    cell = (np.floor(p[0] / mesh.dx) + nx * np.floor(p[1] / mesh.dy)).astype(int)
    p = np.asfortranarray(p[:, np.argsort(cell, kind="stable")])
    st = O.f32_from_f64(p, nx, ny, mesh.dx, mesh.dy)
    w = np.float32(p[4, 0])
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny)) * 0.3) for _ in range(2))
    ctx = pic.Context(mesh, capacity=n)
    ctx.set_state_f32(st, float(w))
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    e32 = np.empty(2 * nx * ny, np.float32)
    e32[0::2] = ex.ravel(order="F")
    e32[1::2] = ey.ravel(order="F")
    dt = 0.1
    kc = pic._lib.PIC_CIC if kernel == "cic" else pic._lib.PIC_M6
    ctx.set_bz(pic._lib.PIC_BZ_NONE)
    raw = np.zeros((nx, ny), order="F")
    pic._lib.call("pic_step_pass", ctx.h, kc, 1, dt, raw.ctypes.data)
    got = ctx.get_state_f32()
    acc = O.f32_step(st, nx, ny, kernel, e32, np.float32(dt / 2), np.float32(dt / mesh.dx),
                     np.float32(dt / mesh.dy), w)
    for a, b, name in zip(got, st, ("cell", "ox", "oy", "vx", "vy")):
        assert np.array_equal(a, b), name
    assert nrel(raw, acc) < 1e-12
    # deposit-only pass of the new state
    pic._lib.call("pic_step_pass", ctx.h, kc, 3, 0.0, raw.ctypes.data)
    assert nrel(raw, O.f32_deposit(st, nx, ny, kernel, w)) < 1e-12
    ctx.close()


# stencil of each shape: nodes per direction S, nodes left of the cell HL,
# and QMAX, the bound on the product of two deposit weights the library's
# fixed-point quantum is sized by (pic_internal.h Shape<K>)
SHAPE_STENCIL = {"cic": (2, 0, 1.0), "ngp": (2, 0, 1.0), "m6": (6, 2, 0.31),
                 "tsc": (4, 1, 0.5625), "m4": (4, 1, 0.445)}
SHAPE_CODE = {"cic": 0, "m6": 1, "ngp": 2, "tsc": 3, "m4": 4}


def fixed_point_bound(cell, nx, ny, kernel, w, n):
    """Largest error the int32 fixed-point tiles (accumulation mode 2) may
    add to each node, against the oracle's FP64 sum of FP32 contributions.

    A block of m particles works in units of u = 2^e, e = max(ceil(log2(qmax
    w (m + 5))) - 30, ceil(log2(qmax w)) - 22), qmax the shape's bound on a
    product of two weights; the block of all n particles gives an upper
    bound U on every block's u.  Each particle adds R = round(w/u) units,
    worth w/R each (a scale off u by eps <= u/(2w)): every node but one of
    its stencil takes its share rounded to the nearest unit, and stencil node
    (HL | 1, HL) takes the remainder (the cell's own node for TSC and M4, the
    next one along x for CIC, NGP and M6).  So each share is off by at most
    a = U/2 (1 + eps) + eps qmax w + 2^-22 qmax w (the last term for the
    FP32 roundings of the products), and each remainder by the S^2 - 1
    others' errors plus eps w plus 4e-6 w (the FP32 weights' sum misses 1 by
    ~1e-6).  Returns a per node times the shares it receives plus that for
    the particles whose remainder node it is."""
    s, hl, qmax = SHAPE_STENCIL[kernel]
    aw = qmax * abs(float(w))
    e = max(math.ceil(math.log2(aw * (n + 5))) - 30, math.ceil(math.log2(aw)) - 22)
    U = 2.0 ** e
    eps = U / (2 * abs(float(w)) - U)
    counts = np.bincount(cell, minlength=nx * ny).reshape(ny, nx).T  # (i, j)
    contrib = np.zeros((nx, ny))
    for du in range(s):
        for dv in range(s):
            # a particle in cell (i, j) reaches node (i - hl + du, j - hl + dv)
            contrib += np.roll(np.roll(counts, du - hl, axis=0), dv - hl, axis=1)
    a = U / 2 * (1 + eps) + eps * aw + 2.0 ** -22 * aw
    b = (s * s - 1) * a + (eps + 4e-6) * abs(float(w))
    rem = np.roll(counts, (hl | 1) - hl, axis=0)  # remainder node of each cell's particles
    return contrib * a + rem * b


@pytest.mark.parametrize("acc", [0, 1, 2])
@pytest.mark.parametrize("kernel", ["cic", "m6"])
@pytest.mark.parametrize("nx,ny", [(128, 128), (64, 32), (32, 64)])
def test_f32_segment_pass_bitwise(gpu, O, kernel, acc, nx, ny):
    """The segment-persistent FP32 pass of pic_run (after pic_sort_lanes):
    cells, offsets, velocities bit-identical to the oracle.  The charge of
    the compensated FP32 pairs (acc 0) or FP64 windows (acc 1) within 1e-12
    of the oracle's FP64 sum of the same FP32 contributions; that of the
    int32 fixed-point windows (acc 2) within the rounding bound of every
    node.  Particles are spread wide enough that some leave their window."""
    _check_f32_segment_pass(gpu, O, kernel, acc, nx, ny, None)


@pytest.mark.parametrize("tiles", [1, 2, 4])
def test_f32_segment_pass_tile_copies(gpu, O, tiles):
    """The M6 fixed-point pass with 1, 2 or 4 copies of the charge tile
    (pic_set_f32_tiles): the same bound against the oracle for every count."""
    _check_f32_segment_pass(gpu, O, "m6", 2, 128, 128, tiles)


@pytest.mark.parametrize("tiles", [1, 4])
@pytest.mark.parametrize("kernel", ["cic", "ngp", "tsc", "m4"])
def test_f32_segment_pass_shapes(gpu, O, kernel, tiles):
    """The CIC, NGP, TSC and M4 fixed-point passes on 128 x 128, where the
    windows are narrower than the grid and the deferred-particle kernel
    runs, with one and four tile copies: cells, offsets and velocities
    bit-identical to the oracle's FP32 step of the same shape, charge within
    the fixed-point bound of every node."""
    _check_f32_segment_pass(gpu, O, kernel, 2, 128, 128, tiles)


@pytest.mark.parametrize("window", [40, 80, 160])
def test_f32_segment_pass_windows(gpu, O, window):
    """The M6 fixed-point pass with LDS windows of 40, 80 and 160 KiB
    (pic_set_f32_window): the same bound against the oracle for every size
    (the narrower windows send more particles to the global-memory path)."""
    _check_f32_segment_pass(gpu, O, "m6", 2, 128, 128, 4, window)


def test_f32_window_rejects_other_sizes(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(4 * np.pi, 16, 4 * np.pi, 16)
    ctx = pic.Context(mesh, capacity=16)
    for bad in (0, 39, 161, 1024):
        with pytest.raises(pic._lib.PicError, match="window"):
            pic._lib.call("pic_set_f32_window", ctx.h, bad)
    ctx.close()


def test_f32_tiles_rejects_other_counts(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(4 * np.pi, 16, 4 * np.pi, 16)
    ctx = pic.Context(mesh, capacity=16)
    for bad in (0, 3, 8):
        with pytest.raises(pic._lib.PicError, match="tile copies"):
            pic._lib.call("pic_set_f32_tiles", ctx.h, bad)
    ctx.close()


def _check_f32_segment_pass(gpu, O, kernel, acc, nx, ny, tiles, window=None):
    pic = gpu
    rng = np.random.default_rng(41 + nx)
    n = 200000
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    p[2:4] *= 3.0
    st0 = O.f32_from_f64(p, nx, ny, mesh.dx, mesh.dy)
    w = np.float32(p[4, 0])
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny)) * 0.3) for _ in range(2))
    ctx = pic.Context(mesh, capacity=n)
    ctx.set_state_f32(st0, float(w))
    perm = np.zeros(n, np.int32)
    pic._lib.call("pic_sort_lanes", ctx.h, perm.ctypes.data)
    st = tuple(np.ascontiguousarray(a[perm]) for a in st0)
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    ctx.set_bz(pic._lib.PIC_BZ_NONE)
    pic._lib.call("pic_set_f32_accumulation", ctx.h, acc)
    if tiles is not None:
        pic._lib.call("pic_set_f32_tiles", ctx.h, tiles)
    if window is not None:
        pic._lib.call("pic_set_f32_window", ctx.h, window)
    e32 = np.empty(2 * nx * ny, np.float32)
    e32[0::2] = ex.ravel(order="F")
    e32[1::2] = ey.ravel(order="F")
    kc = SHAPE_CODE[kernel]
    raw = np.zeros((nx, ny), order="F")
    dt = 0.5
    slow0 = slow_count(pic, ctx)
    slow_before = slow0
    for _ in range(3):  # the slots stay; particles drift out of their windows
        pic._lib.call("pic_step_pass", ctx.h, kc, 1, dt, raw.ctypes.data)
        acc_o = O.f32_step(st, nx, ny, kernel, e32, np.float32(dt / 2),
                           np.float32(dt / mesh.dx), np.float32(dt / mesh.dy), w)
        got = ctx.get_state_f32()
        for a, b, name in zip(got, st, ("cell", "ox", "oy", "vx", "vy")):
            assert np.array_equal(a, b), name
        if acc == 2:
            bound = fixed_point_bound(st[0], nx, ny, kernel, w, n)
            assert np.all(np.abs(raw - acc_o) <= bound + 1e-12 * np.abs(acc_o))
            assert nrel(raw, acc_o) < 1e-6
            # every particle on the LDS path deposits exactly w; the deferred
            # ones add their FP32 shares (within 4e-6 w of w each)
            ndef = slow_count(pic, ctx) - slow_before
            assert abs(raw.sum() - n * float(w)) <= 4e-6 * float(w) * ndef + 1e-12 * n * float(w)
        else:
            assert nrel(raw, acc_o) < 1e-12
        slow_before = slow_count(pic, ctx)
    if nx * ny > 64 * 64:
        # windows smaller than the grid: the global-memory fallback of
        # out-of-window particles ran
        assert slow_count(pic, ctx) > slow0
    ctx.close()


@pytest.mark.parametrize("dt", [0.2, 1.0])
@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_f64_segment_pass(gpu, O, kernel, dt):
    """The segment-persistent FP64 pass after pic_sort_lanes: CIC bit-identical
    to the reference sequence, M6 within rounding; deposit within 1e-12.  At
    dt = 1 (about 10 v cells of drift) the fast particles leave their
    windows' 20-cell margin and take the deferred kernel."""
    pic = gpu
    rng = np.random.default_rng(42)
    nx, ny, n = 128, 64, 150000
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(2))
    ctx = pic.Context(mesh, capacity=n)
    ctx.set_particles(p)
    perm = np.zeros(n, np.int32)
    pic._lib.call("pic_sort_lanes", ctx.h, perm.ctypes.data)
    q = np.asfortranarray(p[:, perm])
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    ctx.set_bz(pic._lib.PIC_BZ_NONE)
    kc = pic._lib.PIC_CIC if kernel == "cic" else pic._lib.PIC_M6
    raw = np.zeros((nx, ny), order="F")
    slow0 = slow_count(pic, ctx)
    pic._lib.call("pic_step_pass", ctx.h, kc, 1, dt, raw.ctypes.data)
    if dt > 0.5:
        assert slow_count(pic, ctx) > slow0  # some deposits left their windows
    got = ctx.get_particles(np.zeros((5, n), order="F"))
    O.push_v(q, kernel, nx, ny, mesh.dx, mesh.dy, ex, ey, None, dt)
    O.push_x(q, mesh.dimx, mesh.dimy, dt / 2)
    O.push_x(q, mesh.dimx, mesh.dimy, dt / 2)
    if kernel == "cic":
        assert np.array_equal(got, q)
    else:
        assert np.max(np.abs(got - q)) < 1e-13
    rho_o = O.compute_rho(got, kernel, nx, ny, mesh.dx, mesh.dy, mesh.dimx, mesh.dimy)
    rho = raw / (mesh.dx * mesh.dy)
    rho -= rho.mean()
    assert nrel(rho, rho_o) < RTOL
    ctx.close()


@pytest.mark.parametrize("order", ["any", "lanes"])
@pytest.mark.parametrize("bz", ["none", "uniform", "grid"])
@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_f32_step_bz_bitwise(gpu, O, kernel, bz, order):
    """The FP32 pass with a magnetic field (uniform, as the strong-B
    configuration, or a grid), through the chunk kernels (any particle order)
    and the segment kernels (after pic_sort_lanes, FP64 windows): cells,
    offsets, velocities bit-identical, the deposit within 1e-12."""
    pic = gpu
    rng = np.random.default_rng(39)
    nx, ny, n = 32, 64, 60000
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    st = O.f32_from_f64(p, nx, ny, mesh.dx, mesh.dy)
    w = np.float32(p[4, 0])
    ex, ey, bzg = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(3))
    ctx = pic.Context(mesh, capacity=n)
    ctx.set_state_f32(st, float(w))
    if order == "lanes":
        perm = np.zeros(n, np.int32)
        pic._lib.call("pic_sort_lanes", ctx.h, perm.ctypes.data)
        st = tuple(np.ascontiguousarray(a[perm]) for a in st)
        pic._lib.call("pic_set_f32_accumulation", ctx.h, 1)
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    e32 = np.empty(2 * nx * ny, np.float32)
    e32[0::2] = ex.ravel(order="F")
    e32[1::2] = ey.ravel(order="F")
    mode = {"none": 0, "uniform": 1, "grid": 2}[bz]
    if bz == "grid":
        ctx.set_grid(pic._lib.PIC_GRID_BZ, bzg)
    ctx.set_bz(mode, 20.0)
    dt = 0.0125
    kc = pic._lib.PIC_CIC if kernel == "cic" else pic._lib.PIC_M6
    raw = np.zeros((nx, ny), order="F")
    pic._lib.call("pic_step_pass", ctx.h, kc, 1, dt, raw.ctypes.data)
    got = ctx.get_state_f32()
    acc = O.f32_step(st, nx, ny, kernel, e32, np.float32(dt / 2), np.float32(dt / mesh.dx),
                     np.float32(dt / mesh.dy), w, bz_mode=mode, bz0=20.0,
                     bz32=bzg.ravel(order="F").astype(np.float32))
    for a, b, name in zip(got, st, ("cell", "ox", "oy", "vx", "vy")):
        assert np.array_equal(a, b), name
    assert nrel(raw, acc) < 1e-12
    ctx.close()


@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_f64_fused_pass(gpu, O, kernel):
    """The fused FP64 pass = push_v!; push_x!(dt/2); push_x!(dt/2);
    compute_rho (raw): positions and velocities bit-identical for CIC, within
    rounding for M6; the deposit within 1e-12."""
    pic = gpu
    rng = np.random.default_rng(40)
    nx, ny, n = 64, 64, 120000
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    p = random_particles(rng, n, mesh.dimx, mesh.dimy)
    ex, ey = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(2))
    ctx = pic.Context(mesh, capacity=n)
    ctx.set_particles(p)
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    ctx.set_bz(pic._lib.PIC_BZ_NONE)
    kc = pic._lib.PIC_CIC if kernel == "cic" else pic._lib.PIC_M6
    raw = np.zeros((nx, ny), order="F")
    dt = 0.1
    pic._lib.call("pic_step_pass", ctx.h, kc, 1, dt, raw.ctypes.data)
    got = ctx.get_particles(np.zeros((5, n), order="F"))
    q = p.copy(order="F")
    O.push_v(q, kernel, nx, ny, mesh.dx, mesh.dy, ex, ey, None, dt)
    O.push_x(q, mesh.dimx, mesh.dimy, dt / 2)
    O.push_x(q, mesh.dimx, mesh.dimy, dt / 2)
    if kernel == "cic":
        assert np.array_equal(got, q)
    else:
        assert np.max(np.abs(got - q)) < 1e-13
    rho_o = O.compute_rho(got, kernel, nx, ny, mesh.dx, mesh.dy, mesh.dimx, mesh.dimy)
    rho = raw / (mesh.dx * mesh.dy)
    rho -= rho.mean()
    assert nrel(rho, rho_o) < RTOL
    ctx.close()


# ------------------------------------------------------------ time loop

@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_run_matches_reference_loop(gpu, O, kernel, fused):
    """pic_run against the oracle's restatement of the reference loop
    (examples/landau_damping_poisson_cic_m6.jl:70-83), FP64, 10 steps, in
    its one-launch form and as separate kernels: the energies agree to 1e-10
    (atomics reorder the deposit sums, the transforms round differently, so
    trajectories are equal to rounding, not bit for bit)."""
    pic = gpu
    mesh, p = landau_particles(pic, 100000, 64, 64)
    sim = pic.Simulation(mesh, p, pic.CloudInCell() if kernel == "cic" else pic.M6())
    sim.set_fused_step(fused)
    e = sim.run(0.1, 10)
    got = sim.particles()
    sim.close()
    q = p.copy(order="F")
    eo = O.landau_loop(q, kernel, 64, 64, mesh.dimx, mesh.dimy, 0.1, 10)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
    # the separate-kernel loop sorts the particles by cell, so match them as a set: every
    # GPU particle's nearest oracle particle in (x, y, vx, vy) is within
    # 1e-10 and the matching is one to one
    from scipy.spatial import cKDTree
    dist, idx = cKDTree(q[:4].T).query(got[:4].T)
    assert np.max(dist) < 1e-10
    assert np.unique(idx).size == q.shape[1]


def test_run_nonuniform_weights(gpu, O):
    """pic_run with per-particle weights (FP64 state, re-sorted every 2 steps,
    so the weights travel with the particles through the sorts): energies
    match the oracle loop and the particles, weights included, are the
    oracle's as a set."""
    pic = gpu
    rng = np.random.default_rng(29)
    nx = ny = 32
    mesh = pic.TwoDGrid(4 * np.pi, nx, 4 * np.pi, ny)
    n = 40000
    p = random_particles(rng, n, mesh.dimx, mesh.dimy,
                         w=rng.uniform(0.5, 1.5, n) * mesh.dimx * mesh.dimy / n)
    sim = pic.Simulation(mesh, p, pic.M6())
    e = sim.run(0.1, 7, sort_every=2)
    got = sim.particles()
    sim.close()
    q = p.copy(order="F")
    eo = O.landau_loop(q, "m6", nx, ny, mesh.dimx, mesh.dimy, 0.1, 7)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
    from scipy.spatial import cKDTree
    dist, idx = cKDTree(q.T).query(got.T)
    assert np.max(dist) < 1e-9
    assert np.unique(idx).size == n


def test_run_sorted_matches_unsorted(gpu):
    """Sorting every step only permutes the particles: the energy history is
    the same to rounding."""
    pic = gpu
    mesh, p = landau_particles(pic, 200000, 128, 128)
    a = pic.Simulation(mesh, p, pic.M6())
    ea = a.run(0.05, 12, sort_every=0)
    a.close()
    b = pic.Simulation(mesh, p, pic.M6())
    eb = b.run(0.05, 12, sort_every=1)
    b.close()
    assert np.max(np.abs(ea - eb) / ea) < 1e-9


def test_run_after_lane_sort_skips_its_sort(gpu, O):
    """pic_run skips its first lane sort when the particles were lane-sorted
    and have not moved since: a run after pic_sort_lanes matches a run that
    sorts at its start (FP32 state) to the rounding of the FP64 sums of the
    blocks' charge, which arrive at the grid in any order, with the
    particles in the same slots.  After pic_push_x the particles have moved,
    the run sorts again, and its energies still match the oracle loop (FP64)."""
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    mesh, p = landau_particles(pic, 200_000, 128, 128)
    a = pic.Simulation(mesh, p, pic.M6(), state=F32)
    a.sort(lanes=True)
    ea = a.run(0.05, 12, sort_every=8)
    sa = a.state_f32()
    a.close()
    b = pic.Simulation(mesh, p, pic.M6(), state=F32)
    eb = b.run(0.05, 12, sort_every=8)
    sb = b.state_f32()
    b.close()
    assert np.max(np.abs(ea - eb) / np.abs(eb)) < 1e-10
    # same slots: the cell of (nearly) every slot agrees, offsets to FP32 rounding
    same = sa[0] == sb[0]
    assert same.mean() > 0.999
    assert np.max(np.abs(sa[1][same] - sb[1][same])) < 1e-5
    c = pic.Simulation(mesh, p, pic.M6())
    c.sort(lanes=True)
    pic._lib.call("pic_push_x", c.ctx.h, 0.3)
    ec = c.run(0.05, 6, sort_every=0)
    c.close()
    q = p.copy(order="F")
    O.push_x(q, mesh.dimx, mesh.dimy, 0.3)
    eo = O.landau_loop(q, "m6", 128, 128, mesh.dimx, mesh.dimy, 0.05, 6)
    assert np.max(np.abs(np.sqrt(ec[:, 0]) - eo) / eo) < 1e-10


@pytest.mark.parametrize("grid", [64, 128])
def test_run_timing_every(gpu, grid):
    """pic_set_timing(ctx, k): events around every k-th step's pass (the
    one-launch loop on 64^2, the separate passes on 128^2), so 12 steps count
    12 timed passes with k = 1 and 3 with k = 4, each of positive duration;
    the energies do not depend on it."""
    pic = gpu
    mesh, p = landau_particles(pic, 100_000, grid, grid)
    res = {}
    for every in (1, 4):
        sim = pic.Simulation(mesh, p, pic.M6(), state=pic._lib.PIC_STATE_F32)
        sim.set_timing(True, every=every)
        e = sim.run(0.05, 12, sort_every=4)
        ms, n, _ = sim.timing()
        sim.close()
        res[every] = (e, ms, n)
    assert res[1][2] == 12 and res[4][2] == 3
    assert res[1][1] > 0 and res[4][1] > 0
    assert np.max(np.abs(res[1][0] - res[4][0]) / res[4][0]) < 1e-10


def test_run_empty(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(4 * np.pi, 16, 4 * np.pi, 16)
    sim = pic.Simulation(mesh, np.zeros((5, 0), order="F"), pic.CloudInCell())
    e = sim.run(0.1, 3)
    assert np.all(e == 0)
    sim.close()
