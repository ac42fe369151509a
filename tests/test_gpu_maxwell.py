"""The GPU Vlasov-Maxwell path (FDTD fields of src/fdtd.jl, the CIC current
deposit of src/cloud_in_cell.jl, the loop of
examples/landau_damping_maxwell_fdtd.jl) against the oracle and the
reference's own tests (test/test_fdtd.jl, test/test_deposition.jl)."""
import math

import numpy as np
import pytest

from conftest import random_particles

pytestmark = pytest.mark.gpu


def _periodic_fdtd(pic, O, mesh, rng):
    """Random FDTD fields whose ghost row of ex and ghost column of ey repeat
    the first, as every field the loop produces does."""
    nx, ny = mesh.nx, mesh.ny
    f = pic.FDTD(mesh)
    f.ex[...] = rng.normal(size=f.ex.shape)
    f.ex[:, ny] = f.ex[:, 0]
    f.ey[...] = rng.normal(size=f.ey.shape)
    f.ey[nx, :] = f.ey[0, :]
    f.bz[...] = rng.normal(size=f.bz.shape)
    o = O.FDTD(nx, ny)
    o.ex[...], o.ey[...], o.bz[...] = f.ex, f.ey, f.bz
    return f, o


def _periodic_nodes(rng, nx, ny):
    a = np.asfortranarray(rng.normal(size=(nx + 1, ny + 1)))
    a[nx, :] = a[0, :]
    a[:, ny] = a[:, 0]
    return a


@pytest.mark.parametrize("nx,ny", [(32, 8), (128, 16), (96, 40)])
def test_fdtd_kernels_bit_exact(gpu, O, nx, ny):
    """faraday!, ampere_maxwell!, update_fields! and compute_energy follow the
    reference's arithmetic term by term: equal bit for bit."""
    pic = gpu
    rng = np.random.default_rng(nx + ny)
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, nx, 0.0, 1.3, ny)
    dx, dy = mesh.dx, mesh.dy
    f, o = _periodic_fdtd(pic, O, mesh, rng)
    pic.faraday_(f, mesh, 0.013)
    O.faraday(o, nx, ny, dx, dy, 0.013)
    assert np.array_equal(f.bz, o.bz)
    jx, jy = _periodic_nodes(rng, nx, ny), _periodic_nodes(rng, nx, ny)
    pic.ampere_maxwell_(f, mesh, jx, jy, 0.021)
    O.ampere_maxwell(o, nx, ny, dx, dy, jx, jy, 0.021)
    assert np.array_equal(f.ex, o.ex) and np.array_equal(f.ey, o.ey)
    # ampere_maxwell! keeps the ghost row / column periodic
    assert np.array_equal(f.ex[:, ny], f.ex[:, 0]) and np.array_equal(f.ey[nx, :], f.ey[0, :])
    ex, ey, bz = (np.zeros((nx + 1, ny + 1), order="F") for _ in range(3))
    pic.update_fields_(ex, ey, bz, mesh, f)
    oex, oey, obz = O.update_fields(o, nx, ny)
    for a, b in ((ex, oex), (ey, oey), (bz, obz)):
        assert np.array_equal(a, b)
    # the energy: a sum in another order than the oracle's pairwise one
    assert pic.compute_energy(f, mesh) == pytest.approx(O.fdtd_energy(o, nx, ny, dx, dy),
                                                        rel=1e-14)


def test_fdtd_analytic_mode(gpu):
    """test/test_fdtd.jl:5-60 on the GPU: the TE mode of the periodic box
    over 8 steps against the exact solution, to the reference's bounds."""
    pic = gpu
    dimx = dimy = 2 * math.pi
    nx = ny = 128
    dt, nstep = 1e-4, 8
    mesh = pic.TwoDGrid(dimx, nx, dimy, ny)
    f = pic.FDTD(mesh)
    w = math.sqrt(2)
    xn = np.linspace(0, dimx, nx + 1)[:, None]
    yn = np.linspace(0, dimy, ny + 1)[None, :]
    xc = (xn[:-1] + xn[1:]) / 2
    yc = (yn[:, :-1] + yn[:, 1:]) / 2
    t = 0.5 * dt
    f.bz[...] = -np.cos(xc) * np.cos(yc) * math.cos(w * t)
    jx = np.zeros((nx + 1, ny + 1), order="F")
    jy = np.zeros((nx + 1, ny + 1), order="F")
    ex, ey, bz = (np.zeros((nx + 1, ny + 1), order="F") for _ in range(3))
    for _ in range(nstep):
        pic.ampere_maxwell_(f, mesh, jx, jy, dt)
        t += 0.5 * dt
        assert np.max(np.abs(f.ex - np.cos(xc) * np.sin(yn) / w * math.sin(w * t))) < 1e-6
        assert np.max(np.abs(f.ey + np.sin(xn) * np.cos(yc) / w * math.sin(w * t))) < 1e-6
        pic.update_fields_(ex, ey, bz, mesh, f)
        assert np.max(np.abs(ex - np.cos(xn) * np.sin(yn) / w * math.sin(w * t))) < 1e-6
        assert np.max(np.abs(ey + np.sin(xn) * np.cos(yn) / w * math.sin(w * t))) < 1e-6
        pic.faraday_(f, mesh, dt)
        t += 0.5 * dt
        assert np.max(np.abs(f.bz + np.cos(xc) * np.cos(yc) * math.cos(w * t))) < 1e-6
        pic.update_fields_(ex, ey, bz, mesh, f)
        assert np.max(np.abs(bz + np.cos(xn) * np.cos(yn) * math.cos(w * t))) < 1e-3


@pytest.mark.parametrize("nx,ny,n", [(32, 8, 25600), (128, 16, 204800), (128, 128, 100000),
                                     (64, 48, 1), (16, 16, 0)])
def test_compute_current_matches_oracle(gpu, O, nx, ny, n):
    """compute_current! on random particles, including ones on the upper
    boundary x = dimx (push_x! leaves them there) and on cell faces: the
    GPU's atomic sums reorder the reference's, so within 1e-13 relative."""
    pic = gpu
    rng = np.random.default_rng(n + nx)
    dimx, dimy = 4 * math.pi, 1.0
    mesh = pic.TwoDGrid(0.0, dimx, nx, 0.0, dimy, ny)
    p = random_particles(rng, n, dimx, dimy, w=None if n else 1.0)
    if n > 100:
        p[0, :10] = dimx
        p[1, 10:20] = dimy
        p[0, 20:30] = mesh.dx * 3
        p[4] = rng.uniform(0.5, 1.5, n) * dimx * dimy / n
    jx, jy = (np.zeros((nx + 1, ny + 1), order="F") for _ in range(2))
    pic.compute_current_(jx, jy, mesh, pic.CloudInCell(), p)
    ox, oy = O.compute_current(p, nx, ny, mesh.dx, mesh.dy)
    if n == 0:
        assert not jx.any() and not jy.any()
        return
    for a, b in ((jx, ox), (jy, oy)):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
        # folded ghosts: the last row and column repeat the first
        assert np.array_equal(a[nx, :], a[0, :]) and np.array_equal(a[:, ny], a[:, 0])


def test_compute_current_known_total(gpu):
    """test/test_deposition.jl: 1e6 particles around 2pi on 128^2 cells of
    [0, 4pi]^2 with v = w = (4pi)^2/1e6: sum(jx) = sum(jy) =
    2.5872575761191685 (numpy's normal samples stand in for Julia's, so the
    sums agree to sampling, 1.5e-8)."""
    pic = gpu
    rng = np.random.default_rng(1234)
    n, nx = 1_000_000, 128
    dim = 4 * math.pi
    mesh = pic.TwoDGrid(dim, nx, dim, nx)
    p = np.zeros((5, n), order="F")
    p[0:2] = np.clip(rng.normal(size=(2, n)) + 2 * math.pi, 0, dim - 1e-9)
    p[2:5] = dim * dim / n
    jx, jy = (np.zeros((nx + 1, nx + 1), order="F") for _ in range(2))
    pic.compute_current_(jx, jy, mesh, pic.CloudInCell(), p)
    assert math.isclose(jx.sum(), 2.5872575761191685, rel_tol=1.5e-8)
    assert math.isclose(jy.sum(), 2.5872575761191685, rel_tol=1.5e-8)


def test_compute_current_rejects_other_shapes(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(1.0, 8, 1.0, 8)
    p = random_particles(np.random.default_rng(0), 100, 1.0, 1.0)
    jx, jy = (np.zeros((9, 9), order="F") for _ in range(2))
    with pytest.raises(pic._lib.PicError):
        pic.compute_current_(jx, jy, mesh, pic.M6(), p)


def _vm_setup(pic, O, nx, ny, npm, midpoint):
    """The initial state of examples/vm2d.jl (midpoint=True: ex at the cell
    midpoints) or of examples/landau_damping_maxwell_fdtd.jl (ex at the
    nodes): Landau alpha 0.1, kx 0.5, dimy 1."""
    alpha, kx = 0.1, 0.5
    mesh = pic.TwoDGrid(2 * math.pi / kx, nx, 1.0, ny)
    n = npm * nx * ny
    p = np.asfortranarray(pic.sampling.landau_damping(n, mesh, alpha, kx,
                                                      sobol=pic.sampling.sobol3(n)))
    f = pic.FDTD(mesh)
    x = np.linspace(0, mesh.dimx, nx + 1)
    xs = (x[:-1] + x[1:]) / 2 if midpoint else x[:-1]
    f.ex[...] = (alpha / kx * np.sin(kx * xs))[:, None]
    o = O.FDTD(nx, ny)
    o.ex[...], o.ey[...], o.bz[...] = f.ex, f.ey, f.bz
    return mesh, p, f, o


def test_run_maxwell_first_step_bit_exact(gpu, O):
    """One iteration: the particle update precedes the current's effect on
    the fields, so positions, velocities and faraday!'s bz equal the
    oracle's bit for bit; ex, ey carry the reordered current sums."""
    pic = gpu
    mesh, p, f, o = _vm_setup(pic, O, 128, 16, 100, midpoint=False)
    sim = pic.Simulation(mesh, p, pic.CloudInCell())
    sim.set_fdtd(f)
    e = sim.run_maxwell(0.01, 1)
    g = sim.fdtd()
    q = sim.particles()
    sim.close()
    po = p.copy(order="F")
    eo = O.maxwell_loop(po, o, mesh.nx, mesh.ny, mesh.dimx, mesh.dimy, 0.01, 1)
    assert np.array_equal(q, po)
    assert np.array_equal(g.bz, o.bz)
    for a, b in ((g.ex, o.ex), (g.ey, o.ey)):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    assert e[0] == pytest.approx(eo[0], rel=1e-13)


@pytest.mark.parametrize("nx,ny,npm,steps,midpoint,path", [
    (128, 16, 100, 100, False, 0), (128, 16, 100, 100, False, 2), (128, 16, 100, 100, False, 3),
    (32, 8, 50, 200, True, 0), (32, 8, 50, 3, True, 0), (32, 8, 50, 200, True, 2),
    (64, 40, 20, 40, True, 0), (96, 40, 20, 40, True, 0), (128, 64, 10, 40, True, 0)])
def test_run_maxwell_matches_oracle_loop(gpu, O, nx, ny, npm, steps, midpoint, path):
    """pic_run_maxwell against the oracle's loop: the README's 204,800
    particles on 128 x 16 cells for 100 steps (the loop of
    examples/landau_damping_maxwell_fdtd.jl) in each launch form (0: one
    launch per iteration, 2: a particle pass and a one-workgroup field
    kernel, 3: one kernel per reference function), a small grid for 200
    and for 3 steps (every current buffer of the one-launch form last), a
    64 x 40 grid (two launches: the one-launch LDS does not fit), 96 x 40 and
    128 x 64 (one kernel per function).  The current's reordered sums are
    the only difference, so energies and particles agree to 1e-10
    relative."""
    pic = gpu
    mesh, p, f, o = _vm_setup(pic, O, nx, ny, npm, midpoint)
    sim = pic.Simulation(mesh, p, pic.CloudInCell())
    sim.set_maxwell_path(path)
    sim.set_fdtd(f)
    e = sim.run_maxwell(0.01, steps)
    g = sim.fdtd()
    q = sim.particles()
    sim.close()
    eo = O.maxwell_loop(p, o, nx, ny, mesh.dimx, mesh.dimy, 0.01, steps)
    assert np.max(np.abs(e - eo)) < 1e-10 * np.max(np.abs(eo))
    for a, b in ((g.ex, o.ex), (g.ey, o.ey), (g.bz, o.bz)):
        assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1e-300)
    assert np.max(np.abs(q[:4] - p[:4])) < 1e-10 * np.max(np.abs(p[:4]))
    assert np.array_equal(q[4], p[4])
    # the Landau mode is damped: the field energy falls over the run
    if steps > 10:
        assert e[-1] < e[0]


def test_run_maxwell_matches_per_call_api(gpu):
    """The fused loop against the same loop written with the per-call drop-ins
    (faraday_, update_fields_, push_v_, push_x_, compute_current_,
    ampere_maxwell_, compute_energy), 20 iterations."""
    pic = gpu
    nx, ny, dt, steps = 64, 8, 0.01, 20
    mesh = pic.TwoDGrid(4 * math.pi, nx, 1.0, ny)
    n = 50 * nx * ny
    p = np.asfortranarray(pic.sampling.landau_damping(n, mesh, 0.1, 0.5))
    f = pic.FDTD(mesh)
    f.ex[...] = (0.2 * np.sin(0.5 * np.linspace(0, mesh.dimx, nx + 1)[:-1]))[:, None]
    sim = pic.Simulation(mesh, p, pic.CloudInCell())
    sim.set_fdtd(f)
    e = sim.run_maxwell(dt, steps)
    g = sim.fdtd()
    q = sim.particles()
    sim.close()
    grp = pic.ParticleGroup(n)
    grp.array[...] = p
    ex, ey, bz, jx, jy = (np.zeros((nx + 1, ny + 1), order="F") for _ in range(5))
    k = pic.CloudInCell()
    e2 = []
    for s in range(1, steps + 1):
        if s > 1:
            pic.faraday_(f, mesh, 0.5 * dt)
        pic.update_fields_(ex, ey, bz, mesh, f)
        pic.push_v_(grp, k, mesh, ex, ey, bz, dt)
        pic.push_x_(grp, mesh, 0.5 * dt)
        pic.compute_current_(jx, jy, mesh, k, grp)
        pic.push_x_(grp, mesh, 0.5 * dt)
        pic.faraday_(f, mesh, 0.5 * dt)
        pic.ampere_maxwell_(f, mesh, jx, jy, dt)
        e2.append(pic.compute_energy(f, mesh))
    assert np.max(np.abs(e - np.array(e2))) < 1e-11 * np.max(np.abs(e))
    assert np.max(np.abs(q - grp.array)) < 1e-11 * np.max(np.abs(q))
    assert np.max(np.abs(g.ex - f.ex)) < 1e-11 * np.max(np.abs(f.ex))


def test_run_maxwell_needs_fp64_cic(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(1.0, 8, 1.0, 8)
    p = random_particles(np.random.default_rng(1), 1000, 1.0, 1.0)
    sim = pic.Simulation(mesh, p, pic.M6())
    with pytest.raises(pic._lib.PicError):
        sim.run_maxwell(0.01, 2)
    sim.close()
    sim = pic.Simulation(mesh, p, pic.CloudInCell(), state=pic._lib.PIC_STATE_F32)
    with pytest.raises(pic._lib.PicError):
        sim.run_maxwell(0.01, 2)
    sim.close()


def test_run_maxwell_continues(gpu, O):
    """Seven iterations, faraday!(dt/2) (the leading half step the loop does
    from its second iteration on), then five more equal twelve iterations:
    the one-launch form leaves the context's fields and current as the
    reference's arrays after its last iteration."""
    pic = gpu
    mesh, p, f, o = _vm_setup(pic, O, 32, 8, 50, True)
    sim = pic.Simulation(mesh, p, pic.CloudInCell())
    sim.set_fdtd(f)
    e7 = sim.run_maxwell(0.01, 7)
    pic._lib.call("pic_faraday", sim.ctx.h, 0.005)
    e5 = sim.run_maxwell(0.01, 5)
    g = sim.fdtd()
    sim.close()
    eo = O.maxwell_loop(p, o, 32, 8, mesh.dimx, mesh.dimy, 0.01, 12)
    e = np.concatenate([e7, e5])
    assert np.max(np.abs(e - eo)) < 1e-10 * np.max(np.abs(eo))
    for a, b in ((g.ex, o.ex), (g.ey, o.ey), (g.bz, o.bz)):
        assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b))
