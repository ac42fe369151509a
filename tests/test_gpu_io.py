"""Initial conditions made on the GPU (sample! of src/landau_damping.jl:77-97)
against the host sampler, and particle files (read_particles of
src/read_particles.jl:3-35 and its writer) against numpy."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _landau_mesh(pic, nx=128, ny=128, kx=0.5):
    return pic.TwoDGrid(0.0, 2 * math.pi / kx, nx, 0.0, 2 * math.pi / kx, ny)


@pytest.mark.parametrize("n", [1, 1000, 300_000])
def test_sample_landau_matches_host_sampler(gpu, n):
    """The Sobol points are exact dyadic fractions on both sides, so y (a
    linear map of the second coordinate) agrees bit for bit; x (Newton with
    sin, cos) and v (log, sqrt, cos, sin) agree to the device library's
    rounding."""
    pic = gpu
    mesh = _landau_mesh(pic)
    alpha, kx = 0.05, 0.5
    sim = pic.Simulation.landau(mesh, n, alpha, kx, pic.CloudInCell())
    q = sim.particles()
    sim.close()
    p = pic.sampling.landau_damping(n, mesh, alpha, kx)
    assert np.array_equal(q[1], p[1])
    assert np.array_equal(q[4], p[4])
    assert np.max(np.abs(q[0] - p[0])) < 1e-13 * mesh.dimx
    assert np.max(np.abs(q[2:4] - p[2:4])) < 1e-14 * np.max(np.abs(p[2:4]))


def test_sample_landau_into_fp32_state(gpu):
    """Sampled straight into the FP32 cell-relative state: the same cells and
    offsets as the host sample converted by pic_set_state."""
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    n = 200_000
    mesh = _landau_mesh(pic)
    sim = pic.Simulation.landau(mesh, n, 0.05, 0.5, pic.M6(), state=F32)
    a = sim.state_f32()
    sim.close()
    p = pic.sampling.landau_damping(n, mesh, 0.05, 0.5)
    sim = pic.Simulation(mesh, p, pic.M6(), state=F32)
    b = sim.state_f32()
    sim.close()
    # the positions differ by at most an ulp: a particle may round onto the
    # other side of an FP32 offset step, very rarely across a cell face
    assert np.mean(a[0] != b[0]) < 1e-4
    same = a[0] == b[0]
    for u, w in zip(a[1:], b[1:]):
        assert np.max(np.abs(u[same] - w[same])) <= 2e-7


@pytest.mark.parametrize("n,seed", [(1, 5), (1000, 1234), (300_000, 2**40 + 7)])
def test_sample_maxwellian_matches_restatement(gpu, n, seed):
    """pic_sample_maxwellian against its numpy restatement (tests/philox.py,
    whose Philox4x32-10 meets the generator's known answers): x is a linear
    map of the same 53-bit uniform, so it agrees bit for bit; y (Newton with
    sin, cos) and v (log, sqrt, cos, sin) agree to the device library's
    rounding.  Offset mesh origin, two periods in y."""
    from philox import maxwellian
    pic = gpu
    mesh = pic.TwoDGrid(-1.0, -1.0 + 4 * math.pi, 64, 0.5, 0.5 + 4 * math.pi, 32)
    sim = pic.Simulation.maxwellian(mesh, n, 0.05, 0.5, 1.0, pic.CloudInCell(), seed=seed,
                                    vth=1.5)
    q = sim.particles()
    sim.close()
    x, y, vx, vy = maxwellian(n, 0.05, 0.5, 1.0, 1.5, seed, mesh.xmin, mesh.dimx, mesh.ymin,
                              mesh.dimy)
    assert np.array_equal(q[0], x)
    assert np.max(np.abs(q[1] - y)) < 1e-13 * mesh.dimy
    assert np.max(np.abs(q[2] - vx)) < 1e-14 * max(1.0, np.max(np.abs(vx)))
    assert np.max(np.abs(q[3] - vy)) < 1e-14 * max(1.0, np.max(np.abs(vy)))
    assert np.all(q[4] == mesh.dimx * mesh.dimy / n)


def test_sample_maxwellian_rejects_partial_periods(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 32, 0.0, 10.0, 32)
    sim = pic.Simulation(mesh, None, pic.CloudInCell())
    with pytest.raises(pic._lib.PicError, match="whole periods"):
        sim.sample_maxwellian(100, 0.05, 0.5, 0.5)
    with pytest.raises(pic._lib.PicError, match="alpha"):
        sim.sample_maxwellian(100, 1.5, 0.5, 2 * math.pi / 10.0)
    sim.close()


def test_sample_landau_rejects_a_wrong_box(gpu):
    pic = gpu
    mesh = pic.TwoDGrid(0.0, 10.0, 32, 0.0, 10.0, 32)
    with pytest.raises(pic._lib.PicError):
        pic.Simulation.landau(mesh, 100, 0.05, 0.5, pic.CloudInCell())


def _write_lines(path, ix, iy, dpx, dpy, vx, vy, blank_every=0):
    with open(path, "w") as f:
        for k in range(len(ix)):
            f.write(f"{int(ix[k])} {int(iy[k])} {float(dpx[k])!r} {float(dpy[k])!r} "
                    f"{float(vx[k])!r} {float(vy[k])!r}\n")
            if blank_every and k % blank_every == 0:
                f.write("\n")


def _random_lines(rng, n, nx, ny):
    return (rng.integers(0, nx, n), rng.integers(0, ny, n), rng.uniform(0, 1, n),
            rng.uniform(0, 1, n), rng.normal(size=n), rng.normal(size=n))


def test_read_particles_fp64(gpu, tmp_path):
    """x = (dpx + ix) dx, (dpy + iy) dy as read_particles.jl:28-31 computes
    it, bit for bit; a multi-megabyte file is parsed in parallel chunks and
    keeps its line order; blank lines are skipped."""
    pic = gpu
    rng = np.random.default_rng(1)
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 128, 0.0, 2 * math.pi, 64)
    n = 120_000
    ix, iy, dpx, dpy, vx, vy = _random_lines(rng, n, 128, 64)
    path = str(tmp_path / "particles.dat")
    _write_lines(path, ix, iy, dpx, dpy, vx, vy, blank_every=9973)
    assert os.path.getsize(path) > 4 << 20
    p = pic.read_particles(path, mesh)
    assert p.nbpart == n
    assert p.w == mesh.dimx * mesh.dimy / n
    assert np.array_equal(p.x[0], (dpx + ix) * mesh.dx)
    assert np.array_equal(p.x[1], (dpy + iy) * mesh.dy)
    assert np.array_equal(p.v[0], vx) and np.array_equal(p.v[1], vy)


def test_read_particles_count(gpu, tmp_path):
    """test/test_uapic.jl:5-14: a file of 204,800 lines on Mesh(0, 4pi, 128,
    0, 2pi, 64) gives nbpart == 204800 (the reference's particles.dat is not
    in its repository; the file here is synthetic)."""
    pic = gpu
    rng = np.random.default_rng(2)
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 128, 0.0, 2 * math.pi, 64)
    path = str(tmp_path / "particles.dat")
    _write_lines(path, *_random_lines(rng, 204_800, 128, 64))
    assert pic.read_particles(path, mesh).nbpart == 204_800


def test_read_particles_into_fp32_state_and_round_trip(gpu, tmp_path):
    """The format maps onto the FP32 state one to one: cell = ix + nx iy and
    the offsets and velocities rounded to FP32; written back from that state
    and read again, the state is bit-identical."""
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    rng = np.random.default_rng(3)
    nx, ny, n = 64, 32, 50_000
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, nx, 0.0, 2 * math.pi, ny)
    ix, iy, dpx, dpy, vx, vy = _random_lines(rng, n, nx, ny)
    path = str(tmp_path / "a.dat")
    _write_lines(path, ix, iy, dpx, dpy, vx, vy)
    sim = pic.Simulation.from_file(mesh, path, pic.M6(), state=F32)
    cell, ox, oy, fvx, fvy = sim.state_f32()
    assert np.array_equal(cell, ix + nx * iy)
    for a, b in ((ox, dpx), (oy, dpy), (fvx, vx), (fvy, vy)):
        assert np.array_equal(a, b.astype(np.float32))
    path2 = str(tmp_path / "b.dat")
    sim.write_particles(path2)
    sim.close()
    sim = pic.Simulation.from_file(mesh, path2, pic.M6(), state=F32)
    again = sim.state_f32()
    sim.close()
    for a, b in zip((cell, ox, oy, fvx, fvy), again):
        assert np.array_equal(a, b)


def test_write_particles_fp64_round_trip(gpu, tmp_path):
    """From the FP64 state: ix = floor(x/dx), dpx = x/dx - ix with 17 digits;
    read back, positions agree to rounding and velocities exactly."""
    pic = gpu
    rng = np.random.default_rng(4)
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 64, 0.0, 2 * math.pi, 32)
    n = 20_000
    p = pic.Particles(n, mesh.dimx * mesh.dimy / n)
    p.x[0] = rng.uniform(0, mesh.dimx, n)
    p.x[1] = rng.uniform(0, mesh.dimy, n)
    p.v[...] = rng.normal(size=(2, n))
    path = str(tmp_path / "c.dat")
    pic.write_particles(path, p, mesh)
    q = pic.read_particles(path, mesh)
    assert np.max(np.abs(q.x - p.x)) <= 1e-15 * mesh.dimx
    assert np.array_equal(q.v, p.v)


def test_read_particles_errors(gpu, tmp_path):
    pic = gpu
    mesh = pic.TwoDGrid(0.0, 1.0, 8, 0.0, 1.0, 8)
    with pytest.raises(pic._lib.PicError, match="cannot open"):
        pic.read_particles(str(tmp_path / "missing.dat"), mesh)
    bad = tmp_path / "bad.dat"
    bad.write_text("1 2 0.5 0.5 0.1 0.2\n3 4 0.5 oops 0.1 0.2\n")
    with pytest.raises(pic._lib.PicError, match=":2:"):
        pic.read_particles(str(bad), mesh)
    empty = tmp_path / "empty.dat"
    empty.write_text("")
    assert pic.read_particles(str(empty), mesh).nbpart == 0
    out = tmp_path / "cell.dat"
    out.write_text("9 0 0.5 0.5 0.0 0.0\n")
    with pytest.raises(pic._lib.PicError):
        pic.Simulation.from_file(mesh, str(out), pic.M6(), state=pic._lib.PIC_STATE_F32)


def test_operator_t_drift_without_wrap(gpu):
    """operator_t!(split, dt) (poisson_2d.jl:265-277): x += dt v, no periodic
    wrap, bit for bit with numpy."""
    pic = gpu
    rng = np.random.default_rng(6)
    mesh = pic.TwoDGrid(0.0, 1.0, 16, 0.0, 1.0, 16)
    n = 5000
    g = pic.ParticleGroup(n)
    g.array[0:2] = rng.uniform(0, 1, (2, n))
    g.array[2:4] = rng.normal(size=(2, n)) * 10
    g.array[4] = 1.0 / n
    want = g.array.copy(order="F")
    want[0:2] += 0.37 * want[2:4]
    pic.operator_t_(g, mesh, 0.37)
    assert np.array_equal(g.array, want)
    assert np.max(g.array[0]) > 1.0 and np.min(g.array[0]) < 0.0
