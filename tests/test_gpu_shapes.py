"""The NearestGridPoint, TriangularShapeCloud and M4 shapes (pushers_ngp.jl,
pushers_tsc.jl, pushers_m4.jl) on the GPU against the oracle's restatement:
deposits, pushes (TSC gathers with the M4 weights, as the reference does)
and the time loop of examples/landau_damping_poisson_cic_m6.jl:88-103."""
import numpy as np
import pytest

from conftest import landau_particles, random_particles

pytestmark = pytest.mark.gpu

SHAPES = ["ngp", "tsc", "m4"]


def _shape(pic, name):
    return {"ngp": pic.NearestGridPoint(), "tsc": pic.TriangularShapeCloud(),
            "m4": pic.M4()}[name]


def nrel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", SHAPES)
def test_shape_compute_rho(gpu, O, name):
    """compute_rho on the GPU (FP64, any particle order): normwise 1e-12 of
    the oracle (the atomics reorder the sums); total charge removed."""
    pic = gpu
    rng = np.random.default_rng(41)
    nx, ny = 64, 32
    mesh = pic.TwoDGrid(4 * np.pi, nx, 2.0, ny)
    p = pic.ParticleGroup(60000)
    p.array[...] = random_particles(rng, 60000, mesh.dimx, mesh.dimy)
    rho = pic.compute_rho(p, _shape(pic, name), mesh)
    ro = O.compute_rho(p.array, name, nx, ny, mesh.dx, mesh.dy, mesh.dimx, mesh.dimy)
    assert nrel(rho, ro) < 1e-12
    assert abs(rho.sum()) * mesh.dx * mesh.dy < 1e-9


@pytest.mark.parametrize("name", SHAPES)
@pytest.mark.parametrize("bz", ["none", "grid"])
def test_shape_push_v(gpu, O, name, bz):
    """push_v! on the GPU against the oracle: positions untouched, velocities
    to 1e-14 (the sums run in another association)."""
    pic = gpu
    rng = np.random.default_rng(42)
    nx, ny = 32, 64
    mesh = pic.TwoDGrid(4 * np.pi, nx, 2 * np.pi, ny)
    p = pic.ParticleGroup(50000)
    p.array[...] = random_particles(rng, 50000, mesh.dimx, mesh.dimy)
    ex, ey, bzg = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(3))
    q = p.array.copy(order="F")
    pic.push_v_(p, _shape(pic, name), mesh, ex, ey, bzg if bz == "grid" else None, 0.05)
    O.push_v(q, name, nx, ny, mesh.dx, mesh.dy, ex, ey, bzg if bz == "grid" else None, 0.05)
    assert np.array_equal(p.array[:2], q[:2])
    assert np.max(np.abs(p.array[2:4] - q[2:4])) < 1e-14


def test_ngp_ties_go_to_even_nodes(gpu, O):
    """Particles exactly half-way between nodes: the GPU deposit picks the
    even node, as the reference's round(Int, x) does."""
    pic = gpu
    nx = ny = 16
    mesh = pic.TwoDGrid(16.0, nx, 16.0, ny)
    xs = np.array([0.5, 1.5, 2.5, 3.5, 6.5, 7.25, 9.75, 14.5])
    p = pic.ParticleGroup(xs.size)
    p.array[0], p.array[1], p.array[4] = xs, xs[::-1], 1.0
    rho = pic.compute_rho(p, pic.NearestGridPoint(), mesh)
    ro = O.compute_rho(p.array, "ngp", nx, ny, mesh.dx, mesh.dy, mesh.dimx, mesh.dimy)
    assert np.max(np.abs(rho - ro)) < 1e-14


@pytest.mark.parametrize("name", SHAPES)
@pytest.mark.parametrize("sort_every", [0, 3])
def test_shape_run_matches_reference_loop(gpu, O, name, sort_every):
    """pic_run with each shape against the oracle loop, FP64, 10 steps:
    energies to 1e-10, particles as a set to 1e-10."""
    pic = gpu
    mesh, p = landau_particles(pic, 60000, 64, 64)
    sim = pic.Simulation(mesh, p, _shape(pic, name))
    e = sim.run(0.1, 10, sort_every)
    got = sim.particles()
    sim.close()
    q = p.copy(order="F")
    eo = O.landau_loop(q, name, 64, 64, mesh.dimx, mesh.dimy, 0.1, 10)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
    from scipy.spatial import cKDTree
    dist, idx = cKDTree(q[:4].T).query(got[:4].T)
    assert np.max(dist) < 1e-10
    assert np.unique(idx).size == q.shape[1]


@pytest.mark.parametrize("name", SHAPES)
def test_shape_f32_loop_tracks_f64(gpu, name):
    """The FP32 cell-relative loop with each shape against the FP64 loop from
    the same particles over 40 steps: field energies within FP32 rounding
    (NGP's node choice flips for particles within rounding of a half-node,
    hence its looser bound)."""
    pic = gpu
    mesh, p = landau_particles(pic, 200000, 64, 64)
    out = []
    for state in (pic._lib.PIC_STATE_F64, pic._lib.PIC_STATE_F32):
        sim = pic.Simulation(mesh, p, _shape(pic, name), state=state)
        out.append(sim.run(0.1, 40, 4))
        sim.close()
    e64, e32 = out[0][:, 0], out[1][:, 0]
    tol = 2e-3 if name == "ngp" else 1e-4
    assert np.max(np.abs(e32 - e64) / e64) < tol


@pytest.mark.parametrize("name", ["tsc", "m4"])
def test_shape_f32_tile_copies_agree(gpu, name):
    """The FP32 loop with 1 and with 4 copies of the fixed-point charge tile
    (pic_set_f32_tiles): the copies sum exactly, so the runs differ only by
    the order of the FP64 flush atomics (which can flip the FP32 rounding of
    a field value, and so move a particle by an ulp); field energies over 10
    steps agree to 1e-6, where a lost or doubled copy would be off by more
    than 1e-2."""
    pic = gpu
    mesh, p = landau_particles(pic, 200000, 64, 64)
    out = []
    for tiles in (1, 4):
        sim = pic.Simulation(mesh, p, _shape(pic, name), state=pic._lib.PIC_STATE_F32)
        sim.set_f32_tiles(tiles)
        out.append(sim.run(0.1, 10, 4))
        sim.close()
    e1, e4 = out[0][:, 0], out[1][:, 0]
    assert np.max(np.abs(e4 - e1) / e1) < 1e-6
