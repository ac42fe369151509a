"""pic_run's one-launch iteration on small grids (power-of-two nx, ny <= 64,
pic_small.hip) against the oracle's restatement of the reference loop
(examples/landau_damping_poisson_cic_m6.jl:70-83, the shapes of :88-103) and
against the library's separate-kernel loop (pic_set_fused_step(ctx, 0)).

The one-launch loop keeps the particles in their order (no sort), so the
FP64 state is compared with the oracle particle by particle."""
import numpy as np
import pytest

from conftest import landau_particles, random_particles

pytestmark = pytest.mark.gpu

SHAPES = ["cic", "m6", "ngp", "tsc", "m4"]
GRIDS = [(64, 64), (16, 32), (64, 8)]


def _shape(pic, name):
    return {"cic": pic.CloudInCell(), "m6": pic.M6(), "ngp": pic.NearestGridPoint(),
            "tsc": pic.TriangularShapeCloud(), "m4": pic.M4()}[name]


def _pdist(a, b, period):
    d = np.abs(a - b) % period
    return np.minimum(d, period - d)


def _run(pic, mesh, p, name, steps, dt, state=0, fused=True, bz=None):
    sim = pic.Simulation(mesh, p, _shape(pic, name), state=state, bz=bz)
    sim.set_fused_step(fused)
    e = sim.run(dt, steps)
    got = sim.particles()
    rho = sim.grid(pic._lib.PIC_GRID_RHO)
    ex = sim.grid(pic._lib.PIC_GRID_EX)
    sim.close()
    return e, got, rho, ex


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("name", SHAPES)
def test_fused_step_f64_matches_oracle(gpu, O, name, grid):
    """FP64, 20 steps of 20,000 particles: energies within 1e-10 of the
    oracle loop, every particle within 1e-9 (positions, periodic) and
    velocities (FP64 atomics reorder the deposit sums and the transforms
    round differently, so trajectories agree to rounding); rho and ex of the
    last step within 1e-9 normwise of the separate-kernel loop."""
    pic = gpu
    nx, ny = grid
    mesh, p = landau_particles(pic, 20_000, nx, ny)
    print(f"fused step f64 {name} {nx}x{ny}", flush=True)
    e, got, rho, ex = _run(pic, mesh, p, name, 20, 0.1)
    q = p.copy(order="F")
    eo = O.landau_loop(q, name, nx, ny, mesh.dimx, mesh.dimy, 0.1, 20)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
    assert np.max(_pdist(got[0], q[0], mesh.dimx)) < 1e-9
    assert np.max(_pdist(got[1], q[1], mesh.dimy)) < 1e-9
    assert np.max(np.abs(got[2:4] - q[2:4])) < 1e-9
    assert np.array_equal(got[4], q[4])
    e2, _, rho2, ex2 = _run(pic, mesh, p, name, 20, 0.1, fused=False)
    assert np.max(np.abs(e - e2) / e2) < 1e-10
    assert np.linalg.norm(rho - rho2) < 1e-9 * np.linalg.norm(rho2)
    assert np.linalg.norm(ex - ex2) < 1e-9 * np.linalg.norm(ex2)


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("name", SHAPES)
def test_fused_step_f32(gpu, O, name, grid):
    """FP32 state, 20 steps: energies within 1e-4 of the FP64 oracle loop and
    within 1e-5 of the separate-kernel FP32 loop (both round every charge
    share to a power-of-two quantum, each its own), positions within 1e-4;
    a second run from the same state is bit-identical (int64 charge sums,
    fixed-order energy sums)."""
    pic = gpu
    nx, ny = grid
    mesh, p = landau_particles(pic, 20_000, nx, ny)
    print(f"fused step f32 {name} {nx}x{ny}", flush=True)
    e, got, rho, _ = _run(pic, mesh, p, name, 20, 0.1, state=1)
    q = p.copy(order="F")
    eo = O.landau_loop(q, name, nx, ny, mesh.dimx, mesh.dimy, 0.1, 20)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-4
    assert np.max(_pdist(got[0], q[0], mesh.dimx)) < 1e-4
    assert np.max(_pdist(got[1], q[1], mesh.dimy)) < 1e-4
    e2, _, _, _ = _run(pic, mesh, p, name, 20, 0.1, state=1, fused=False)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - np.sqrt(e2[:, 0])) / np.sqrt(e2[:, 0])) < 1e-5
    e3, got3, rho3, _ = _run(pic, mesh, p, name, 20, 0.1, state=1)
    assert np.array_equal(e, e3)
    assert np.array_equal(got, got3)
    assert np.array_equal(rho, rho3)


@pytest.mark.parametrize("name", ["cic", "m6"])
def test_fused_step_uniform_bz(gpu, O, name):
    """Uniform bz = 2 (the rotation of push_v!, pushers_cic.jl:38-49): FP64
    against the oracle loop with the same bz, particle by particle."""
    pic = gpu
    mesh, p = landau_particles(pic, 20_000, 32, 32)
    e, got, _, _ = _run(pic, mesh, p, name, 15, 0.1, bz=2.0)
    q = p.copy(order="F")
    eo = O.landau_loop(q, name, 32, 32, mesh.dimx, mesh.dimy, 0.1, 15, bz0=2.0)
    assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
    assert np.max(_pdist(got[0], q[0], mesh.dimx)) < 1e-9
    assert np.max(np.abs(got[2:4] - q[2:4])) < 1e-9


def test_fused_step_weights_and_few_particles(gpu, O):
    """Per-particle weights, and 300 particles (one block, most lanes idle):
    FP64 against the oracle loop particle by particle."""
    pic = gpu
    rng = np.random.default_rng(31)
    mesh = pic.TwoDGrid(4 * np.pi, 32, 4 * np.pi, 32)
    for n in (40_000, 300):
        w = rng.uniform(0.5, 1.5, n) * mesh.dimx * mesh.dimy / n
        p = random_particles(rng, n, mesh.dimx, mesh.dimy, w=w)
        e, got, _, _ = _run(pic, mesh, p, "m6", 7, 0.1)
        q = p.copy(order="F")
        eo = O.landau_loop(q, "m6", 32, 32, mesh.dimx, mesh.dimy, 0.1, 7)
        assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
        assert np.max(_pdist(got[0], q[0], mesh.dimx)) < 1e-9
        assert np.array_equal(got[4], q[4])


def test_fused_step_one_step_and_rerun(gpu, O):
    """nsteps = 1 (the prologue launch, then the last one alone) and two runs
    in a row on one context (the charge buffers rotate from a cleared
    start every call): energies against the oracle loop."""
    pic = gpu
    mesh, p = landau_particles(pic, 20_000, 64, 64)
    sim = pic.Simulation(mesh, p, pic.CloudInCell())
    e1 = sim.run(0.1, 1)
    e2 = sim.run(0.1, 4)
    got = sim.particles()
    sim.close()
    q = p.copy(order="F")
    eo1 = O.landau_loop(q, "cic", 64, 64, mesh.dimx, mesh.dimy, 0.1, 1)
    eo2 = O.landau_loop(q, "cic", 64, 64, mesh.dimx, mesh.dimy, 0.1, 4)
    assert abs(np.sqrt(e1[0, 0]) - eo1[0]) < 1e-10 * eo1[0]
    assert np.max(np.abs(np.sqrt(e2[:, 0]) - eo2) / eo2) < 1e-10
    assert np.max(_pdist(got[0], q[0], mesh.dimx)) < 1e-9


@pytest.mark.parametrize("grid", [(4, 4), (8, 4), (4, 64)])
def test_fused_step_tiny_grids(gpu, O, grid):
    """Grids whose transforms span 4 or 8 lanes of a wave (16 or 8
    transforms per wave at once): FP64 CIC and M6 (a stencil wider than the
    grid, which wraps onto itself) against the oracle loop, particle by
    particle."""
    pic = gpu
    nx, ny = grid
    mesh, p = landau_particles(pic, 4096, nx, ny)
    for name in ("cic", "m6"):
        e, got, _, _ = _run(pic, mesh, p, name, 10, 0.1)
        q = p.copy(order="F")
        eo = O.landau_loop(q, name, nx, ny, mesh.dimx, mesh.dimy, 0.1, 10)
        assert np.max(np.abs(np.sqrt(e[:, 0]) - eo) / eo) < 1e-10
        assert np.max(_pdist(got[0], q[0], mesh.dimx)) < 1e-9
        assert np.max(np.abs(got[2:4] - q[2:4])) < 1e-9
