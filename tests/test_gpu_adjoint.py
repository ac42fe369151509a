"""The GPU adjoint sweep (pic_adjoint) against the oracle's adjoint and
against central finite differences of GPU forward runs (FP64)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _landau(pic, n, nx, alpha, kx=0.5, sob=None):
    mesh = pic.TwoDGrid(0.0, 2 * math.pi / kx, nx, 0.0, 2 * math.pi / kx, nx)
    if sob is None:
        sob = pic.sampling.sobol3(n)
    return mesh, sob, pic.sampling.landau_damping(n, mesh, alpha, kx, sobol=sob)


@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_adjoint_matches_oracle(gpu, O, kernel):
    pic = gpu
    n, nx, dt, steps, ck = 20000, 32, 0.1, 12, 5
    mesh, sob, p = _landau(pic, n, nx, 0.05)
    sim = pic.Simulation(mesh, p, pic.CloudInCell() if kernel == "cic" else pic.M6())
    J, gx, gv = sim.adjoint(dt, steps, ck)
    after = sim.particles()
    sim.close()
    Jo, gxo, gvo, _ = O.adjoint(p, kernel, nx, nx, mesh.dimx, mesh.dimy, dt, steps, ck)
    assert J == pytest.approx(Jo, rel=1e-10)
    for a, b in ((gx, gxo), (gv, gvo)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-8
    # the particles are left as they were
    assert np.array_equal(after, p)


def test_adjoint_checkpoint_independent(gpu):
    """Checkpoints every 1, 4 or all steps: the recomputation with the stored
    fields reproduces the forward, so the gradients agree closely (the
    deposits of different checkpoint runs differ only in summation order)."""
    pic = gpu
    n, nx, dt, steps = 16384, 32, 0.1, 8
    mesh, sob, p = _landau(pic, n, nx, 0.05)
    res = []
    for ck in (1, 4, 8):
        sim = pic.Simulation(mesh, p, pic.CloudInCell())
        res.append(sim.adjoint(dt, steps, ck))
        sim.close()
    for J, gx, gv in res[1:]:
        assert J == pytest.approx(res[0][0], rel=1e-12)
        assert np.linalg.norm(gx - res[0][1]) / np.linalg.norm(res[0][1]) < 1e-10
        assert np.linalg.norm(gv - res[0][2]) / np.linalg.norm(res[0][2]) < 1e-10


@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_adjoint_finite_differences(gpu, kernel):
    """dJ/dalpha (through the sampler's dx0/dalpha) and a random directional
    derivative in vx against central differences of GPU forward runs,
    h = 1e-6: within 1e-4 relative."""
    pic = gpu
    n, nx, alpha, dt, steps = 65536, 32, 0.05, 0.1, 20
    k = pic.CloudInCell() if kernel == "cic" else pic.M6()
    mesh, sob, p = _landau(pic, n, nx, alpha)
    sim = pic.Simulation(mesh, p, k)
    J, gx, gv = sim.adjoint(dt, steps, 5)
    sim.close()

    def fwd(q):
        s = pic.Simulation(mesh, q, k)
        e = s.run(dt, steps)
        s.close()
        return e[-1, 0] + e[-1, 1]

    assert fwd(p) == pytest.approx(J, rel=1e-12)
    h = 1e-6
    dxa = pic.sampling.dnewton_dalpha(p[0], alpha, 0.5)
    fd = (fwd(_landau(pic, n, nx, alpha + h, sob=sob)[2])
          - fwd(_landau(pic, n, nx, alpha - h, sob=sob)[2])) / (2 * h)
    assert float(np.sum(gx[0] * dxa)) == pytest.approx(fd, rel=1e-4)
    u = np.random.default_rng(7).choice([-1.0, 1.0], size=n)
    qp, qm = p.copy(order="F"), p.copy(order="F")
    qp[2] += h * u
    qm[2] -= h * u
    fdv = (fwd(qp) - fwd(qm)) / (2 * h)
    assert float(np.sum(gv[0] * u)) == pytest.approx(fdv, rel=1e-4)


# ------------------------------------------------------------ FP32 state
# The FP32 adjoint differentiates the FP32 loop (cell-relative FP32 particle
# arithmetic, FP64 charge sums and Poisson solve); its reference values are
# the FP64 adjoint and central differences of FP64 runs.  Over 20 steps the
# two trajectories stay within a few FP32 roundings of each other, so the
# gradients agree to about 1e-6.


@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_adjoint_f32_against_f64_and_fd(gpu, kernel):
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    n, nx, alpha, dt, steps, ck = 65536, 32, 0.05, 0.1, 20, 5
    k = pic.CloudInCell() if kernel == "cic" else pic.M6()
    mesh, sob, p = _landau(pic, n, nx, alpha)
    sim = pic.Simulation(mesh, p, k)
    J64, gx64, gv64 = sim.adjoint(dt, steps, ck)
    sim.close()
    sim = pic.Simulation(mesh, p, k, state=F32)
    before = sim.state_f32()
    J32, gx32, gv32 = sim.adjoint(dt, steps, ck)
    after = sim.state_f32()
    sim.close()
    # the FP32 state is left bit for bit as it was
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    # J is the FP32 loop's own field energy (same sorts, same charge sums):
    # the separate-kernel loop, whose segment passes the sweep's forward
    # pass runs (on this grid pic_run's default is the one-launch loop, whose
    # fixed-point charge rounds with another quantum)
    sim = pic.Simulation(mesh, p, k, state=F32)
    sim.set_fused_step(False)
    e = sim.run(dt, steps, ck)
    sim.close()
    assert J32 == pytest.approx(e[-1, 0] + e[-1, 1], rel=1e-12)
    assert J32 == pytest.approx(J64, rel=1e-5)
    for a, b in ((gx32, gx64), (gv32, gv64)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 2e-5
    dxa = pic.sampling.dnewton_dalpha(p[0], alpha, 0.5)

    def fwd(q):
        s = pic.Simulation(mesh, q, k)
        en = s.run(dt, steps)
        s.close()
        return en[-1, 0] + en[-1, 1]

    h = 1e-6
    fd = (fwd(_landau(pic, n, nx, alpha + h, sob=sob)[2])
          - fwd(_landau(pic, n, nx, alpha - h, sob=sob)[2])) / (2 * h)
    assert float(np.sum(gx32[0] * dxa)) == pytest.approx(fd, rel=1e-5)
    u = np.random.default_rng(7).choice([-1.0, 1.0], size=n)
    qp, qm = p.copy(order="F"), p.copy(order="F")
    qp[2] += h * u
    qm[2] -= h * u
    fdv = (fwd(qp) - fwd(qm)) / (2 * h)
    assert float(np.sum(gv32[0] * u)) == pytest.approx(fdv, rel=1e-4)


def test_adjoint_f32_checkpoint_independent(gpu):
    """FP32: checkpoints every 1, 3 or all steps give the same gradient up to
    the order of the FP64 charge sums (the recomputed positions are
    bit-identical to the forward's)."""
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    n, nx, dt, steps = 16384, 32, 0.1, 8
    mesh, sob, p = _landau(pic, n, nx, 0.05)
    res = []
    for ck in (1, 3, 8):
        sim = pic.Simulation(mesh, p, pic.M6(), state=F32)
        res.append(sim.adjoint(dt, steps, ck))
        sim.close()
    for J, gx, gv in res[1:]:
        assert J == pytest.approx(res[0][0], rel=1e-6)
        assert np.linalg.norm(gx - res[0][1]) / np.linalg.norm(res[0][1]) < 1e-5
        assert np.linalg.norm(gv - res[0][2]) / np.linalg.norm(res[0][2]) < 1e-5


@pytest.mark.parametrize("kernel,sort_every", [("cic", 0), ("cic", 2), ("m6", 3)])
def test_adjoint_f32_windows_and_sorts(gpu, kernel, sort_every):
    """FP32 adjoint on a 256 x 32 grid, where the reverse passes' LDS windows
    cover a few cells around each 64-cell segment: particles that drift past
    them take the deferred global-memory path (sort_every 0), and sorts
    between checkpoints (2, or 3, which does not divide the checkpoint
    interval) are replayed by the recomputation.  Reference: the FP64
    adjoint (checkpoint sorts only), itself pinned to the oracle's adjoint
    on this grid by test_adjoint_f64_windows_and_sorts."""
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    n, dt, steps, ck = 65536, 0.1, 20, 5
    k = pic.CloudInCell() if kernel == "cic" else pic.M6()
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 256, 0.0, 4 * math.pi, 32)
    p = pic.sampling.landau_damping(n, mesh, 0.05, 0.5, sobol=pic.sampling.sobol3(n))
    sim = pic.Simulation(mesh, p, k)
    J64, gx64, gv64 = sim.adjoint(dt, steps, ck)
    sim.close()
    sim = pic.Simulation(mesh, p, k, state=F32)
    before = sim.state_f32()
    slow0 = sim.slow_chunks()
    J32, gx32, gv32 = sim.adjoint(dt, steps, ck, sort_every)
    deferred = sim.slow_chunks() - slow0
    after = sim.state_f32()
    sim.close()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    if sort_every == 0:
        assert deferred > 0  # the global-memory fallback was exercised
    assert J32 == pytest.approx(J64, rel=1e-5)
    # CIC's shape derivative jumps at cell faces: a particle whose FP32 and
    # FP64 trajectories sit on either side of a face takes a different
    # derivative, more often on this grid's narrow cells
    tol = 1e-4 if kernel == "cic" else 3e-5
    for a, b in ((gx32, gx64), (gv32, gv64)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < tol
    if sort_every == 2:
        # J is the loop's own: pic_run sorting every 2 steps sorts at the
        # same steps but the odd checkpoint steps 5 and 15, which changes
        # only the rounding of the charge sums
        sim = pic.Simulation(mesh, p, k, state=F32)
        e = sim.run(dt, steps, sort_every)
        sim.close()
        assert J32 == pytest.approx(e[-1, 0] + e[-1, 1], rel=1e-6)


@pytest.mark.parametrize("kernel,sort_every", [("cic", 0), ("cic", 2), ("m6", 0), ("m6", 3)])
def test_adjoint_f64_windows_and_sorts(gpu, O, kernel, sort_every):
    """FP64 adjoint on a 256 x 32 grid, where the reverse passes' LDS windows
    cover a few cells around each 64-cell segment: with sorts at checkpoints
    only, particles drift past their windows and take the deferred
    global-memory path; sorts between checkpoints (2, or 3, which does not
    divide the checkpoint interval) are replayed by the recomputation and
    the adjoints carried back through their permutations.  Against the
    oracle's adjoint: J to 1e-10, gradients to 1e-8 normwise; the particles
    are left as they were."""
    pic = gpu
    n, dt, steps, ck = 65536, 0.1, 20, 5
    k = pic.CloudInCell() if kernel == "cic" else pic.M6()
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 256, 0.0, 4 * math.pi, 32)
    p = pic.sampling.landau_damping(n, mesh, 0.05, 0.5, sobol=pic.sampling.sobol3(n))
    sim = pic.Simulation(mesh, p, k)
    slow0 = sim.slow_chunks()
    J, gx, gv = sim.adjoint(dt, steps, ck, sort_every)
    deferred = sim.slow_chunks() - slow0
    after = sim.particles()
    sim.close()
    assert np.array_equal(after, p)
    if sort_every == 0:
        assert deferred > 0  # the global-memory fallback was exercised
    Jo, gxo, gvo, _ = O.adjoint(p, kernel, 256, 32, mesh.dimx, mesh.dimy, dt, steps, ck)
    assert J == pytest.approx(Jo, rel=1e-10)
    for a, b in ((gx, gxo), (gv, gvo)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-8


def test_adjoint_f32_two_level_matches_one_level(gpu):
    """Restart points inside the checkpoint segments (positions of 1 or 3
    steps kept at once, restart points at sorted and unsorted steps) give the
    one-level sweep's gradient up to the order of its FP64 sums."""
    pic = gpu
    F32 = pic._lib.PIC_STATE_F32
    n, dt, steps, ck, se = 32768, 0.1, 20, 10, 2
    mesh = pic.TwoDGrid(0.0, 4 * math.pi, 256, 0.0, 4 * math.pi, 32)
    p = pic.sampling.landau_damping(n, mesh, 0.05, 0.5, sobol=pic.sampling.sobol3(n))
    res = []
    for block in (0, 3, 1):
        sim = pic.Simulation(mesh, p, pic.M6(), state=F32)
        sim.set_adjoint_block(block)
        res.append(sim.adjoint(dt, steps, ck, se))
        used, nbytes = sim.adjoint_plan()
        assert used == (block if block else ck) and nbytes > 0
        sim.close()
    for J, gx, gv in res[1:]:
        assert J == pytest.approx(res[0][0], rel=1e-12)
        assert np.linalg.norm(gx - res[0][1]) / np.linalg.norm(res[0][1]) < 1e-6
        assert np.linalg.norm(gv - res[0][2]) / np.linalg.norm(res[0][2]) < 1e-6
