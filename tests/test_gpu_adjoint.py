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
