"""Pins of the oracle's adjoint sweep: the transpose identities of its parts
and the gradient against central finite differences (CPU only)."""
import math

import numpy as np
import pytest


def test_poisson_adjoint_dot_product(O):
    """<P rho, mu> = <rho, P^T mu> for random rho, mu (P: rho -> (ex, ey))."""
    rng = np.random.default_rng(50)
    for nx, ny, dimx, dimy in ((32, 16, 4 * np.pi, 2.0), (16, 64, 1.0, 3.0)):
        rho = np.asfortranarray(rng.normal(size=(nx, ny)))
        mux, muy = (np.asfortranarray(rng.normal(size=(nx, ny))) for _ in range(2))
        ex, ey = O.solve_e(rho, dimx, dimy)
        lhs = np.sum(ex * mux) + np.sum(ey * muy)
        rhs = np.sum(rho * O.poisson_adjoint(mux, muy, dimx, dimy))
        assert lhs == pytest.approx(rhs, rel=1e-12)


def _setup(pic, n, nx, alpha=0.05, kx=0.5):
    mesh = pic.TwoDGrid(0.0, 2 * math.pi / kx, nx, 0.0, 2 * math.pi / kx, nx)
    sob = pic.sampling.sobol3(n)
    return mesh, sob


def _particles(pic, mesh, sob, alpha, kx=0.5):
    return pic.sampling.landau_damping(sob.shape[0], mesh, alpha, kx, sobol=sob)


@pytest.mark.parametrize("kernel", ["cic", "m6"])
def test_adjoint_matches_finite_differences(O, pic, kernel):
    """dJ/dalpha through the sampler (dx0/dalpha of the Newton inversion) and
    a directional derivative in the initial velocities, against central
    differences with h = 1e-6 in FP64: within 1e-4 relative."""
    n, nx, alpha, kx, dt, steps = 8192, 16, 0.05, 0.5, 0.1, 12
    mesh, sob = _setup(pic, n, nx, alpha, kx)
    p = _particles(pic, mesh, sob, alpha, kx)
    J, gx, gv, _ = O.adjoint(p, kernel, nx, nx, mesh.dimx, mesh.dimy, dt, steps, 4)
    assert J > 0
    dxa = pic.sampling.dnewton_dalpha(p[0], alpha, kx)
    dJ_da = float(np.sum(gx[0] * dxa))
    h = 1e-6
    jp = O.energy_objective(_particles(pic, mesh, sob, alpha + h, kx), kernel, nx, nx,
                            mesh.dimx, mesh.dimy, dt, steps)
    jm = O.energy_objective(_particles(pic, mesh, sob, alpha - h, kx), kernel, nx, nx,
                            mesh.dimx, mesh.dimy, dt, steps)
    fd = (jp - jm) / (2 * h)
    assert dJ_da == pytest.approx(fd, rel=1e-4), (dJ_da, fd)
    # directional derivative along a random direction in (vx)
    rng = np.random.default_rng(51)
    u = rng.choice([-1.0, 1.0], size=n)
    qp, qm = p.copy(order="F"), p.copy(order="F")
    qp[2] += h * u
    qm[2] -= h * u
    fdv = (O.energy_objective(qp, kernel, nx, nx, mesh.dimx, mesh.dimy, dt, steps)
           - O.energy_objective(qm, kernel, nx, nx, mesh.dimx, mesh.dimy, dt, steps)) / (2 * h)
    assert float(np.sum(gv[0] * u)) == pytest.approx(fdv, rel=1e-4)


def test_adjoint_checkpointing_is_exact(O, pic):
    """The reverse sweep with checkpoints every 1, 3 or all steps gives the
    same gradient bit for bit (recomputation with the stored fields)."""
    n, nx = 4096, 16
    mesh, sob = _setup(pic, n, nx)
    p = _particles(pic, mesh, sob, 0.05)
    ref = O.adjoint(p, "cic", nx, nx, mesh.dimx, mesh.dimy, 0.1, 9, 9)
    for ck in (1, 3, 4):
        got = O.adjoint(p, "cic", nx, nx, mesh.dimx, mesh.dimy, 0.1, 9, ck)
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])


def test_adjoint_energy_history_matches_loop(O, pic):
    n, nx = 4096, 16
    mesh, sob = _setup(pic, n, nx)
    p = _particles(pic, mesh, sob, 0.05)
    J, _, _, eh = O.adjoint(p, "m6", nx, nx, mesh.dimx, mesh.dimy, 0.1, 6, 2)
    q = p.copy(order="F")
    e = O.landau_loop(q, "m6", nx, nx, mesh.dimx, mesh.dimy, 0.1, 6)
    # the loop records sqrt(sum ex^2 dx dy); the adjoint sums ex^2 + ey^2
    assert np.all(eh >= e ** 2 * (1 - 1e-12))
    assert J == eh[-1]
