"""Initial particle distributions of the hot-path configurations (host side).

landau_damping(): the quiet start of sample!(pg::ParticleGroup{2,2}, mesh,
LandauDamping(alpha, kx)) (landau_damping.jl:77-97):
    x  = newton(r1)  with P(x) = x + alpha*sin(kx*x)/kx = r1*2pi/kx
    y  = ymin + r2*dimy
    |v| = sqrt(-2 log((i - 1/2)/N)),  theta = 2pi*r3,  v = |v|(cos, sin)
    w  = dimx*dimy/N
with (r1, r2, r3) the i-th point of the 3-dimensional Sobol sequence.  The
Sobol points come from scipy's unscrambled Sobol generator (Joe-Kuo
direction numbers) with the initial all-zero point skipped; Sobol.jl's exact
sequence is not checked here (parity unpinned, the same arrays feed the GPU
and the CPU oracle).

newton() restates landau_damping.jl:19-28 exactly, including its interleaved
pair update (x0, x1) <- (x1, x0 - (P(x0) - r)/P'(x0)), element by element.
"""
import math
import warnings

import numpy as np


def newton(r, alpha, kx):
    r = np.asarray(r, dtype=np.float64) * (2 * math.pi / kx)
    x0 = np.zeros_like(r)
    x1 = np.ones_like(r)
    act = np.abs(x1 - x0) > 1e-12
    while act.any():
        a0 = x0[act]
        p = a0 + alpha * np.sin(kx * a0) / kx
        f = 1 + alpha * np.cos(kx * a0)
        nx1 = a0 - (p - r[act]) / f
        x0[act] = x1[act]
        x1[act] = nx1
        act = np.abs(x1 - x0) > 1e-12
    return x1


def dnewton_dalpha(x, alpha, kx):
    """Derivative of the sampled position with respect to alpha at fixed r:
    P(x; alpha) = const gives dx/dalpha = -(sin(kx x)/kx) / (1 + alpha cos(kx x))."""
    x = np.asarray(x, dtype=np.float64)
    return -(np.sin(kx * x) / kx) / (1 + alpha * np.cos(kx * x))


def sobol3(n):
    from scipy.stats import qmc
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        s = qmc.Sobol(d=3, scramble=False).random(n + 1)
    return s[1:]


def landau_damping(n, mesh, alpha, kx, sobol=None):
    """Returns the 5 x n particle array of the Landau quiet start."""
    if not math.isclose(mesh.dimx, 2 * math.pi / kx, rel_tol=1e-12):
        raise ValueError("the Landau sampler needs dimx == 2pi/kx (landau_damping.jl:81)")
    s = sobol3(n) if sobol is None else sobol
    p = np.zeros((5, n), order="F")
    i = np.arange(1, n + 1, dtype=np.float64)
    v = np.sqrt(-2 * np.log((i - 0.5) / n))
    theta = s[:, 2] * 2 * math.pi
    p[0] = newton(s[:, 0], alpha, kx)
    p[1] = mesh.ymin + s[:, 1] * mesh.dimy
    p[2] = v * np.cos(theta)
    p[3] = v * np.sin(theta)
    p[4] = (mesh.dimx * mesh.dimy) / n
    return p


def perturbed_maxwellian(n, mesh, alpha, kx, ky, seed, vth=1.0):
    """Rejection sample of density 1 + alpha*cos(kx x)*cos(ky y) on the mesh
    with Maxwellian velocities of width vth (seeded, the non-separable
    strong-B configuration; the accept/reject loop follows plasma.jl:17-31)."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    need = n
    while need > 0:
        m = int(need * 1.2) + 1024
        x = rng.random(m) * mesh.dimx + mesh.xmin
        y = rng.random(m) * mesh.dimy + mesh.ymin
        z = (1.0 + alpha) * rng.random(m)
        keep = (1.0 + alpha * np.cos(kx * x) * np.cos(ky * y)) >= z
        xs.append(x[keep][:need])
        ys.append(y[keep][:need])
        need -= xs[-1].size
    p = np.zeros((5, n), order="F")
    p[0] = np.concatenate(xs)[:n]
    p[1] = np.concatenate(ys)[:n]
    p[2] = rng.normal(0.0, vth, n)
    p[3] = rng.normal(0.0, vth, n)
    p[4] = (mesh.dimx * mesh.dimy) / n
    return p


# The large workloads sample on the device, through the library:
# Simulation.landau (pic_sample_landau, the same quiet start as
# landau_damping above, made on the GPU) and Simulation.maxwellian
# (pic_sample_maxwellian, a seeded perturbed Maxwellian like
# perturbed_maxwellian above), so that their particle arrays never exist in
# host memory.
