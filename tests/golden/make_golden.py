"""Writes the golden fixtures of tests/golden from the reference's own data.

  f90_push.npz        inputs and outputs of the reference's Fortran twin
                      src/pic.f90 (push_v, push_x, deposition), built by
                      oracle/Makefile into oracle/_ref/libpic.so and run here;
                      the reference's tests assert these equal Julia's
                      push_v!/push_x! bit for bit (test/test_push.jl:41,
                      test/test_landau_damping.jl:62-73)
  poisson_test2.npz   the data file test/test2.dat of the reference (FFTW
                      output of the second test/test_poisson.jl case), columns
                      x y ex ey rho at 65 x 129 nodes, as arrays

Run from the repository root where /root/reference exists:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

REF_TEST2 = "/root/reference/test/test2.dat"


def f90_push():
    rng = np.random.default_rng(20240601)
    nx, ny = 32, 16
    dimx, dimy = 4 * np.pi, 2.0
    dx, dy = dimx / nx, dimy / ny
    n = 2000
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(0, dimx, n)
    p[1] = rng.uniform(0, dimy, n)
    p[2] = rng.normal(size=n)
    p[3] = rng.normal(size=n)
    p[4] = dimx * dimy / n
    ex, ey, bz = (np.asfortranarray(rng.normal(size=(nx + 1, ny + 1))) for _ in range(3))
    for g in (ex, ey, bz):  # periodic ghosts, as MeshFields carries them
        g[nx, :] = g[0, :]
        g[:, ny] = g[:, 0]
    dt = 0.1
    pv = p.copy(order="F")
    O.ref_push_v(pv, nx, ny, dx, dy, ex, ey, bz, dt)
    px = pv.copy(order="F")
    O.ref_push_x(px, dimx, dimy, dt)
    # deposition with v = w and n = nx*ny (factor 1): jx is the raw CIC charge
    m = nx * ny
    q = np.zeros((5, m), order="F")
    q[0] = rng.uniform(0, dimx, m)
    q[1] = rng.uniform(0, dimy, m)
    q[2] = q[3] = q[4] = dimx * dimy / m
    jx, jy = O.ref_deposition(q, nx, ny, dx, dy)
    # test_push.jl: one particle, 50 x (push_v, push_x), 100 x 100 on [0,1]^2
    k = np.zeros((5, 1), order="F")
    k[0, 0] = k[1, 0] = 0.1
    k[4, 0] = 1.0
    g1 = np.ones((101, 101), order="F")
    g0 = np.zeros((101, 101), order="F")
    traj = []
    for _ in range(50):
        O.ref_push_v(k, 100, 100, 0.01, 0.01, g1, g0, g1, 0.1)
        O.ref_push_x(k, 1.0, 1.0, 0.1)
        traj.append(k[:, 0].copy())
    np.savez_compressed(os.path.join(HERE, "f90_push.npz"), nx=nx, ny=ny, dimx=dimx, dimy=dimy,
                        dt=dt, p_in=p, ex=ex, ey=ey, bz=bz, p_push_v=pv, p_push_x=px,
                        dep_in=q, dep_jx=jx, dep_jy=jy, single_traj=np.array(traj))


def poisson_test2():
    d = np.loadtxt(REF_TEST2)
    nx, ny = 64, 128
    shape = (nx + 1, ny + 1)  # written i-outer, j-inner (src/gnuplot.jl:13-22)
    np.savez_compressed(os.path.join(HERE, "poisson_test2.npz"), nx=nx, ny=ny,
                        xmax=2 * np.pi / 0.5, ymax=2 * np.pi,
                        x=d[:, 0].reshape(shape), y=d[:, 1].reshape(shape),
                        ex=d[:, 2].reshape(shape), ey=d[:, 3].reshape(shape),
                        rho=d[:, 4].reshape(shape))


if __name__ == "__main__":
    if not O.ref_available():
        O.build()
    f90_push()
    poisson_test2()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
