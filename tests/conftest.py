import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def O():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def pic():
    """The package; GPU tests fail (never skip) when the library is absent."""
    return entry.load_package()


@pytest.fixture(scope="session")
def gpu(pic):
    n = pic._lib.device_count()
    assert n >= 1, "no GPU visible to libpicmi355x"
    return pic


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def landau_particles(pic, n, nx, ny, alpha=0.05, kx=0.5, dimy=None):
    import math
    dimx = 2 * math.pi / kx
    mesh = pic.TwoDGrid(0.0, dimx, nx, 0.0, dimy if dimy else dimx, ny)
    return mesh, pic.sampling.landau_damping(n, mesh, alpha, kx)


def random_particles(rng, n, dimx, dimy, w=None):
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(0, dimx, n)
    p[1] = rng.uniform(0, dimy, n)
    p[2] = rng.normal(size=n)
    p[3] = rng.normal(size=n)
    p[4] = dimx * dimy / n if w is None else w
    return p
