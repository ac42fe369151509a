"""pic_compute_rho_noalloc, the allocation-free variant of pic_compute_rho:
the caller's device arrays in and out, the library's static scratch array in
between.  Checked against pic_compute_rho (through a context) on the same
particles, to the FP64 bar of 1e-12 normwise (atomics order the sums
differently), on consecutive calls with shrinking inputs, so that the
scratch keeps the previous call's contents when the next one starts."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import random_particles

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip(pic):
    """The HIP runtime the library itself loaded (its device allocations
    stand for the caller's own)."""
    pic._lib.load()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 not mapped"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0

    def put(self, a):
        a = np.asfortranarray(a)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(pic, hip, kernel, p, mesh):
    n = p.shape[1]
    dp, drho = DevBuf(hip, p.nbytes), DevBuf(hip, 8 * mesh.nx * mesh.ny)
    dp.put(p)
    pic._lib.call("pic_compute_rho_noalloc", kernel.code, dp.p, n, mesh.nx, mesh.ny, mesh.xmin,
                  mesh.xmax, mesh.ymin, mesh.ymax, drho.p, None)
    rho = drho.get(np.zeros((mesh.nx, mesh.ny), order="F"))  # hipMemcpy waits for the stream
    dp.free()
    drho.free()
    return rho


def _nrel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_compute_rho_noalloc_matches_compute_rho(gpu):
    pic = gpu
    hip = _hip(pic)
    rng = np.random.default_rng(91)
    cases = [  # shrinking inputs: particles, then the grid
        (pic.M6(), 128, 128, 300_000),
        (pic.M6(), 128, 128, 40_000),
        (pic.CloudInCell(), 64, 32, 20_000),
        (pic.TriangularShapeCloud(), 32, 32, 5_000),
    ]
    for kernel, nx, ny, n in cases:
        mesh = pic.TwoDGrid(4 * math.pi, nx, 4 * math.pi, ny)
        p = random_particles(rng, n, mesh.dimx, mesh.dimy,
                             w=rng.uniform(0.5, 1.5, n) * mesh.dimx * mesh.dimy / n)
        got = _noalloc(pic, hip, kernel, p, mesh)
        pg = pic.ParticleGroup(n)
        pg.array[...] = p
        want = pic.compute_rho(pg, kernel, mesh)
        assert _nrel(got, want) < 1e-12, (type(kernel).__name__, nx, ny, n)


def test_compute_rho_noalloc_empty_and_limits(gpu):
    pic = gpu
    hip = _hip(pic)
    mesh = pic.TwoDGrid(4 * math.pi, 16, 4 * math.pi, 16)
    got = _noalloc(pic, hip, pic.CloudInCell(), np.zeros((5, 0), order="F"), mesh)
    assert np.all(got == 0)
    with pytest.raises(pic._lib.PicError, match="exceeds"):
        pic._lib.call("pic_compute_rho_noalloc", pic._lib.PIC_CIC, None, 0, 1024, 512, 0.0, 1.0,
                      0.0, 1.0, C.c_void_p(1), None)
