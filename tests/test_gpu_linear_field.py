"""The gathers of every push on the GPU reproduce linear fields exactly.

B-splines of order two and up (CIC, TSC, M4, M6) reproduce linear
functions: the gather of e = a + b x + c y at a particle is e(x_p), so
push_v! with bz = 0 kicks it by dt e(x_p) (the invariant test/test_uapic.jl
:43-74 checks on the MeshFields gather).  A stencil shifted by one node is
off by b dx (0.33 here), ex and ey swapped or the grid's two indices
transposed by far more, so these tests pin the alignment of the FP64 M6
push (pushers_m6.jl:24-207), of the fused segment passes of the time loop
and of the FP32 state's gather, independently of the oracle.  Particles
stay 4 cells inside the periodic seam, where the linear node values jump.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, DX, DY = 32, 16, 0.3, 0.45
SHAPES = {"cic": 0, "m6": 1, "tsc": 3, "m4": 4}


def linear_e(x, y):
    return 0.3 + 1.1 * x - 0.4 * y, -0.2 + 0.7 * y + 0.05 * x


def linear_b(x, y):
    return 1.5 - 0.2 * x + 0.3 * y


def _setup(pic, n, seed):
    rng = np.random.default_rng(seed)
    mesh = pic.TwoDGrid(NX * DX, NX, NY * DY, NY)
    p = np.zeros((5, n), order="F")
    p[0] = rng.uniform(4 * DX, (NX - 5) * DX, n)
    p[1] = rng.uniform(4 * DY, (NY - 5) * DY, n)
    p[2] = rng.normal(size=n)
    p[3] = rng.normal(size=n)
    p[4] = mesh.dimx * mesh.dimy / n
    X, Y = np.meshgrid(np.arange(NX) * DX, np.arange(NY) * DY, indexing="ij")
    ex, ey = (np.asfortranarray(a) for a in linear_e(X, Y))
    bz = np.asfortranarray(linear_b(X, Y))
    return mesh, p, ex, ey, bz


def _boris(v1, v2, e1, e2, b3, dt):
    v1 = v1 + 0.5 * dt * e1
    v2 = v2 + 0.5 * dt * e2
    tn = 0.5 * dt * b3
    sn = 2 * tn / (1 + tn * tn)
    v1 = v1 + v2 * tn
    v2 = v2 - v1 * sn
    v1 = v1 + v2 * tn
    return v1 + 0.5 * dt * e1, v2 + 0.5 * dt * e2


@pytest.mark.parametrize("kernel", ["m6", "tsc", "m4", "cic"])
@pytest.mark.parametrize("with_bz", [False, True])
def test_push_v_f64_linear_field(gpu, kernel, with_bz):
    """push_v!(p, kernel, mesh, ex, ey, bz, dt) through the per-call API:
    the velocity is the Boris update with e and b taken at the particle,
    to 1e-12."""
    pic = gpu
    mesh, p, ex, ey, bz = _setup(pic, 5000, 81)
    dt = 0.1
    k = {"m6": pic.M6(), "tsc": pic.TriangularShapeCloud(), "m4": pic.M4(),
         "cic": pic.CloudInCell()}[kernel]
    q = pic.ParticleGroup(p.shape[1])
    q.array[...] = p
    pic.push_v_(q, k, mesh, ex, ey, bz if with_bz else None, dt)
    e1, e2 = linear_e(p[0], p[1])
    b3 = linear_b(p[0], p[1]) if with_bz else np.zeros_like(e1)
    v1, v2 = _boris(p[2], p[3], e1, e2, b3, dt)
    assert np.max(np.abs(q.array[2] - v1)) < 1e-12
    assert np.max(np.abs(q.array[3] - v2)) < 1e-12
    assert np.array_equal(q.array[:2], p[:2])


@pytest.mark.parametrize("kernel", ["m6", "cic"])
def test_segment_pass_f64_linear_field(gpu, kernel):
    """The time loop's FP64 segment pass (kick, then drift dt/2: mode 2 of
    pic_step_pass after pic_sort_lanes): velocities kicked by dt e(x_p)."""
    pic = gpu
    mesh, p, ex, ey, _ = _setup(pic, 20000, 82)
    ctx = pic.Context(mesh, capacity=p.shape[1])
    ctx.set_particles(p)
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    ctx.set_bz(pic._lib.PIC_BZ_NONE)
    perm = np.zeros(p.shape[1], np.int32)
    pic._lib.call("pic_sort_lanes", ctx.h, perm.ctypes.data)
    dt = 0.1
    pic._lib.call("pic_step_pass", ctx.h, SHAPES[kernel], 2, dt, None)
    got = ctx.get_particles(np.zeros((5, p.shape[1]), order="F"))
    ctx.close()
    src = p[:, perm]
    e1, e2 = linear_e(src[0], src[1])
    assert np.max(np.abs((got[2] - src[2]) / dt - e1)) < 1e-11
    assert np.max(np.abs((got[3] - src[3]) / dt - e2)) < 1e-11


@pytest.mark.parametrize("lanes", [False, True])
@pytest.mark.parametrize("kernel", ["m6", "tsc", "m4", "cic"])
def test_f32_gather_linear_field(gpu, O, kernel, lanes):
    """The FP32 state's gather (interp32), in the passes for any particle
    order and in the segment passes after a lane sort: kick then drift
    dt/2 (pic_step_pass mode 2), velocities kicked by dt e(x_p) to FP32
    rounding (2e-4 of a field up to 15)."""
    pic = gpu
    mesh, p, ex, ey, _ = _setup(pic, 20000, 83)
    st = O.f32_from_f64(p, NX, NY, DX, DY)
    x32 = O.f32_to_f64(st, NX, DX, DY, p[4, 0])  # positions as the FP32 state holds them
    ctx = pic.Context(mesh, capacity=p.shape[1])
    ctx.set_state_f32(st, float(p[4, 0]))
    ctx.set_grid(pic._lib.PIC_GRID_EX, ex)
    ctx.set_grid(pic._lib.PIC_GRID_EY, ey)
    ctx.set_bz(pic._lib.PIC_BZ_NONE)
    order = np.arange(p.shape[1])
    if lanes:
        perm = np.zeros(p.shape[1], np.int32)
        pic._lib.call("pic_sort_lanes", ctx.h, perm.ctypes.data)
        order = perm
    dt = 0.1
    pic._lib.call("pic_step_pass", ctx.h, SHAPES[kernel], 2, dt, None)
    got = ctx.get_state_f32()
    ctx.close()
    e1, e2 = linear_e(x32[0, order], x32[1, order])
    v0x, v0y = st[3][order].astype(np.float64), st[4][order].astype(np.float64)
    assert np.max(np.abs((got[3] - v0x) / dt - e1)) < 2e-4
    assert np.max(np.abs((got[4] - v0y) / dt - e2)) < 2e-4
