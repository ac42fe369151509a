"""The reference's particle-mesh API, backed by libpicmi355x on the GPU.

Mirrors ParticleInCell.jl's calls on the hot path with the same names,
argument order and array layouts (a Julia `f!` is written `f_` here):

    TwoDGrid(xmin, xmax, nx, ymin, ymax, ny) / TwoDGrid(dimx, nx, dimy, ny)
    ParticleGroup(n, charge=1.0, mass=1.0, n_weights=1)   .array is (5, n)
    CloudInCell(), M6()                                    kernel tags
    compute_rho(p, kernel, mesh)            pushers_cic.jl:75, pushers_m6.jl:226
    push_v_(p, kernel, mesh, ex, ey, bz, dt) pushers_cic.jl:3, pushers_m6.jl:5
    push_x_(p, mesh, dt)                     pushers_cic.jl:55
    operator_t_(pg, mesh, dt)                poisson_2d.jl:265 (splitting drift)
    interpolate_cic_(epx, epy, mesh, p, ex, ey)   interpolate_cic.jl:3
    TwoDPoissonPeriodic(grid)                poisson_2d.jl:28
    solve_(phi, poisson, rho)                poisson_2d.jl:78
    solve_(ex, ey, poisson, rho)             poisson_2d.jl:97
    MeshFields2D(mesh), Poisson2D(mesh), Particles(n, w)
    solve_(fields, poisson2d) -> energy      meshfields.jl:87
    compute_rho_m6_(fields, particles)       compute_rho_m6.jl:181
    interpol_eb_m6_(particles, fields)       interpolation_m6.jl:125
    F90.push_v_/push_x_/compute_current_     pic.jl:18-114 (Fortran-twin ABI)
    FDTD(mesh), faraday_, ampere_maxwell_, update_fields_, compute_energy
                                             fdtd.jl:3-80
    compute_current_(jx, jy, mesh, CloudInCell(), p)   cloud_in_cell.jl:10
    UA(ntau, eps, n), preparation_, update_particles_e_, compute_f_, ua_step_,
    update_particles_x_, compute_v_             ua_type.jl, ua_steps.jl
    read_particles(filename, mesh)           read_particles.jl:3
    Simulation.landau(mesh, n, alpha, kx, kernel)   landau_damping.jl:77 (on the GPU)
    Simulation.maxwellian(mesh, n, alpha, kx, ky, kernel, seed)  (seeded, on the GPU)

Arrays are numpy float64 in Fortran order, so the memory layout is the Julia
one.  Each call copies its inputs to the GPU and its outputs back (the
drop-in semantics of a host array API); `Simulation` keeps particles and
fields resident for the time loop.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _lib as L


# ------------------------------------------------------------------ types

class TwoDGrid:
    """GEMPIC's TwoDGrid: periodic nx x ny cells, dx = (xmax-xmin)/nx,
    x = LinRange(xmin, xmax, nx+1)."""

    def __init__(self, *args):
        if len(args) == 4:  # TwoDGrid(dimx, nx, dimy, ny)
            dimx, nx, dimy, ny = args
            xmin, xmax, ymin, ymax = 0.0, float(dimx), 0.0, float(dimy)
        elif len(args) == 6:  # TwoDGrid(xmin, xmax, nx, ymin, ymax, ny)
            xmin, xmax, nx, ymin, ymax, ny = args
        else:
            raise TypeError("TwoDGrid(dimx, nx, dimy, ny) or TwoDGrid(xmin, xmax, nx, ymin, ymax, ny)")
        self.xmin, self.xmax, self.nx = float(xmin), float(xmax), int(nx)
        self.ymin, self.ymax, self.ny = float(ymin), float(ymax), int(ny)
        self.dimx = self.xmax - self.xmin
        self.dimy = self.ymax - self.ymin
        self.dx = self.dimx / self.nx
        self.dy = self.dimy / self.ny
        self.x = np.linspace(self.xmin, self.xmax, self.nx + 1)
        self.y = np.linspace(self.ymin, self.ymax, self.ny + 1)

    def key(self):
        return (self.nx, self.ny, self.xmin, self.xmax, self.ymin, self.ymax)


class ParticleGroup:
    """ParticleGroup{2,2}: .array is 5 x n (x, y, vx, vy, w), column-major."""

    def __init__(self, n, charge=1.0, mass=1.0, n_weights=1):
        self.n_particles = int(n)
        self.charge, self.mass, self.n_weights = charge, mass, n_weights
        self.array = np.zeros((5, self.n_particles), order="F")


class CloudInCell:
    code = L.PIC_CIC


class M6:
    code = L.PIC_M6


class NearestGridPoint:
    """pushers_ngp.jl: the nearest node (round half to even)."""
    code = L.PIC_NGP


class TriangularShapeCloud:
    """pushers_tsc.jl: TSC deposit; its push_v! gathers with the M4 weights,
    as the reference's does."""
    code = L.PIC_TSC


class M4:
    """pushers_m4.jl: the cubic B-spline."""
    code = L.PIC_M4


_SHAPES = (CloudInCell, M6, NearestGridPoint, TriangularShapeCloud, M4)


def _kcode(kernel):
    k = kernel if isinstance(kernel, type) else type(kernel)
    if k in _SHAPES:
        return k.code
    raise TypeError(f"unsupported kernel {kernel!r}: CloudInCell(), M6(), NearestGridPoint(), "
                    "TriangularShapeCloud() or M4()")


# --------------------------------------------------------- device context




class Context:
    """One libpicmi355x context: mesh, particles and grids on one GPU."""

    def __init__(self, mesh, capacity=0, device=0):
        self.mesh = mesh
        h = C.c_void_p()
        L.call("pic_ctx_create", C.byref(h), device, mesh.nx, mesh.ny, mesh.xmin, mesh.xmax,
               mesh.ymin, mesh.ymax, int(capacity))
        self.h = h

    def close(self):
        if self.h:
            L.load().pic_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # particles
    def set_particles(self, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        assert a.ndim == 2 and a.shape[0] == 5
        L.call("pic_set_particles", self.h, a.ctypes.data, a.shape[1])
        return a

    def get_particles(self, out):
        assert out.flags.f_contiguous and out.dtype == np.float64
        L.call("pic_get_particles", self.h, out.ctypes.data)
        return out

    def num_particles(self):
        n = C.c_int64()
        L.call("pic_num_particles", self.h, C.byref(n))
        return n.value

    def set_state(self, state):
        L.call("pic_set_state", self.h, state)

    def get_state_f32(self):
        n = self.num_particles()
        cell = np.zeros(n, np.int32)
        ox, oy, vx, vy = (np.zeros(n, np.float32) for _ in range(4))
        L.call("pic_get_state_f32", self.h, cell.ctypes.data, ox.ctypes.data, oy.ctypes.data,
               vx.ctypes.data, vy.ctypes.data)
        return cell, ox, oy, vx, vy

    def set_state_f32(self, state, w):
        cell, ox, oy, vx, vy = (np.ascontiguousarray(a) for a in state)
        assert cell.dtype == np.int32 and ox.dtype == np.float32
        L.call("pic_set_state_f32", self.h, cell.ctypes.data, ox.ctypes.data, oy.ctypes.data,
               vx.ctypes.data, vy.ctypes.data, float(w), cell.size)

    # grids
    def set_grid(self, which, arr, stride=1):
        a = np.asarray(arr, dtype=np.float64)
        if stride == 1:
            a = np.asfortranarray(a)
            ld = a.shape[0]
        else:
            ld = a.shape[1]
        L.call("pic_set_grid", self.h, which, a.ctypes.data, ld, stride)

    def get_grid(self, which, ghost=False):
        nx, ny = self.mesh.nx, self.mesh.ny
        out = np.zeros((nx + 1, ny + 1) if ghost else (nx, ny), order="F")
        L.call("pic_get_grid", self.h, which, out.ctypes.data, out.shape[0], 1, 1 if ghost else 0)
        return out

    def set_bz(self, mode, bz0=0.0):
        L.call("pic_set_bz", self.h, mode, float(bz0))

    def sync(self):
        L.call("pic_synchronize", self.h)


_ctx_cache = {}


def _context(mesh):
    """A cached context per mesh for the per-call API (plans are reused)."""
    k = mesh.key()
    c = _ctx_cache.get(k)
    if c is None:
        c = Context(mesh)
        _ctx_cache[k] = c
    return c


def _fields_in(ctx, ex, ey, bz):
    ctx.set_grid(L.PIC_GRID_EX, ex)
    ctx.set_grid(L.PIC_GRID_EY, ey)
    if bz is None:
        ctx.set_bz(L.PIC_BZ_NONE)
    else:
        b = np.asarray(bz, dtype=np.float64)
        if b.ndim == 0:
            ctx.set_bz(L.PIC_BZ_UNIFORM, float(b))
        else:
            ctx.set_grid(L.PIC_GRID_BZ, b)
            ctx.set_bz(L.PIC_BZ_GRID)


def _check_grid(a, mesh, name):
    a = np.asarray(a)
    if a.shape[0] < mesh.nx or a.shape[1] < mesh.ny:
        raise ValueError(f"{name} has shape {a.shape}, the mesh needs at least "
                         f"({mesh.nx}, {mesh.ny})")


# ------------------------------------------------------ the reference calls

def compute_rho(p, kernel, mesh):
    """compute_rho(p, kernel, mesh): returns the nx x ny charge density,
    normalised and with its mean removed (pushers_cic.jl:75-114,
    pushers_m6.jl:226-348)."""
    ctx = _context(mesh)
    ctx.set_particles(p.array)
    L.call("pic_compute_rho", ctx.h, _kcode(kernel))
    return ctx.get_grid(L.PIC_GRID_RHO)


def push_v_(p, kernel, mesh, ex, ey, bz, dt):
    """push_v!(p, kernel, mesh, ex, ey, bz, dt), in place on p.array.
    bz may be an array, a scalar (uniform field) or None (no field)."""
    for a, nm in ((ex, "ex"), (ey, "ey")):
        _check_grid(a, mesh, nm)
    ctx = _context(mesh)
    ctx.set_particles(p.array)
    _fields_in(ctx, ex, ey, bz)
    L.call("pic_push_v", ctx.h, _kcode(kernel), float(dt))
    ctx.get_particles(p.array)


def push_x_(p, mesh, dt):
    """push_x!(p, mesh, dt), in place on p.array."""
    ctx = _context(mesh)
    ctx.set_particles(p.array)
    L.call("pic_push_x", ctx.h, float(dt))
    ctx.get_particles(p.array)


def operator_t_(pg, mesh, dt):
    """operator_t!(split, dt) (poisson_2d.jl:265-277) on the splitting
    operator's particle group: x += dt v, no wrap, in place on pg.array."""
    ctx = _context(mesh)
    ctx.set_particles(pg.array)
    L.call("pic_operator_t", ctx.h, float(dt))
    ctx.get_particles(pg.array)


def interpolate_cic_(epx, epy, mesh, p, ex, ey):
    """interpolate_cic!(epx, epy, mesh, p, ex, ey), fills epx, epy."""
    ctx = _context(mesh)
    ctx.set_particles(p.array)
    _fields_in(ctx, ex, ey, None)
    n = p.array.shape[1]
    if epx.shape[0] < n or epy.shape[0] < n:
        raise ValueError("epx/epy shorter than the particle count")
    ax = np.zeros(n)
    ay = np.zeros(n)
    L.call("pic_interpolate_cic", ctx.h, ax.ctypes.data, ay.ctypes.data)
    epx[:n] = ax
    epy[:n] = ay


class TwoDPoissonPeriodic:
    """TwoDPoissonPeriodic(grid) (poisson_2d.jl:28-61): the tables live in the
    device context of the grid."""

    def __init__(self, grid):
        self.grid = grid


class Poisson2D:
    """Poisson2D(mesh) (meshfields.jl:27-73)."""

    def __init__(self, mesh):
        self.mesh = mesh


class MeshFields2D:
    """MeshFields2D(mesh): e is 2 x (nx+1) x (ny+1), rho (nx+1) x (ny+1)."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.e = np.zeros((2, mesh.nx + 1, mesh.ny + 1), order="F")
        self.rho = np.zeros((mesh.nx + 1, mesh.ny + 1), order="F")


class Particles:
    """Particles(nbpart, w) (particles.jl:12-35): x, v, e are 2 x n."""

    def __init__(self, nbpart, w):
        self.nbpart = int(nbpart)
        self.x = np.zeros((2, self.nbpart), order="F")
        self.v = np.zeros((2, self.nbpart), order="F")
        self.e = np.zeros((2, self.nbpart), order="F")
        self.b = np.zeros(self.nbpart)
        self.t = np.zeros(self.nbpart)
        self.w = float(w)


def solve_(*args):
    """solve!(phi, poisson, rho); solve!(ex, ey, poisson, rho);
    solve!(fields, poisson::Poisson2D) -> field energy."""
    if len(args) == 2 and isinstance(args[1], Poisson2D):
        fields, ps = args
        ctx = _context(ps.mesh)
        ctx.set_grid(L.PIC_GRID_RHO, fields.rho)
        en = C.c_double()
        L.call("pic_solve_fields", ctx.h, C.byref(en))
        fields.e[0] = ctx.get_grid(L.PIC_GRID_EX, ghost=True)
        fields.e[1] = ctx.get_grid(L.PIC_GRID_EY, ghost=True)
        return en.value
    if len(args) == 3:
        phi, ps, rho = args
        ctx = _context(ps.grid)
        ctx.set_grid(L.PIC_GRID_RHO, rho)
        L.call("pic_solve_phi", ctx.h)
        phi[...] = ctx.get_grid(L.PIC_GRID_PHI)
        return phi
    if len(args) == 4:
        ex, ey, ps, rho = args
        ctx = _context(ps.grid)
        ctx.set_grid(L.PIC_GRID_RHO, rho)
        L.call("pic_solve_e", ctx.h)
        ex[...] = ctx.get_grid(L.PIC_GRID_EX)
        ey[...] = ctx.get_grid(L.PIC_GRID_EY)
        return ex, ey
    raise TypeError("solve_(phi, poisson, rho) | solve_(ex, ey, poisson, rho) | "
                    "solve_(fields, poisson2d)")


def _set_xv(ctx, particles):
    x = np.asfortranarray(particles.x)
    v = np.asfortranarray(particles.v)
    L.call("pic_set_particles_xv", ctx.h, x.ctypes.data, v.ctypes.data, particles.w,
           particles.nbpart)


def compute_rho_m6_(fields, particles, compat=True):
    """compute_rho_m6!(fields, particles) (compute_rho_m6.jl:181-316).
    compat=True reproduces the reference's swapped bounds and neighbour
    indices; compat=False the corrected stencil (see the C ABI header)."""
    ctx = _context(fields.mesh)
    _set_xv(ctx, particles)
    L.call("pic_compute_rho_m6_fields", ctx.h, 1 if compat else 0)
    fields.rho[...] = ctx.get_grid(L.PIC_GRID_RHO, ghost=True)
    x = np.zeros((2, particles.nbpart), order="F")
    L.call("pic_get_particles_xv", ctx.h, x.ctypes.data, None)
    particles.x[...] = x


def interpol_eb_m6_(particles, fields):
    """interpol_eb_m6!(particles, fields) (interpolation_m6.jl:125-247)."""
    ctx = _context(fields.mesh)
    _set_xv(ctx, particles)
    ctx.set_grid(L.PIC_GRID_EX, fields.e[0])
    ctx.set_grid(L.PIC_GRID_EY, fields.e[1])
    L.call("pic_interpol_eb_m6", ctx.h)
    e = np.zeros((2, particles.nbpart), order="F")
    L.call("pic_get_particle_fields", ctx.h, e.ctypes.data)
    particles.e[...] = e
    x = np.zeros((2, particles.nbpart), order="F")
    L.call("pic_get_particles_xv", ctx.h, x.ctypes.data, None)
    particles.x[...] = x


class F90:
    """ParticleInCell.F90 (src/pic.jl): the Fortran twin's entry points with its
    exact argument conventions, served by pic_f90_* of the HIP library."""

    @staticmethod
    def push_v_(p, kernel, m, ex, ey, bz, dt):
        _kcode(kernel)
        a = p.array
        n = C.c_int32(a.shape[1])
        ex, ey, bz = (np.asfortranarray(g, dtype=np.float64) for g in (ex, ey, bz))
        L.call("pic_f90_push_v", C.byref(n), a.ctypes.data, C.byref(C.c_double(dt)),
               C.byref(C.c_int32(m.nx)), C.byref(C.c_int32(m.ny)), C.byref(C.c_double(m.dx)),
               C.byref(C.c_double(m.dy)), ex.ctypes.data, ey.ctypes.data, bz.ctypes.data)

    @staticmethod
    def push_x_(p, mesh, dt):
        a = p.array
        L.call("pic_f90_push_x", C.byref(C.c_int32(a.shape[1])), C.byref(C.c_double(mesh.dimx)),
               C.byref(C.c_double(mesh.dimy)), a.ctypes.data, C.byref(C.c_double(dt)))

    @staticmethod
    def compute_current_(jx, jy, m, kernel, p):
        """F90.compute_current! (pic.jl:18-58): deposition, then the Julia-side
        fold of the jx ghosts (pic.jl:49-56; jy is not folded there)."""
        _kcode(kernel)
        a = p.array
        assert jx.flags.f_contiguous and jy.flags.f_contiguous
        L.call("pic_f90_deposition", C.byref(C.c_int32(a.shape[1])), C.byref(C.c_int32(m.nx)),
               C.byref(C.c_int32(m.ny)), C.byref(C.c_double(m.dx)), C.byref(C.c_double(m.dy)),
               a.ctypes.data, jx.ctypes.data, jy.ctypes.data)
        nx, ny = m.nx, m.ny
        for i in range(nx + 1):
            jx[i, 0] += jx[i, ny]
            jx[i, ny] = jx[i, 0]
        for j in range(ny + 1):
            jx[0, j] += jx[nx, j]
            jx[nx, j] = jx[0, j]


# ------------------------------------------------------- Vlasov-Maxwell

class FDTD:
    """FDTD(mesh) (fdtd.jl:3-20): the Yee-grid fields ex nx x (ny+1),
    ey (nx+1) x ny, bz nx x ny, float64 in Fortran order, zero."""

    def __init__(self, mesh):
        nx, ny = mesh.nx, mesh.ny
        self.ex = np.zeros((nx, ny + 1), order="F")
        self.ey = np.zeros((nx + 1, ny), order="F")
        self.bz = np.zeros((nx, ny), order="F")

    def _check(self, mesh):
        nx, ny = mesh.nx, mesh.ny
        for a, shp, nm in ((self.ex, (nx, ny + 1), "ex"), (self.ey, (nx + 1, ny), "ey"),
                           (self.bz, (nx, ny), "bz")):
            if a.shape != shp or a.dtype != np.float64 or not a.flags.f_contiguous:
                raise ValueError(f"fdtd.{nm} must be a float64 Fortran array of shape {shp}")


def _fdtd_in(ctx, fdtd, mesh):
    fdtd._check(mesh)
    L.call("pic_fdtd_set", ctx.h, fdtd.ex.ctypes.data, fdtd.ey.ctypes.data, fdtd.bz.ctypes.data)


def _fdtd_out(ctx, fdtd):
    L.call("pic_fdtd_get", ctx.h, fdtd.ex.ctypes.data, fdtd.ey.ctypes.data, fdtd.bz.ctypes.data)


def _node_array(a, mesh, name):
    if (a.shape != (mesh.nx + 1, mesh.ny + 1) or a.dtype != np.float64
            or not a.flags.f_contiguous):
        raise ValueError(f"{name} must be a float64 Fortran array of shape "
                         f"({mesh.nx + 1}, {mesh.ny + 1})")


def faraday_(fdtd, mesh, dt):
    """faraday!(fdtd, m, dt) (fdtd.jl:24-36), in place on fdtd.bz."""
    ctx = _context(mesh)
    _fdtd_in(ctx, fdtd, mesh)
    L.call("pic_faraday", ctx.h, float(dt))
    _fdtd_out(ctx, fdtd)


def ampere_maxwell_(fdtd, mesh, jx, jy, dt):
    """ampere_maxwell!(fdtd, m, jx, jy, dt) (fdtd.jl:39-54), in place on
    fdtd.ex, fdtd.ey; jx, jy are (nx+1) x (ny+1) node currents."""
    _node_array(jx, mesh, "jx")
    _node_array(jy, mesh, "jy")
    ctx = _context(mesh)
    _fdtd_in(ctx, fdtd, mesh)
    L.call("pic_set_current", ctx.h, jx.ctypes.data, jy.ctypes.data)
    L.call("pic_ampere_maxwell", ctx.h, float(dt))
    _fdtd_out(ctx, fdtd)


def update_fields_(ex, ey, bz, mesh, fdtd):
    """update_fields!(ex, ey, bz, m, fdtd) (fdtd.jl:58-77): fills the
    (nx+1) x (ny+1) node arrays; the ghost row and column are the periodic
    images of the first."""
    for a, nm in ((ex, "ex"), (ey, "ey"), (bz, "bz")):
        _node_array(a, mesh, nm)
    ctx = _context(mesh)
    _fdtd_in(ctx, fdtd, mesh)
    L.call("pic_update_fields", ctx.h)
    for a, which in ((ex, L.PIC_GRID_EX), (ey, L.PIC_GRID_EY), (bz, L.PIC_GRID_BZ)):
        a[...] = ctx.get_grid(which, ghost=True)


def compute_current_(jx, jy, mesh, kernel, p):
    """compute_current!(jx, jy, m, ::CloudInCell, p) (cloud_in_cell.jl:10-81):
    fills the (nx+1) x (ny+1) arrays, ghosts folded."""
    _node_array(jx, mesh, "jx")
    _node_array(jy, mesh, "jy")
    ctx = _context(mesh)
    ctx.set_particles(p.array if isinstance(p, ParticleGroup) else p)
    L.call("pic_compute_current", ctx.h, _kcode(kernel), jx.ctypes.data, jy.ctypes.data)


def compute_energy(fdtd, mesh):
    """compute_energy(fdtd, m) = 0.5 log(sum(fdtd.ex .^ 2) dx dy) (fdtd.jl:79-80)."""
    ctx = _context(mesh)
    _fdtd_in(ctx, fdtd, mesh)
    e = C.c_double()
    L.call("pic_fdtd_energy", ctx.h, C.byref(e))
    return e.value


# ---------------------------------------------------------- particle files

def read_particles(filename, mesh):
    """read_particles(filename, mesh) (read_particles.jl:3-35): Particles
    with x = (dpx + ix) dx, (dpy + iy) dy, v and w = dimx dimy / n."""
    ctx = Context(mesh)
    try:
        n = C.c_int64()
        L.call("pic_read_particles", ctx.h, os.fsencode(filename), L.PIC_STATE_F64, C.byref(n))
        p = Particles(n.value, (mesh.dimx * mesh.dimy) / n.value if n.value else 0.0)
        if n.value:
            x = np.zeros((2, n.value), order="F")
            v = np.zeros((2, n.value), order="F")
            L.call("pic_get_particles_xv", ctx.h, x.ctypes.data, v.ctypes.data)
            p.x[...], p.v[...] = x, v
        return p
    finally:
        ctx.close()


def write_particles(filename, particles, mesh):
    """Writes Particles in read_particles' format (no reference counterpart)."""
    ctx = Context(mesh, capacity=particles.nbpart)
    try:
        _set_xv(ctx, particles)
        L.call("pic_write_particles", ctx.h, os.fsencode(filename))
    finally:
        ctx.close()


# ------------------------------------------------------------------- UA

class UA:
    """UA(ntau, eps, nbpart) (ua_type.jl:16-41): tau, ltau, and the pl, ql
    tables ((ntau, nbpart) complex) that preparation_ fills."""

    def __init__(self, ntau, eps, nbpart):
        self.ntau, self.eps, self.nbpart = int(ntau), float(eps), int(nbpart)
        dtau = 2 * math.pi / self.ntau
        self.ltau = np.array(list(range(self.ntau // 2)) + list(range(-(self.ntau // 2), 0)),
                             dtype=np.float64)
        self.tau = np.array([i * dtau for i in range(self.ntau)])
        self.pl = np.zeros((self.ntau, self.nbpart), np.complex128, order="F")
        self.ql = np.zeros((self.ntau, self.nbpart), np.complex128, order="F")


# the context of the UA calls that take no mesh (their kernels use none)
_PARTICLE_MESH = TwoDGrid(0.0, 1.0, 8, 0.0, 1.0, 8)


def _ua_context(ua, particles, mesh=None):
    ctx = _context(mesh if mesh is not None else _PARTICLE_MESH)
    _set_xv(ctx, particles)
    L.call("pic_set_particle_fields", ctx.h, np.asfortranarray(particles.e).ctypes.data)
    L.call("pic_ua_init", ctx.h, ua.ntau, ua.eps)
    return ctx


def _ua_array(a, ua, n, name, dtype=np.complex128):
    if a.shape != (ua.ntau, 2, n) or a.dtype != dtype or not a.flags.f_contiguous:
        raise ValueError(f"{name} must be a {np.dtype(dtype).name} Fortran array of shape "
                         f"({ua.ntau}, 2, {n})")
    return a.ctypes.data


def preparation_(ua, dt, particles, xt, yt):
    """preparation!(ua, dt, particles, xt, yt) (ua_steps.jl:3-83): fills
    particles.b, particles.t, ua.pl, ua.ql, xt and yt."""
    n = particles.nbpart
    ctx = _ua_context(ua, particles)
    b, t = np.zeros(n), np.zeros(n)
    L.call("pic_ua_preparation", ctx.h, float(dt), b.ctypes.data, t.ctypes.data,
           ua.pl.ctypes.data, ua.ql.ctypes.data, _ua_array(xt, ua, n, "xt"),
           _ua_array(yt, ua, n, "yt"))
    particles.b[...] = b
    particles.t[...] = t


def update_particles_e_(particles, et, fields, ua, xt):
    """update_particles_e!(particles, et, fields, ua, xt) (ua_steps.jl:87-93):
    the M6 gather of fields.e at every real(xt[m, :, k])."""
    n = particles.nbpart
    ctx = _ua_context(ua, particles, fields.mesh)
    ctx.set_grid(L.PIC_GRID_EX, fields.e[0])
    ctx.set_grid(L.PIC_GRID_EY, fields.e[1])
    L.call("pic_ua_interpol", ctx.h, _ua_array(xt, ua, n, "xt"),
           _ua_array(et, ua, n, "et", np.float64))


def compute_f_(fx, fy, ua, particles, xt, yt, et):
    """compute_f!(fx, fy, ua, particles, xt, yt, et) (ua_steps.jl:107-148)."""
    n = particles.nbpart
    ctx = _ua_context(ua, particles)
    b = np.ascontiguousarray(particles.b, dtype=np.float64)
    L.call("pic_ua_compute_f", ctx.h, b.ctypes.data, _ua_array(xt, ua, n, "xt"),
           _ua_array(yt, ua, n, "yt"), _ua_array(et, ua, n, "et", np.float64),
           _ua_array(fx, ua, n, "fx"), _ua_array(fy, ua, n, "fy"))


def ua_step_(xt, xtilde, ua, particles, fx, gx=None):
    """ua_step!(xt, x~t, ua, particles, fx[, gx]) (ua_steps.jl:152-202)."""
    n = particles.nbpart
    ctx = _ua_context(ua, particles)
    t = np.ascontiguousarray(particles.t, dtype=np.float64)
    L.call("pic_ua_step", ctx.h, t.ctypes.data, _ua_array(xtilde, ua, n, "xtilde"),
           _ua_array(fx, ua, n, "fx"), None if gx is None else _ua_array(gx, ua, n, "gx"),
           _ua_array(xt, ua, n, "xt"))


def update_particles_x_(particles, fields, ua, xt):
    """update_particles_x!(particles, fields, ua, xt) = compute_rho_m6!(fields,
    particles, xt, ua) (compute_rho_m6.jl:29-179): particles.x at t and
    fields.rho."""
    n = particles.nbpart
    ctx = _ua_context(ua, particles, fields.mesh)
    t = np.ascontiguousarray(particles.t, dtype=np.float64)
    L.call("pic_ua_compute_rho", ctx.h, t.ctypes.data, _ua_array(xt, ua, n, "xt"))
    fields.rho[...] = ctx.get_grid(L.PIC_GRID_RHO, ghost=True)
    x = np.zeros((2, n), order="F")
    L.call("pic_get_particles_xv", ctx.h, x.ctypes.data, None)
    particles.x[...] = x


def compute_v_(yt, particles, ua):
    """compute_v!(yt, particles, ua) (ua_steps.jl:206-224): particles.v."""
    n = particles.nbpart
    ctx = _ua_context(ua, particles)
    t = np.ascontiguousarray(particles.t, dtype=np.float64)
    L.call("pic_ua_compute_v", ctx.h, t.ctypes.data, _ua_array(yt, ua, n, "yt"))
    v = np.zeros((2, n), order="F")
    L.call("pic_get_particles_xv", ctx.h, None, v.ctypes.data)
    particles.v[...] = v


class UASimulation:
    """The UA loop of test/bupdate.jl on device-resident Particles: the
    prologue (compute_rho_m6!, solve!, interpol_eb_m6!, :56-60) at
    construction, then run(dt, nsteps) (pic_ua_run)."""

    def __init__(self, mesh, particles, ntau, eps, device=0):
        self.mesh, self.n = mesh, particles.nbpart
        self.ctx = Context(mesh, capacity=self.n, device=device)
        _set_xv(self.ctx, particles)
        L.call("pic_compute_rho_m6_fields", self.ctx.h, 1)
        en = C.c_double()
        L.call("pic_solve_fields", self.ctx.h, C.byref(en))
        L.call("pic_interpol_eb_m6", self.ctx.h)
        L.call("pic_ua_init", self.ctx.h, int(ntau), float(eps))
        self.energy0 = en.value

    def run(self, dt, nsteps):
        """Returns the solve! energies, (2, nsteps): prediction, correction."""
        e = np.zeros((2, nsteps), order="F")
        L.call("pic_ua_run", self.ctx.h, float(dt), int(nsteps), e.ctypes.data)
        return e

    def set_timing(self, on=True, every=1):
        """HIP events around the loop's kernels, at every `every`-th step
        (pic_set_timing)."""
        L.call("pic_set_timing", self.ctx.h, (max(1, int(every)) if on else 0))

    def timing(self):
        """{kernel: (total_ms, launches)} of the prediction and correction
        kernels since timing was enabled (pic_get_ua_timing)."""
        ms = (C.c_double * 2)()
        cnt = (C.c_int64 * 2)()
        L.call("pic_get_ua_timing", self.ctx.h, ms, cnt)
        return {"predict": (ms[0], cnt[0]), "correct": (ms[1], cnt[1])}

    def particles(self):
        """(x, v), each 2 x n."""
        x, v = np.zeros((2, self.n), order="F"), np.zeros((2, self.n), order="F")
        L.call("pic_get_particles_xv", self.ctx.h, x.ctypes.data, v.ctypes.data)
        return x, v

    def fields_e(self):
        """MeshFields2D.e (2, nx+1, ny+1) of the last solve."""
        e = np.zeros((2, self.mesh.nx + 1, self.mesh.ny + 1), order="F")
        e[0] = self.ctx.get_grid(L.PIC_GRID_EX, ghost=True)
        e[1] = self.ctx.get_grid(L.PIC_GRID_EY, ghost=True)
        return e

    def close(self):
        self.ctx.close()


# ------------------------------------------------- device-resident run

class Simulation:
    """Particles and fields resident on the GPU for the hot-path time loop
    (examples/landau_damping_poisson_cic_m6.jl:70-83)."""

    def __init__(self, mesh, p, kernel, state=L.PIC_STATE_F64, bz=None, device=0):
        self.mesh = mesh
        self.kernel = _kcode(kernel)
        if p is None:  # the classmethods below fill the particles
            self.ctx = Context(mesh, capacity=0, device=device)
            self.ctx.set_bz(L.PIC_BZ_NONE)
            self.n = 0
            return
        arr = p.array if isinstance(p, ParticleGroup) else p
        n = arr.shape[1]
        self.ctx = Context(mesh, capacity=n, device=device)
        self.ctx.set_particles(arr)
        self._set_bz(bz)
        self.ctx.set_state(state)
        self.n = n

    def _set_bz(self, bz):
        if bz is None:
            self.ctx.set_bz(L.PIC_BZ_NONE)
        elif np.ndim(bz) == 0:
            self.ctx.set_bz(L.PIC_BZ_UNIFORM, float(bz))
        else:
            self.ctx.set_grid(L.PIC_GRID_BZ, bz)
            self.ctx.set_bz(L.PIC_BZ_GRID)

    @classmethod
    def landau(cls, mesh, n, alpha, kx, kernel, state=L.PIC_STATE_F64, device=0):
        """The quiet start of sample!(pg, mesh, LandauDamping(alpha, kx))
        (landau_damping.jl:77-97) made on the device (pic_sample_landau)."""
        sim = cls(mesh, None, kernel, device=device)
        sim.sample_landau(n, alpha, kx, state)
        return sim

    @classmethod
    def maxwellian(cls, mesh, n, alpha, kx, ky, kernel, seed=0, vth=1.0,
                   state=L.PIC_STATE_F64, bz=None, device=0):
        """n particles of density 1 + alpha cos(kx x) cos(ky y) with
        Maxwellian velocities of width vth, made on the device from a seed
        (pic_sample_maxwellian); bz as in Simulation()."""
        sim = cls(mesh, None, kernel, device=device)
        sim._set_bz(bz)
        sim.sample_maxwellian(n, alpha, kx, ky, seed, vth, state)
        return sim

    def sample_maxwellian(self, n, alpha, kx, ky, seed=0, vth=1.0, state=L.PIC_STATE_F64):
        """Replaces the particles by the device-made sample of
        Simulation.maxwellian (pic_sample_maxwellian)."""
        L.call("pic_sample_maxwellian", self.ctx.h, int(n), float(alpha), float(kx), float(ky),
               float(vth), int(seed), int(state))
        self.n = int(n)

    def sample_landau(self, n, alpha, kx, state=L.PIC_STATE_F64):
        """Replaces the particles by the device-made quiet start of
        Simulation.landau (pic_sample_landau)."""
        L.call("pic_sample_landau", self.ctx.h, int(n), float(alpha), float(kx), int(state))
        self.n = int(n)

    @classmethod
    def from_file(cls, mesh, path, kernel, state=L.PIC_STATE_F64, device=0):
        """Particles of a read_particles file (read_particles.jl:3-35),
        parsed by the library straight into `state` (pic_read_particles)."""
        sim = cls(mesh, None, kernel, device=device)
        n = C.c_int64()
        L.call("pic_read_particles", sim.ctx.h, os.fsencode(path), int(state), C.byref(n))
        sim.n = n.value
        return sim

    def write_particles(self, path):
        """The particles in read_particles' format (pic_write_particles)."""
        L.call("pic_write_particles", self.ctx.h, os.fsencode(path))

    def sort(self, lanes=False):
        """Stable sort by cell (pic_sort), or that followed by the lane-spread
        order of the fused kernels (pic_sort_lanes); returns the permutation."""
        perm = np.zeros(self.n, np.int32)
        L.call("pic_sort_lanes" if lanes else "pic_sort", self.ctx.h, perm.ctypes.data)
        return perm

    def cell_offsets(self):
        off = np.zeros(self.mesh.nx * self.mesh.ny + 1, np.int64)
        L.call("pic_cell_offsets", self.ctx.h, off.ctypes.data)
        return off

    def run(self, dt, nsteps, sort_every=0):
        """Runs nsteps loop iterations; returns (sum ex^2 dxdy, sum ey^2 dxdy)
        per step as an (nsteps, 2) array."""
        e = np.zeros((nsteps, 2))
        L.call("pic_run", self.ctx.h, self.kernel, float(dt), int(nsteps), int(sort_every),
               e.ctypes.data)
        return e

    def particles(self):
        out = np.zeros((5, self.n), order="F")
        return self.ctx.get_particles(out)

    def state_f32(self):
        return self.ctx.get_state_f32()

    def grid(self, which):
        return self.ctx.get_grid(which)

    def adjoint(self, dt, nsteps, ckpt, sort_every=0):
        """(J, gx, gv): J = sum(ex^2+ey^2) dx dy of the last step's field after
        nsteps loop iterations, gx = dJ/d(x0, y0), gv = dJ/d(vx0, vy0), each
        2 x n in the particle order of the context (pic_adjoint).  In the FP32
        state the sweeps also lane-sort every sort_every steps."""
        J = C.c_double()
        gx = np.zeros((2, self.n), order="F")
        gv = np.zeros((2, self.n), order="F")
        L.call("pic_adjoint", self.ctx.h, self.kernel, float(dt), int(nsteps), int(ckpt),
               int(sort_every), C.byref(J), gx.ctypes.data, gv.ctypes.data)
        return J.value, gx, gv

    def step_pass(self, mode, dt):
        """One particle pass of the loop (pic_step_pass): 0 drift dt/2 +
        deposit, 1 kick + drift + drift + deposit, 2 kick + drift dt/2,
        3 deposit only.  The raw deposit stays on the device."""
        L.call("pic_step_pass", self.ctx.h, self.kernel, int(mode), float(dt), None)

    def set_adjoint_block(self, block):
        """Steps whose positions the FP32 adjoint sweep keeps at once (0: the
        most that fits the free device memory) (pic_set_adjoint_block)."""
        L.call("pic_set_adjoint_block", self.ctx.h, int(block))

    def adjoint_plan(self):
        """(block, scratch bytes) of the last adjoint sweep (pic_get_adjoint_plan)."""
        b, nb = C.c_int(), C.c_int64()
        L.call("pic_get_adjoint_plan", self.ctx.h, C.byref(b), C.byref(nb))
        return b.value, nb.value

    def adjoint_timing(self):
        """{pass: (total_ms, count)} of the last adjoint sweeps since timing
        was enabled, for pass in forward, gather_adjoint, deposit_adjoint,
        recompute (pic_get_adjoint_timing)."""
        ms = (C.c_double * 4)()
        cnt = (C.c_int64 * 4)()
        L.call("pic_get_adjoint_timing", self.ctx.h, ms, cnt)
        names = ("forward", "gather_adjoint", "deposit_adjoint", "recompute")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(names)}

    def set_fdtd(self, fdtd):
        """Loads the FDTD fields (fdtd.jl:3-20) of the Vlasov-Maxwell loop."""
        _fdtd_in(self.ctx, fdtd, self.mesh)

    def fdtd(self):
        """The resident FDTD fields, as an FDTD."""
        f = FDTD(self.mesh)
        _fdtd_out(self.ctx, f)
        return f

    def set_maxwell_path(self, path):
        """0: the most fused launch form the grid allows; 2: at most two
        launches per iteration; 3: one kernel per reference function
        (pic_set_maxwell_path)."""
        L.call("pic_set_maxwell_path", self.ctx.h, int(path))

    def run_maxwell(self, dt, nsteps):
        """nsteps iterations of the Vlasov-Maxwell loop
        (examples/landau_damping_maxwell_fdtd.jl:58-83, pic_run_maxwell);
        returns compute_energy after each iteration."""
        e = np.zeros(nsteps)
        L.call("pic_run_maxwell", self.ctx.h, self.kernel, float(dt), int(nsteps), e.ctypes.data)
        return e

    def set_timing(self, on=True, every=1):
        """HIP events around the loop's kernels, at every `every`-th step
        (pic_set_timing)."""
        L.call("pic_set_timing", self.ctx.h, (max(1, int(every)) if on else 0))

    def set_fused_solve(self, on=True):
        """E solve as three kernels (power-of-two grids up to 512: transforms
        across the lanes of a wave from 64 nodes a side, in LDS below) or as
        rocFFT transforms (pic_set_fused_solve)."""
        L.call("pic_set_fused_solve", self.ctx.h, 1 if on else 0)

    def set_fused_step(self, on=True):
        """pic_run on power-of-two grids of at most 64 x 64 nodes as one
        launch per iteration (True, the default) or as the separate solve,
        pass and sort kernels (pic_set_fused_step)."""
        L.call("pic_set_fused_step", self.ctx.h, 1 if on else 0)

    def set_f32_accumulation(self, mode):
        """Charge accumulation of the FP32 state: 0 = compensated FP32
        (hi, lo) tile pairs, 1 = FP64 tile sums, 2 = int32 fixed-point tile
        sums (pic_set_f32_accumulation)."""
        L.call("pic_set_f32_accumulation", self.ctx.h, int(mode))

    def set_f32_tiles(self, tiles):
        """Copies (1, 2 or 4, default 4) of the fixed-point charge tile in the
        FP32 drift-and-deposit passes (pic_set_f32_tiles)."""
        L.call("pic_set_f32_tiles", self.ctx.h, int(tiles))

    def set_f32_window(self, kib):
        """LDS window per block (KiB, 40 to 160, default 160) of the M6, TSC and
        M4 FP32 drift-and-deposit passes (pic_set_f32_window)."""
        L.call("pic_set_f32_window", self.ctx.h, int(kib))

    def timing(self):
        """(pass_ms_total, passes, solve_ms_total) of pic_run since timing
        was enabled (HIP events on the context stream)."""
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        L.call("pic_get_timing", self.ctx.h, C.byref(a), C.byref(n), C.byref(b))
        return a.value, n.value, b.value

    def set_particles(self, p, state=L.PIC_STATE_F64):
        arr = p.array if isinstance(p, ParticleGroup) else p
        self.ctx.set_particles(arr)
        self.ctx.set_state(state)
        self.n = arr.shape[1]

    def slow_chunks(self):
        v = C.c_uint64()
        L.call("pic_slow_chunks", self.ctx.h, C.byref(v))
        return v.value

    def close(self):
        self.ctx.close()


def energy(ex, mesh):
    """sqrt(sum(ex.^2) * dx * dy) (examples/landau_damping_poisson_cic_m6.jl:65)."""
    return math.sqrt(float(np.sum(np.asarray(ex) ** 2)) * mesh.dx * mesh.dy)
