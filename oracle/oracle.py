"""ctypes front end of the CPU oracle (liboracle.so) and of the reference's
Fortran twin (oracle/_ref/libpic.so, built from src/pic.f90).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The HIP library never calls it.

Arrays follow the reference's Julia layouts, held in numpy with Fortran order:
a particle group is ``p.array`` of shape (5, n) (rows x, y, vx, vy, w), grids
are (nx, ny).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpic.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_dpc = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")
_i64, _i32, _d, _f = C.c_int64, C.c_int, C.c_double, C.c_float


def build():
    """Compile liboracle.so (and _ref/libpic.so where the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_f_m6": (_d, [_d]),
            "or_push_v_cic": (None, [_i64, _dp, _d, _i32, _i32, _d, _d, _dp, _dp, C.c_void_p, _i32]),
            "or_push_x": (None, [_i64, _dp, _d, _d, _d]),
            "or_compute_rho_cic": (None, [_i64, _dp, _i32, _i32, _d, _d, _dp]),
            "or_interpolate_cic": (None, [_i64, _dp, _i32, _i32, _d, _d, _dp, _dp, _dpc, _dpc]),
            "or_push_v_m6": (None, [_i64, _dp, _d, _i32, _i32, _d, _d, _dp, _dp, C.c_void_p, _i32]),
            "or_compute_rho_m6": (None, [_i64, _dp, _i32, _i32, _d, _d, _d, _d, _dp]),
            "or_faraday": (None, [_i32, _i32, _d, _d, _dp, _dp, _dp, _d]),
            "or_ampere_maxwell": (None, [_i32, _i32, _d, _d, _dp, _dp, _dp, _dp, _dp, _d]),
            "or_update_fields": (None, [_i32, _i32, _dp, _dp, _dp, _dp, _dp, _dp]),
            "or_compute_current": (None, [_i64, _dp, _i32, _i32, _d, _d, _dp, _dp]),
            "or_fdtd_energy": (_d, [_i32, _i32, _d, _d, _dp]),
            "or_maxwell_loop": (_i32, [_i64, _dp, _i32, _i32, _d, _d, _d, _i32, _dp, _dp, _dp,
                                       _dpc]),
            "or_fft_c": (None, [_cp, _i32, _i32]),
            "or_ua_preparation": (None, [_i32, _d, _d, _i64, _dp, _dp, _dp, _dp, _dp, _cp, _cp,
                                         _cp, _cp]),
            "or_ua_interpol": (None, [_i32, _i64, _cp, _i32, _i32, _d, _d, _d, _d, _dp, _dp]),
            "or_ua_compute_f": (None, [_i32, _d, _i64, _dp, _cp, _cp, _dp, _cp, _cp]),
            "or_ua_step": (None, [_i32, _d, _i64, _dp, _cp, C.c_void_p, _cp, _cp, C.c_void_p,
                                  _cp]),
            "or_ua_compute_rho": (None, [_i32, _d, _i64, _dp, _cp, _d, _i32, _i32, _d, _d, _d,
                                         _d, _dp, _dp]),
            "or_ua_compute_v": (None, [_i32, _d, _i64, _dp, _cp, _dp]),
            "or_ua_run": (_i32, [_i32, _d, _d, _i32, _i64, _dp, _dp, _dp, _d, _i32, _i32, _d, _d,
                                 _d, _d, _dp, _dp]),
            "or_f_tsc": (_d, [_d]),
            "or_f_m4": (_d, [_d]),
            "or_push_v_ngp": (None, [_i64, _dp, _d, _i32, _i32, _d, _d, _dp, _dp, C.c_void_p, _i32]),
            "or_push_v_m4": (None, [_i64, _dp, _d, _i32, _i32, _d, _d, _dp, _dp, C.c_void_p, _i32]),
            "or_compute_rho_ngp": (None, [_i64, _dp, _i32, _i32, _d, _d, _d, _d, _dp]),
            "or_compute_rho_shape4": (None, [_i64, _dp, _i32, _i32, _d, _d, _d, _d, _i32, _dp]),
            "or_compute_rho_m6_fields": (None, [_i64, _dp, _d, _i32, _i32, _d, _d, _d, _d, _i32, _dp]),
            "or_interpol_eb_m6": (None, [_i64, _dp, _i32, _i32, _d, _d, _d, _d, _dp, _dp]),
            "or_rfft2": (_i32, [_i32, _i32, _dp, _i32, _cp]),
            "or_irfft2": (_i32, [_i32, _i32, _cp, _dp, _i32]),
            "or_poisson_tables": (None, [_i32, _i32, _d, _d, _dp, _dp, _dp]),
            "or_solve_phi": (_i32, [_i32, _i32, _d, _d, _dp, _dp]),
            "or_solve_e": (_i32, [_i32, _i32, _d, _d, _dp, _dp, _dp]),
            "or_solve_fields": (_i32, [_i32, _i32, _d, _d, _d, _d, _dp, _dp, C.POINTER(_d)]),
            "or_energy": (_d, [_i32, _i32, _d, _d, _dp, C.c_void_p, C.c_void_p]),
            "or_counting_sort": (None, [_i64, _ip, C.c_int32, _ip, _lp]),
            "or_landau_loop": (_i32, [_i64, _dp, _i32, _i32, _d, _d, _d, _i32, _i32, _d, _dpc]),
            "or_num_threads": (_i32, []),
            "orf_m6_weights": (None, [_f, _fp]),
            "orf_shape_weights": (None, [_i32, _i32, _f, _i32, _fp]),
            "orf_step": (None, [_i64, _ip, _fp, _fp, _fp, _fp, _i32, _i32, _i32, _fp, _i32, _f,
                                C.c_void_p, _f, _f, _f, _f, _i32, _dp]),
            "orf_deposit": (None, [_i64, _ip, _fp, _fp, _i32, _i32, _i32, _f, _dp]),
            "orf_from_f64": (None, [_i64, _dp, _i32, _i32, _d, _d, _ip, _fp, _fp, _fp, _fp]),
            "orf_to_f64": (None, [_i64, _ip, _fp, _fp, _fp, _fp, _i32, _d, _d, _d, _dp]),
            "or_poisson_adjoint": (_i32, [_i32, _i32, _d, _d, _dp, _dp, _dp]),
            "or_adjoint": (_i32, [_i64, _dp, _i32, _i32, _d, _d, _d, _i32, _i32, _i32,
                                  C.POINTER(_d), _dp, _dp, C.c_void_p]),
            "or_energy_objective": (_i32, [_i64, _dp, _i32, _i32, _d, _d, _d, _i32, _i32,
                                           C.POINTER(_d)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr_or_none(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _F(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


# ------------------------------------------------------------------ CIC / M6

def f_m6(q):
    return lib().or_f_m6(float(q))


# kernel names and the loop's kernel codes
KERNELS = {"cic": 0, "m6": 1, "ngp": 2, "tsc": 3, "m4": 4}


def f_tsc(x):
    return lib().or_f_tsc(float(x))


def f_m4(x):
    return lib().or_f_m4(float(x))


def push_v(p, kernel, nx, ny, dx, dy, ex, ey, bz, dt):
    """push_v!(p, kernel, mesh, ex, ey, bz, dt) on a (5, n) array, in place.
    TSC gathers with the M4 weights, as the reference's method does."""
    ld = ex.shape[0]
    assert p.flags.f_contiguous and p.dtype == np.float64
    f = {"cic": lib().or_push_v_cic, "m6": lib().or_push_v_m6, "ngp": lib().or_push_v_ngp,
         "tsc": lib().or_push_v_m4, "m4": lib().or_push_v_m4}[kernel]
    bzf = None if bz is None else _F(bz)
    f(p.shape[1], p, dt, nx, ny, dx, dy, _F(ex), _F(ey), _ptr_or_none(bzf), ld)


def push_x(p, dimx, dimy, dt):
    lib().or_push_x(p.shape[1], p, dt, dimx, dimy)


def compute_rho(p, kernel, nx, ny, dx, dy, dimx=None, dimy=None):
    rho = np.zeros((nx, ny), order="F")
    if kernel == "cic":
        lib().or_compute_rho_cic(p.shape[1], p, nx, ny, dx, dy, rho)
    elif kernel == "m6":
        lib().or_compute_rho_m6(p.shape[1], p, nx, ny, dx, dy, dimx, dimy, rho)
    elif kernel == "ngp":
        lib().or_compute_rho_ngp(p.shape[1], p, nx, ny, dx, dy, dimx, dimy, rho)
    else:
        lib().or_compute_rho_shape4(p.shape[1], p, nx, ny, dx, dy, dimx, dimy,
                                    1 if kernel == "m4" else 0, rho)
    return rho


def interpolate_cic(p, nx, ny, dx, dy, ex, ey):
    n = p.shape[1]
    epx, epy = np.zeros(n), np.zeros(n)
    lib().or_interpolate_cic(n, p, nx, ny, dx, dy, _F(ex), _F(ey), epx, epy)
    return epx, epy


def compute_rho_m6_fields(x, w, nx, ny, xmin, xmax, ymin, ymax, compat=True):
    """compute_rho_m6!(fields, particles): returns rho ((nx+1, ny+1)); x is
    rewritten in place as the reference does."""
    rho = np.zeros((nx + 1, ny + 1), order="F")
    lib().or_compute_rho_m6_fields(x.shape[1], x, w, nx, ny, xmin, xmax, ymin, ymax,
                                   1 if compat else 0, rho)
    return rho


def interpol_eb_m6(x, e_fields, nx, ny, xmin, xmax, ymin, ymax):
    """interpol_eb_m6!(particles, fields): e_fields is (2, nx+1, ny+1);
    returns particles.e (2, n); x is rewritten in place."""
    n = x.shape[1]
    ep = np.zeros((2, n), order="F")
    lib().or_interpol_eb_m6(n, x, nx, ny, xmin, xmax, ymin, ymax, _F(e_fields), ep)
    return ep


# ------------------------------------------------------------------ Poisson

def rfft2(a):
    nx, ny = a.shape
    out = np.zeros((nx // 2 + 1, ny), dtype=np.complex128, order="F")
    rc = lib().or_rfft2(nx, ny, _F(a), nx, out)
    assert rc == 0
    return out


def irfft2(y, nx):
    ny = y.shape[1]
    out = np.zeros((nx, ny), order="F")
    rc = lib().or_irfft2(nx, ny, np.asfortranarray(y, dtype=np.complex128), out, nx)
    assert rc == 0
    return out


def poisson_tables(nx, ny, dimx, dimy):
    h = nx // 2 + 1
    kx, ky, k2 = (np.zeros((h, ny), order="F") for _ in range(3))
    lib().or_poisson_tables(nx, ny, dimx, dimy, kx, ky, k2)
    return kx, ky, k2


def solve_phi(rho, dimx, dimy):
    nx, ny = rho.shape
    phi = np.zeros((nx, ny), order="F")
    assert lib().or_solve_phi(nx, ny, dimx, dimy, _F(rho), phi) == 0
    return phi


def solve_e(rho, dimx, dimy):
    nx, ny = rho.shape
    ex, ey = np.zeros((nx, ny), order="F"), np.zeros((nx, ny), order="F")
    assert lib().or_solve_e(nx, ny, dimx, dimy, _F(rho), ex, ey) == 0
    return ex, ey


def solve_fields(rho_g, xmin, xmax, ymin, ymax):
    """Poisson2D solve!(fields): rho_g is (nx+1, ny+1); returns (e, energy)
    with e of shape (2, nx+1, ny+1)."""
    nx, ny = rho_g.shape[0] - 1, rho_g.shape[1] - 1
    e = np.zeros((2, nx + 1, ny + 1), order="F")
    en = C.c_double(0.0)
    assert lib().or_solve_fields(nx, ny, xmin, xmax, ymin, ymax, _F(rho_g), e, C.byref(en)) == 0
    return e, en.value


def energy(ex, dx, dy):
    nx, ny = ex.shape
    return lib().or_energy(nx, ny, dx, dy, _F(ex), None, None)


def counting_sort(key, nkeys):
    key = np.ascontiguousarray(key, dtype=np.int32)
    perm = np.zeros(key.size, dtype=np.int32)
    offs = np.zeros(nkeys + 1, dtype=np.int64)
    lib().or_counting_sort(key.size, key, nkeys, perm, offs)
    return perm, offs


def lane_spread(cell, ncells):
    """The lane-spread order of the fused kernels (build addition, no
    reference counterpart): stable sort by cell, then within each segment of
    64 consecutive cells order by (rank within the cell, cell).  Returns the
    source index of each slot."""
    cell = np.asarray(cell, dtype=np.int64)
    perm = np.argsort(cell, kind="stable")
    sc = cell[perm]
    first = np.searchsorted(sc, np.arange(ncells + 1), side="left")
    rank = np.arange(sc.size) - first[sc]
    order = np.lexsort((sc % 64, rank, sc // 64))
    return perm[order].astype(np.int32)


def landau_loop(p, kernel, nx, ny, dimx, dimy, dt, nsteps, bz0=0.0):
    e = np.zeros(nsteps)
    rc = lib().or_landau_loop(p.shape[1], p, nx, ny, dimx, dimy, dt, nsteps,
                              KERNELS[kernel], bz0, e)
    assert rc == 0
    return e


def num_threads():
    return lib().or_num_threads()


# --------------------------------------------------------------- FP32 state

def f32_m6_weights(t):
    w = np.zeros(6, dtype=np.float32)
    lib().orf_m6_weights(np.float32(t), w)
    return w


# shape codes of the FP32 step (the library's PIC_CIC .. PIC_M4)
F32_SHAPES = {"cic": 0, "m6": 1, "ngp": 2, "tsc": 3, "m4": 4}


def f32_shape_code(kernel):
    try:
        return F32_SHAPES[kernel]
    except KeyError:
        raise ValueError(f"unknown shape {kernel!r}: one of {sorted(F32_SHAPES)}") from None


def f32_shape_weights(kernel, t, cell=0, gather=True):
    """FP32 weights of the S nodes (i - HL ..) of a separable shape at
    in-cell offset t of 0-based cell `cell` (NGP's tie rule reads its
    parity); gather=False gives compute_rho's (TSC: f_tsc)."""
    s = {"m6": 6, "ngp": 2, "tsc": 4, "m4": 4}[kernel]
    w = np.zeros(6, dtype=np.float32)
    lib().orf_shape_weights(f32_shape_code(kernel), 1 if gather else 0, np.float32(t), int(cell),
                            w)
    return w[:s]


def f32_from_f64(p, nx, ny, dx, dy):
    n = p.shape[1]
    cell = np.zeros(n, np.int32)
    ox, oy, vx, vy = (np.zeros(n, np.float32) for _ in range(4))
    lib().orf_from_f64(n, p, nx, ny, dx, dy, cell, ox, oy, vx, vy)
    return cell, ox, oy, vx, vy


def f32_to_f64(state, nx, dx, dy, w):
    cell, ox, oy, vx, vy = state
    n = cell.size
    p = np.zeros((5, n), order="F")
    lib().orf_to_f64(n, cell, ox, oy, vx, vy, nx, dx, dy, w, p)
    return p


def f32_step(state, nx, ny, kernel, e32, hdt, dtdx, dtdy, w, bz_mode=0, bz0=0.0,
             bz32=None, deposit=True):
    """One fused FP32 step in place on state = (cell, ox, oy, vx, vy); returns
    the FP64 accumulation of the deposited FP32 contributions (nx*ny, raw)."""
    cell, ox, oy, vx, vy = state
    rho = np.zeros((nx, ny), order="F")
    b = None if bz32 is None else np.ascontiguousarray(bz32, dtype=np.float32)
    lib().orf_step(cell.size, cell, ox, oy, vx, vy, nx, ny, f32_shape_code(kernel),
                   np.ascontiguousarray(e32, dtype=np.float32), bz_mode, np.float32(bz0),
                   _ptr_or_none(b), np.float32(hdt), np.float32(dtdx), np.float32(dtdy),
                   np.float32(w), 1 if deposit else 0, rho)
    return rho


def f32_deposit(state, nx, ny, kernel, w):
    cell, ox, oy = state[:3]
    rho = np.zeros((nx, ny), order="F")
    lib().orf_deposit(cell.size, cell, ox, oy, nx, ny, f32_shape_code(kernel),
                      np.float32(w), rho)
    return rho


# ---------------------------------------------------------------- adjoint

def poisson_adjoint(mux, muy, dimx, dimy):
    nx, ny = mux.shape
    out = np.zeros((nx, ny), order="F")
    assert lib().or_poisson_adjoint(nx, ny, dimx, dimy, _F(mux), _F(muy), out) == 0
    return out


def adjoint(p0, kernel, nx, ny, dimx, dimy, dt, nsteps, ckpt):
    """Forward loop + reverse sweep: returns (J, gx (2, n), gv (2, n), e_hist)
    with J = sum(ex^2+ey^2) dx dy of the last step's field."""
    n = p0.shape[1]
    J = C.c_double()
    gx = np.zeros((2, n), order="F")
    gv = np.zeros((2, n), order="F")
    eh = np.zeros(nsteps)
    rc = lib().or_adjoint(n, _F(p0), nx, ny, dimx, dimy, dt, nsteps, 0 if kernel == "cic" else 1,
                          ckpt, C.byref(J), gx, gv, eh.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    return J.value, gx, gv, eh


def energy_objective(p0, kernel, nx, ny, dimx, dimy, dt, nsteps):
    J = C.c_double()
    rc = lib().or_energy_objective(p0.shape[1], _F(p0), nx, ny, dimx, dimy, dt, nsteps,
                                   0 if kernel == "cic" else 1, C.byref(J))
    assert rc == 0
    return J.value


# ------------------------------------------------- reference Fortran twin

_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """src/pic.f90 (push_v, push_x, deposition) as built by the Makefile."""
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        pi, pd = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        L.push_v.argtypes = [pi, _dp, pd, pi, pi, pd, pd, _dp, _dp, _dp]
        L.push_x.argtypes = [pi, pd, pd, _dp, pd]
        L.deposition.argtypes = [pi, pi, pi, pd, pd, _dp, _dp, _dp]
        for f in (L.push_v, L.push_x, L.deposition):
            f.restype = None
        _ref = L
    return _ref


def _i(v):
    return C.byref(C.c_int32(v))


def _r(v):
    return C.byref(C.c_double(v))


def ref_push_v(p, nx, ny, dx, dy, ex, ey, bz, dt):
    """F90.push_v! (src/pic.jl:79-114): fields carry the ghost row/column,
    (nx+1, ny+1)."""
    ref().push_v(_i(p.shape[1]), p, _r(dt), _i(nx), _i(ny), _r(dx), _r(dy), _F(ex), _F(ey), _F(bz))


def ref_push_x(p, dimx, dimy, dt):
    ref().push_x(_i(p.shape[1]), _r(dimx), _r(dimy), p, _r(dt))


def ref_deposition(p, nx, ny, dx, dy):
    jx = np.zeros((nx + 1, ny + 1), order="F")
    jy = np.zeros((nx + 1, ny + 1), order="F")
    ref().deposition(_i(p.shape[1]), _i(nx), _i(ny), _r(dx), _r(dy), p, jx, jy)
    return jx, jy


# ------------------------------------------------------- Vlasov-Maxwell

class FDTD:
    """fdtd.jl's FDTD struct: ex nx x (ny+1), ey (nx+1) x ny, bz nx x ny."""

    def __init__(self, nx, ny):
        self.ex = np.zeros((nx, ny + 1), order="F")
        self.ey = np.zeros((nx + 1, ny), order="F")
        self.bz = np.zeros((nx, ny), order="F")


def faraday(f, nx, ny, dx, dy, dt):
    lib().or_faraday(nx, ny, dx, dy, f.ex, f.ey, f.bz, dt)


def ampere_maxwell(f, nx, ny, dx, dy, jx, jy, dt):
    lib().or_ampere_maxwell(nx, ny, dx, dy, f.ex, f.ey, f.bz, _F(jx), _F(jy), dt)


def update_fields(f, nx, ny):
    ex, ey, bz = (np.zeros((nx + 1, ny + 1), order="F") for _ in range(3))
    lib().or_update_fields(nx, ny, f.ex, f.ey, f.bz, ex, ey, bz)
    return ex, ey, bz


def compute_current(p, nx, ny, dx, dy):
    jx, jy = (np.zeros((nx + 1, ny + 1), order="F") for _ in range(2))
    lib().or_compute_current(p.shape[1], _F(p), nx, ny, dx, dy, jx, jy)
    return jx, jy


def fdtd_energy(f, nx, ny, dx, dy):
    return lib().or_fdtd_energy(nx, ny, dx, dy, f.ex)


def maxwell_loop(p, f, nx, ny, dimx, dimy, dt, nsteps):
    """The loop of examples/landau_damping_maxwell_fdtd.jl; p and f are
    updated in place; returns compute_energy after every step."""
    e = np.zeros(nsteps)
    rc = lib().or_maxwell_loop(p.shape[1], p, nx, ny, dimx, dimy, dt, nsteps, f.ex, f.ey, f.bz,
                               e)
    assert rc == 0
    return e


# ------------------------------------------------------------------ UA

def _C(a):
    return np.asfortranarray(a, dtype=np.complex128)


def fft_c(a, sign=-1):
    """The oracle's radix-2 transform of a 1-D complex array (sign -1
    forward, +1 backward unnormalised)."""
    out = _C(np.array(a, dtype=np.complex128))
    lib().or_fft_c(out, out.shape[0], sign)
    return out


def ua_preparation(ntau, eps, dt, x, v, e):
    """preparation!(ua, dt, particles, xt, yt): x, v, e are Particles' (2, n)
    arrays; returns b, t, pl, ql (ntau, n), xt, yt (ntau, 2, n)."""
    n = x.shape[1]
    b, t = np.zeros((1, n), order="F"), np.zeros((1, n), order="F")
    pl, ql = (np.zeros((ntau, n), np.complex128, order="F") for _ in range(2))
    xt, yt = (np.zeros((ntau, 2, n), np.complex128, order="F") for _ in range(2))
    lib().or_ua_preparation(ntau, eps, dt, n, _F(x), _F(v), _F(e), b, t, pl, ql, xt, yt)
    return b[0], t[0], pl, ql, xt, yt


def ua_interpol(xt, e_fields, nx, ny, xmin, xmax, ymin, ymax):
    """update_particles_e!: et (ntau, 2, n) from the real parts of xt."""
    ntau, _, n = xt.shape
    et = np.zeros((ntau, 2, n), order="F")
    lib().or_ua_interpol(ntau, n, _C(xt), nx, ny, xmin, xmax, ymin, ymax, _F(e_fields), et)
    return et


def ua_compute_f(eps, b, xt, yt, et):
    ntau, _, n = xt.shape
    fx, fy = (np.zeros((ntau, 2, n), np.complex128, order="F") for _ in range(2))
    lib().or_ua_compute_f(ntau, eps, n, _F(np.reshape(b, (1, n))), _C(xt), _C(yt), _F(et), fx, fy)
    return fx, fy


def ua_step(eps, t, pl, xtilde, fx, ql=None, gx=None):
    ntau, _, n = xtilde.shape
    out = np.zeros((ntau, 2, n), np.complex128, order="F")
    qlc = _C(ql) if ql is not None else None
    gxc = _C(gx) if gx is not None else None
    lib().or_ua_step(ntau, eps, n, _F(np.reshape(t, (1, n))), _C(pl),
                     qlc.ctypes.data if qlc is not None else None, _C(xtilde), _C(fx),
                     gxc.ctypes.data if gxc is not None else None, out)
    return out


def ua_compute_rho(eps, t, xt, w, nx, ny, xmin, xmax, ymin, ymax):
    """update_particles_x!: returns (Particles.x (2, n), fields.rho (nx+1, ny+1))."""
    ntau, _, n = xt.shape
    x = np.zeros((2, n), order="F")
    rho = np.zeros((nx + 1, ny + 1), order="F")
    lib().or_ua_compute_rho(ntau, eps, n, _F(np.reshape(t, (1, n))), _C(xt), w, nx, ny, xmin,
                            xmax, ymin, ymax, x, rho)
    return x, rho


def ua_compute_v(eps, t, yt):
    ntau, _, n = yt.shape
    v = np.zeros((2, n), order="F")
    lib().or_ua_compute_v(ntau, eps, n, _F(np.reshape(t, (1, n))), _C(yt), v)
    return v


def ua_run(ntau, eps, dt, nsteps, x, v, e, w, e_fields, xmin, xmax, ymin, ymax):
    """The loop of test/bupdate.jl; x, v and e_fields are updated in place;
    returns the solve energies (2, nsteps): prediction, correction."""
    n = x.shape[1]
    nx, ny = e_fields.shape[1] - 1, e_fields.shape[2] - 1
    en = np.zeros((2, nsteps), order="F")
    rc = lib().or_ua_run(ntau, eps, dt, nsteps, n, x, v, _F(e), w, nx, ny, xmin, xmax, ymin,
                         ymax, e_fields, en)
    assert rc == 0
    return en
