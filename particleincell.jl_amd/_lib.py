"""ctypes bindings of libpicmi355x.so (the C ABI declared in include/pic_mi355x.h).

The library is loaded from this package directory.  There is no fallback: if
it is missing or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libpicmi355x.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

PIC_OK = 0
PIC_CIC, PIC_M6, PIC_NGP, PIC_TSC, PIC_M4 = 0, 1, 2, 3, 4
PIC_STATE_F64, PIC_STATE_F32 = 0, 1
PIC_GRID_RHO, PIC_GRID_EX, PIC_GRID_EY, PIC_GRID_BZ, PIC_GRID_PHI = range(5)
PIC_BZ_NONE, PIC_BZ_UNIFORM, PIC_BZ_GRID = range(3)

_vp = C.c_void_p
_i, _i64, _d = C.c_int, C.c_int64, C.c_double
_pi32 = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)

# name -> argtypes (every function returns int status except pic_last_error)
SIGNATURES = {
    "pic_version": [],
    "pic_device_count": [C.POINTER(C.c_int)],
    "pic_ctx_create": [C.POINTER(_vp), _i, _i, _i, _d, _d, _d, _d, _i64],
    "pic_ctx_destroy": [_vp],
    "pic_synchronize": [_vp],
    "pic_stream": [_vp, C.POINTER(_vp)],
    "pic_set_particles": [_vp, _vp, _i64],
    "pic_get_particles": [_vp, _vp],
    "pic_set_particles_device": [_vp, _vp, _i64],
    "pic_get_particles_device": [_vp, _vp],
    "pic_num_particles": [_vp, C.POINTER(C.c_int64)],
    "pic_set_particles_xv": [_vp, _vp, _vp, _d, _i64],
    "pic_get_particles_xv": [_vp, _vp, _vp],
    "pic_get_particle_fields": [_vp, _vp],
    "pic_set_state": [_vp, _i],
    "pic_get_state_f32": [_vp, _vp, _vp, _vp, _vp, _vp],
    "pic_set_state_f32": [_vp, _vp, _vp, _vp, _vp, _vp, _d, _i64],
    "pic_set_grid": [_vp, _i, _vp, _i, _i],
    "pic_get_grid": [_vp, _i, _vp, _i, _i, _i],
    "pic_grid_device_ptr": [_vp, _i, C.POINTER(_vp)],
    "pic_set_bz": [_vp, _i, _d],
    "pic_compute_rho": [_vp, _i],
    "pic_compute_rho_noalloc": [_i, _vp, _i64, _i, _i, _d, _d, _d, _d, _vp, _vp],
    "pic_push_v": [_vp, _i, _d],
    "pic_push_x": [_vp, _d],
    "pic_operator_t": [_vp, _d],
    "pic_interpolate_cic": [_vp, _vp, _vp],
    "pic_solve_e": [_vp],
    "pic_solve_phi": [_vp],
    "pic_field_energy": [_vp, _vp],
    "pic_compute_rho_m6_fields": [_vp, _i],
    "pic_interpol_eb_m6": [_vp],
    "pic_solve_fields": [_vp, _pd],
    "pic_sort": [_vp, _vp],
    "pic_sort_lanes": [_vp, _vp],
    "pic_cell_offsets": [_vp, _vp],
    "pic_run": [_vp, _i, _d, _i, _i, _vp],
    "pic_step_pass": [_vp, _i, _i, _d, _vp],
    "pic_adjoint": [_vp, _i, _d, _i, _i, _i, _pd, _vp, _vp],
    "pic_set_adjoint_block": [_vp, _i],
    "pic_get_adjoint_plan": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int64)],
    "pic_get_adjoint_timing": [_vp, _pd, C.POINTER(C.c_int64)],
    "pic_set_timing": [_vp, _i],
    "pic_set_f32_accumulation": [_vp, _i],
    "pic_set_f32_tiles": [_vp, _i],
    "pic_set_f32_window": [_vp, _i],
    "pic_set_fused_solve": [_vp, _i],
    "pic_set_fused_step": [_vp, _i],
    "pic_get_timing": [_vp, _pd, C.POINTER(C.c_int64), _pd],
    "pic_slow_chunks": [_vp, C.POINTER(C.c_uint64)],
    "pic_fdtd_set": [_vp, _vp, _vp, _vp],
    "pic_fdtd_get": [_vp, _vp, _vp, _vp],
    "pic_faraday": [_vp, _d],
    "pic_ampere_maxwell": [_vp, _d],
    "pic_update_fields": [_vp],
    "pic_compute_current": [_vp, _i, _vp, _vp],
    "pic_set_current": [_vp, _vp, _vp],
    "pic_fdtd_energy": [_vp, _pd],
    "pic_run_maxwell": [_vp, _i, _d, _i, _vp],
    "pic_set_maxwell_path": [_vp, _i],
    "pic_ua_init": [_vp, _i, _d],
    "pic_ua_preparation": [_vp, _d, _vp, _vp, _vp, _vp, _vp, _vp],
    "pic_ua_interpol": [_vp, _vp, _vp],
    "pic_ua_compute_f": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pic_ua_step": [_vp, _vp, _vp, _vp, _vp, _vp],
    "pic_ua_compute_rho": [_vp, _vp, _vp],
    "pic_ua_compute_v": [_vp, _vp, _vp],
    "pic_ua_run": [_vp, _d, _i, _vp],
    "pic_set_particle_fields": [_vp, _vp],
    "pic_get_ua_timing": [_vp, _pd, C.POINTER(C.c_int64)],
    "pic_sample_landau": [_vp, _i64, _d, _d, _i],
    "pic_sample_maxwellian": [_vp, _i64, _d, _d, _d, _d, C.c_uint64, _i],
    "pic_read_particles": [_vp, C.c_char_p, _i, C.POINTER(C.c_int64)],
    "pic_write_particles": [_vp, C.c_char_p],
    "pic_f90_push_v": [_pi32, _vp, _pd, _pi32, _pi32, _pd, _pd, _vp, _vp, _vp],
    "pic_f90_push_x": [_pi32, _pd, _pd, _vp, _pd],
    "pic_f90_deposition": [_pi32, _pi32, _pi32, _pd, _pd, _vp, _vp, _vp],
}

_lib = None


class PicError(RuntimeError):
    """A libpicmi355x call returned an error status."""


def load():
    """Loads libpicmi355x.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run __graft_entry__.build() (the HIP library has "
            "no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.pic_last_error.restype = C.c_char_p
    L.pic_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = args
    _lib = L
    return L


def last_error():
    return load().pic_last_error().decode(errors="replace")


def call(name, *args):
    """Calls `name`; raises PicError with the library's message on failure."""
    rc = getattr(load(), name)(*args)
    if rc != PIC_OK:
        raise PicError(f"{name} failed ({rc}): {last_error()}")
    return rc


def device_count():
    n = C.c_int(0)
    call("pic_device_count", C.byref(n))
    return n.value
