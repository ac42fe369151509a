// pic_calls.hip -- the reference's per-call operations on device-resident
// particles: compute_rho, push_v!, push_x!, interpolate_cic!, the MeshFields
// M6 deposit and gather, and the Fortran twin's current deposit.
#include <cmath>
#include <vector>

#include "pic_internal.h"

namespace pic {

int launch_step64(Ctx *c, int kernel, int mode, double dt, double dtd);
int launch_step32(Ctx *c, int kernel, int mode, double dt, double dtd);

static inline dim3 grid1(int64_t n, int t = 256) {
    int64_t b = (n + t - 1) / t;
    return dim3((unsigned)(b > 0 ? b : 1));
}

// push_x!: pushers_cic.jl:55-72.
__global__ void k_push_x(int64_t n, double *x, double *y, const double *vx, const double *vy,
                         double dt, double dimx, double dimy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double p1 = x[k] + dt * vx[k];
    double p2 = y[k] + dt * vy[k];
    if (p1 > dimx) p1 -= dimx;
    if (p2 > dimy) p2 -= dimy;
    if (p1 < 0.0) p1 += dimx;
    if (p2 < 0.0) p2 += dimy;
    x[k] = p1;
    y[k] = p2;
}

// operator_t!(split, dt): poisson_2d.jl:265-277, x += dt v with no wrap.
__global__ void k_operator_t(int64_t n, double *x, double *y, const double *vx, const double *vy,
                             double dt) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    x[k] += dt * vx[k];
    y[k] += dt * vy[k];
}

// interpolate_cic!: interpolate_cic.jl:3-33 (trunc index, mod1 neighbours).
__global__ void k_interpolate_cic(int64_t n, const double *x, const double *y, int nx, int ny,
                                  double dx, double dy, const double *ex, const double *ey,
                                  double *epx, double *epy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double xp = x[k] / dx, yp = y[k] / dy;
    long long i = (long long)xp + 1, j = (long long)yp + 1;
    double dxp = (xp - (double)i) + 1, dyp = (yp - (double)j) + 1;
    double a1 = (1 - dxp) * (1 - dyp);
    double a2 = dxp * (1 - dyp);
    double a3 = dxp * dyp;
    double a4 = (1 - dxp) * dyp;
    long long ip = picd::mod1z(i + 1, nx), jp = picd::mod1z(j + 1, ny);
    long long i0 = picd::mod1z(i, nx), j0 = picd::mod1z(j, ny);
    size_t c00 = i0 + (size_t)nx * j0, c10 = ip + (size_t)nx * j0;
    size_t c11 = ip + (size_t)nx * jp, c01 = i0 + (size_t)nx * jp;
    epx[k] = a1 * ex[c00] + a2 * ex[c10] + a3 * ex[c11] + a4 * ex[c01];
    epy[k] = a1 * ey[c00] + a2 * ey[c10] + a3 * ey[c11] + a4 * ey[c01];
}

using picd::jl_fmod;

// compute_rho_m6!(fields, particles): compute_rho_m6.jl:181-316, raw
// accumulation into acc (nx*ny, the ghost copies are implied).  compat
// selects the reference's swapped bounds and mod1 neighbours (see the C ABI
// header).  Weights are the M6 pieces in FP64.
__global__ void k_mf_rho_m6(int64_t n, double *x, double *y, double w, int nx, int ny,
                            double x0, double y0, double dimx, double dimy, double dx, double dy,
                            int compat, double *acc) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double xt1 = jl_fmod(x[k] - x0, dimx);
    double xt2 = jl_fmod(y[k] - y0, dimy);
    x[k] = xt1 + x0;
    y[k] = xt2 + y0;
    double px = xt1 / dx, py = xt2 / dy;
    int i = (int)px, j = (int)py;  // trunc(Int32, .)
    double dpx = px - (double)i, dpy = py - (double)j;
    double wx[6], wy[6];
    picd::m6_weights_d(dpx, wx);
    picd::m6_weights_d(dpy, wy);
    int ii[6], jj[6];
    for (int a = 0; a < 6; ++a) {
        int off = a - 2;  // node offsets -2..3
        if (off == 0) {
            ii[a] = picd::wrapi(i, nx);
            jj[a] = picd::wrapi(j, ny);
        } else if (compat) {
            ii[a] = picd::wrapi(i + off - 1, nx);
            jj[a] = picd::wrapi(j + off - 1, ny);
        } else {
            ii[a] = picd::wrapi(i + off, nx);
            jj[a] = picd::wrapi(j + off, ny);
        }
    }
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b)
            unsafeAtomicAdd(&acc[ii[a] + (size_t)nx * jj[b]], wx[a] * wy[b] * w);
}

// interpol_eb_m6!(particles, fields): interpolation_m6.jl:125-247.
__global__ void k_mf_gather_m6(int64_t n, double *x, double *y, int nx, int ny, double xmin,
                               double ymin, double dimx, double dimy, double dx, double dy,
                               const double *ex, const double *ey, double *epx, double *epy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double xn = jl_fmod(x[k] - xmin, dimx);
    double yn = jl_fmod(y[k] - ymin, dimy);
    double px = xn / dx, py = yn / dy;
    int i = (int)px, j = (int)py;
    double dpx = px - (double)i, dpy = py - (double)j;
    x[k] = xmin + xn;
    y[k] = ymin + yn;
    double e1, e2;
    picd::mf_gather_m6(i, j, dpx, dpy, nx, ny, ex, ey, e1, e2);
    epx[k] = e1;
    epy[k] = e2;
}

// deposition (pic.f90:102-149): CIC current into ghosted (nx+1) x (ny+1)
// grids, floor index, no wrap, weights vx*factor and vy*factor with
// factor = real(nx*ny)/nbpart evaluated in single precision as the Fortran
// default REAL does.
__global__ void k_f90_deposition(int64_t n, const double *__restrict__ p, int nx, double dx,
                                 double dy, double factor, double *jx, double *jy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *q = p + 5 * k;
    double xp = q[0] / dx, yp = q[1] / dy;
    long long i = (long long)floor(xp) + 1, j = (long long)floor(yp) + 1;
    double dxp = (xp - (double)i) + 1, dyp = (yp - (double)j) + 1;
    double a1 = (1 - dxp) * (1 - dyp);
    double a2 = dxp * (1 - dyp);
    double a3 = dxp * dyp;
    double a4 = (1 - dxp) * dyp;
    double w1 = q[2] * factor, w2 = q[3] * factor;
    const size_t ld = (size_t)nx + 1;
    size_t c00 = (i - 1) + ld * (j - 1), c10 = i + ld * (j - 1);
    size_t c11 = i + ld * j, c01 = (i - 1) + ld * j;
    unsafeAtomicAdd(&jx[c00], a1 * w1);
    unsafeAtomicAdd(&jx[c10], a2 * w1);
    unsafeAtomicAdd(&jx[c11], a3 * w1);
    unsafeAtomicAdd(&jx[c01], a4 * w1);
    unsafeAtomicAdd(&jy[c00], a1 * w2);
    unsafeAtomicAdd(&jy[c10], a2 * w2);
    unsafeAtomicAdd(&jy[c11], a3 * w2);
    unsafeAtomicAdd(&jy[c01], a4 * w2);
}

// Largest |x|/dx and |y|/dy cell indices must stay inside the ghosted grid for
// the Fortran deposition; checked on the host before the launch.
static int check_f90_bounds(const double *p, int64_t n, int nx, int ny, double dx, double dy) {
    for (int64_t k = 0; k < n; ++k) {
        double xp = p[5 * k] / dx, yp = p[5 * k + 1] / dy;
        PIC_CHECK(xp >= 0.0 && yp >= 0.0 && floor(xp) < nx && floor(yp) < ny, PIC_ERR_ARG,
                  "particle %lld at (%g, %g) lies outside the grid the Fortran deposition "
                  "addresses (pic.f90:125-127)",
                  (long long)k, p[5 * k], p[5 * k + 1]);
    }
    return PIC_OK;
}

static int need_f64(Ctx *c) {
    PIC_CHECK(c->state == PIC_STATE_F64, PIC_ERR_STATE,
              "the per-call operations need the FP64 state (pic_set_state)");
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_compute_rho(pic_ctx *ctx, int kernel) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(valid_shape(kernel), PIC_ERR_ARG, "kernel %d", kernel);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, (int64_t)c->nx * c->ny));
    if (c->state == PIC_STATE_F64)
        PIC_TRY(launch_step64(c, kernel, 3, 0.0, 0.0));
    else
        PIC_TRY(launch_step32(c, kernel, 3, 0.0, 0.0));
    PIC_TRY(finalize_rho(c, kernel));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_push_v(pic_ctx *ctx, int kernel, double dt) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(valid_shape(kernel), PIC_ERR_ARG, "kernel %d", kernel);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_f64(c));
    PIC_TRY(launch_step64(c, kernel, 4, dt, 0.0));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_push_x(pic_ctx *ctx, double dt) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_f64(c));
    if (c->n) {
        hipLaunchKernelGGL(k_push_x, grid1(c->n), dim3(256), 0, c->stream, c->n, c->x, c->y, c->vx,
                           c->vy, dt, c->dimx, c->dimy);
        PIC_HIP(hipGetLastError());
    }
    c->sort_fresh = false;
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_operator_t(pic_ctx *ctx, double dt) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_f64(c));
    if (c->n) {
        hipLaunchKernelGGL(k_operator_t, grid1(c->n), dim3(256), 0, c->stream, c->n, c->x, c->y,
                           c->vx, c->vy, dt);
        PIC_HIP(hipGetLastError());
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->seg_valid = false;
    c->sort_fresh = false;
    return PIC_OK;
}

int pic_interpolate_cic(pic_ctx *ctx, double *epx, double *epy) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_f64(c));
    PIC_TRY(ensure_pfields(c));
    if (c->n) {
        hipLaunchKernelGGL(k_interpolate_cic, grid1(c->n), dim3(256), 0, c->stream, c->n, c->x,
                           c->y, c->nx, c->ny, c->dx, c->dy, c->ex, c->ey, c->epx, c->epy);
        PIC_HIP(hipGetLastError());
        if (epx)
            PIC_HIP(hipMemcpyAsync(epx, c->epx, (size_t)c->n * 8, hipMemcpyDeviceToHost,
                                   c->stream));
        if (epy)
            PIC_HIP(hipMemcpyAsync(epy, c->epy, (size_t)c->n * 8, hipMemcpyDeviceToHost,
                                   c->stream));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_compute_rho_m6_fields(pic_ctx *ctx, int compat) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_f64(c));
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, (int64_t)c->nx * c->ny));
    if (c->n) {
        double x0 = compat ? c->xmax : c->xmin, y0 = compat ? c->ymax : c->ymin;
        hipLaunchKernelGGL(k_mf_rho_m6, grid1(c->n), dim3(256), 0, c->stream, c->n, c->x, c->y,
                           c->w0, c->nx, c->ny, x0, y0, c->dimx, c->dimy, c->dx, c->dy,
                           compat ? 1 : 0, c->rho_acc);
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(finalize_rho(c, PIC_M6));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_interpol_eb_m6(pic_ctx *ctx) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_f64(c));
    PIC_TRY(ensure_pfields(c));
    if (c->n) {
        hipLaunchKernelGGL(k_mf_gather_m6, grid1(c->n), dim3(256), 0, c->stream, c->n, c->x, c->y,
                           c->nx, c->ny, c->xmin, c->ymin, c->dimx, c->dimy, c->dx, c->dy, c->ex,
                           c->ey, c->epx, c->epy);
        PIC_HIP(hipGetLastError());
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_f90_deposition(const int32_t *nbpart, const int32_t *nx, const int32_t *ny,
                       const double *dx, const double *dy, const double *p, double *jx,
                       double *jy) {
    PIC_CHECK(nbpart && nx && ny && dx && dy && p && jx && jy, PIC_ERR_ARG, "null argument");
    PIC_CHECK(*nbpart > 0 && *nx > 0 && *ny > 0, PIC_ERR_ARG, "empty input");
    const int64_t n = *nbpart;
    PIC_TRY(check_f90_bounds(p, n, *nx, *ny, *dx, *dy));
    const size_t g = (size_t)(*nx + 1) * (*ny + 1);
    double *dp = nullptr, *djx = nullptr, *djy = nullptr;
    hipStream_t s = nullptr;
    auto cleanup = [&]() {
        if (dp) (void)hipFree(dp);
        if (djx) (void)hipFree(djx);
        if (djy) (void)hipFree(djy);
        if (s) (void)hipStreamDestroy(s);
    };
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&dp, (size_t)n * 5 * 8);
    if (e == hipSuccess) e = hipMalloc(&djx, g * 8);
    if (e == hipSuccess) e = hipMalloc(&djy, g * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, p, (size_t)n * 5 * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(djx, 0, g * 8, s);
    if (e == hipSuccess) e = hipMemsetAsync(djy, 0, g * 8, s);
    if (e == hipSuccess) {
        const double factor = (double)((float)(*nx * *ny) / (float)n);
        hipLaunchKernelGGL(k_f90_deposition, grid1(n), dim3(256), 0, s, n, dp, *nx, *dx, *dy,
                           factor, djx, djy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(jx, djx, g * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(jy, djy, g * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return fail(PIC_ERR_HIP, "pic_f90_deposition: %s", hipGetErrorString(e));
    return PIC_OK;
}

}  // extern "C"
