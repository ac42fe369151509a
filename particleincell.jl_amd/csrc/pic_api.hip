// pic_api.hip -- context management, particle/grid transfers and the
// Fortran-twin drop-ins of libpicmi355x.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <vector>

#include "pic_internal.h"

namespace pic {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

Ctx *ctx_of(pic_ctx *c) { return reinterpret_cast<Ctx *>(c); }

template <typename T>
static int dalloc(T **p, size_t count) {
    if (*p) return PIC_OK;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    if (e != hipSuccess)
        return fail(PIC_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    *p = static_cast<T *>(q);
    return PIC_OK;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

static void free_particles(Ctx *c) {
    dfree(c->x); dfree(c->y); dfree(c->vx); dfree(c->vy); dfree(c->w);
    dfree(c->epx); dfree(c->epy);
    dfree(c->cell); dfree(c->ov32);
    dfree(c->swap);
    dfree(c->defer);
    c->swap_bytes = 0;
    for (int k = 0; k < Ctx::NALT; ++k) {
        if (c->alt[k]) (void)hipFree(c->alt[k]);
        c->alt[k] = nullptr;
        c->alt_of[k] = nullptr;
    }
    for (int k = 0; k < 2; ++k) { dfree(c->sort_keys[k]); dfree(c->sort_vals[k]); }
    // the histograms and their scan block sums are sized by the particle
    // capacity too: both are reallocated together by the next ensure_sort
    dfree(c->sort_hist);
    dfree(c->sort_blk);
    c->sort_hist_cap = 0;
    dfree(c->perm2);
    c->sort_cap = 0;
}

int ensure_particles(Ctx *c, int64_t n) {
    PIC_CHECK(n >= 0 && n < (int64_t)INT32_MAX, PIC_ERR_ARG, "particle count %lld out of range",
              (long long)n);
    if (n > c->cap) {
        free_particles(c);
        c->cap = n;
    }
    PIC_TRY(dalloc(&c->x, c->cap));
    PIC_TRY(dalloc(&c->y, c->cap));
    PIC_TRY(dalloc(&c->vx, c->cap));
    PIC_TRY(dalloc(&c->vy, c->cap));
    return PIC_OK;
}

int ensure_f32(Ctx *c) {
    PIC_TRY(dalloc(&c->cell, c->cap));
    PIC_TRY(dalloc(&c->ov32, c->cap));
    return PIC_OK;
}

int ensure_swap(Ctx *c, int64_t bytes) {
    if (bytes > c->swap_bytes) {
        dfree(c->swap);
        c->swap_bytes = 0;
    }
    if (!c->swap) {
        PIC_TRY(dalloc(reinterpret_cast<char **>(&c->swap), (size_t)bytes));
        c->swap_bytes = bytes;
    }
    return PIC_OK;
}

int ensure_defer(Ctx *c) { return dalloc(&c->defer, c->cap); }

int ensure_pfields(Ctx *c) {
    PIC_TRY(dalloc(&c->epx, c->cap));
    PIC_TRY(dalloc(&c->epy, c->cap));
    return PIC_OK;
}

int ensure_energy(Ctx *c, int64_t n) {
    if (n > c->energy_cap) {
        dfree(c->energy_hist);
        c->energy_cap = 0;
    }
    if (!c->energy_hist) {
        int64_t cap = n < 1024 ? 1024 : n;
        PIC_TRY(dalloc(&c->energy_hist, (size_t)cap * 2));
        c->energy_cap = cap;
    }
    return PIC_OK;
}

// ------------------------------------------------------------- kernels

__global__ void k_fill(double *p, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

int launch_fill(Ctx *c, double *p, double v, int64_t n) {
    int blocks = (int)((n + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, c->stream, p, v, n);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// 5 x n AoS (x, y, vx, vy, w) -> SoA.  Each thread moves one particle; the
// 40-byte records are read with 8-byte loads, adjacent lanes adjacent records.
__global__ void k_aos_to_soa(const double *__restrict__ p, int64_t n, double *x, double *y,
                             double *vx, double *vy, double *w) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *q = p + 5 * k;
    x[k] = q[0];
    y[k] = q[1];
    vx[k] = q[2];
    vy[k] = q[3];
    if (w) w[k] = q[4];
}

__global__ void k_soa_to_aos(double *__restrict__ p, int64_t n, const double *x, const double *y,
                             const double *vx, const double *vy, const double *w, double w0) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double *q = p + 5 * k;
    q[0] = x[k];
    q[1] = y[k];
    q[2] = vx[k];
    q[3] = vy[k];
    q[4] = w ? w[k] : w0;
}

// Checks whether all weights equal w[0]: flag stays 0 if so.
__global__ void k_weights_differ(const double *__restrict__ p, int64_t n, int *flag) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (p[5 * k + 4] != p[4]) atomicOr(flag, 1);
}

// FP64 (x, y, vx, vy) -> FP32 cell-relative state.  xp = x/dx, i = floor(xp),
// ox = (float)(xp - i); an offset that rounds to 1.0f moves to 0 of the next
// cell; cells wrap periodically.
__global__ void k_f64_to_f32(int64_t n, const double *x, const double *y, const double *vx,
                             const double *vy, int nx, int ny, double dx, double dy,
                             int32_t *cell, float4 *ov) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double xp = x[k] / dx, yp = y[k] / dy;
    double fi = floor(xp), fj = floor(yp);
    float fx = (float)(xp - fi), fy = (float)(yp - fj);
    int i = (int)fi, j = (int)fj;
    if (fx >= 1.0f) { fx = 0.0f; i += 1; }
    if (fy >= 1.0f) { fy = 0.0f; j += 1; }
    i = picd::wrapi(i, nx);
    j = picd::wrapi(j, ny);
    cell[k] = i + nx * j;
    ov[k] = make_float4(fx, fy, (float)vx[k], (float)vy[k]);
}

__global__ void k_f32_to_f64(int64_t n, const int32_t *cell, const float4 *ov, int nx, double dx,
                             double dy, double *x, double *y, double *vx, double *vy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int c = cell[k];
    int i = c % nx, j = c / nx;
    const float4 o = ov[k];
    x[k] = ((double)i + (double)o.x) * dx;
    y[k] = ((double)j + (double)o.y) * dy;
    vx[k] = (double)o.z;
    vy[k] = (double)o.w;
}

// (ox, oy, vx, vy) records <-> four planes of n floats
__global__ void k_ov_unpack(int64_t n, const float4 *ov, float *planes) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 o = ov[k];
    planes[k] = o.x;
    planes[n + k] = o.y;
    planes[2 * n + k] = o.z;
    planes[3 * n + k] = o.w;
}

__global__ void k_ov_pack(int64_t n, const float *planes, float4 *ov) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    ov[k] = make_float4(planes[k], planes[n + k], planes[2 * n + k], planes[3 * n + k]);
}

static inline dim3 grid1(int64_t n, int t = 256) {
    int64_t b = (n + t - 1) / t;
    return dim3((unsigned)(b > 0 ? b : 1));
}

int to_state(Ctx *c, int state) {
    if (c->state == state) return PIC_OK;
    if (c->n == 0) {
        c->state = state;
        return PIC_OK;
    }
    if (state == PIC_STATE_F32) {
        PIC_CHECK(c->uniform_w, PIC_ERR_STATE,
                  "the FP32 cell-relative state needs uniform particle weights");
        PIC_CHECK((c->nx & (c->nx - 1)) == 0 && (c->ny & (c->ny - 1)) == 0, PIC_ERR_STATE,
                  "the FP32 cell-relative state needs nx and ny powers of two (%dx%d)", c->nx,
                  c->ny);
        PIC_TRY(ensure_f32(c));
        hipLaunchKernelGGL(k_f64_to_f32, grid1(c->n), dim3(256), 0, c->stream, c->n, c->x, c->y,
                           c->vx, c->vy, c->nx, c->ny, c->dx, c->dy, c->cell, c->ov32);
    } else {
        hipLaunchKernelGGL(k_f32_to_f64, grid1(c->n), dim3(256), 0, c->stream, c->n, c->cell,
                           c->ov32, c->nx, c->dx, c->dy, c->x, c->y, c->vx, c->vy);
    }
    PIC_HIP(hipGetLastError());
    c->state = state;
    return PIC_OK;
}

static int set_particles_impl(Ctx *c, const double *d_p, int64_t n) {
    PIC_TRY(ensure_particles(c, n));
    c->n = n;
    c->state = PIC_STATE_F64;
    c->seg_valid = false;
    c->sort_fresh = false;
    if (n == 0) {
        c->uniform_w = true;
        c->w0 = 0.0;
        return PIC_OK;
    }
    int *flag = nullptr;
    PIC_HIP(hipMalloc(&flag, sizeof(int)));
    PIC_HIP(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_weights_differ, grid1(n), dim3(256), 0, c->stream, d_p, n, flag);
    int h = 0;
    double w0 = 0;
    PIC_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipMemcpyAsync(&w0, d_p + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_HIP(hipFree(flag));
    c->uniform_w = (h == 0);
    c->w0 = w0;
    if (!c->uniform_w) PIC_TRY(dalloc(&c->w, c->cap));
    hipLaunchKernelGGL(k_aos_to_soa, grid1(n), dim3(256), 0, c->stream, d_p, n, c->x, c->y,
                       c->vx, c->vy, c->uniform_w ? nullptr : c->w);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int get_particles_impl(Ctx *c, double *d_p) {
    if (c->n == 0) return PIC_OK;
    int prev = c->state;
    PIC_TRY(to_state(c, PIC_STATE_F64));
    hipLaunchKernelGGL(k_soa_to_aos, grid1(c->n), dim3(256), 0, c->stream, d_p, c->n, c->x, c->y,
                       c->vx, c->vy, c->uniform_w ? nullptr : c->w, c->w0);
    PIC_HIP(hipGetLastError());
    (void)prev;
    return PIC_OK;
}

// A device staging buffer for host transfers, released after each call.
struct Staging {
    void *p = nullptr;
    ~Staging() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace pic

using namespace pic;

extern "C" {

const char *pic_last_error(void) { return g_err; }

int pic_version(void) { return 100; }

int pic_device_count(int *count) {
    PIC_CHECK(count, PIC_ERR_ARG, "null pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return PIC_OK;
}

int pic_ctx_create(pic_ctx **out, int device, int nx, int ny, double xmin, double xmax,
                   double ymin, double ymax, int64_t capacity) {
    PIC_CHECK(out, PIC_ERR_ARG, "null context pointer");
    *out = nullptr;
    PIC_CHECK(nx >= 4 && ny >= 4, PIC_ERR_ARG, "nx=%d, ny=%d: need at least 4 cells each", nx, ny);
    PIC_CHECK((int64_t)nx * ny <= (1 << 24), PIC_ERR_ARG, "grid %dx%d too large", nx, ny);
    PIC_CHECK(xmax > xmin && ymax > ymin, PIC_ERR_ARG, "empty domain");
    PIC_CHECK(capacity >= 0 && capacity < (int64_t)INT32_MAX, PIC_ERR_ARG,
              "capacity %lld out of range", (long long)capacity);
    int ndev = 0;
    PIC_HIP(hipGetDeviceCount(&ndev));
    PIC_CHECK(device >= 0 && device < ndev, PIC_ERR_ARG, "device %d of %d", device, ndev);
    PIC_HIP(hipSetDevice(device));
    Ctx *c = new Ctx();
    c->device = device;
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        c->ncu < 1)
        c->ncu = 1;
    c->nx = nx;
    c->ny = ny;
    c->xmin = xmin;
    c->xmax = xmax;
    c->ymin = ymin;
    c->ymax = ymax;
    c->dimx = xmax - xmin;
    c->dimy = ymax - ymin;
    c->dx = c->dimx / nx;
    c->dy = c->dimy / ny;
    c->cap = capacity;
    int rc = PIC_OK;
    auto chk = [&](int r) {
        if (rc == PIC_OK) rc = r;
    };
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(PIC_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    size_t m = (size_t)nx * ny;
    chk(dalloc(&c->rho, m));
    chk(dalloc(&c->ex, 2 * m));  // ex and ey contiguous: one batched c2r fills both
    if (rc == PIC_OK) c->ey = c->ex + m;
    chk(dalloc(&c->bz, m));
    chk(dalloc(&c->phi, m));
    chk(dalloc(&c->rho_acc, m));
    chk(dalloc(&c->e32, 2 * m));
    chk(dalloc(&c->bz32, m));
    chk(dalloc(&c->red_part, (size_t)RED_BLOCKS * 4));
    chk(dalloc(&c->red_count, 4));
    chk(dalloc(&c->step_counter, 4));
    chk(dalloc(&c->slow_chunks, 1));
    chk(dalloc(&c->defer_count, 2));
    if (rc == PIC_OK && capacity > 0) chk(ensure_particles(c, capacity));
    if (rc == PIC_OK) chk(poisson_init(c));
    if (rc == PIC_OK) {
        e = hipMemsetAsync(c->red_count, 0, 4 * sizeof(unsigned), c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->step_counter, 0, 4 * sizeof(int), c->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync(c->slow_chunks, 0, sizeof(unsigned long long), c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->defer_count, 0, 2 * sizeof(unsigned), c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->rho, 0, m * 8, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->ex, 0, 2 * m * 8, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->bz, 0, m * 8, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->bz32, 0, m * 4, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->rho_acc, 0, m * 8, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(PIC_ERR_HIP, "context init: %s", hipGetErrorString(e));
    }
    if (rc != PIC_OK) {
        std::string msg = g_err;
        pic_ctx_destroy(reinterpret_cast<pic_ctx *>(c));
        set_error("%s", msg.c_str());
        return rc;
    }
    *out = reinterpret_cast<pic_ctx *>(c);
    return PIC_OK;
}

int pic_ctx_destroy(pic_ctx *ctx) {
    if (!ctx) return PIC_OK;
    Ctx *c = ctx_of(ctx);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    poisson_free(c);
    free_particles(c);
    dfree(c->rho); dfree(c->ex); c->ey = nullptr; dfree(c->bz); dfree(c->phi);
    dfree(c->rho_acc); dfree(c->e32); dfree(c->bz32); dfree(c->es_buf);
    dfree(c->red_part); dfree(c->red_count); dfree(c->energy_hist); dfree(c->step_counter);
    dfree(c->sort_hist); dfree(c->cell_off);
    dfree(c->slow_chunks); dfree(c->defer_count);
    dfree(c->seg_k); dfree(c->seg_tbl); dfree(c->seg_A); dfree(c->seg_M);
    dfree(c->fd_ex); dfree(c->fd_ey); dfree(c->fd_bz); dfree(c->jx); dfree(c->jy);
    dfree(c->fd2_ex); dfree(c->fd2_ey); dfree(c->fd2_bz);
    for (double *&p : c->jb) dfree(p);
    dfree(c->ua_b); dfree(c->ua_t); dfree(c->ua_x0); dfree(c->ua_y0); dfree(c->ua_fhat);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PIC_OK;
}

int pic_synchronize(pic_ctx *ctx) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_HIP(hipStreamSynchronize(ctx_of(ctx)->stream));
    return PIC_OK;
}

int pic_stream(pic_ctx *ctx, void **stream) {
    PIC_CHECK(ctx && stream, PIC_ERR_ARG, "null pointer");
    *stream = (void *)ctx_of(ctx)->stream;
    return PIC_OK;
}

int pic_set_particles_device(pic_ctx *ctx, const double *d_p, int64_t n) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(d_p || n == 0, PIC_ERR_ARG, "null particle pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    return set_particles_impl(c, d_p, n);
}

int pic_get_particles_device(pic_ctx *ctx, double *d_p) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_CHECK(d_p || c->n == 0, PIC_ERR_ARG, "null particle pointer");
    PIC_HIP(hipSetDevice(c->device));
    return get_particles_impl(c, d_p);
}

int pic_set_particles(pic_ctx *ctx, const double *p, int64_t n) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(p || n == 0, PIC_ERR_ARG, "null particle pointer");
    PIC_CHECK(n >= 0, PIC_ERR_ARG, "negative particle count");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    if (n == 0) return set_particles_impl(c, nullptr, 0);
    Staging s;
    PIC_HIP(hipMalloc(&s.p, (size_t)n * 5 * sizeof(double)));
    PIC_HIP(hipMemcpyAsync(s.p, p, (size_t)n * 5 * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    PIC_TRY(set_particles_impl(c, (const double *)s.p, n));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_get_particles(pic_ctx *ctx, double *p) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_CHECK(p || c->n == 0, PIC_ERR_ARG, "null particle pointer");
    PIC_HIP(hipSetDevice(c->device));
    if (c->n == 0) return PIC_OK;
    Staging s;
    PIC_HIP(hipMalloc(&s.p, (size_t)c->n * 5 * sizeof(double)));
    int prev = c->state;
    PIC_TRY(get_particles_impl(c, (double *)s.p));
    PIC_HIP(hipMemcpyAsync(p, s.p, (size_t)c->n * 5 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_TRY(to_state(c, prev));
    return PIC_OK;
}

int pic_num_particles(pic_ctx *ctx, int64_t *n) {
    PIC_CHECK(ctx && n, PIC_ERR_ARG, "null pointer");
    *n = ctx_of(ctx)->n;
    return PIC_OK;
}

int pic_set_particles_xv(pic_ctx *ctx, const double *x, const double *v, double w, int64_t n) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(x || n == 0, PIC_ERR_ARG, "null position pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    std::vector<double> aos((size_t)n * 5);
    for (int64_t k = 0; k < n; ++k) {
        aos[5 * k + 0] = x[2 * k];
        aos[5 * k + 1] = x[2 * k + 1];
        aos[5 * k + 2] = v ? v[2 * k] : 0.0;
        aos[5 * k + 3] = v ? v[2 * k + 1] : 0.0;
        aos[5 * k + 4] = w;
    }
    return pic_set_particles(ctx, aos.data(), n);
}

int pic_get_particles_xv(pic_ctx *ctx, double *x, double *v) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    std::vector<double> aos((size_t)c->n * 5);
    PIC_TRY(pic_get_particles(ctx, aos.data()));
    for (int64_t k = 0; k < c->n; ++k) {
        if (x) {
            x[2 * k] = aos[5 * k];
            x[2 * k + 1] = aos[5 * k + 1];
        }
        if (v) {
            v[2 * k] = aos[5 * k + 2];
            v[2 * k + 1] = aos[5 * k + 3];
        }
    }
    return PIC_OK;
}

int pic_get_particle_fields(pic_ctx *ctx, double *e) {
    PIC_CHECK(ctx && e, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    PIC_CHECK(c->epx, PIC_ERR_STATE, "no particle fields gathered yet");
    PIC_HIP(hipSetDevice(c->device));
    std::vector<double> a((size_t)c->n), b((size_t)c->n);
    PIC_HIP(hipMemcpyAsync(a.data(), c->epx, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipMemcpyAsync(b.data(), c->epy, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < c->n; ++k) {
        e[2 * k] = a[k];
        e[2 * k + 1] = b[k];
    }
    return PIC_OK;
}

int pic_set_particle_fields(pic_ctx *ctx, const double *e) {
    PIC_CHECK(ctx && e, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_pfields(c));
    std::vector<double> a((size_t)c->n), b((size_t)c->n);
    for (int64_t k = 0; k < c->n; ++k) {
        a[k] = e[2 * k];
        b[k] = e[2 * k + 1];
    }
    PIC_HIP(hipMemcpyAsync(c->epx, a.data(), (size_t)c->n * 8, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(c->epy, b.data(), (size_t)c->n * 8, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_set_state(pic_ctx *ctx, int state) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(state == PIC_STATE_F64 || state == PIC_STATE_F32, PIC_ERR_ARG, "state %d", state);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    return to_state(c, state);
}

int pic_get_state_f32(pic_ctx *ctx, int32_t *cell, float *ox, float *oy, float *vx, float *vy) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_CHECK(c->state == PIC_STATE_F32, PIC_ERR_STATE, "context is not in FP32 state");
    PIC_HIP(hipSetDevice(c->device));
    const int64_t n = c->n;
    size_t b = (size_t)n * 4;
    if (cell) PIC_HIP(hipMemcpyAsync(cell, c->cell, b, hipMemcpyDeviceToHost, c->stream));
    if (n > 0 && (ox || oy || vx || vy)) {
        PIC_TRY(ensure_swap(c, (int64_t)n * 16));
        float *planes = static_cast<float *>(c->swap);
        hipLaunchKernelGGL(k_ov_unpack, grid1(n), dim3(256), 0, c->stream, n, c->ov32, planes);
        PIC_HIP(hipGetLastError());
        float *dst[4] = {ox, oy, vx, vy};
        for (int q = 0; q < 4; ++q)
            if (dst[q])
                PIC_HIP(hipMemcpyAsync(dst[q], planes + (size_t)q * n, b, hipMemcpyDeviceToHost,
                                       c->stream));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_set_state_f32(pic_ctx *ctx, const int32_t *cell, const float *ox, const float *oy,
                      const float *vx, const float *vy, double w, int64_t n) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(n == 0 || (cell && ox && oy && vx && vy), PIC_ERR_ARG, "null state pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_CHECK((c->nx & (c->nx - 1)) == 0 && (c->ny & (c->ny - 1)) == 0, PIC_ERR_STATE,
              "the FP32 cell-relative state needs nx and ny powers of two (%dx%d)", c->nx, c->ny);
    int ncell = c->nx * c->ny;
    for (int64_t k = 0; k < n; ++k) {
        PIC_CHECK(cell[k] >= 0 && cell[k] < ncell, PIC_ERR_ARG, "cell[%lld] = %d out of range",
                  (long long)k, cell[k]);
        PIC_CHECK(ox[k] >= 0.0f && ox[k] < 1.0f && oy[k] >= 0.0f && oy[k] < 1.0f, PIC_ERR_ARG,
                  "offset of particle %lld outside [0,1)", (long long)k);
    }
    PIC_TRY(ensure_particles(c, n));
    PIC_TRY(ensure_f32(c));
    c->n = n;
    c->uniform_w = true;
    c->w0 = w;
    size_t b = (size_t)n * 4;
    if (n > 0) {
        PIC_HIP(hipMemcpyAsync(c->cell, cell, b, hipMemcpyHostToDevice, c->stream));
        PIC_TRY(ensure_swap(c, (int64_t)n * 16));
        float *planes = static_cast<float *>(c->swap);
        const float *src[4] = {ox, oy, vx, vy};
        for (int q = 0; q < 4; ++q)
            PIC_HIP(hipMemcpyAsync(planes + (size_t)q * n, src[q], b, hipMemcpyHostToDevice,
                                   c->stream));
        hipLaunchKernelGGL(k_ov_pack, grid1(n), dim3(256), 0, c->stream, n, planes, c->ov32);
        PIC_HIP(hipGetLastError());
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->state = PIC_STATE_F32;
    c->seg_valid = false;
    c->sort_fresh = false;
    return PIC_OK;
}

static double *grid_ptr(Ctx *c, int which) {
    switch (which) {
        case PIC_GRID_RHO: return c->rho;
        case PIC_GRID_EX: return c->ex;
        case PIC_GRID_EY: return c->ey;
        case PIC_GRID_BZ: return c->bz;
        case PIC_GRID_PHI: return c->phi;
        default: return nullptr;
    }
}

int pic_set_grid(pic_ctx *ctx, int which, const double *host, int ld, int stride) {
    PIC_CHECK(ctx && host, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    double *d = grid_ptr(c, which);
    PIC_CHECK(d, PIC_ERR_ARG, "unknown grid %d", which);
    PIC_CHECK(ld >= c->nx && stride >= 1, PIC_ERR_ARG, "ld=%d stride=%d", ld, stride);
    PIC_HIP(hipSetDevice(c->device));
    std::vector<double> t((size_t)c->nx * c->ny);
    for (int j = 0; j < c->ny; ++j)
        for (int i = 0; i < c->nx; ++i)
            t[(size_t)i + (size_t)c->nx * j] = host[(size_t)stride * ((size_t)i + (size_t)ld * j)];
    PIC_HIP(hipMemcpyAsync(d, t.data(), t.size() * 8, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    if (which == PIC_GRID_BZ) {
        std::vector<float> f(t.size());
        for (size_t k = 0; k < t.size(); ++k) f[k] = (float)t[k];
        PIC_HIP(hipMemcpyAsync(c->bz32, f.data(), f.size() * 4, hipMemcpyHostToDevice,
                               c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
    }
    if (which == PIC_GRID_EX || which == PIC_GRID_EY) {
        // keep the FP32 interleaved copy in step with host-provided fields
        std::vector<double> o((size_t)c->nx * c->ny);
        PIC_HIP(hipMemcpyAsync(o.data(), which == PIC_GRID_EX ? c->ey : c->ex, o.size() * 8,
                               hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        std::vector<float> f(2 * t.size());
        for (size_t k = 0; k < t.size(); ++k) {
            double vx = which == PIC_GRID_EX ? t[k] : o[k];
            double vy = which == PIC_GRID_EX ? o[k] : t[k];
            f[2 * k] = (float)vx;
            f[2 * k + 1] = (float)vy;
        }
        PIC_HIP(hipMemcpyAsync(c->e32, f.data(), f.size() * 4, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
    }
    return PIC_OK;
}

int pic_get_grid(pic_ctx *ctx, int which, double *host, int ld, int stride, int ghost) {
    PIC_CHECK(ctx && host, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    double *d = grid_ptr(c, which);
    PIC_CHECK(d, PIC_ERR_ARG, "unknown grid %d", which);
    PIC_CHECK(ld >= c->nx + (ghost ? 1 : 0) && stride >= 1, PIC_ERR_ARG, "ld=%d stride=%d", ld,
              stride);
    PIC_HIP(hipSetDevice(c->device));
    std::vector<double> t((size_t)c->nx * c->ny);
    PIC_HIP(hipMemcpyAsync(t.data(), d, t.size() * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    int nx = c->nx, ny = c->ny;
    auto at = [&](int i, int j) -> double & {
        return host[(size_t)stride * ((size_t)i + (size_t)ld * j)];
    };
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) at(i, j) = t[(size_t)i + (size_t)nx * j];
    if (ghost) {
        for (int j = 0; j < ny; ++j) at(nx, j) = at(0, j);
        for (int i = 0; i <= nx; ++i) at(i, ny) = at(i, 0);
    }
    return PIC_OK;
}

int pic_grid_device_ptr(pic_ctx *ctx, int which, double **d_ptr) {
    PIC_CHECK(ctx && d_ptr, PIC_ERR_ARG, "null pointer");
    double *d = grid_ptr(ctx_of(ctx), which);
    PIC_CHECK(d, PIC_ERR_ARG, "unknown grid %d", which);
    *d_ptr = d;
    return PIC_OK;
}

int pic_set_bz(pic_ctx *ctx, int mode, double bz0) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(mode == PIC_BZ_NONE || mode == PIC_BZ_UNIFORM || mode == PIC_BZ_GRID, PIC_ERR_ARG,
              "bz mode %d", mode);
    Ctx *c = ctx_of(ctx);
    c->bz_mode = mode;
    c->bz0 = mode == PIC_BZ_UNIFORM ? bz0 : 0.0;
    return PIC_OK;
}

int pic_slow_chunks(pic_ctx *ctx, uint64_t *count) {
    PIC_CHECK(ctx && count, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    unsigned long long h = 0;
    PIC_HIP(hipMemcpyAsync(&h, c->slow_chunks, sizeof h, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    *count = h;
    return PIC_OK;
}

// ---------------------------------------------- Fortran-twin drop-ins

// One private context per process, rebuilt when the mesh changes.
static pic_ctx *g_f90 = nullptr;
static int g_f90_nx = 0, g_f90_ny = 0;
static double g_f90_dx = 0, g_f90_dy = 0;

static int f90_ctx(int nx, int ny, double dx, double dy) {
    if (g_f90 && g_f90_nx == nx && g_f90_ny == ny && g_f90_dx == dx && g_f90_dy == dy)
        return PIC_OK;
    if (g_f90) pic_ctx_destroy(g_f90);
    g_f90 = nullptr;
    PIC_TRY(pic_ctx_create(&g_f90, 0, nx, ny, 0.0, dx * nx, 0.0, dy * ny, 0));
    // the mesh spacing the caller passes is used verbatim (pic.f90 reads dx, dy)
    Ctx *c = ctx_of(g_f90);
    c->dx = dx;
    c->dy = dy;
    g_f90_nx = nx;
    g_f90_ny = ny;
    g_f90_dx = dx;
    g_f90_dy = dy;
    return PIC_OK;
}

int pic_f90_push_v(const int32_t *nbpart, double *p, const double *dt, const int32_t *nx,
                   const int32_t *ny, const double *dx, const double *dy, const double *ex,
                   const double *ey, const double *bz) {
    PIC_CHECK(nbpart && p && dt && nx && ny && dx && dy && ex && ey && bz, PIC_ERR_ARG,
              "null argument");
    PIC_TRY(f90_ctx(*nx, *ny, *dx, *dy));
    PIC_TRY(pic_set_particles(g_f90, p, *nbpart));
    PIC_TRY(pic_set_grid(g_f90, PIC_GRID_EX, ex, *nx + 1, 1));
    PIC_TRY(pic_set_grid(g_f90, PIC_GRID_EY, ey, *nx + 1, 1));
    PIC_TRY(pic_set_grid(g_f90, PIC_GRID_BZ, bz, *nx + 1, 1));
    PIC_TRY(pic_set_bz(g_f90, PIC_BZ_GRID, 0.0));
    PIC_TRY(pic_push_v(g_f90, PIC_CIC, *dt));
    return pic_get_particles(g_f90, p);
}

int pic_f90_push_x(const int32_t *nbpart, const double *dimx, const double *dimy, double *p,
                   const double *dt) {
    PIC_CHECK(nbpart && dimx && dimy && p && dt, PIC_ERR_ARG, "null argument");
    // push_x only needs the domain size: reuse the private context whatever
    // its mesh (push_v never reads dimx, dimy), or make a nominal 8x8 one
    if (!g_f90) PIC_TRY(f90_ctx(8, 8, *dimx / 8, *dimy / 8));
    Ctx *c = ctx_of(g_f90);
    c->dimx = *dimx;
    c->dimy = *dimy;
    PIC_TRY(pic_set_particles(g_f90, p, *nbpart));
    PIC_TRY(pic_push_x(g_f90, *dt));
    return pic_get_particles(g_f90, p);
}

}  // extern "C"

extern "C" int pic_set_f32_accumulation(pic_ctx *ctx, int mode) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(mode >= 0 && mode <= 2, PIC_ERR_ARG, "accumulation mode %d", mode);
    ctx_of(ctx)->f32_acc = mode;
    return PIC_OK;
}

extern "C" int pic_set_f32_window(pic_ctx *ctx, int kib) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(kib >= 40 && kib <= 160, PIC_ERR_ARG, "window of %d KiB (40 to 160)", kib);
    ctx_of(ctx)->f32_window_kib = kib;
    return PIC_OK;
}

extern "C" int pic_set_f32_tiles(pic_ctx *ctx, int tiles) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(tiles == 1 || tiles == 2 || tiles == 4, PIC_ERR_ARG, "tile copies %d (1, 2 or 4)",
              tiles);
    ctx_of(ctx)->f32_tiles = tiles;
    return PIC_OK;
}
