// pic_poisson.hip -- periodic 2D Poisson solve on rocFFT plus a hand-written
// spectral kernel, and the FP64 grid reductions (charge, field energy).
//
// E = -grad(phi), -lap(phi) = rho (poisson_2d.jl:97-113):
//   rhat = rfft(rho)                      rocFFT r2c, lengths {nx, ny}
//   ehat_x = -i kx/k2 rhat, ehat_y = -i ky/k2 rhat, rhat[0,0] = 0
//   ex, ey = irfft(ehat)                  one batched rocFFT c2r of 2
// The spectral kernel also folds in the 1/(nx*ny) normalisation and the
// charge normalisation 1/(dx*dy) of the raw deposit, and replaces the columns
// ik = 0 and ik = nx/2 (self-conjugate in x) by their Hermitian part
// (F[jk] + conj(F[-jk]))/2.  Multiplying a Hermitian column by -i*k makes it
// anti-Hermitian there; FFTW's c2r (and numpy's) keep only the real part of
// the inverse along y for those columns, which is exactly this projection, so
// rocFFT receives data it is defined on and returns what FFTW returns.
#include <cstdio>
#include <cmath>
#include <vector>

#include "pic_internal.h"

namespace pic {

#define PIC_FFT(call)                                                                    \
    do {                                                                                 \
        rocfft_status s_ = (call);                                                       \
        if (s_ != rocfft_status_success)                                                 \
            return fail(PIC_ERR_FFT, "%s failed: rocfft status %d", #call, (int)s_);     \
    } while (0)

static bool g_rocfft_ready = false;

static int make_plan(rocfft_plan *plan, rocfft_transform_type t, int nx, int ny, int batch) {
    size_t lengths[2] = {(size_t)nx, (size_t)ny};  // fastest dimension first
    PIC_FFT(rocfft_plan_create(plan, rocfft_placement_notinplace, t, rocfft_precision_double, 2,
                               lengths, (size_t)batch, nullptr));
    return PIC_OK;
}

int poisson_init(Ctx *c) {
    if (!g_rocfft_ready) {
        PIC_FFT(rocfft_setup());
        g_rocfft_ready = true;
    }
    const int nx = c->nx, ny = c->ny, h = nx / 2 + 1;
    const size_t m = (size_t)h * ny;
    // TwoDPoissonPeriodic tables (poisson_2d.jl:28-61), same arithmetic as the
    // reference: kx1 = (ik-1)*kx0, negative ky from jk = ny/2+1, kx[1,1] = 1.
    std::vector<double> kx(m), ky(m), k2(m);
    const double kx0 = 2 * M_PI / c->dimx, ky0 = 2 * M_PI / c->dimy;
    for (int ik = 1; ik <= h; ++ik) {
        double kx1 = (ik - 1) * kx0;
        for (int jk = 1; jk <= ny; ++jk) {
            size_t k = (size_t)(ik - 1) + (size_t)h * (jk - 1);
            kx[k] = kx1;
            ky[k] = jk <= ny / 2 ? (jk - 1) * ky0 : (jk - 1 - ny) * ky0;
        }
    }
    kx[0] = 1.0;
    for (size_t k = 0; k < m; ++k) k2[k] = kx[k] * kx[k] + ky[k] * ky[k];
    for (size_t k = 0; k < m; ++k) {
        kx[k] = kx[k] / k2[k];
        ky[k] = ky[k] / k2[k];
    }
    PIC_HIP(hipMalloc(&c->kx_k2, m * 8));
    PIC_HIP(hipMalloc(&c->ky_k2, m * 8));
    PIC_HIP(hipMalloc(&c->k2, m * 8));
    PIC_HIP(hipMalloc(&c->rhat, 2 * m * sizeof(double2)));
    PIC_HIP(hipMalloc(&c->ehat, 2 * m * sizeof(double2)));
    PIC_HIP(hipMemcpy(c->kx_k2, kx.data(), m * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->ky_k2, ky.data(), m * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->k2, k2.data(), m * 8, hipMemcpyHostToDevice));
    PIC_TRY(make_plan(&c->plan_r2c, rocfft_transform_type_real_forward, nx, ny, 1));
    PIC_TRY(make_plan(&c->plan_r2c2, rocfft_transform_type_real_forward, nx, ny, 2));
    PIC_TRY(make_plan(&c->plan_c2r2, rocfft_transform_type_real_inverse, nx, ny, 2));
    PIC_TRY(make_plan(&c->plan_c2r1, rocfft_transform_type_real_inverse, nx, ny, 1));
    size_t wb = 0;
    for (rocfft_plan p : {c->plan_r2c, c->plan_r2c2, c->plan_c2r2, c->plan_c2r1}) {
        size_t s = 0;
        PIC_FFT(rocfft_plan_get_work_buffer_size(p, &s));
        if (s > wb) wb = s;
    }
    c->fft_work_bytes = wb;
    if (wb) PIC_HIP(hipMalloc(&c->fft_work, wb));
    PIC_FFT(rocfft_execution_info_create(&c->fft_info));
    PIC_FFT(rocfft_execution_info_set_stream(c->fft_info, c->stream));
    if (wb) PIC_FFT(rocfft_execution_info_set_work_buffer(c->fft_info, c->fft_work, wb));
    return PIC_OK;
}

void poisson_free(Ctx *c) {
    for (rocfft_plan *p : {&c->plan_r2c, &c->plan_r2c2, &c->plan_c2r2, &c->plan_c2r1})
        if (*p) {
            rocfft_plan_destroy(*p);
            *p = nullptr;
        }
    if (c->fft_info) rocfft_execution_info_destroy(c->fft_info);
    c->fft_info = nullptr;
    for (void **q : {(void **)&c->kx_k2, (void **)&c->ky_k2, (void **)&c->k2, (void **)&c->rhat,
                     (void **)&c->ehat, &c->fft_work})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
}

static int exec(Ctx *c, rocfft_plan plan, void *in, void *out) {
    void *ib[1] = {in}, *ob[1] = {out};
    PIC_FFT(rocfft_execute(plan, ib, ob, c->fft_info));
    return PIC_OK;
}

// -i * m * v * scale for complex v = (a, b): (m*b*scale, -m*a*scale).
__device__ __forceinline__ double2 mul_mi(double m, double2 v, double scale) {
    return make_double2(m * v.y * scale, -(m * v.x) * scale);
}

// mode 0: E spectra (x into out[0..m), y into out[m..2m)), DC zeroed.
// mode 1: phi spectrum rhat/k2 into out[0..m), DC passed through.
// sign multiplies the result (-1 for the adjoint: the E operator is a real
// convolution with a purely imaginary, odd symbol, so its transpose is its
// negative).
__global__ void k_spectral(const double2 *__restrict__ rhat, double2 *__restrict__ out,
                           const double *__restrict__ kxk2, const double *__restrict__ kyk2,
                           const double *__restrict__ k2, int nx, int ny, double scale,
                           int mode) {
    const int h = nx / 2 + 1;
    const int64_t m = (int64_t)h * ny;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int ik = (int)(t % h), jk = (int)(t / h);
    const bool selfconj = (ik == 0) || (ik == nx / 2);
    const int jm = (ny - jk) % ny;
    const int64_t tm = (int64_t)ik + (int64_t)h * jm;
    double2 r = rhat[t];
    double2 rm = rhat[tm];
    if (mode == 0) {
        if (t == 0) r = make_double2(0.0, 0.0);
        if (tm == 0) rm = make_double2(0.0, 0.0);
        double2 fx = mul_mi(kxk2[t], r, scale), fy = mul_mi(kyk2[t], r, scale);
        if (selfconj) {
            double2 gx = mul_mi(kxk2[tm], rm, scale), gy = mul_mi(kyk2[tm], rm, scale);
            fx = make_double2(0.5 * (fx.x + gx.x), 0.5 * (fx.y - gx.y));
            fy = make_double2(0.5 * (fy.x + gy.x), 0.5 * (fy.y - gy.y));
        }
        out[t] = fx;
        out[m + t] = fy;
    } else {
        double2 f = make_double2(r.x / k2[t] * scale, r.y / k2[t] * scale);
        if (selfconj) {
            double2 g = make_double2(rm.x / k2[tm] * scale, rm.y / k2[tm] * scale);
            f = make_double2(0.5 * (f.x + g.x), 0.5 * (f.y - g.y));
        }
        out[t] = f;
    }
}

// Adjoint spectral step: out = -(mx * X + my * Y) with X, Y the spectra of the
// two field adjoints (in[0..m), in[m..2m)), projected like k_spectral.
__global__ void k_spectral_adj(const double2 *__restrict__ in, double2 *__restrict__ out,
                               const double *__restrict__ kxk2, const double *__restrict__ kyk2,
                               int nx, int ny, double scale) {
    const int h = nx / 2 + 1;
    const int64_t m = (int64_t)h * ny;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int ik = (int)(t % h), jk = (int)(t / h);
    const bool selfconj = (ik == 0) || (ik == nx / 2);
    const int jm = (ny - jk) % ny;
    const int64_t tm = (int64_t)ik + (int64_t)h * jm;
    auto val = [&](int64_t q) {
        if (q == 0) return make_double2(0.0, 0.0);
        double2 a = mul_mi(kxk2[q], in[q], scale), b = mul_mi(kyk2[q], in[m + q], scale);
        return make_double2(-(a.x + b.x), -(a.y + b.y));
    };
    double2 f = val(t);
    if (selfconj) {
        double2 g = val(tm);
        f = make_double2(0.5 * (f.x + g.x), 0.5 * (f.y - g.y));
    }
    out[t] = f;
}

// Block-wide sum of two doubles in a fixed order (deterministic).
template <int T>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *sh) {
    const int tid = threadIdx.x;
    sh[tid] = a;
    sh[T + tid] = b;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (tid < s) {
            sh[tid] += sh[tid + s];
            sh[T + tid] += sh[T + tid + s];
        }
        __syncthreads();
    }
    a = sh[0];
    b = sh[T];
    __syncthreads();
}

// After the c2r: optional FP32 interleaved copy of (ex, ey), optional zeroing
// of the raw deposit accumulator for the next step, and the energy sums
// sum(ex^2)*dxdy, sum(ey^2)*dxdy written by the last block to arrive into
// hist[2*slot] (slot >= 0), hist[2*(*counter)] with *counter incremented
// (slot == -2) or nowhere (slot == -1).  Partial sums are per block over a
// fixed node subset and combined in block order, so the result is the same
// on every run.
__global__ __launch_bounds__(RED_THREADS) void k_post(const double *__restrict__ ex,
                                                      const double *__restrict__ ey, int64_t m,
                                                      float2 *__restrict__ e32,
                                                      double *__restrict__ zero_acc,
                                                      double *part, unsigned *count, double *hist,
                                                      int slot, int *counter, double dxdy) {
    __shared__ double sh[2 * RED_THREADS];
    __shared__ bool last;
    double sx = 0.0, sy = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * RED_THREADS) {
        double a = ex[k], b = ey[k];
        sx += a * a;
        sy += b * b;
        if (e32) e32[k] = make_float2((float)a, (float)b);
        if (zero_acc) zero_acc[k] = 0.0;
    }
    block_sum2<RED_THREADS>(sx, sy, sh);
    if (slot == -1) return;
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sx;
        part[2 * blockIdx.x + 1] = sy;
        __threadfence();
        unsigned ticket = atomicAdd(count, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += RED_THREADS) {
        a += __hip_atomic_load(&part[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b += __hip_atomic_load(&part[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum2<RED_THREADS>(a, b, sh);
    if (threadIdx.x == 0) {
        int s = slot;
        if (slot == -2) {
            s = *counter;
            *counter = s + 1;
        }
        hist[2 * s] = a * dxdy;
        hist[2 * s + 1] = b * dxdy;
        *count = 0u;
    }
}

// post: bit 0 writes the FP32 interleaved field copy, bit 1 clears the raw
// deposit accumulator (after the r2c has read it).
int poisson_solve_e(Ctx *c, const double *rho_in, double scale, int post, int energy_slot) {
    const int h = c->nx / 2 + 1;
    const int64_t m = (int64_t)h * c->ny;
    const int64_t mm = (int64_t)c->nx * c->ny;
    PIC_TRY(exec(c, c->plan_r2c, (void *)rho_in, c->rhat));
    hipLaunchKernelGGL(k_spectral, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                       c->rhat, c->ehat, c->kx_k2, c->ky_k2, c->k2, c->nx, c->ny, scale, 0);
    PIC_HIP(hipGetLastError());
    PIC_TRY(exec(c, c->plan_c2r2, c->ehat, c->ex));
    if (post || energy_slot != -1) {
        int blocks = (int)((mm + RED_THREADS - 1) / RED_THREADS);
        if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
        hipLaunchKernelGGL(k_post, dim3(blocks), dim3(RED_THREADS), 0, c->stream, c->ex, c->ey, mm,
                           (post & 1) ? (float2 *)c->e32 : nullptr,
                           (post & 2) ? c->rho_acc : nullptr, c->red_part, c->red_count,
                           c->energy_hist, energy_slot, c->step_counter, c->dx * c->dy);
        PIC_HIP(hipGetLastError());
    }
    return PIC_OK;
}

int poisson_solve_phi(Ctx *c) {
    const int h = c->nx / 2 + 1;
    const int64_t m = (int64_t)h * c->ny;
    PIC_TRY(exec(c, c->plan_r2c, c->rho, c->rhat));
    double scale = 1.0 / ((double)c->nx * (double)c->ny);
    hipLaunchKernelGGL(k_spectral, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                       c->rhat, c->ehat, c->kx_k2, c->ky_k2, c->k2, c->nx, c->ny, scale, 1);
    PIC_HIP(hipGetLastError());
    PIC_TRY(exec(c, c->plan_c2r1, c->ehat, c->phi));
    return PIC_OK;
}

// lrho = E^T (lex, ley): two r2c, the transposed spectral symbol, one c2r.
// lex and ley must be contiguous (ley = lex + nx*ny).
int poisson_adjoint(Ctx *c, const double *lex, const double *ley, double *lrho, double scale) {
    const int h = c->nx / 2 + 1;
    const int64_t m = (int64_t)h * c->ny;
    PIC_CHECK(ley == lex + (int64_t)c->nx * c->ny, PIC_ERR_ARG, "adjoint fields not contiguous");
    PIC_TRY(exec(c, c->plan_r2c2, (void *)lex, c->ehat));
    hipLaunchKernelGGL(k_spectral_adj, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                       c->ehat, c->rhat, c->kx_k2, c->ky_k2, c->nx, c->ny, scale);
    PIC_HIP(hipGetLastError());
    PIC_TRY(exec(c, c->plan_c2r1, c->rhat, lrho));
    return PIC_OK;
}

// ------------------------------------------------------ charge reductions

// Sum of a (nx*ny) into *out via the same two-level fixed-order scheme.
__global__ __launch_bounds__(RED_THREADS) void k_sum(const double *__restrict__ a, int64_t m,
                                                     double *part, unsigned *count,
                                                     double *out) {
    __shared__ double sh[2 * RED_THREADS];
    __shared__ bool last;
    double s = 0.0, z = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * RED_THREADS)
        s += a[k];
    block_sum2<RED_THREADS>(s, z, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
        __threadfence();
        unsigned ticket = atomicAdd(count, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += RED_THREADS)
        t += __hip_atomic_load(&part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    block_sum2<RED_THREADS>(t, z, sh);
    if (threadIdx.x == 0) {
        *out = t;
        *count = 0u;
    }
}

// rho = acc / (dx*dy) - shift, with shift from the normalised total:
// CIC: mean(rho) (pushers_cic.jl:109-111); M6: (sum(rho)*dx*dy)/dimx/dimy
// (pushers_m6.jl:342-346).
__global__ void k_finalize_rho(const double *__restrict__ acc, double *__restrict__ rho, int64_t m,
                               double dxdy, const double *total, int kernel, double dx, double dy,
                               double dimx, double dimy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    // *total is the sum of the raw accumulator; the reference sums the
    // normalised grid, sum(acc/dxdy) == sum(acc)/dxdy up to rounding.
    double tot = *total / dxdy;
    double shift = kernel == PIC_CIC ? tot / (double)m : (tot * dx * dy) / dimx / dimy;
    rho[k] = acc[k] / dxdy - shift;
}

int finalize_rho(Ctx *c, int kernel) {
    const int64_t m = (int64_t)c->nx * c->ny;
    int blocks = (int)((m + RED_THREADS - 1) / RED_THREADS);
    if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
    double *tot = c->red_part + 2 * RED_BLOCKS;  // spare slot after the partials
    hipLaunchKernelGGL(k_sum, dim3(blocks), dim3(RED_THREADS), 0, c->stream, c->rho_acc, m,
                       c->red_part, c->red_count, tot);
    hipLaunchKernelGGL(k_finalize_rho, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                       c->rho_acc, c->rho, m, c->dx * c->dy, tot, kernel, c->dx, c->dy, c->dimx,
                       c->dimy);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int field_energy(Ctx *c, double *out2) {
    PIC_TRY(ensure_energy(c, 1));
    const int64_t mm = (int64_t)c->nx * c->ny;
    int blocks = (int)((mm + RED_THREADS - 1) / RED_THREADS);
    if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
    hipLaunchKernelGGL(k_post, dim3(blocks), dim3(RED_THREADS), 0, c->stream, c->ex, c->ey, mm,
                       nullptr, nullptr, c->red_part, c->red_count, c->energy_hist, 0,
                       c->step_counter, c->dx * c->dy);
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipMemcpyAsync(out2, c->energy_hist, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_solve_e(pic_ctx *ctx) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(poisson_solve_e(c, c->rho, 1.0 / ((double)c->nx * (double)c->ny), 1, -1));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_solve_phi(pic_ctx *ctx) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(poisson_solve_phi(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_field_energy(pic_ctx *ctx, double out[2]) {
    PIC_CHECK(ctx && out, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    return field_energy(c, out);
}

int pic_solve_fields(pic_ctx *ctx, double *energy) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(poisson_solve_e(c, c->rho, 1.0 / ((double)c->nx * (double)c->ny), 1, -1));
    if (energy) {
        // meshfields.jl:105-106 sums over the (nx+1)(ny+1) ghosted nodes: the
        // interior, plus column i = 0 and row j = 0 again, plus node (0,0) a
        // third time (the ghost corner).
        double s[2];
        PIC_TRY(field_energy(c, s));
        const int nx = c->nx, ny = c->ny;
        std::vector<double> ex((size_t)nx * ny), ey((size_t)nx * ny);
        PIC_HIP(hipMemcpyAsync(ex.data(), c->ex, ex.size() * 8, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipMemcpyAsync(ey.data(), c->ey, ey.size() * 8, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        double g = 0.0;
        for (int j = 0; j < ny; ++j) {
            size_t k = (size_t)nx * j;
            g += ex[k] * ex[k] + ey[k] * ey[k];
        }
        for (int i = 0; i < nx; ++i) g += ex[i] * ex[i] + ey[i] * ey[i];
        g += ex[0] * ex[0] + ey[0] * ey[0];
        *energy = s[0] + s[1] + g * c->dx * c->dy;
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // extern "C"
