// pic_poisson.hip -- periodic 2D Poisson solve on rocFFT plus a hand-written
// spectral kernel, and the FP64 grid reductions (charge, field energy).
//
// E = -grad(phi), -lap(phi) = rho (poisson_2d.jl:97-113):
//   R = fft2(rho)                           rocFFT complex-to-complex forward
//   half spectrum (ik <= nx/2): Fx = -i kx/k2 R, Fy = -i ky/k2 R, R[0,0] = 0
//   (ex, ey) = (Re, Im) ifft2(Fx + i Fy)    rocFFT complex-to-complex inverse
// The spectral kernel extends the half spectrum of the reference's rfft to
// the full grid by Hermitian symmetry, F(ik, jk) = conj(F(nx-ik, -jk)) for
// ik > nx/2, which is the array irfft inverts; because both extended spectra
// are Hermitian, ifft2(Fx + i Fy) carries ex in its real and ey in its
// imaginary part.  On the columns ik = 0 and ik = nx/2 (self-conjugate in x)
// the half spectrum is first replaced by its Hermitian part
// (F[jk] + conj(F[-jk]))/2: multiplying a Hermitian column by -i*k makes it
// anti-Hermitian there, and FFTW's c2r keeps only the real part of the
// inverse along y for those columns, which is this projection.  The kernel
// also folds in the 1/(nx*ny) normalisation and, in the time loop, the
// 1/(dx*dy) charge normalisation of the raw deposit.
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <type_traits>
#include <vector>

#include "pic_fftlanes.h"
#include "pic_internal.h"

namespace pic {

#define PIC_FFT(call)                                                                    \
    do {                                                                                 \
        rocfft_status s_ = (call);                                                       \
        if (s_ != rocfft_status_success)                                                 \
            return fail(PIC_ERR_FFT, "%s failed: rocfft status %d", #call, (int)s_);     \
    } while (0)

static bool g_rocfft_ready = false;

static int make_plan(rocfft_plan *plan, rocfft_transform_type t, int nx, int ny) {
    size_t lengths[2] = {(size_t)nx, (size_t)ny};  // fastest dimension first
    PIC_FFT(rocfft_plan_create(plan, rocfft_placement_notinplace, t, rocfft_precision_double, 2,
                               lengths, 1, nullptr));
    return PIC_OK;
}

static int fft_tw_init(Ctx *c);

int poisson_init(Ctx *c) {
    if (!g_rocfft_ready) {
        PIC_FFT(rocfft_setup());
        g_rocfft_ready = true;
    }
    const int nx = c->nx, ny = c->ny, h = nx / 2 + 1;
    const size_t m = (size_t)h * ny, mm = (size_t)nx * ny;
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
    PIC_HIP(hipMalloc(&c->zf, mm * sizeof(double2)));
    PIC_HIP(hipMalloc(&c->zg, mm * sizeof(double2)));
    PIC_HIP(hipMemcpy(c->kx_k2, kx.data(), m * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->ky_k2, ky.data(), m * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->k2, k2.data(), m * 8, hipMemcpyHostToDevice));
    PIC_TRY(fft_tw_init(c));
    PIC_TRY(make_plan(&c->plan_fwd, rocfft_transform_type_complex_forward, nx, ny));
    PIC_TRY(make_plan(&c->plan_inv, rocfft_transform_type_complex_inverse, nx, ny));
    size_t wb = 0;
    for (rocfft_plan p : {c->plan_fwd, c->plan_inv}) {
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
    for (rocfft_plan *p : {&c->plan_fwd, &c->plan_inv})
        if (*p) {
            rocfft_plan_destroy(*p);
            *p = nullptr;
        }
    if (c->fft_info) rocfft_execution_info_destroy(c->fft_info);
    c->fft_info = nullptr;
    for (void **q : {(void **)&c->kx_k2, (void **)&c->ky_k2, (void **)&c->k2, (void **)&c->zf,
                     (void **)&c->zg, (void **)&c->fft_tw, &c->fft_work})
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

// z = (a, b) (b may be NULL: imaginary part 0).
__global__ void k_pack(const double *__restrict__ a, const double *__restrict__ b, int64_t m,
                       double2 *__restrict__ z) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) z[k] = make_double2(a[k], b ? b[k] : 0.0);
}

__global__ void k_real_part(const double2 *__restrict__ z, int64_t m, double *__restrict__ out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) out[k] = z[k].x;
}

// -i * m * v * scale for complex v = (a, b): (m*b*scale, -m*a*scale).
__device__ __forceinline__ double2 mul_mi(double m, double2 v, double scale) {
    return make_double2(m * v.y * scale, -(m * v.x) * scale);
}

__device__ __forceinline__ double2 cconj(double2 v) { return make_double2(v.x, -v.y); }

constexpr int SPEC_E = 0, SPEC_PHI = 1, SPEC_ADJ = 2;

// Spectrum multiplier on the half spectrum index (ih <= nx/2, jh) of the
// full forward transform S (nx x ny, x fastest); tables are (nx/2+1) x ny.
//   SPEC_E:   (Fx, Fy) = -i (kx, ky)/k2 R, R[0,0] = 0
//   SPEC_PHI:  Fphi = R / k2                         (no DC zeroing)
//   SPEC_ADJ:  -(-i kx/k2 Mx - i ky/k2 My) with S = fft(mux + i muy),
//              Mx = (S(k) + conj S(-k))/2, My = (S(k) - conj S(-k))/(2i)
template <int MODE>
__device__ __forceinline__ void half_value(const double2 *__restrict__ S, int nx, int ny, int ih,
                                           int jh, const double *__restrict__ kxk2,
                                           const double *__restrict__ kyk2,
                                           const double *__restrict__ k2, double scale,
                                           double2 &f, double2 &g) {
    const int h = nx / 2 + 1;
    const int64_t q = (int64_t)ih + (int64_t)h * jh;
    const double2 R = S[(int64_t)ih + (int64_t)nx * jh];
    if (MODE == SPEC_E) {
        if (q == 0) {
            f = g = make_double2(0.0, 0.0);
        } else {
            f = mul_mi(kxk2[q], R, scale);
            g = mul_mi(kyk2[q], R, scale);
        }
    } else if (MODE == SPEC_PHI) {
        f = make_double2(R.x / k2[q] * scale, R.y / k2[q] * scale);
        g = make_double2(0.0, 0.0);
    } else {
        const int im = (nx - ih) % nx, jm = (ny - jh) % ny;
        const double2 Rm = cconj(S[(int64_t)im + (int64_t)nx * jm]);
        const double2 mx = make_double2(0.5 * (R.x + Rm.x), 0.5 * (R.y + Rm.y));
        // (R - Rm) / (2i) = (Im, -Re) / 2 of the difference
        const double2 my = make_double2(0.5 * (R.y - Rm.y), -0.5 * (R.x - Rm.x));
        if (q == 0) {
            f = make_double2(0.0, 0.0);
        } else {
            const double2 a = mul_mi(kxk2[q], mx, scale), b = mul_mi(kyk2[q], my, scale);
            f = make_double2(-(a.x + b.x), -(a.y + b.y));
        }
        g = make_double2(0.0, 0.0);
    }
}

// Full-grid spectrum Z for the inverse transform (see the file comment).
template <int MODE>
__global__ void k_spectral(const double2 *__restrict__ S, double2 *__restrict__ Z,
                           const double *__restrict__ kxk2, const double *__restrict__ kyk2,
                           const double *__restrict__ k2, int nx, int ny, double scale) {
    const int64_t mm = (int64_t)nx * ny;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mm) return;
    const int ik = (int)(t % nx), jk = (int)(t / nx);
    const bool mirror = ik > nx / 2;
    const int ih = mirror ? nx - ik : ik;
    const int jh = mirror ? (ny - jk) % ny : jk;
    double2 f, g;
    half_value<MODE>(S, nx, ny, ih, jh, kxk2, kyk2, k2, scale, f, g);
    if (ih == 0 || 2 * ih == nx) {  // project the self-conjugate column
        double2 fm, gm;
        half_value<MODE>(S, nx, ny, ih, (ny - jh) % ny, kxk2, kyk2, k2, scale, fm, gm);
        f = make_double2(0.5 * (f.x + fm.x), 0.5 * (f.y - fm.y));
        g = make_double2(0.5 * (g.x + gm.x), 0.5 * (g.y - gm.y));
    }
    if (mirror) {
        f = cconj(f);
        g = cconj(g);
    }
    Z[t] = make_double2(f.x - g.y, f.y + g.x);  // f + i g
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

// After the inverse transform: (ex, ey) from z (real, imaginary parts) when z
// is given, else read from ex, ey; optional FP32 interleaved copy, optional
// zeroing of the raw deposit accumulator for the next step, and the energy
// sums sum(ex^2)*dxdy, sum(ey^2)*dxdy written by the last block to arrive
// into hist[2*slot] (slot >= 0), hist[2*(*counter)] with *counter
// incremented (slot == -2) or nowhere (slot == -1).  Partial sums are per
// block over a fixed node subset and combined in block order, so the result
// is the same on every run.
__global__ __launch_bounds__(RED_THREADS) void k_post(const double2 *__restrict__ z,
                                                      double *__restrict__ ex,
                                                      double *__restrict__ ey, int64_t m,
                                                      float2 *__restrict__ e32,
                                                      double *__restrict__ zero_acc,
                                                      double *part, unsigned *count, double *hist,
                                                      int slot, int *counter, double dxdy) {
    __shared__ double sh[2 * RED_THREADS];
    __shared__ bool last;
    double sx = 0.0, sy = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * RED_THREADS) {
        double a, b;
        if (z) {
            const double2 v = z[k];
            a = v.x;
            b = v.y;
            ex[k] = a;
            ey[k] = b;
        } else {
            a = ex[k];
            b = ey[k];
        }
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

static inline dim3 grid_of(int64_t m) { return dim3((unsigned)((m + 255) / 256)); }

// ------------------------------------------- fused solve for small grids
//
// For power-of-two nx, ny <= 512 the E solve runs as three kernels instead of
// the rocFFT sequence (seven launches of a few microseconds each at these
// sizes); with 64 <= nx, ny the transforms run across the lanes of a wave
// (k_wrows_fwd, k_wcols_e, k_wrows_inv below), on smaller grids in LDS:
//   k_fft_rows_fwd   per row: forward FFT along x of rho (imaginary part 0);
//                    keeps the half spectrum ik <= nx/2, column-major
//   k_fft_cols_e     per half-spectrum column ih: forward FFT along y, the
//                    spectral values of k_spectral<SPEC_E> for column ih and
//                    its mirror nx - ih (which depends on column ih only), and
//                    the inverse FFTs along y of both
//   k_fft_rows_inv   per row: inverse FFT along x; ex = Re, ey = Im, with
//                    k_post's FP32 copy, accumulator clearing and energy sums
// Same transforms (unnormalised, exp(-+2 pi i jk/N)) and the same spectral
// arithmetic as the rocFFT path; results agree to rounding.
constexpr int FFT_T = 64;
constexpr int FFT_MAX_LOG = 9;

// w[k] = exp(sign 2 pi i k / N), k < N/2, computed once per context into
// Ctx::fft_tw (k_fft_tw_init) and copied into LDS by each block.
__device__ void fft_twiddles(double2 *w, int logN, int sign) {
    const int N = 1 << logN;
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x) {
        double sn, cs;
        sincospi((double)(sign * 2 * k) / (double)N, &sn, &cs);
        w[k] = make_double2(cs, sn);
    }
}
__global__ void k_fft_tw_init(double2 *tw, int lognx, int logny) {
    const int nx = 1 << lognx, ny = 1 << logny;
    fft_twiddles(tw, lognx, -1);
    fft_twiddles(tw + nx / 2, lognx, 1);
    fft_twiddles(tw + nx, logny, -1);
    fft_twiddles(tw + nx + ny / 2, logny, 1);
}
__device__ void fft_twiddles_from(double2 *w, const double2 *__restrict__ tab, int N) {
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x) w[k] = tab[k];
}

// In-place radix-2 decimation-in-time FFT of x[0, N) in LDS by the block's
// threads, with twiddles w[k] = exp(sign 2 pi i k / N).  Ends synchronised.
__device__ void fft_lds(double2 *x, const double2 *w, int logN) {
    const int N = 1 << logN, tid = threadIdx.x, T = blockDim.x;
    for (int i = tid; i < N; i += T) {
        const int r = (int)(__brev((unsigned)i) >> (32 - logN));
        if (i < r) {
            const double2 t = x[i];
            x[i] = x[r];
            x[r] = t;
        }
    }
    __syncthreads();
    for (int s = 1; s <= logN; ++s) {
        const int h = 1 << (s - 1), stride = N >> s;
        for (int b = tid; b < N / 2; b += T) {
            const int j = b & (h - 1);
            const int k = ((b >> (s - 1)) << s) + j;
            const double2 tw = w[j * stride];
            const double2 u = x[k], v = x[k + h];
            const double2 vt = make_double2(v.x * tw.x - v.y * tw.y, v.x * tw.y + v.y * tw.x);
            x[k] = make_double2(u.x + vt.x, u.y + vt.y);
            x[k + h] = make_double2(u.x - vt.x, u.y - vt.y);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(FFT_T) void k_fft_rows_fwd(const double *__restrict__ rho, int nx,
                                                        int ny, int lognx,
                                                        const double2 *__restrict__ tw,
                                                        double2 *__restrict__ At) {
    extern __shared__ double2 fsm[];
    double2 *x = fsm, *w = fsm + nx;
    const int j = blockIdx.x;
    fft_twiddles_from(w, tw, nx);
    for (int i = threadIdx.x; i < nx; i += FFT_T)
        x[i] = make_double2(rho[(int64_t)j * nx + i], 0.0);
    __syncthreads();
    fft_lds(x, w, lognx);
    for (int ih = threadIdx.x; ih <= nx / 2; ih += FFT_T) At[(int64_t)ih * ny + j] = x[ih];
}

__global__ __launch_bounds__(FFT_T) void k_fft_cols_e(const double2 *__restrict__ At, int nx,
                                                      int ny, int logny,
                                                      const double2 *__restrict__ tw,
                                                      const double *__restrict__ kxk2,
                                                      const double *__restrict__ kyk2,
                                                      double scale, double2 *__restrict__ B) {
    extern __shared__ double2 fsm[];
    double2 *S = fsm, *Z0 = fsm + ny, *Z1 = fsm + 2 * ny, *wf = fsm + 3 * ny,
            *wi = fsm + 3 * ny + ny / 2;
    const int ih = blockIdx.x, h = nx / 2 + 1, tid = threadIdx.x;
    fft_twiddles_from(wf, tw + nx, ny);
    fft_twiddles_from(wi, tw + nx + ny / 2, ny);
    for (int j = tid; j < ny; j += FFT_T) S[j] = At[(int64_t)ih * ny + j];
    __syncthreads();
    fft_lds(S, wf, logny);
    const bool self = ih == 0 || 2 * ih == nx;  // self-conjugate column
    // half_value<SPEC_E> of (ih, jh) from the column in LDS
    auto hv = [&](int jh, double2 &f, double2 &g) {
        const int64_t q = (int64_t)ih + (int64_t)h * jh;
        if (q == 0) {
            f = g = make_double2(0.0, 0.0);
        } else {
            f = mul_mi(kxk2[q], S[jh], scale);
            g = mul_mi(kyk2[q], S[jh], scale);
        }
    };
    for (int jk = tid; jk < ny; jk += FFT_T) {
        double2 f, g;
        hv(jk, f, g);
        if (self) {
            double2 fm, gm;
            hv((ny - jk) % ny, fm, gm);
            f = make_double2(0.5 * (f.x + fm.x), 0.5 * (f.y - fm.y));
            g = make_double2(0.5 * (g.x + gm.x), 0.5 * (g.y - gm.y));
        }
        Z0[jk] = make_double2(f.x - g.y, f.y + g.x);
        if (!self) {  // column nx - ih: the conjugate half values at -jk
            double2 f1, g1;
            hv((ny - jk) % ny, f1, g1);
            f1 = cconj(f1);
            g1 = cconj(g1);
            Z1[jk] = make_double2(f1.x - g1.y, f1.y + g1.x);
        }
    }
    __syncthreads();
    fft_lds(Z0, wi, logny);
    if (!self) fft_lds(Z1, wi, logny);
    for (int j = tid; j < ny; j += FFT_T) {
        B[(int64_t)j * nx + ih] = Z0[j];
        if (!self) B[(int64_t)j * nx + (nx - ih)] = Z1[j];
    }
}

// The field energy of a 64-thread block's partial sums (sx, sy) over its
// rows: each block's sums stored, and the last block to arrive adds them in
// block order and appends (sx, sy) dx dy to the history (slot -2: the next
// entry of the step counter; -1: none).
__device__ void energy_tail(double sx, double sy, double *sh, bool &last, double *part,
                            unsigned *count, double *hist, int slot, int *counter, double dxdy) {
    block_sum2<FFT_T>(sx, sy, sh);
    if (slot == -1) return;
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sx;
        part[2 * blockIdx.x + 1] = sy;
        __threadfence();
        const unsigned ticket = atomicAdd(count, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += FFT_T) {
        a += __hip_atomic_load(&part[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b += __hip_atomic_load(&part[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum2<FFT_T>(a, b, sh);
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

// Inverse transform along x and k_post's work, row by row; partial sums per
// block over its rows in a fixed order, combined in block order by the last
// block to arrive.
__global__ __launch_bounds__(FFT_T) void k_fft_rows_inv(
    const double2 *__restrict__ B, int nx, int ny, int lognx, const double2 *__restrict__ tw,
    double *__restrict__ ex,
    double *__restrict__ ey, float2 *__restrict__ e32, double *__restrict__ zero_acc,
    double *part, unsigned *count, double *hist, int slot, int *counter, double dxdy) {
    extern __shared__ double2 fsm[];
    __shared__ double sh[2 * FFT_T];
    __shared__ bool last;
    double2 *x = fsm, *w = fsm + nx;
    fft_twiddles_from(w, tw + nx / 2, nx);
    double sx = 0.0, sy = 0.0;
    for (int j = blockIdx.x; j < ny; j += gridDim.x) {
        __syncthreads();
        for (int i = threadIdx.x; i < nx; i += FFT_T) x[i] = B[(int64_t)j * nx + i];
        __syncthreads();
        fft_lds(x, w, lognx);
        for (int i = threadIdx.x; i < nx; i += FFT_T) {
            const int64_t k = (int64_t)j * nx + i;
            const double a = x[i].x, b = x[i].y;
            ex[k] = a;
            ey[k] = b;
            sx += a * a;
            sy += b * b;
            if (e32) e32[k] = make_float2((float)a, (float)b);
            if (zero_acc) zero_acc[k] = 0.0;
        }
    }
    energy_tail(sx, sy, sh, last, part, count, hist, slot, counter, dxdy);
}

// The same three kernels with the transforms across the lanes of one wave
// (pic_fftlanes.h) for 64 <= nx, ny <= 512: no LDS round trip or barrier
// between the stages.  PX = nx / 64 and PY = ny / 64 points per lane; a
// transform's output sits at bit-reversed positions and is stored there.
template <int PX>
__global__ __launch_bounds__(64) void k_wrows_fwd(const double *__restrict__ rho, int nx, int ny,
                                                  int lognx, const double2 *__restrict__ tw,
                                                  double2 *__restrict__ At) {
    __shared__ double2 w[256];
    const int l = threadIdx.x, j = blockIdx.x;
    for (int k = l; k < nx / 2; k += 64) w[k] = tw[k];
    double2 z[PX];
#pragma unroll
    for (int r = 0; r < PX; ++r) z[r] = make_double2(rho[(int64_t)j * nx + l + 64 * r], 0.0);
    __syncthreads();
    fft_wave<PX>(z, l, lognx, w, false);
#pragma unroll
    for (int r = 0; r < PX; ++r) {
        const int ik = brev_bits(l + 64 * r, lognx);
        if (ik <= nx / 2) At[(int64_t)ik * ny + j] = z[r];
    }
}

template <int PY>
__global__ __launch_bounds__(64) void k_wcols_e(const double2 *__restrict__ At, int nx, int ny,
                                                int logny, const double2 *__restrict__ tw,
                                                const double *__restrict__ kxk2,
                                                const double *__restrict__ kyk2, double scale,
                                                double2 *__restrict__ B) {
    __shared__ double2 w[256];
    __shared__ double2 S[512];
    const int l = threadIdx.x, ih = blockIdx.x, h = nx / 2 + 1;
    for (int k = l; k < ny / 2; k += 64) w[k] = tw[k];
    double2 z[PY];
#pragma unroll
    for (int r = 0; r < PY; ++r) z[r] = At[(int64_t)ih * ny + l + 64 * r];
    __syncthreads();
    fft_wave<PY>(z, l, logny, w, false);
#pragma unroll
    for (int r = 0; r < PY; ++r) S[brev_bits(l + 64 * r, logny)] = z[r];
    __syncthreads();
    const bool self = ih == 0 || 2 * ih == nx;  // self-conjugate column
    auto hv = [&](int jh, double2 &f, double2 &g) {
        const int64_t q = (int64_t)ih + (int64_t)h * jh;
        if (q == 0) {
            f = g = make_double2(0.0, 0.0);
        } else {
            f = mul_mi(kxk2[q], S[jh], scale);
            g = mul_mi(kyk2[q], S[jh], scale);
        }
    };
    double2 z0[PY], z1[PY];
#pragma unroll
    for (int r = 0; r < PY; ++r) {
        const int jk = l + 64 * r;
        double2 f, g;
        hv(jk, f, g);
        if (self) {
            double2 fm, gm;
            hv((ny - jk) % ny, fm, gm);
            f = make_double2(0.5 * (f.x + fm.x), 0.5 * (f.y - fm.y));
            g = make_double2(0.5 * (g.x + gm.x), 0.5 * (g.y - gm.y));
            z1[r] = make_double2(0.0, 0.0);
        } else {  // column nx - ih: the conjugate half values at -jk
            double2 f1, g1;
            hv((ny - jk) % ny, f1, g1);
            f1 = cconj(f1);
            g1 = cconj(g1);
            z1[r] = make_double2(f1.x - g1.y, f1.y + g1.x);
        }
        z0[r] = make_double2(f.x - g.y, f.y + g.x);
    }
    fft_wave<PY>(z0, l, logny, w, true);
    if (!self) fft_wave<PY>(z1, l, logny, w, true);  // uniform over the block
#pragma unroll
    for (int r = 0; r < PY; ++r) {
        const int j = brev_bits(l + 64 * r, logny);
        B[(int64_t)j * nx + ih] = z0[r];
        if (!self) B[(int64_t)j * nx + (nx - ih)] = z1[r];
    }
}

template <int PX>
__global__ __launch_bounds__(64) void k_wrows_inv(
    const double2 *__restrict__ B, int nx, int ny, int lognx, const double2 *__restrict__ tw,
    double *__restrict__ ex, double *__restrict__ ey, float2 *__restrict__ e32,
    double *__restrict__ zero_acc, double *part, unsigned *count, double *hist, int slot,
    int *counter, double dxdy) {
    __shared__ double2 w[256];
    __shared__ double sh[2 * FFT_T];
    __shared__ bool last;
    const int l = threadIdx.x;
    for (int k = l; k < nx / 2; k += 64) w[k] = tw[k];
    __syncthreads();
    double sx = 0.0, sy = 0.0;
    for (int j = blockIdx.x; j < ny; j += gridDim.x) {
        double2 z[PX];
#pragma unroll
        for (int r = 0; r < PX; ++r) z[r] = B[(int64_t)j * nx + l + 64 * r];
        fft_wave<PX>(z, l, lognx, w, true);
#pragma unroll
        for (int r = 0; r < PX; ++r) {
            const int64_t k = (int64_t)j * nx + brev_bits(l + 64 * r, lognx);
            const double a = z[r].x, b = z[r].y;
            ex[k] = a;
            ey[k] = b;
            sx += a * a;
            sy += b * b;
            if (e32) e32[k] = make_float2((float)a, (float)b);
            if (zero_acc) zero_acc[k] = 0.0;
        }
    }
    energy_tail(sx, sy, sh, last, part, count, hist, slot, counter, dxdy);
}

static int log2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}

// The fused path covers power-of-two grids from 4 to 512 nodes per side.
static bool fused_solve_ok(const Ctx *c) {
    const int lx = log2_exact(c->nx), ly = log2_exact(c->ny);
    return lx >= 2 && ly >= 2 && lx <= FFT_MAX_LOG && ly <= FFT_MAX_LOG;
}

// nx, ny from 64 to 512: the wave-transform kernels, PX = nx / 64 and
// PY = ny / 64 as template arguments.
template <class F>
static void with_points(int n, F &&f) {
    switch (n / 64) {
        case 1: f(std::integral_constant<int, 1>()); break;
        case 2: f(std::integral_constant<int, 2>()); break;
        case 4: f(std::integral_constant<int, 4>()); break;
        default: f(std::integral_constant<int, 8>()); break;
    }
}

static int solve_e_wave(Ctx *c, const double *rho_in, double scale, int post, int energy_slot) {
    const int nx = c->nx, ny = c->ny, lx = log2_exact(nx), ly = log2_exact(ny);
    double2 *At = c->zf, *B = c->zg;  // (nx/2+1) x ny column-major; nx x ny
    const double2 *twx = (const double2 *)c->fft_tw, *twy = twx + nx;
    with_points(nx, [&](auto px) {
        hipLaunchKernelGGL(k_wrows_fwd<decltype(px)::value>, dim3((unsigned)ny), dim3(64), 0,
                           c->stream, rho_in, nx, ny, lx, twx, At);
    });
    with_points(ny, [&](auto py) {
        hipLaunchKernelGGL(k_wcols_e<decltype(py)::value>, dim3((unsigned)(nx / 2 + 1)), dim3(64),
                           0, c->stream, (const double2 *)At, nx, ny, ly, twy,
                           (const double *)c->kx_k2, (const double *)c->ky_k2, scale, B);
    });
    const int blocks = std::min(ny, RED_BLOCKS);
    with_points(nx, [&](auto px) {
        hipLaunchKernelGGL(k_wrows_inv<decltype(px)::value>, dim3((unsigned)blocks), dim3(64), 0,
                           c->stream, (const double2 *)B, nx, ny, lx, twx, c->ex, c->ey,
                           (post & 1) ? (float2 *)c->e32 : nullptr,
                           (post & 2) ? c->rho_acc : nullptr, c->red_part, c->red_count,
                           c->energy_hist, energy_slot, c->step_counter, c->dx * c->dy);
    });
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int fft_tw_init(Ctx *c) {
    if (!fused_solve_ok(c)) return PIC_OK;
    const int lx = log2_exact(c->nx), ly = log2_exact(c->ny);
    PIC_HIP(hipMalloc(&c->fft_tw, (size_t)(c->nx + c->ny) * sizeof(double2)));
    hipLaunchKernelGGL(k_fft_tw_init, dim3(1), dim3(256), 0, c->stream, (double2 *)c->fft_tw, lx,
                       ly);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int solve_e_fused(Ctx *c, const double *rho_in, double scale, int post, int energy_slot) {
    const int nx = c->nx, ny = c->ny, lx = log2_exact(nx), ly = log2_exact(ny);
    double2 *At = c->zf, *B = c->zg;  // (nx/2+1) x ny column-major; nx x ny
    if (lx >= 6 && ly >= 6) return solve_e_wave(c, rho_in, scale, post, energy_slot);
    hipLaunchKernelGGL(k_fft_rows_fwd, dim3((unsigned)ny), dim3(FFT_T),
                       (size_t)(nx + nx / 2) * sizeof(double2), c->stream, rho_in, nx, ny, lx,
                       (const double2 *)c->fft_tw, At);
    hipLaunchKernelGGL(k_fft_cols_e, dim3((unsigned)(nx / 2 + 1)), dim3(FFT_T),
                       (size_t)(4 * ny) * sizeof(double2), c->stream, (const double2 *)At, nx, ny,
                       ly, (const double2 *)c->fft_tw, (const double *)c->kx_k2,
                       (const double *)c->ky_k2, scale, B);
    const int blocks = std::min(ny, RED_BLOCKS);
    hipLaunchKernelGGL(k_fft_rows_inv, dim3((unsigned)blocks), dim3(FFT_T),
                       (size_t)(nx + nx / 2) * sizeof(double2), c->stream, (const double2 *)B,
                       nx, ny, lx, (const double2 *)c->fft_tw, c->ex, c->ey,
                       (post & 1) ? (float2 *)c->e32 : nullptr,
                       (post & 2) ? c->rho_acc : nullptr, c->red_part, c->red_count,
                       c->energy_hist, energy_slot, c->step_counter, c->dx * c->dy);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

template <int MODE>
static int spectral(Ctx *c, double scale) {
    const int64_t mm = (int64_t)c->nx * c->ny;
    hipLaunchKernelGGL(k_spectral<MODE>, grid_of(mm), dim3(256), 0, c->stream, c->zg, c->zf,
                       c->kx_k2, c->ky_k2, c->k2, c->nx, c->ny, scale);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// rho_in (nx*ny real) -> ex, ey.  post: bit 0 writes the FP32 interleaved
// field copy, bit 1 clears the raw deposit accumulator (after the forward
// transform has read it).  energy_slot as in k_post.
int poisson_solve_e(Ctx *c, const double *rho_in, double scale, int post, int energy_slot) {
    if (c->fused_solve && fused_solve_ok(c))
        return solve_e_fused(c, rho_in, scale, post, energy_slot);
    const int64_t mm = (int64_t)c->nx * c->ny;
    hipLaunchKernelGGL(k_pack, grid_of(mm), dim3(256), 0, c->stream, rho_in, (const double *)nullptr,
                       mm, c->zf);
    PIC_HIP(hipGetLastError());
    PIC_TRY(exec(c, c->plan_fwd, c->zf, c->zg));
    PIC_TRY(spectral<SPEC_E>(c, scale));
    PIC_TRY(exec(c, c->plan_inv, c->zf, c->zg));
    int blocks = (int)((mm + RED_THREADS - 1) / RED_THREADS);
    if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
    hipLaunchKernelGGL(k_post, dim3(blocks), dim3(RED_THREADS), 0, c->stream, c->zg, c->ex, c->ey,
                       mm, (post & 1) ? (float2 *)c->e32 : nullptr,
                       (post & 2) ? c->rho_acc : nullptr, c->red_part, c->red_count,
                       c->energy_hist, energy_slot, c->step_counter, c->dx * c->dy);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int poisson_solve_phi(Ctx *c) {
    const int64_t mm = (int64_t)c->nx * c->ny;
    hipLaunchKernelGGL(k_pack, grid_of(mm), dim3(256), 0, c->stream, (const double *)c->rho,
                       (const double *)nullptr, mm, c->zf);
    PIC_HIP(hipGetLastError());
    PIC_TRY(exec(c, c->plan_fwd, c->zf, c->zg));
    PIC_TRY(spectral<SPEC_PHI>(c, 1.0 / ((double)c->nx * (double)c->ny)));
    PIC_TRY(exec(c, c->plan_inv, c->zf, c->zg));
    hipLaunchKernelGGL(k_real_part, grid_of(mm), dim3(256), 0, c->stream, c->zg, mm, c->phi);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// lrho = E^T (lex, ley) = -(P_x lex + P_y ley): one forward transform of
// lex + i ley, the transposed symbol, one inverse transform.
int poisson_adjoint(Ctx *c, const double *lex, const double *ley, double *lrho, double scale) {
    const int64_t mm = (int64_t)c->nx * c->ny;
    hipLaunchKernelGGL(k_pack, grid_of(mm), dim3(256), 0, c->stream, lex, ley, mm, c->zf);
    PIC_HIP(hipGetLastError());
    PIC_TRY(exec(c, c->plan_fwd, c->zf, c->zg));
    PIC_TRY(spectral<SPEC_ADJ>(c, scale));
    PIC_TRY(exec(c, c->plan_inv, c->zf, c->zg));
    hipLaunchKernelGGL(k_real_part, grid_of(mm), dim3(256), 0, c->stream, c->zg, mm, lrho);
    PIC_HIP(hipGetLastError());
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

// The ghost-node terms of meshfields.jl:105-106's energy sum: column i = 0,
// row j = 0, then node (0,0) once more, added in that order by one thread
// (the order of a sequential loop over the nodes, so the sum does not depend
// on the launch).
__global__ void k_ghost_energy(const double *__restrict__ ex, const double *__restrict__ ey,
                               int nx, int ny, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double g = 0.0;
    for (int j = 0; j < ny; ++j) {
        const int64_t k = (int64_t)nx * j;
        g += ex[k] * ex[k] + ey[k] * ey[k];
    }
    for (int i = 0; i < nx; ++i) g += ex[i] * ex[i] + ey[i] * ey[i];
    g += ex[0] * ex[0] + ey[0] * ey[0];
    *out = g;
}

int field_energy_async(Ctx *c, int slot, double *ghost) {
    const int64_t mm = (int64_t)c->nx * c->ny;
    int blocks = (int)((mm + RED_THREADS - 1) / RED_THREADS);
    if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
    hipLaunchKernelGGL(k_post, dim3(blocks), dim3(RED_THREADS), 0, c->stream,
                       (const double2 *)nullptr, c->ex, c->ey, mm, nullptr, nullptr, c->red_part,
                       c->red_count, c->energy_hist, slot, c->step_counter, c->dx * c->dy);
    hipLaunchKernelGGL(k_ghost_energy, dim3(1), dim3(64), 0, c->stream, c->ex, c->ey, c->nx,
                       c->ny, ghost);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int field_energy(Ctx *c, double *out2) {
    PIC_TRY(ensure_energy(c, 1));
    const int64_t mm = (int64_t)c->nx * c->ny;
    int blocks = (int)((mm + RED_THREADS - 1) / RED_THREADS);
    if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
    hipLaunchKernelGGL(k_post, dim3(blocks), dim3(RED_THREADS), 0, c->stream,
                       (const double2 *)nullptr, c->ex, c->ey, mm, nullptr, nullptr, c->red_part,
                       c->red_count, c->energy_hist, 0, c->step_counter, c->dx * c->dy);
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
        // third time (the ghost corner).  Interior sums (k_post) and ghost
        // terms (k_ghost_energy) on the device, one copy of three values.
        PIC_TRY(ensure_energy(c, 2));
        PIC_TRY(field_energy_async(c, 0, c->energy_hist + 2));
        double s[3];
        PIC_HIP(hipMemcpyAsync(s, c->energy_hist, 3 * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        *energy = s[0] + s[1] + s[2] * c->dx * c->dy;
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // extern "C"

extern "C" int pic_set_fused_solve(pic_ctx *ctx, int fused) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    ::pic::ctx_of(ctx)->fused_solve = fused != 0;
    return PIC_OK;
}
