// pic_internal.h -- shared declarations of the MI355X particle-mesh library.
//
// One context (`pic_ctx`) owns, on one GPU and one HIP stream:
//   * the mesh (nx x ny periodic cells on [xmin,xmax) x [ymin,ymax)),
//   * the particles, device resident, structure-of-arrays:
//       FP64 state  x, y, vx, vy (+ w when weights differ)       32-40 B each
//       FP32 state  cell (int32), ox, oy, vx, vy (float)          20 B each
//   * the grids rho, ex, ey, bz, phi (FP64, nx*ny, x fastest like Julia's
//     column-major nx x ny arrays) and the FP32 interleaved (ex, ey) copy the
//     FP32 gather reads,
//   * rocFFT plans for the Poisson solve, sort scratch, reduction scratch.
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <stdint.h>
#include <string>

#include "../../include/pic_mi355x.h"

namespace pic {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define PIC_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::pic::fail(PIC_ERR_HIP, "%s failed: %s (%s:%d)", #call,              \
                               hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

#define PIC_CHECK(cond, code, ...)                                                       \
    do {                                                                                 \
        if (!(cond)) return ::pic::fail(code, __VA_ARGS__);                              \
    } while (0)

#define PIC_TRY(call)                                                                    \
    do {                                                                                 \
        int r_ = (call);                                                                 \
        if (r_ != PIC_OK) return r_;                                                     \
    } while (0)

// ------------------------------------------------------------ the context
struct Ctx {
    int device = 0;
    int ncu = 1;                        // compute units of the device
    hipStream_t stream = nullptr;

    // mesh (TwoDGrid: dx = dimx/nx as GEMPIC's grid defines it)
    int nx = 0, ny = 0;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    double dimx = 0, dimy = 0, dx = 0, dy = 0;

    // particles
    int64_t n = 0, cap = 0;
    int state = PIC_STATE_F64;          // which representation is current
    bool uniform_w = true;
    double w0 = 0.0;                    // the common weight when uniform_w
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr, *w = nullptr;
    double *epx = nullptr, *epy = nullptr;  // particle fields (gather results)
    // FP32 cell-relative state: cell index and (ox, oy, vx, vy) per particle
    int32_t *cell = nullptr;
    float4 *ov32 = nullptr;
    void *swap = nullptr;               // 40 B/particle reorder scratch
    int64_t swap_bytes = 0;
    // second set of state arrays: the sort gathers into these and swaps the
    // pointers (keyed by the address of the Ctx member they pair with)
    static constexpr int NALT = 10;
    void **alt_of[NALT] = {};
    void *alt[NALT] = {};

    // grids, nx*ny each, FP64
    double *rho = nullptr, *ex = nullptr, *ey = nullptr, *bz = nullptr, *phi = nullptr;
    double *rho_acc = nullptr;          // raw deposit accumulator (not normalised)
    float *e32 = nullptr;               // (ex, ey) interleaved FP32 per node
    float *bz32 = nullptr;
    int bz_mode = PIC_BZ_NONE;          // none / uniform / grid
    double bz0 = 0.0;

    // Poisson: tables (nx/2+1) x ny, complex work grids nx x ny, rocFFT
    double *kx_k2 = nullptr, *ky_k2 = nullptr, *k2 = nullptr;
    double2 *zf = nullptr, *zg = nullptr;
    // twiddles of the LDS transforms: x forward, x inverse (nx/2 each), y
    // forward, y inverse (ny/2 each)
    double2 *fft_tw = nullptr;
    rocfft_plan plan_fwd = nullptr, plan_inv = nullptr;
    rocfft_execution_info fft_info = nullptr;
    void *fft_work = nullptr;
    size_t fft_work_bytes = 0;

    // reductions: partial sums and results (device), counters
    double *red_part = nullptr;         // [RED_BLOCKS * 4]
    unsigned *red_count = nullptr;      // arrival counter (reset by the last block)
    double *energy_hist = nullptr;      // device energy history [energy_cap * 2]
    int64_t energy_cap = 0;
    int *step_counter = nullptr;        // device index into energy_hist

    // sort scratch
    uint32_t *sort_keys[2] = {nullptr, nullptr};
    uint32_t *sort_vals[2] = {nullptr, nullptr};
    uint32_t *sort_hist = nullptr;
    int64_t sort_hist_cap = 0;
    uint32_t *sort_blk = nullptr;
    int64_t *cell_off = nullptr;        // ncells + 1 exclusive offsets after a sort
    int64_t sort_cap = 0;
    // lane-spread order: per-segment max count, table offsets, tables, perm
    uint32_t *seg_k = nullptr, *seg_tbl = nullptr, *seg_A = nullptr, *perm2 = nullptr;
    unsigned long long *seg_M = nullptr;
    int64_t seg_cap = 0;
    bool lanes_order = false;           // last sort used the lane-spread order
    // after a lane sort, segment s occupies slots [cell_off[64 s],
    // cell_off[min(64 s + 64, ncells)]) until the next sort
    bool seg_valid = false;             // cell_off describes the current slots
    // the particles sit in the order the last lane sort left (or its pending
    // permutation gives) and have not moved since: pic_run can skip its own
    // first sort
    bool sort_fresh = false;
    // a lane sort's permutation not yet applied to the state (see sort_by_cell)
    const uint32_t *pending_perm = nullptr;

    // fused-step tile statistics (diagnostic): chunks that took the slow path
    unsigned long long *slow_chunks = nullptr;
    // particles that the segment passes hand to their global-memory fallback
    // kernel: entry list (capacity entries) and {count, arrivals}
    uint32_t *defer = nullptr;
    unsigned *defer_count = nullptr;

    // captured step graphs
    hipGraphExec_t graph_exec = nullptr;
    int graph_key = -1;

    // FP32 charge accumulation in the LDS windows: 0 compensated FP32
    // (hi, lo) pairs, 1 FP64, 2 int32 fixed point (the default: the fastest
    // on gfx950, whose ds_add_f32 is slow and whose 8-byte LDS atomics
    // conflict on banks twice as often as 4-byte ones)
    int f32_acc = 2;
    int f32_tiles = 4;  // fixed-point charge tiles of the FP32 drift-and-deposit passes
    int f32_window_kib = 160;  // LDS window per block of those passes (pic_set_f32_window)

    // E solve on power-of-two grids up to 512 x 512: three LDS-transform
    // kernels (true) or the rocFFT sequence (false)
    bool fused_solve = true;
    // pic_run on power-of-two grids up to 64 x 64: one launch per iteration
    // (true, pic_set_fused_step) or the separate solve, pass and sort
    // kernels; es_buf holds the launches' three raw charge buffers
    bool fused_step = true;
    double *es_buf = nullptr;

    // Vlasov-Maxwell: FDTD fields on the Yee grid (fdtd.jl:3-20; ex nx x
    // (ny+1), ey (nx+1) x ny, bz nx x ny, column-major) and the periodic CIC
    // current jx, jy (nx x ny: the ghost row and column of compute_current!
    // folded in)
    double *fd_ex = nullptr, *fd_ey = nullptr, *fd_bz = nullptr;
    double *jx = nullptr, *jy = nullptr;
    // the one-launch iteration's second set of FDTD fields and two more
    // current buffers
    double *fd2_ex = nullptr, *fd2_ey = nullptr, *fd2_bz = nullptr;
    double *jb[4] = {nullptr, nullptr, nullptr, nullptr};  // J1 x, y, J2 x, y
    int vm_path = 0;  // pic_set_maxwell_path

    // UA (ua_type.jl): ntau, eps; per particle b, t (Particles.b, .t), the
    // positions at the start of a UA step and, between its prediction and
    // correction, the tau-spectra of fx and fy (ntau x 4 complex)
    int ua_ntau = 0;
    double ua_eps = 0.0;
    int64_t ua_cap = 0;
    double *ua_b = nullptr, *ua_t = nullptr, *ua_x0 = nullptr, *ua_y0 = nullptr;
    double *ua_fhat = nullptr;

    // adjoint reverse sweep: steps whose positions are kept at once (0 =
    // the most the free device memory allows), and the last sweep's plan
    int adj_block = 0;
    int adj_block_used = 0;
    int64_t adj_bytes_used = 0;
    // with timing enabled: summed durations (ms) and counts of the adjoint's
    // forward passes, gather-adjoint, deposit-adjoint and recompute passes
    double t_adj_ms[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t t_adj_n[4] = {0, 0, 0, 0};
    // UA loop kernels: 0 prediction, 1 correction
    double t_ua_ms[2] = {0.0, 0.0};
    int64_t t_ua_n[2] = {0, 0};

    // pic_run kernel timing (HIP events on the context stream), every
    // timing_every-th step (pic_set_timing)
    bool timing = false;
    int timing_every = 1;
    double t_pass_ms = 0.0, t_solve_ms = 0.0;
    int64_t t_passes = 0;
};

Ctx *ctx_of(pic_ctx *c);

// ---------------------------------------------------------- allocation
int ensure_particles(Ctx *c, int64_t n);
int ensure_f32(Ctx *c);
int ensure_swap(Ctx *c, int64_t bytes);
int ensure_pfields(Ctx *c);
int ensure_sort(Ctx *c);
int ensure_energy(Ctx *c, int64_t n);
int ensure_defer(Ctx *c);

// ------------------------------------------------------ module entry points
int poisson_init(Ctx *c);
void poisson_free(Ctx *c);
int poisson_solve_e(Ctx *c, const double *rho_in, double scale, int make_e32, int energy_slot);
int poisson_solve_phi(Ctx *c);
int poisson_adjoint(Ctx *c, const double *lex, const double *ley, double *lrho, double scale);
int finalize_rho(Ctx *c, int kernel);
int field_energy(Ctx *c, double *out2);
// Queues meshfields.jl:105-106's energy sums of the current ex, ey: the
// interior sums times dx dy into energy_hist[2 slot], [2 slot + 1] and the
// ghost-node terms (without dx dy) into the device double *ghost.  No wait.
int field_energy_async(Ctx *c, int slot, double *ghost);

// Stable sort by cell (lanes: then the lane-spread order of the segment
// passes).  apply = false with lanes leaves the permutation pending in
// c->pending_perm for the next segment pass to apply while it reads the
// particles (no host permutation requested).
int sort_by_cell(Ctx *c, int32_t *perm_host, bool lanes, bool apply = true);
int settle_state(Ctx *c);
int alt_f32(Ctx *c, int32_t **cell, float4 **ov);
void swap_alt_f32(Ctx *c);
int alt_f64(Ctx *c, double **x, double **y, double **vx, double **vy, double **w);
void swap_alt_f64(Ctx *c);

int launch_fill(Ctx *c, double *p, double v, int64_t n);

// pic_run's loop with one launch per iteration on small grids (pic_small.hip).
bool es_step_ok(const Ctx *c);
int es_run(Ctx *c, int kernel, double dt, int nsteps);

// The adjoint sweep of the FP32 cell-relative state (pic_adjoint32.hip).
int adjoint_f32(Ctx *c, int kernel, double dt, int nsteps, int ckpt, int sort_every, double *J,
                double *gx, double *gv);

// Device scratch owned by one call, freed when it goes out of scope.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count) {
        PIC_HIP(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
        return PIC_OK;
    }
};

// ------------------------------------------------------------ constants
constexpr int RED_BLOCKS = 256;
constexpr int RED_THREADS = 256;

// Charge reductions of pic_poisson.hip, also launched by the allocation-free
// compute_rho (pic_noalloc.hip): the fixed-order two-level sum of a grid
// (partials per block, the last block to arrive sums them and clears
// *count), and rho = acc / (dx dy) minus the reference's shift.
__global__ __launch_bounds__(RED_THREADS) void k_sum(const double *__restrict__ a, int64_t m,
                                                     double *part, unsigned *count,
                                                     double *out);
__global__ void k_finalize_rho(const double *__restrict__ acc, double *__restrict__ rho, int64_t m,
                               double dxdy, const double *total, int kernel, double dx, double dy,
                               double dimx, double dimy);

}  // namespace pic

// ------------------------------------------------------- device helpers
namespace picd {

__host__ __device__ inline int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// i mod n for i in [-2n, 3n) with two conditional corrections.
__host__ __device__ inline int wrap2(int i, int n) {
    if (i < 0) i += n;
    if (i < 0) i += n;
    if (i >= n) i -= n;
    if (i >= n) i -= n;
    return i;
}

// Julia mod(x, y) for Float64: fmod shifted into the sign of y.
__device__ __forceinline__ double jl_fmod(double x, double y) {
    double r = fmod(x, y);
    if (r == 0.0) return copysign(r, y);
    if ((r > 0) != (y > 0)) return r + y;
    return r;
}

// jl_fmod(x, y) for y > 0, with the common cases -y <= x < 2y done without
// fmod: there fmod(x, y) is x, x or x - y (exact by Sterbenz's lemma), so the
// result is the same bits.
__device__ __forceinline__ double jl_mod_pos(double x, double y) {
    if (x > 0.0 && x < y) return x;
    if (x < 0.0 && x >= -y) return x + y;
    if (x >= y && x < y + y) {
        const double r = x - y;
        return r == 0.0 ? 0.0 : r;
    }
    return jl_fmod(x, y);
}

// Julia mod1(x, n) - 1, i.e. the 0-based periodic index of 1-based x.
__host__ __device__ inline long long mod1z(long long x, long long n) {
    long long r = (x - 1) % n;
    return r < 0 ? r + n : r;
}

// The reference's cell of a coordinate, i = fl + 1 with fl = floor(xp) (or
// trunc(xp)), as the 0-based periodic cell mod1(i, n) - 1 = fl mod n, and
// the in-cell offset d = (xp - i) + 1.  32-bit arithmetic: fl is clamped to
// +-2^30 first (the reference's index is out of bounds long before; NaN
// becomes -2^30), so the result is a valid cell for every input and equals
// the 64-bit mod1 for |fl| < 2^30.
__device__ __forceinline__ int cell_of(double xp, double fl, int n, double &d) {
    d = (xp - (fl + 1.0)) + 1;
    int i = (int)fmin(fmax(fl, -1073741824.0), 1073741824.0);
    if ((unsigned)i >= (unsigned)n) {
        i %= n;
        if (i < 0) i += n;
    }
    return i;
}

// Quintic B-spline (M6) weights of nodes i-2..i+3 at offset t in [0,1), FP32.
// The pieces of f_m6 (compute_rho_m6.jl:10-25) are 120 w5 = t^5, 120 w4 =
// P4(t) = 1+5t+10t^2+10t^3+5t^4-5t^5, 120 w3 = P3(t) =
// 26+50t+20t^2-20t^3-20t^4+10t^5 and w_k(t) = w_(5-k)(1-t).  In z = t - 1/2
// each mirrored pair is w_(5-k), w_k = E_k(z^2) +/- z Q_k(z^2) with quadratic
// E_k, Q_k (exact rational coefficients rounded to float), so the six weights
// cost 20 operations: z, z^2, then per pair two Horner steps for E and Q and
// two fused multiply-adds.
#define PIC_F(num, den) ((float)((double)(num) / (double)(den)))
__device__ __forceinline__ void m6_weights_f(float t, float w[6]) {
    const float z = t - 0.5f, z2 = z * z, mz = -z;
    float e, q;
    e = __builtin_fmaf(__builtin_fmaf(PIC_F(1, 48), z2, PIC_F(1, 96)), z2, PIC_F(1, 3840));
    q = __builtin_fmaf(__builtin_fmaf(PIC_F(1, 120), z2, PIC_F(1, 48)), z2, PIC_F(1, 384));
    w[5] = __builtin_fmaf(z, q, e);
    w[0] = __builtin_fmaf(mz, q, e);
    e = __builtin_fmaf(__builtin_fmaf(PIC_F(-1, 16), z2, PIC_F(7, 32)), z2, PIC_F(79, 1280));
    q = __builtin_fmaf(__builtin_fmaf(PIC_F(-1, 24), z2, PIC_F(1, 16)), z2, PIC_F(25, 128));
    w[4] = __builtin_fmaf(z, q, e);
    w[1] = __builtin_fmaf(mz, q, e);
    e = __builtin_fmaf(__builtin_fmaf(PIC_F(1, 24), z2, PIC_F(-11, 48)), z2, PIC_F(841, 1920));
    q = __builtin_fmaf(__builtin_fmaf(PIC_F(1, 12), z2, PIC_F(-7, 24)), z2, PIC_F(77, 192));
    w[3] = __builtin_fmaf(z, q, e);
    w[2] = __builtin_fmaf(mz, q, e);
}

// The same six weights in FP64, exact integer coefficients, divided by 120 at
// the end (Horner, fused multiply-adds).
__device__ __forceinline__ double m6_p3d(double u) {
    double r = 10.0;
    r = __builtin_fma(r, u, -20.0);
    r = __builtin_fma(r, u, -20.0);
    r = __builtin_fma(r, u, 20.0);
    r = __builtin_fma(r, u, 50.0);
    r = __builtin_fma(r, u, 26.0);
    return r;
}
__device__ __forceinline__ double m6_p4d(double u) {
    double r = -5.0;
    r = __builtin_fma(r, u, 5.0);
    r = __builtin_fma(r, u, 10.0);
    r = __builtin_fma(r, u, 10.0);
    r = __builtin_fma(r, u, 5.0);
    r = __builtin_fma(r, u, 1.0);
    return r;
}
__device__ __forceinline__ void m6_weights_d(double t, double w[6]) {
    const double c = 1.0 / 120.0;
    double s = 1.0 - t;
    double t2 = t * t, s2 = s * s;
    w[0] = (s2 * s2 * s) * c;
    w[1] = m6_p4d(s) * c;
    w[2] = m6_p3d(s) * c;
    w[3] = m6_p3d(t) * c;
    w[4] = m6_p4d(t) * c;
    w[5] = (t2 * t2 * t) * c;
}

// ------------------------------------------------ NGP, TSC, M4 weights
//
// Stencils of four nodes i-1 .. i+2 around the cell i of the particle, at the
// in-cell offset d, in the reference's order cm1, c, cp1, cp2: the weights
// f(1+d), f(d), f(1-d), f(2-d) with f = f_tsc (pushers_tsc.jl:3-12) or f_m4
// (pushers_m4.jl:8-19), evaluated with the reference's operations.
__device__ __forceinline__ double f_tsc_d(double x) {
    if (x <= 0.5) return 0.75 - x * x;
    if (x <= 1.5) return 0.5 * ((1.5 - x) * (1.5 - x));
    return 0.0;
}
__device__ __forceinline__ double f_m4_d(double x) {
    double f = 0.0;
    if (x > 2) {
        f = 0.0;
    } else if (x >= 1) {
        const double t = 2 - x;
        f = 0.25 * ((t * t) * t);
    } else {
        f = (1 - 1.5 * (x * x)) + 0.75 * ((x * x) * x);
    }
    return f * 2 / 3;
}
__device__ __forceinline__ float f_tsc_f(float x) {
    if (x <= 0.5f) return 0.75f - x * x;
    if (x <= 1.5f) return 0.5f * ((1.5f - x) * (1.5f - x));
    return 0.0f;
}
__device__ __forceinline__ float f_m4_f(float x) {
    if (x > 2.0f) return 0.0f;
    if (x >= 1.0f) {
        const float t = 2.0f - x;
        return (t * t) * t * PIC_F(1, 6);
    }
    return __builtin_fmaf(x * x, __builtin_fmaf(0.5f, x, -1.0f), PIC_F(2, 3));
}

// NGP as a two-node stencil i, i+1 with weights (1, 0) or (0, 1): the node
// of round(Int, xp) (ties to even, pushers_ngp.jl:22-23) with xp = i + d.
// The parity of the 0-based cell stands for that of floor(xp) (equal on
// grids with an even number of cells).
template <typename T>
__device__ __forceinline__ void ngp_weights(T d, int cell, T w[2]) {
    const bool up = d > T(0.5) || (d == T(0.5) && (cell & 1));
    w[0] = up ? T(0) : T(1);
    w[1] = up ? T(1) : T(0);
}

// The M6 gather of interpol_eb_m6! (interpolation_m6.jl:125-247, and its
// UA form :3-123) at cell (i, j) with offsets (dpx, dpy): nodes i-2..i+3 and
// j-2..j+3, wrapped; ex, ey are nx x ny; i in [0, nx], j in [0, ny] (cells of
// a coordinate brought into the domain).  Sums with FMAs, y inner.
__device__ __forceinline__ void mf_gather_m6(int i, int j, double dpx, double dpy, int nx, int ny,
                                             const double *ex, const double *ey, double &e1,
                                             double &e2) {
    double wx[6], wy[6];
    m6_weights_d(dpx, wx);
    m6_weights_d(dpy, wy);
    e1 = 0.0;
    e2 = 0.0;
    for (int a = 0; a < 6; ++a) {
        int ia = wrap2(i - 2 + a, nx);
        double t1 = 0.0, t2 = 0.0;
        for (int b = 0; b < 6; ++b) {
            size_t g = ia + (size_t)nx * wrap2(j - 2 + b, ny);
            t1 = __builtin_fma(wy[b], ex[g], t1);
            t2 = __builtin_fma(wy[b], ey[g], t2);
        }
        e1 = __builtin_fma(wx[a], t1, e1);
        e2 = __builtin_fma(wx[a], t2, e2);
    }
}

}  // namespace picd

namespace pic {

// Stencil of a particle shape: S nodes per direction starting HL nodes left
// of the particle's cell; QMAX bounds the product of two deposit weights.
template <int K>
struct Shape;
template <>
struct Shape<PIC_CIC> {
    static constexpr int S = 2, HL = 0;
    static constexpr double QMAX = 1.0;
};
template <>
struct Shape<PIC_M6> {
    static constexpr int S = 6, HL = 2;
    static constexpr double QMAX = 0.31;  // (66/120)^2
};
template <>
struct Shape<PIC_NGP> {
    static constexpr int S = 2, HL = 0;
    static constexpr double QMAX = 1.0;
};
template <>
struct Shape<PIC_TSC> {
    static constexpr int S = 4, HL = 1;
    static constexpr double QMAX = 0.5625;  // (3/4)^2
};
template <>
struct Shape<PIC_M4> {
    static constexpr int S = 4, HL = 1;
    static constexpr double QMAX = 0.445;  // (2/3)^2
};

// Gather (push_v!) and deposit (compute_rho) weights of the separable
// shapes, FP64 and FP32, for the S nodes of a direction.  CIC keeps its own
// bilinear code in the kernels.
template <int K>
__device__ __forceinline__ void gather_w(double d, int cell, double *w) {
    if (K == PIC_M6) {
        picd::m6_weights_d(d, w);
    } else if (K == PIC_NGP) {
        picd::ngp_weights(d, cell, w);
    } else {  // TSC and M4 both gather with f_m4
        w[0] = picd::f_m4_d(1 + d);
        w[1] = picd::f_m4_d(d);
        w[2] = picd::f_m4_d(1 - d);
        w[3] = picd::f_m4_d(2 - d);
    }
}
template <int K>
__device__ __forceinline__ void deposit_w(double d, int cell, double *w) {
    if (K == PIC_TSC) {
        w[0] = picd::f_tsc_d(1 + d);
        w[1] = picd::f_tsc_d(d);
        w[2] = picd::f_tsc_d(1 - d);
        w[3] = picd::f_tsc_d(2 - d);
    } else {
        gather_w<K>(d, cell, w);
    }
}
template <int K>
__device__ __forceinline__ void gather_w(float d, int cell, float *w) {
    if (K == PIC_M6) {
        picd::m6_weights_f(d, w);
    } else if (K == PIC_NGP) {
        picd::ngp_weights(d, cell, w);
    } else {
        w[0] = picd::f_m4_f(1.0f + d);
        w[1] = picd::f_m4_f(d);
        w[2] = picd::f_m4_f(1.0f - d);
        w[3] = picd::f_m4_f(2.0f - d);
    }
}
template <int K>
__device__ __forceinline__ void deposit_w(float d, int cell, float *w) {
    if (K == PIC_TSC) {
        w[0] = picd::f_tsc_f(1.0f + d);
        w[1] = picd::f_tsc_f(d);
        w[2] = picd::f_tsc_f(1.0f - d);
        w[3] = picd::f_tsc_f(2.0f - d);
    } else {
        gather_w<K>(d, cell, w);
    }
}

// Runs the statement(s) given after `kernel` with the compile-time constant
// KK set to the run-time shape code `kernel`.
#define PIC_SHAPE_SWITCH(kernel, ...)                                    \
    switch (kernel) {                                                    \
        case PIC_CIC: { constexpr int KK = PIC_CIC; __VA_ARGS__; } break; \
        case PIC_M6: { constexpr int KK = PIC_M6; __VA_ARGS__; } break;   \
        case PIC_NGP: { constexpr int KK = PIC_NGP; __VA_ARGS__; } break; \
        case PIC_TSC: { constexpr int KK = PIC_TSC; __VA_ARGS__; } break; \
        default: { constexpr int KK = PIC_M4; __VA_ARGS__; } break;       \
    }

inline bool valid_shape(int kernel) { return kernel >= PIC_CIC && kernel <= PIC_M4; }

}  // namespace pic
