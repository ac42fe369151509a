// pic_step.hip -- the particle kernels of the particle-mesh step, and the
// device-resident time loop.
//
// One kernel family does every particle pass of the hot path.  Per particle,
// in this order and each part switchable at compile time:
//   GATHER  interpolate (ex, ey[, bz]) at the current position and kick:
//           push_v!  (pushers_cic.jl:3-53, pushers_m6.jl:5-224)
//   NDRIFT  0, 1 or 2 drifts: push_x! (pushers_cic.jl:55-72)
//   DEPOSIT deposit w at the new position: compute_rho
//           (pushers_cic.jl:75-114, pushers_m6.jl:226-348)
// so that one time-loop iteration of the reference
// (examples/landau_damping_poisson_cic_m6.jl:70-83) is one particle pass plus
// the Poisson solve: kick + drift(dt/2) + drift(dt/2) + deposit.
//
// Work decomposition.  A block of 256 threads owns a contiguous chunk of
// 256*PPT particles (PPT per thread, kept in registers).  Particles are kept
// sorted by cell (pic_sort), so a chunk covers a small patch of cells.  The
// block reduces the bounding box of its particles' cells (relative to the
// chunk's first particle, periodic), stages the field tile covering the box
// plus the stencil halo into LDS, gathers from LDS, pushes, reduces the box
// of the new cells, accumulates the deposit in an LDS tile and flushes the
// tile's non-zero nodes into the FP64 grid with global atomics.  A chunk
// whose box does not fit the tile capacity takes the global-memory path
// (direct loads and FP64 global atomics) and is counted in slow_chunks.
//
// FP64 state: reference arithmetic, operation for operation (the CIC
// results are bit-identical to pic.f90's push_v/push_x); the deposit tile
// accumulates with LDS FP64 atomics.
// FP32 cell-relative state: int32 cell, FP32 offsets and velocities; the
// deposit tile holds compensated (hi, lo) FP32 pairs: every LDS atomic add
// returns the old value, the exact rounding error of that add is recovered
// with TwoSum and added to lo, and hi + lo is flushed in FP64.
#include <climits>
#include <cstdio>
#include <vector>

#include "pic_internal.h"

namespace pic {

constexpr int ST = 256;        // threads per block
constexpr int CAP64 = 2048;    // tile capacity (nodes), FP64 kernels
constexpr int CAP32 = 4096;    // tile capacity (nodes), FP32 kernels

struct ArgsF64 {
    int64_t n;
    double *x, *y, *vx, *vy;
    const double *w;
    double w0;
    const double *ex, *ey, *bz;
    double bz0;
    int nx, ny;
    double dx, dy, dimx, dimy;
    double dt;        // kick length
    double dtd;       // length of each drift
    double *acc;      // raw deposit accumulator (FP64, nx*ny)
    unsigned long long *slow;
};

struct ArgsF32 {
    int64_t n;
    int32_t *cell;
    float4 *ov;       // (ox, oy, vx, vy)
    float w;
    const float2 *e32;
    const float *bz;
    float bz0;
    int nx, ny, lognx;  // nx is a power of two in FP32 state
    float hdt;        // dt/2 (kick halves)
    float dtdx, dtdy; // drift length / cell size
    double *acc;
    unsigned long long *slow;
};

// Stencil of shape K: S nodes per direction, HL left and HR right of the
// particle's cell.
template <int K>
struct Stencil {
    static constexpr int S = Shape<K>::S, HL = Shape<K>::HL, HR = S - 1 - HL;
};

// Wrapped difference of cell indices i, iref in [0, n) into [-n/2, n/2).
__device__ __forceinline__ int rel(int i, int iref, int n) {
    int d = i - iref;
    if (d >= n - n / 2) d -= n;
    else if (d < -(n / 2)) d += n;
    return d;
}

// i mod n for i in [-n, 2n) (stencil neighbours of a cell in [0, n)).
__device__ __forceinline__ int wrap1(int i, int n) {
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// i mod n for any i.
__device__ __forceinline__ int wrapm(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

struct Box {
    int x0, y0, tw, th;  // tile origin relative to the reference cell, size
    bool fits;
};

// Block-wide bounding box of (di, dj) over the live particles, extended by
// the stencil halo.  sbox: 4 ints of LDS.
template <int HL, int HR>
__device__ __forceinline__ Box block_box(int mnx, int mxx, int mny, int mxy, int *sbox, int cap) {
    for (int o = 32; o > 0; o >>= 1) {
        mnx = min(mnx, __shfl_xor(mnx, o));
        mxx = max(mxx, __shfl_xor(mxx, o));
        mny = min(mny, __shfl_xor(mny, o));
        mxy = max(mxy, __shfl_xor(mxy, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&sbox[0], mnx);
        atomicMax(&sbox[1], mxx);
        atomicMin(&sbox[2], mny);
        atomicMax(&sbox[3], mxy);
    }
    __syncthreads();
    Box b;
    b.x0 = sbox[0] - HL;
    b.y0 = sbox[2] - HL;
    b.tw = sbox[1] + HR - b.x0 + 1;
    b.th = sbox[3] + HR - b.y0 + 1;
    b.fits = (sbox[0] <= sbox[1]) && b.tw > 0 && b.th > 0 && (int64_t)b.tw * b.th <= cap;
    __syncthreads();
    if (threadIdx.x == 0) {
        sbox[0] = INT_MAX;
        sbox[1] = INT_MIN;
        sbox[2] = INT_MAX;
        sbox[3] = INT_MIN;
    }
    // (the next reader of sbox is behind at least one more __syncthreads)
    return b;
}

__device__ __forceinline__ void init_box(int *sbox) {
    if (threadIdx.x == 0) {
        sbox[0] = INT_MAX;
        sbox[1] = INT_MIN;
        sbox[2] = INT_MAX;
        sbox[3] = INT_MIN;
    }
}

// Global node of tile node (a, b).
__device__ __forceinline__ int64_t tile_node(const Box &b, int a, int bb, int iref, int jref,
                                             int nx, int ny) {
    int gi = wrapm(iref + b.x0 + a, nx);
    int gj = wrapm(jref + b.y0 + bb, ny);
    return (int64_t)gi + (int64_t)nx * gj;
}

// ============================================================ FP64 kernel

// The reference's CIC weights and cell of (x, y) (pushers_cic.jl:18-29):
// i = floor(xp)+1 (trunc for the deposit, :86-87; equal for x >= 0),
// dxp = (xp - i) + 1, 0-based wrapped cell returned.
__device__ __forceinline__ void ref_cell(double x, double dx, int n, bool trunc_, int &iw,
                                         double &d) {
    const double xp = x / dx;
    iw = picd::cell_of(xp, trunc_ ? trunc(xp) : floor(xp), n, d);
}

template <int K, bool GATHER, int NDRIFT, bool DEPOSIT, int BZ, int PPT>
__global__ __launch_bounds__(ST) void k_step_f64(ArgsF64 a) {
    using St = Stencil<K>;
    extern __shared__ double smem64[];
    __shared__ int sbox[4];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * ST * PPT;
    const int nx = a.nx, ny = a.ny;

    double px[PPT], py[PPT], pvx[PPT], pvy[PPT];
    bool live[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        int64_t p = base + (int64_t)k * ST + tid;
        live[k] = p < a.n;
        if (live[k]) {
            px[k] = a.x[p];
            py[k] = a.y[p];
            if (GATHER) {
                pvx[k] = a.vx[p];
                pvy[k] = a.vy[p];
            }
        }
    }
    init_box(sbox);
    // reference cell: the chunk's first particle
    int iref, jref;
    {
        double d;
        ref_cell(a.x[base], a.dx, nx, false, iref, d);
        ref_cell(a.y[base], a.dy, ny, false, jref, d);
    }
    __syncthreads();

    if (GATHER) {
        int ci[PPT], cj[PPT];
        double fx[PPT], fy[PPT];
        int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            // push_v! uses floor (pushers_cic.jl:20-21, pushers_m6.jl:22-23)
            ref_cell(px[k], a.dx, nx, false, ci[k], fx[k]);
            ref_cell(py[k], a.dy, ny, false, cj[k], fy[k]);
            if (live[k]) {
                int di = rel(ci[k], iref, nx), dj = rel(cj[k], jref, ny);
                mnx = min(mnx, di);
                mxx = max(mxx, di);
                mny = min(mny, dj);
                mxy = max(mxy, dj);
            }
        }
        Box b = block_box<St::HL, St::HR>(mnx, mxx, mny, mxy, sbox, CAP64);
        double2 *te = reinterpret_cast<double2 *>(smem64);
        double *tb = smem64 + 2 * CAP64;
        if (b.fits) {
            const int nt = b.tw * b.th;
            for (int t = tid; t < nt; t += ST) {
                int64_t g = tile_node(b, t % b.tw, t / b.tw, iref, jref, nx, ny);
                te[t] = make_double2(a.ex[g], a.ey[g]);
                if (BZ == PIC_BZ_GRID) tb[t] = a.bz[g];
            }
        } else if (tid == 0) {
            atomicAdd(a.slow, 1ull);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!live[k]) continue;
            double e1, e2, b3 = 0.0;
            // field value of stencil node (u, v) (u, v in [0, S))
            int la = rel(ci[k], iref, nx) - (b.x0 + St::HL);
            int lb = rel(cj[k], jref, ny) - (b.y0 + St::HL);
            auto node_e = [&](int u, int v) -> double2 {
                if (b.fits) return te[(la + u) + b.tw * (lb + v)];
                int64_t g = (int64_t)wrap1(ci[k] - St::HL + u, nx) +
                            (int64_t)nx * wrap1(cj[k] - St::HL + v, ny);
                return make_double2(a.ex[g], a.ey[g]);
            };
            auto node_b = [&](int u, int v) -> double {
                if (b.fits) return tb[(la + u) + b.tw * (lb + v)];
                int64_t g = (int64_t)wrap1(ci[k] - St::HL + u, nx) +
                            (int64_t)nx * wrap1(cj[k] - St::HL + v, ny);
                return a.bz[g];
            };
            if (K == PIC_CIC) {
                double dxp = fx[k], dyp = fy[k];
                double a1 = (1 - dxp) * (1 - dyp);
                double a2 = dxp * (1 - dyp);
                double a3 = dxp * dyp;
                double a4 = (1 - dxp) * dyp;
                double2 E00 = node_e(0, 0), E10 = node_e(1, 0), E11 = node_e(1, 1),
                        E01 = node_e(0, 1);
                e1 = a1 * E00.x + a2 * E10.x + a3 * E11.x + a4 * E01.x;
                e2 = a1 * E00.y + a2 * E10.y + a3 * E11.y + a4 * E01.y;
                if (BZ == PIC_BZ_GRID)
                    b3 = a1 * node_b(0, 0) + a2 * node_b(1, 0) + a3 * node_b(1, 1) +
                         a4 * node_b(0, 1);
            } else {
                double wx[6], wy[6];
                gather_w<K>(fx[k], ci[k], wx);
                gather_w<K>(fy[k], cj[k], wy);
                e1 = 0.0;
                e2 = 0.0;
#pragma unroll
                for (int u = 0; u < St::S; ++u) {
                    double2 E0 = node_e(u, 0);
                    double t1 = wy[0] * E0.x, t2 = wy[0] * E0.y;
#pragma unroll
                    for (int v = 1; v < St::S; ++v) {
                        double2 E = node_e(u, v);
                        t1 = __builtin_fma(wy[v], E.x, t1);
                        t2 = __builtin_fma(wy[v], E.y, t2);
                    }
                    e1 = __builtin_fma(wx[u], t1, e1);
                    e2 = __builtin_fma(wx[u], t2, e2);
                }
                if (BZ == PIC_BZ_GRID) {
#pragma unroll
                    for (int u = 0; u < St::S; ++u) {
                        double t3 = wy[0] * node_b(u, 0);
#pragma unroll
                        for (int v = 1; v < St::S; ++v) t3 = __builtin_fma(wy[v], node_b(u, v), t3);
                        b3 = __builtin_fma(wx[u], t3, b3);
                    }
                }
            }
            if (BZ == PIC_BZ_UNIFORM) b3 = a.bz0;
            // kick: pushers_cic.jl:38-49
            const double dt = a.dt;
            double v1 = pvx[k], v2 = pvy[k];
            v1 += 0.5 * dt * e1;
            v2 += 0.5 * dt * e2;
            if (BZ != PIC_BZ_NONE) {
                double tantheta = 0.5 * dt * b3;
                double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
                v1 += v2 * tantheta;
                v2 += -v1 * sintheta;
                v1 += v2 * tantheta;
            }
            pvx[k] = v1 + 0.5 * dt * e1;
            pvy[k] = v2 + 0.5 * dt * e2;
        }
    }
    // drift: pushers_cic.jl:55-72, NDRIFT times with the same length
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!live[k]) continue;
        double vxk = GATHER ? pvx[k] : 0.0, vyk = GATHER ? pvy[k] : 0.0;
        if (!GATHER && NDRIFT > 0) {
            int64_t p = base + (int64_t)k * ST + tid;
            vxk = a.vx[p];
            vyk = a.vy[p];
        }
#pragma unroll
        for (int s = 0; s < NDRIFT; ++s) {
            double p1 = px[k] + a.dtd * vxk;
            double p2 = py[k] + a.dtd * vyk;
            if (p1 > a.dimx) p1 -= a.dimx;
            if (p2 > a.dimy) p2 -= a.dimy;
            if (p1 < 0.0) p1 += a.dimx;
            if (p2 < 0.0) p2 += a.dimy;
            px[k] = p1;
            py[k] = p2;
        }
        int64_t p = base + (int64_t)k * ST + tid;
        if (NDRIFT > 0) {
            a.x[p] = px[k];
            a.y[p] = py[k];
        }
        if (GATHER) {
            a.vx[p] = pvx[k];
            a.vy[p] = pvy[k];
        }
    }
    if (!DEPOSIT) return;

    // deposit: compute_rho (CIC: trunc index, pushers_cic.jl:86-87; M6: floor,
    // pushers_m6.jl:248-249)
    int ci[PPT], cj[PPT];
    double fx[PPT], fy[PPT];
    int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        ref_cell(px[k], a.dx, nx, K == PIC_CIC, ci[k], fx[k]);
        ref_cell(py[k], a.dy, ny, K == PIC_CIC, cj[k], fy[k]);
        if (live[k]) {
            int di = rel(ci[k], iref, nx), dj = rel(cj[k], jref, ny);
            mnx = min(mnx, di);
            mxx = max(mxx, di);
            mny = min(mny, dj);
            mxy = max(mxy, dj);
        }
    }
    __syncthreads();  // gather tile no longer read; sbox reset visible
    Box b = block_box<St::HL, St::HR>(mnx, mxx, mny, mxy, sbox, CAP64);
    double *tr = smem64;
    if (b.fits) {
        const int nt = b.tw * b.th;
        for (int t = tid; t < nt; t += ST) tr[t] = 0.0;
    } else if (tid == 0) {
        atomicAdd(a.slow, 1ull);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!live[k]) continue;
        int64_t p = base + (int64_t)k * ST + tid;
        const double wgt = a.w ? a.w[p] : a.w0;
        int la = rel(ci[k], iref, nx) - (b.x0 + St::HL);
        int lb = rel(cj[k], jref, ny) - (b.y0 + St::HL);
        auto add = [&](int u, int v, double q) {
            if (b.fits) {
                __hip_atomic_fetch_add(&tr[(la + u) + b.tw * (lb + v)], q, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                int64_t g = (int64_t)wrap1(ci[k] - St::HL + u, nx) +
                            (int64_t)nx * wrap1(cj[k] - St::HL + v, ny);
                unsafeAtomicAdd(&a.acc[g], q);
            }
        };
        if (K == PIC_CIC) {
            double dxp = fx[k], dyp = fy[k];
            double a1 = (1 - dxp) * (1 - dyp);
            double a2 = dxp * (1 - dyp);
            double a3 = dxp * dyp;
            double a4 = (1 - dxp) * dyp;
            add(0, 0, a1 * wgt);
            add(1, 0, a2 * wgt);
            add(1, 1, a3 * wgt);
            add(0, 1, a4 * wgt);
        } else {
            double wx[6], wy[6];
            deposit_w<K>(fx[k], ci[k], wx);
            deposit_w<K>(fy[k], cj[k], wy);
#pragma unroll
            for (int u = 0; u < St::S; ++u)
#pragma unroll
                for (int v = 0; v < St::S; ++v) add(u, v, wx[u] * wy[v] * wgt);
        }
    }
    if (!b.fits) return;
    __syncthreads();
    const int nt = b.tw * b.th;
    for (int t = tid; t < nt; t += ST) {
        double q = tr[t];
        if (q != 0.0) unsafeAtomicAdd(&a.acc[tile_node(b, t % b.tw, t / b.tw, iref, jref, nx, ny)], q);
    }
}

// ============================================================ FP32 kernel

__device__ __forceinline__ int drift1(float o, float v, float dtd, float &onew) {
    float t = __builtin_fmaf(v, dtd, o);
    float s = floorf(t);
    float r = t - s;
    if (r >= 1.0f) {
        r = 0.0f;
        s = s + 1.0f;
    }
    onew = r;
    return (int)s;
}

// Compensated add of q into the LDS pair (hi[t], lo[t]).
__device__ __forceinline__ void kahan_add(float *hi, float *lo, int t, float q) {
    float old = __hip_atomic_fetch_add(&hi[t], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    float s = old + q;
    float bp = s - old;
    float err = (old - (s - bp)) + (q - bp);
    __hip_atomic_fetch_add(&lo[t], err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int K, bool GATHER, bool DRIFT, bool DEPOSIT, int BZ, int PPT>
__global__ __launch_bounds__(ST) void k_step_f32(ArgsF32 a) {
    using St = Stencil<K>;
    extern __shared__ float smem32[];
    __shared__ int sbox[4];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * ST * PPT;
    const int nx = a.nx, ny = a.ny;

    int ci[PPT], cj[PPT];
    float fx[PPT], fy[PPT], ux[PPT], uy[PPT];
    bool live[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        int64_t p = base + (int64_t)k * ST + tid;
        live[k] = p < a.n;
        int c = 0;
        if (live[k]) {
            c = a.cell[p];
            const float4 o = a.ov[p];
            fx[k] = o.x;
            fy[k] = o.y;
            ux[k] = o.z;
            uy[k] = o.w;
        }
        ci[k] = c & (nx - 1);
        cj[k] = c >> a.lognx;
    }
    init_box(sbox);
    const int c0 = a.cell[base];
    const int iref = c0 & (nx - 1), jref = c0 >> a.lognx;
    __syncthreads();

    if (GATHER) {
        int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (live[k]) {
                int di = rel(ci[k], iref, nx), dj = rel(cj[k], jref, ny);
                mnx = min(mnx, di);
                mxx = max(mxx, di);
                mny = min(mny, dj);
                mxy = max(mxy, dj);
            }
        }
        Box b = block_box<St::HL, St::HR>(mnx, mxx, mny, mxy, sbox, CAP32);
        float2 *te = reinterpret_cast<float2 *>(smem32);
        float *tb = smem32 + 2 * CAP32;
        if (b.fits) {
            const int nt = b.tw * b.th;
            for (int t = tid; t < nt; t += ST) {
                int64_t g = tile_node(b, t % b.tw, t / b.tw, iref, jref, nx, ny);
                te[t] = a.e32[g];
                if (BZ == PIC_BZ_GRID) tb[t] = a.bz[g];
            }
        } else if (tid == 0) {
            atomicAdd(a.slow, 1ull);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!live[k]) continue;
            int la = rel(ci[k], iref, nx) - (b.x0 + St::HL);
            int lb = rel(cj[k], jref, ny) - (b.y0 + St::HL);
            auto node_e = [&](int u, int v) -> float2 {
                if (b.fits) return te[(la + u) + b.tw * (lb + v)];
                int64_t g = (int64_t)wrap1(ci[k] - St::HL + u, nx) +
                            (int64_t)nx * wrap1(cj[k] - St::HL + v, ny);
                return a.e32[g];
            };
            auto node_b = [&](int u, int v) -> float {
                if (b.fits) return tb[(la + u) + b.tw * (lb + v)];
                int64_t g = (int64_t)wrap1(ci[k] - St::HL + u, nx) +
                            (int64_t)nx * wrap1(cj[k] - St::HL + v, ny);
                return a.bz[g];
            };
            float e1, e2, b3 = 0.0f;
            if (K == PIC_CIC) {
                float sx = 1.0f - fx[k], sy = 1.0f - fy[k];
                float a0 = sx * sy, a1 = fx[k] * sy, a2 = fx[k] * fy[k], a3 = sx * fy[k];
                float2 E00 = node_e(0, 0), E10 = node_e(1, 0), E11 = node_e(1, 1),
                       E01 = node_e(0, 1);
                e1 = a0 * E00.x;
                e1 = __builtin_fmaf(a1, E10.x, e1);
                e1 = __builtin_fmaf(a2, E11.x, e1);
                e1 = __builtin_fmaf(a3, E01.x, e1);
                e2 = a0 * E00.y;
                e2 = __builtin_fmaf(a1, E10.y, e2);
                e2 = __builtin_fmaf(a2, E11.y, e2);
                e2 = __builtin_fmaf(a3, E01.y, e2);
                if (BZ == PIC_BZ_GRID) {
                    b3 = a0 * node_b(0, 0);
                    b3 = __builtin_fmaf(a1, node_b(1, 0), b3);
                    b3 = __builtin_fmaf(a2, node_b(1, 1), b3);
                    b3 = __builtin_fmaf(a3, node_b(0, 1), b3);
                }
            } else {
                float wx[6], wy[6];
                gather_w<K>(fx[k], ci[k], wx);
                gather_w<K>(fy[k], cj[k], wy);
                e1 = 0.0f;
                e2 = 0.0f;
#pragma unroll
                for (int u = 0; u < St::S; ++u) {
                    float2 E0 = node_e(u, 0);
                    float t1 = wy[0] * E0.x, t2 = wy[0] * E0.y;
#pragma unroll
                    for (int v = 1; v < St::S; ++v) {
                        float2 E = node_e(u, v);
                        t1 = __builtin_fmaf(wy[v], E.x, t1);
                        t2 = __builtin_fmaf(wy[v], E.y, t2);
                    }
                    if (u == 0) {
                        e1 = wx[0] * t1;
                        e2 = wx[0] * t2;
                    } else {
                        e1 = __builtin_fmaf(wx[u], t1, e1);
                        e2 = __builtin_fmaf(wx[u], t2, e2);
                    }
                }
                if (BZ == PIC_BZ_GRID) {
#pragma unroll
                    for (int u = 0; u < St::S; ++u) {
                        float t3 = wy[0] * node_b(u, 0);
#pragma unroll
                        for (int v = 1; v < St::S; ++v) t3 = __builtin_fmaf(wy[v], node_b(u, v), t3);
                        b3 = u == 0 ? wx[0] * t3 : __builtin_fmaf(wx[u], t3, b3);
                    }
                }
            }
            if (BZ == PIC_BZ_UNIFORM) b3 = a.bz0;
            float v1 = __builtin_fmaf(a.hdt, e1, ux[k]);
            float v2 = __builtin_fmaf(a.hdt, e2, uy[k]);
            if (BZ != PIC_BZ_NONE) {
                float tn = a.hdt * b3;
                float sn = (2.0f * tn) / __builtin_fmaf(tn, tn, 1.0f);
                v1 = __builtin_fmaf(v2, tn, v1);
                v2 = __builtin_fmaf(-v1, sn, v2);
                v1 = __builtin_fmaf(v2, tn, v1);
            }
            ux[k] = __builtin_fmaf(a.hdt, e1, v1);
            uy[k] = __builtin_fmaf(a.hdt, e2, v2);
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!live[k]) continue;
        int64_t p = base + (int64_t)k * ST + tid;
        if (DRIFT) {
            float nfx, nfy;
            int si = drift1(fx[k], ux[k], a.dtdx, nfx);
            int sj = drift1(fy[k], uy[k], a.dtdy, nfy);
            ci[k] = picd::wrapi(ci[k] + si, nx);
            cj[k] = picd::wrapi(cj[k] + sj, ny);
            fx[k] = nfx;
            fy[k] = nfy;
            a.cell[p] = ci[k] + nx * cj[k];
        }
        if (GATHER || DRIFT) a.ov[p] = make_float4(fx[k], fy[k], ux[k], uy[k]);
    }
    if (!DEPOSIT) return;

    int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (live[k]) {
            int di = rel(ci[k], iref, nx), dj = rel(cj[k], jref, ny);
            mnx = min(mnx, di);
            mxx = max(mxx, di);
            mny = min(mny, dj);
            mxy = max(mxy, dj);
        }
    }
    __syncthreads();
    Box b = block_box<St::HL, St::HR>(mnx, mxx, mny, mxy, sbox, CAP32);
    float *hi = smem32, *lo = smem32 + CAP32;
    if (b.fits) {
        const int nt = b.tw * b.th;
        for (int t = tid; t < nt; t += ST) {
            hi[t] = 0.0f;
            lo[t] = 0.0f;
        }
    } else if (tid == 0) {
        atomicAdd(a.slow, 1ull);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!live[k]) continue;
        int la = rel(ci[k], iref, nx) - (b.x0 + St::HL);
        int lb = rel(cj[k], jref, ny) - (b.y0 + St::HL);
        auto add = [&](int u, int v, float q) {
            if (b.fits) {
                kahan_add(hi, lo, (la + u) + b.tw * (lb + v), q);
            } else {
                int64_t g = (int64_t)wrap1(ci[k] - St::HL + u, nx) +
                            (int64_t)nx * wrap1(cj[k] - St::HL + v, ny);
                unsafeAtomicAdd(&a.acc[g], (double)q);
            }
        };
        if (K == PIC_CIC) {
            float sx = 1.0f - fx[k], sy = 1.0f - fy[k];
            add(0, 0, (sx * sy) * a.w);
            add(1, 0, (fx[k] * sy) * a.w);
            add(1, 1, (fx[k] * fy[k]) * a.w);
            add(0, 1, (sx * fy[k]) * a.w);
        } else {
            float wx[6], wy[6];
            deposit_w<K>(fx[k], ci[k], wx);
            deposit_w<K>(fy[k], cj[k], wy);
#pragma unroll
            for (int v = 0; v < St::S; ++v) wy[v] = wy[v] * a.w;
#pragma unroll
            for (int u = 0; u < St::S; ++u)
#pragma unroll
                for (int v = 0; v < St::S; ++v) add(u, v, wx[u] * wy[v]);
        }
    }
    if (!b.fits) return;
    __syncthreads();
    const int nt = b.tw * b.th;
    for (int t = tid; t < nt; t += ST) {
        double q = (double)hi[t] + (double)lo[t];
        if (q != 0.0) unsafeAtomicAdd(&a.acc[tile_node(b, t % b.tw, t / b.tw, iref, jref, nx, ny)], q);
    }
}

// ============================================================== dispatch

constexpr int PPT64 = 4;
constexpr int PPT32 = 4;

static size_t smem64_bytes(int bz) { return (size_t)CAP64 * (bz == PIC_BZ_GRID ? 24 : 16); }
static size_t smem32_bytes(int bz) { return (size_t)CAP32 * (bz == PIC_BZ_GRID ? 12 : 8); }

template <int K, bool G, int ND, bool D>
static void launch64_bz(Ctx *c, const ArgsF64 &a, int bz, dim3 grid) {
    size_t sm = smem64_bytes(bz);
    switch (bz) {
        case PIC_BZ_NONE:
            hipLaunchKernelGGL((k_step_f64<K, G, ND, D, PIC_BZ_NONE, PPT64>), grid, dim3(ST), sm,
                               c->stream, a);
            break;
        case PIC_BZ_UNIFORM:
            hipLaunchKernelGGL((k_step_f64<K, G, ND, D, PIC_BZ_UNIFORM, PPT64>), grid, dim3(ST), sm,
                               c->stream, a);
            break;
        default:
            hipLaunchKernelGGL((k_step_f64<K, G, ND, D, PIC_BZ_GRID, PPT64>), grid, dim3(ST), sm,
                               c->stream, a);
    }
}

// mode: 0 prologue (drift, deposit), 1 step (kick, 2 drifts, deposit),
// 2 last (kick, drift), 3 deposit only, 4 kick only
template <int K>
static int launch64_k(Ctx *c, const ArgsF64 &a, int mode, int bz) {
    dim3 grid((unsigned)((a.n + ST * PPT64 - 1) / (ST * PPT64)));
    switch (mode) {
        case 0: launch64_bz<K, false, 1, true>(c, a, PIC_BZ_NONE, grid); break;
        case 1: launch64_bz<K, true, 2, true>(c, a, bz, grid); break;
        case 2: launch64_bz<K, true, 1, false>(c, a, bz, grid); break;
        case 3: launch64_bz<K, false, 0, true>(c, a, PIC_BZ_NONE, grid); break;
        case 4: launch64_bz<K, true, 0, false>(c, a, bz, grid); break;
        default: return fail(PIC_ERR_ARG, "step mode %d", mode);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int launch_step64(Ctx *c, int kernel, int mode, double dt, double dtd) {
    if (c->n == 0) return PIC_OK;
    if (mode <= 2) c->sort_fresh = false;  // the pass moves the particles
    ArgsF64 a;
    a.n = c->n;
    a.x = c->x; a.y = c->y; a.vx = c->vx; a.vy = c->vy;
    a.w = c->uniform_w ? nullptr : c->w;
    a.w0 = c->w0;
    a.ex = c->ex; a.ey = c->ey; a.bz = c->bz; a.bz0 = c->bz0;
    a.nx = c->nx; a.ny = c->ny;
    a.dx = c->dx; a.dy = c->dy; a.dimx = c->dimx; a.dimy = c->dimy;
    a.dt = dt; a.dtd = dtd;
    a.acc = c->rho_acc;
    a.slow = c->slow_chunks;
    int rc = PIC_OK;
    PIC_SHAPE_SWITCH(kernel, rc = launch64_k<KK>(c, a, mode, c->bz_mode));
    return rc;
}

template <int K, bool G, bool DR, bool D>
static void launch32_bz(Ctx *c, const ArgsF32 &a, int bz, dim3 grid) {
    size_t sm = smem32_bytes(bz);
    switch (bz) {
        case PIC_BZ_NONE:
            hipLaunchKernelGGL((k_step_f32<K, G, DR, D, PIC_BZ_NONE, PPT32>), grid, dim3(ST), sm,
                               c->stream, a);
            break;
        case PIC_BZ_UNIFORM:
            hipLaunchKernelGGL((k_step_f32<K, G, DR, D, PIC_BZ_UNIFORM, PPT32>), grid, dim3(ST), sm,
                               c->stream, a);
            break;
        default:
            hipLaunchKernelGGL((k_step_f32<K, G, DR, D, PIC_BZ_GRID, PPT32>), grid, dim3(ST), sm,
                               c->stream, a);
    }
}

template <int K>
static int launch32_k(Ctx *c, const ArgsF32 &a, int mode, int bz) {
    dim3 grid((unsigned)((a.n + ST * PPT32 - 1) / (ST * PPT32)));
    switch (mode) {
        case 0: launch32_bz<K, false, true, true>(c, a, PIC_BZ_NONE, grid); break;
        case 1: launch32_bz<K, true, true, true>(c, a, bz, grid); break;
        case 2: launch32_bz<K, true, true, false>(c, a, bz, grid); break;
        case 3: launch32_bz<K, false, false, true>(c, a, PIC_BZ_NONE, grid); break;
        case 4: launch32_bz<K, true, false, false>(c, a, bz, grid); break;
        default: return fail(PIC_ERR_ARG, "step mode %d", mode);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// dt: kick length; dtd: drift length (one full drift per step in FP32).
int launch_step32(Ctx *c, int kernel, int mode, double dt, double dtd) {
    if (c->n == 0) return PIC_OK;
    if (mode <= 2) c->sort_fresh = false;  // the pass moves the particles
    ArgsF32 a;
    a.n = c->n;
    a.cell = c->cell; a.ov = c->ov32;
    a.w = (float)c->w0;
    a.e32 = reinterpret_cast<const float2 *>(c->e32);
    a.bz = c->bz32;
    a.bz0 = (float)c->bz0;
    a.nx = c->nx; a.ny = c->ny;
    a.lognx = 0;
    while ((1 << a.lognx) < c->nx) ++a.lognx;
    a.hdt = (float)(0.5 * dt);
    a.dtdx = (float)(dtd / c->dx);
    a.dtdy = (float)(dtd / c->dy);
    a.acc = c->rho_acc;
    a.slow = c->slow_chunks;
    int rc = PIC_OK;
    PIC_SHAPE_SWITCH(kernel, rc = launch32_k<KK>(c, a, mode, c->bz_mode));
    return rc;
}

}  // namespace pic
