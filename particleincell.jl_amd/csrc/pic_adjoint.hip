// pic_adjoint.hip -- adjoint (reverse-time) sweep of the device time loop.
//
// The reference has no adjoint.  This is the exact discrete adjoint of its
// loop (examples/landau_damping_poisson_cic_m6.jl:70-83) with bz = 0,
//   x' = x + h v;  E_s = P D(x');  v <- v + dt G(x') E_s;  x <- x' + h v,
// for J = sum(ex^2 + ey^2) dx dy of the last step's field (h = dt/2).  With
// lx, lv the adjoints of x, v at the start of step s+1:
//   last step:  mu = 2 E dx dy,  lx' = 0,  lvt = 0
//   other s:    lvt = lv + h lx,  lx' = lx + dt (lvt . grad G(x') E_s),
//               mu = dt G(x')^T lvt                  (the gather's adjoint: a scatter)
//   every s:    lrho = P^T mu = -(P_x mu_x + P_y mu_y)  (spectral, pic_poisson)
//               lx' += (w / dx dy) grad W(x') . lrho   (the deposit's adjoint: a gather)
//               lx = lx',  lv = lvt + h lx'
//
// Forward pass: per step a drift kernel (x' and, in the reverse sweep, its
// copy), the deposit (the time loop's segment pass), the solve (E_s kept,
// 2 nx*ny doubles per step) and a kick + drift kernel.  The particles are
// lane-sorted every `ckpt` steps, where the state, the sort permutation and
// the cell offsets are stored, and every `sort_every` steps between.  The
// reverse sweep recomputes each segment's x'_s from its checkpoint with the
// same kernels, sorts and stored E_s, so the recomputed trajectory and its
// slot orders are bit-identical, then runs the adjoint passes backwards:
// the gather adjoint (a scatter of mu) and the deposit adjoint (a gather of
// lrho) over segment windows in LDS, as the time loop's passes, with the
// particles that left their window completed from global memory.  lx, lv
// are carried through every sort's permutation (and finally into the
// caller's particle order).
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_internal.h"
#include "pic_segcore.h"

namespace pic {

int launch_seg64(Ctx *c, int kernel, int mode, double dt, double dtd);

constexpr int AT = 256;

// Shape weights and derivatives (per cell length) of nodes i0+o .. i0+o+S-1,
// o = 0 (CIC) or -2 (M6), at the reference's in-cell offset f.
template <int K>
__device__ __forceinline__ void shape_d(double f, double *w, double *dw) {
    if (K == PIC_CIC) {
        w[0] = 1 - f;
        w[1] = f;
        dw[0] = -1.0;
        dw[1] = 1.0;
    } else {
        picd::m6_weights_d(f, w);
        // d/dt of 120 w3 = 26+50t+20t^2-20t^3-20t^4+10t^5,
        // 120 w4 = 1+5t+10t^2+10t^3+5t^4-5t^5, 120 w5 = t^5; w_k(t) = w_{5-k}(1-t)
        const double t = f, s = 1.0 - f, c = 1.0 / 120.0;
        const double d3 = (50 + t * (40 + t * (-60 + t * (-80 + t * 50)))) * c;
        const double d4 = (5 + t * (20 + t * (30 + t * (20 - t * 25)))) * c;
        const double d5 = (5 * t * t * t * t) * c;
        const double e3 = (50 + s * (40 + s * (-60 + s * (-80 + s * 50)))) * c;
        const double e4 = (5 + s * (20 + s * (30 + s * (20 - s * 25)))) * c;
        const double e5 = (5 * s * s * s * s) * c;
        dw[0] = -e5;
        dw[1] = -e4;
        dw[2] = -e3;
        dw[3] = d3;
        dw[4] = d4;
        dw[5] = d5;
    }
}

__device__ __forceinline__ int ref_cell_a(double x, double dx, int n, double &d) {
    const double xp = x / dx;
    return picd::cell_of(xp, floor(xp), n, d);
}

// x' = x + h v (push_x!, pushers_cic.jl:55-72); xs receives a copy.
__global__ void k_adj_drift(int64_t n, double *x, double *y, const double *vx, const double *vy,
                            double h, double dimx, double dimy, double *xs, double *ys) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double p1 = x[k] + h * vx[k];
    double p2 = y[k] + h * vy[k];
    if (p1 > dimx) p1 -= dimx;
    if (p2 > dimy) p2 -= dimy;
    if (p1 < 0.0) p1 += dimx;
    if (p2 < 0.0) p2 += dimy;
    x[k] = p1;
    y[k] = p2;
    if (xs) {
        xs[k] = p1;
        ys[k] = p2;
    }
}

// v <- v + dt G(x) E (the two half kicks of push_v! with bz = 0, same
// arithmetic as the time loop's passes), then x <- x + h v.
template <int K>
__global__ void k_adj_kick(int64_t n, double *x, double *y, double *vx, double *vy,
                           const double *__restrict__ ex, const double *__restrict__ ey, int nx,
                           int ny, double dx, double dy, double dt, double h, double dimx,
                           double dimy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
    double fx, fy;
    const int ci = ref_cell_a(x[k], dx, nx, fx), cj = ref_cell_a(y[k], dy, ny, fy);
    double e1, e2;
    if (K == PIC_CIC) {
        const double a1 = (1 - fx) * (1 - fy), a2 = fx * (1 - fy), a3 = fx * fy,
                     a4 = (1 - fx) * fy;
        const int ip = picd::wrap2(ci + 1, nx), jp = picd::wrap2(cj + 1, ny);
        const int64_t c00 = ci + (int64_t)nx * cj, c10 = ip + (int64_t)nx * cj,
                      c11 = ip + (int64_t)nx * jp, c01 = ci + (int64_t)nx * jp;
        e1 = a1 * ex[c00] + a2 * ex[c10] + a3 * ex[c11] + a4 * ex[c01];
        e2 = a1 * ey[c00] + a2 * ey[c10] + a3 * ey[c11] + a4 * ey[c01];
    } else {
        double wx[6], wy[6];
        picd::m6_weights_d(fx, wx);
        picd::m6_weights_d(fy, wy);
        e1 = 0.0;
        e2 = 0.0;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int iu = picd::wrap2(ci - HL + u, nx);
            const int64_t g0 = iu + (int64_t)nx * picd::wrap2(cj - HL, ny);
            double t1 = wy[0] * ex[g0], t2 = wy[0] * ey[g0];
#pragma unroll
            for (int v = 1; v < S; ++v) {
                const int64_t g = iu + (int64_t)nx * picd::wrap2(cj - HL + v, ny);
                t1 = __builtin_fma(wy[v], ex[g], t1);
                t2 = __builtin_fma(wy[v], ey[g], t2);
            }
            e1 = __builtin_fma(wx[u], t1, e1);
            e2 = __builtin_fma(wx[u], t2, e2);
        }
    }
    double v1 = vx[k], v2 = vy[k];
    v1 += 0.5 * dt * e1;
    v2 += 0.5 * dt * e2;
    v1 = v1 + 0.5 * dt * e1;
    v2 = v2 + 0.5 * dt * e2;
    vx[k] = v1;
    vy[k] = v2;
    double p1 = x[k] + h * v1;
    double p2 = y[k] + h * v2;
    if (p1 > dimx) p1 -= dimx;
    if (p2 > dimy) p2 -= dimy;
    if (p1 < 0.0) p1 += dimx;
    if (p2 < 0.0) p2 += dimy;
    x[k] = p1;
    y[k] = p2;
}

// Gather adjoint at x' for one particle (step s < last): lvt = lv + h lx;
// lx += dt (lvt . grad G E_s); mu += dt G^T lvt (add_mu(u, v, qx, qy) for
// stencil node (u, v)); lv <- lvt.  node_e(u, v) is (ex, ey) at the node.
template <int K, class FE, class FA>
__device__ __forceinline__ void adj_gather64_one(double x, double y, double &lx, double &ly,
                                                 double &lvx, double &lvy, int nx, int ny,
                                                 double dx, double dy, double dt, double h,
                                                 int &ci, int &cj, FE node_e, FA add_mu) {
    constexpr int S = K == PIC_CIC ? 2 : 6;
    double fx, fy;
    ci = ref_cell_a(x, dx, nx, fx);
    cj = ref_cell_a(y, dy, ny, fy);
    double wx[6], wy[6], dwx[6], dwy[6];
    shape_d<K>(fx, wx, dwx);
    shape_d<K>(fy, wy, dwy);
    const double ax = lx, ay = ly;
    const double tx = lvx + h * ax, ty = lvy + h * ay;
    double gxx = 0.0, gxy = 0.0, gyx = 0.0, gyy = 0.0;  // d ex/dx, d ex/dy, d ey/dx, d ey/dy
#pragma unroll
    for (int u = 0; u < S; ++u) {
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const double2 e = node_e(u, v);
            gxx += dwx[u] * wy[v] * e.x;
            gxy += wx[u] * dwy[v] * e.x;
            gyx += dwx[u] * wy[v] * e.y;
            gyy += wx[u] * dwy[v] * e.y;
            const double wab = wx[u] * wy[v];
            add_mu(u, v, dt * tx * wab, dt * ty * wab);
        }
    }
    lx = ax + dt * (tx * gxx + ty * gyx) / dx;
    ly = ay + dt * (tx * gxy + ty * gyy) / dy;
    lvx = tx;
    lvy = ty;
}

// Deposit adjoint at x' for one particle: lx += (w/dxdy) grad W . lrho; then
// lv += h lx.  node_r(u, v) is lrho at the node.
template <int K, class FR>
__device__ __forceinline__ void adj_deposit64_one(double x, double y, double &lx, double &ly,
                                                  double &lvx, double &lvy, int nx, int ny,
                                                  double dx, double dy, double wdd, double h,
                                                  int &ci, int &cj, FR node_r) {
    constexpr int S = K == PIC_CIC ? 2 : 6;
    double fx, fy;
    ci = ref_cell_a(x, dx, nx, fx);
    cj = ref_cell_a(y, dy, ny, fy);
    double wx[6], wy[6], dwx[6], dwy[6];
    shape_d<K>(fx, wx, dwx);
    shape_d<K>(fy, wy, dwy);
    double gx = 0.0, gy = 0.0;
#pragma unroll
    for (int u = 0; u < S; ++u) {
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const double r = node_r(u, v);
            gx += dwx[u] * wy[v] * r;
            gy += wx[u] * dwy[v] * r;
        }
    }
    const double ax = lx + wdd * gx / dx, ay = ly + wdd * gy / dy;
    lx = ax;
    ly = ay;
    lvx = lvx + h * ax;
    lvy = lvy + h * ay;
}

// Arguments of the reverse passes over the positions x'_s of one step, in
// the slot order of the lane sort whose offsets `off` are.
struct AdjSeg64 {
    const double *xs, *ys;            // x'_s per slot
    double *lx, *ly, *lvx, *lvy;      // adjoints per slot
    const double *ex, *ey;            // gather adjoint: E_s
    const double *lrho;               // deposit adjoint: P^T mu
    double *mux, *muy;                // gather adjoint: mu accumulators
    int nx, ny, ncells;
    double dx, dy, dt, h, wdd;
    const int64_t *off;
    int bps, margin, cap;
    unsigned long long *slow;
    uint32_t *dlist;
    unsigned *dcount;
};

// The slots [p0, p1) of block (s, bi) and the window of segment s.
struct AdjSlice {
    int64_t p0, p1;
    Win W;
    bool fits;
};

template <int K>
__device__ __forceinline__ AdjSlice adj_slice(const AdjSeg64 &a) {
    constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
    const int s = blockIdx.x / a.bps, bi = blockIdx.x - s * a.bps;
    const int ce = min(64 * s + 64, a.ncells);
    const int64_t s0 = a.off[64 * s], s1 = a.off[ce];
    const int64_t len = s1 - s0;
    AdjSlice r;
    r.p0 = s0 + len * bi / a.bps;
    r.p1 = s0 + len * (bi + 1) / a.bps;
    r.W = seg_window<S, HL>(s, a.nx, a.ny, a.ncells, a.margin, a.cap);
    r.fits = r.W.tw * r.W.th <= a.cap;
    return r;
}

// Gather adjoint over one segment slice: E_s staged in an LDS window, mu
// summed in FP64 LDS tiles (ds_add_f64) and flushed once per window node with
// FP64 global atomics; a particle whose stencil leaves the window goes to
// the deferred list.
template <int K>
__global__ __launch_bounds__(SEG_T) void k_adjseg64_gather(AdjSeg64 a) {
    constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
    extern __shared__ __align__(16) char smem_adj64[];
    const int tid = threadIdx.x, nx = a.nx, ny = a.ny;
    const AdjSlice sl = adj_slice<K>(a);
    const Win &W = sl.W;
    double2 *te = reinterpret_cast<double2 *>(smem_adj64);
    double *tmx = reinterpret_cast<double *>(te + a.cap), *tmy = tmx + a.cap;
    if (sl.fits) {
        const int nt = W.tw * W.th;
        for (int t = tid; t < nt; t += SEG_T) {
            const int u = t % W.tw, v = t / W.tw;
            const int64_t g = (int64_t)picd::wrap2(W.x0 + u, nx) +
                              (int64_t)nx * picd::wrap2(W.y0 + v, ny);
            te[t] = make_double2(a.ex[g], a.ey[g]);
            tmx[t] = 0.0;
            tmy[t] = 0.0;
        }
    }
    __syncthreads();
    for (uint32_t p = (uint32_t)sl.p0 + tid; p < (uint32_t)sl.p1; p += SEG_T) {
        const double x = a.xs[p], y = a.ys[p];
        double dxo;
        const int ci0 = ref_cell_a(x, a.dx, nx, dxo), cj0 = ref_cell_a(y, a.dy, ny, dxo);
        const int lx0 = picd::wrap2(ci0 - HL - W.x0, nx), ly0 = picd::wrap2(cj0 - HL - W.y0, ny);
        const bool in = sl.fits && lx0 + S - 1 < W.tw && ly0 + S - 1 < W.th;
        defer_push(!in, p, a.dlist, a.dcount);
        if (!in) continue;
        double Lx = a.lx[p], Ly = a.ly[p], Lvx = a.lvx[p], Lvy = a.lvy[p];
        int ci, cj;
        adj_gather64_one<K>(
            x, y, Lx, Ly, Lvx, Lvy, nx, ny, a.dx, a.dy, a.dt, a.h, ci, cj,
            [&](int u, int v) { return te[(lx0 + u) + W.tw * (ly0 + v)]; },
            [&](int u, int v, double qa, double qb) {
                const int t = (lx0 + u) + W.tw * (ly0 + v);
                __hip_atomic_fetch_add(&tmx[t], qa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(&tmy[t], qb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            });
        a.lx[p] = Lx;
        a.ly[p] = Ly;
        a.lvx[p] = Lvx;
        a.lvy[p] = Lvy;
    }
    if (!sl.fits) return;
    __syncthreads();
    const int nt = W.tw * W.th;
    for (int t = tid; t < nt; t += SEG_T) {
        const double qa = tmx[t], qb = tmy[t];
        if (qa != 0.0 || qb != 0.0) {
            const int u = t % W.tw, v = t / W.tw;
            const int64_t g = (int64_t)picd::wrap2(W.x0 + u, nx) +
                              (int64_t)nx * picd::wrap2(W.y0 + v, ny);
            unsafeAtomicAdd(&a.mux[g], qa);
            unsafeAtomicAdd(&a.muy[g], qb);
        }
    }
}

// The deferred particles of k_adjseg64_gather, from global memory.
template <int K>
__global__ __launch_bounds__(DEFER_T) void k_adjdefer64_gather(AdjSeg64 a) {
    constexpr int HL = K == PIC_CIC ? 0 : 2;
    const unsigned cnt = __hip_atomic_load(a.dcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nx = a.nx, ny = a.ny;
    for (unsigned i = blockIdx.x * DEFER_T + threadIdx.x; i < cnt; i += gridDim.x * DEFER_T) {
        const uint32_t p = a.dlist[i];
        double Lx = a.lx[p], Ly = a.ly[p], Lvx = a.lvx[p], Lvy = a.lvy[p];
        int ci = 0, cj = 0;
        auto node = [&](int u, int v) {
            return (int64_t)picd::wrap2(ci - HL + u, nx) + (int64_t)nx * picd::wrap2(cj - HL + v, ny);
        };
        adj_gather64_one<K>(
            a.xs[p], a.ys[p], Lx, Ly, Lvx, Lvy, nx, ny, a.dx, a.dy, a.dt, a.h, ci, cj,
            [&](int u, int v) {
                const int64_t g = node(u, v);
                return make_double2(a.ex[g], a.ey[g]);
            },
            [&](int u, int v, double qa, double qb) {
                const int64_t g = node(u, v);
                unsafeAtomicAdd(&a.mux[g], qa);
                unsafeAtomicAdd(&a.muy[g], qb);
            });
        a.lx[p] = Lx;
        a.ly[p] = Ly;
        a.lvx[p] = Lvx;
        a.lvy[p] = Lvy;
    }
    defer_finish(cnt, a.dcount, a.slow);
}

// Deposit adjoint over one segment slice: lrho staged in an LDS window.
template <int K>
__global__ __launch_bounds__(SEG_T) void k_adjseg64_deposit(AdjSeg64 a) {
    constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
    extern __shared__ __align__(16) char smem_adj64[];
    const int tid = threadIdx.x, nx = a.nx, ny = a.ny;
    const AdjSlice sl = adj_slice<K>(a);
    const Win &W = sl.W;
    double *tr = reinterpret_cast<double *>(smem_adj64);
    if (sl.fits) {
        const int nt = W.tw * W.th;
        for (int t = tid; t < nt; t += SEG_T) {
            const int u = t % W.tw, v = t / W.tw;
            tr[t] = a.lrho[(int64_t)picd::wrap2(W.x0 + u, nx) +
                           (int64_t)nx * picd::wrap2(W.y0 + v, ny)];
        }
    }
    __syncthreads();
    for (uint32_t p = (uint32_t)sl.p0 + tid; p < (uint32_t)sl.p1; p += SEG_T) {
        const double x = a.xs[p], y = a.ys[p];
        double dxo;
        const int ci0 = ref_cell_a(x, a.dx, nx, dxo), cj0 = ref_cell_a(y, a.dy, ny, dxo);
        const int lx0 = picd::wrap2(ci0 - HL - W.x0, nx), ly0 = picd::wrap2(cj0 - HL - W.y0, ny);
        const bool in = sl.fits && lx0 + S - 1 < W.tw && ly0 + S - 1 < W.th;
        defer_push(!in, p, a.dlist, a.dcount);
        if (!in) continue;
        double Lx = a.lx[p], Ly = a.ly[p], Lvx = a.lvx[p], Lvy = a.lvy[p];
        int ci, cj;
        adj_deposit64_one<K>(x, y, Lx, Ly, Lvx, Lvy, nx, ny, a.dx, a.dy, a.wdd, a.h, ci, cj,
                             [&](int u, int v) { return tr[(lx0 + u) + W.tw * (ly0 + v)]; });
        a.lx[p] = Lx;
        a.ly[p] = Ly;
        a.lvx[p] = Lvx;
        a.lvy[p] = Lvy;
    }
}

// The deferred particles of k_adjseg64_deposit, from global memory.
template <int K>
__global__ __launch_bounds__(DEFER_T) void k_adjdefer64_deposit(AdjSeg64 a) {
    constexpr int HL = K == PIC_CIC ? 0 : 2;
    const unsigned cnt = __hip_atomic_load(a.dcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nx = a.nx, ny = a.ny;
    for (unsigned i = blockIdx.x * DEFER_T + threadIdx.x; i < cnt; i += gridDim.x * DEFER_T) {
        const uint32_t p = a.dlist[i];
        double Lx = a.lx[p], Ly = a.ly[p], Lvx = a.lvx[p], Lvy = a.lvy[p];
        int ci = 0, cj = 0;
        adj_deposit64_one<K>(a.xs[p], a.ys[p], Lx, Ly, Lvx, Lvy, nx, ny, a.dx, a.dy, a.wdd, a.h,
                             ci, cj, [&](int u, int v) {
                                 return a.lrho[(int64_t)picd::wrap2(ci - HL + u, nx) +
                                               (int64_t)nx * picd::wrap2(cj - HL + v, ny)];
                             });
        a.lx[p] = Lx;
        a.ly[p] = Ly;
        a.lvx[p] = Lvx;
        a.lvy[p] = Lvy;
    }
    defer_finish(cnt, a.dcount, a.slow);
}

// LDS per window node: the gather adjoint holds E (16 B) and two FP64 mu
// sums (16 B) in 80 KiB (two blocks per CU; an M6 window of a 64-cell row
// segment keeps a 9-cell drift margin), the deposit adjoint lrho (8 B) in
// 40 KiB.
constexpr size_t ADJ64_GATHER_NODE = 32, ADJ64_DEPOSIT_NODE = 8;
constexpr size_t ADJ64_GATHER_LDS_BYTES = 80 * 1024;

template <int K, bool GATHER>
static int launch_adjseg64(const Ctx *c, AdjSeg64 a) {
    const size_t per_node = GATHER ? ADJ64_GATHER_NODE : ADJ64_DEPOSIT_NODE;
    a.cap = (int)((GATHER ? ADJ64_GATHER_LDS_BYTES : SEG_LDS_BYTES) / per_node) & ~15;
    a.margin = 32;
    const size_t sm = (size_t)a.cap * per_node;
    const void *k = GATHER ? reinterpret_cast<const void *>(k_adjseg64_gather<K>)
                           : reinterpret_cast<const void *>(k_adjseg64_deposit<K>);
    if (sm > 65536)
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    a.bps = seg_bps(c, k, sm);
    const dim3 grid((unsigned)(((c->nx * c->ny + 63) / 64) * a.bps));
    if (GATHER) {
        hipLaunchKernelGGL(k_adjseg64_gather<K>, grid, dim3(SEG_T), sm, c->stream, a);
        hipLaunchKernelGGL(k_adjdefer64_gather<K>, dim3(DEFER_BLOCKS), dim3(DEFER_T), 0,
                           c->stream, a);
    } else {
        hipLaunchKernelGGL(k_adjseg64_deposit<K>, grid, dim3(SEG_T), sm, c->stream, a);
        hipLaunchKernelGGL(k_adjdefer64_deposit<K>, dim3(DEFER_BLOCKS), dim3(DEFER_T), 0,
                           c->stream, a);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

__global__ void k_adj_seed(int64_t m, const double *ex, const double *ey, double dxdy, double *mux,
                           double *muy) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    mux[k] = 2.0 * ex[k] * dxdy;
    muy[k] = 2.0 * ey[k] * dxdy;
}

__global__ void k_zero2(int64_t n, double *a, double *b) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        a[k] = 0.0;
        b[k] = 0.0;
    }
}

// out[perm[r]] = in[r]: carries per-slot values back into the order the
// permutation was taken from.
__global__ void k_scatter_back(int64_t n, const uint32_t *__restrict__ perm,
                               const double *__restrict__ in, double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[perm[r]] = in[r];
}

static inline dim3 ga(int64_t n) { return dim3((unsigned)((n + AT - 1) / AT)); }

// Steps at which the sweep lane-sorts the particles: every checkpoint step
// and, with sort_every > 0, every multiple of it before the last step.  The
// forward pass and the reverse sweep's recomputation follow the same list,
// and the stable sort of a bit-identical state is the same permutation.
struct SortPlan64 {
    int ckpt, every, nsteps;
    bool at(int s) const {
        return s % ckpt == 0 || (every > 0 && s % every == 0 && s < nsteps - 1);
    }
};

template <int K>
static int adjoint_impl(Ctx *c, double dt, int nsteps, int ckpt, int sort_every, double *J,
                        double *gx, double *gv) {
    const int64_t n = c->n;
    const int nx = c->nx, ny = c->ny;
    const int64_t m = (int64_t)nx * ny;
    const double h = 0.5 * dt, dx = c->dx, dy = c->dy, dxdy = c->dx * c->dy;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    const int len = std::min(ckpt, nsteps);
    const double scale_rho = 1.0 / ((double)nx * (double)ny * dxdy);
    const double scale = 1.0 / ((double)nx * (double)ny);
    const size_t noff = (size_t)m + 1;
    const SortPlan64 sp{ckpt, sort_every, nsteps};
    int nsub = 0;  // the most sorts inside one segment after its checkpoint sort
    for (int s0 = 0; s0 < nsteps; s0 += ckpt) {
        int k = 0;
        for (int s = s0 + 1; s < std::min(nsteps, s0 + ckpt); ++s) k += sp.at(s);
        nsub = std::max(nsub, k);
    }
    PIC_TRY(ensure_energy(c, nsteps));
    PIC_TRY(ensure_defer(c));
    if (c->pending_perm) PIC_TRY(settle_state(c));
    DevBuf<double> Es, ck, xs, lam, mu, lrho;
    DevBuf<uint32_t> perms, subperm;
    DevBuf<int64_t> offs, suboff;
    PIC_TRY(Es.alloc((size_t)2 * m * nsteps));
    PIC_TRY(ck.alloc((size_t)4 * n * nseg));
    PIC_TRY(perms.alloc((size_t)n * nseg));
    PIC_TRY(offs.alloc(noff * nseg));
    PIC_TRY(xs.alloc((size_t)2 * n * len));
    PIC_TRY(lam.alloc((size_t)5 * n));
    PIC_TRY(mu.alloc((size_t)2 * m));
    PIC_TRY(lrho.alloc((size_t)m));
    PIC_TRY(subperm.alloc((size_t)n * nsub));
    PIC_TRY(suboff.alloc(noff * nsub));
    PIC_HIP(hipMemsetAsync(c->step_counter, 0, sizeof(int), c->stream));
    auto dcopy = [&](void *dst, const void *src, size_t bytes) -> int {
        PIC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    };
    double *st[4] = {nullptr, nullptr, nullptr, nullptr};
    auto state = [&]() {
        st[0] = c->x;
        st[1] = c->y;
        st[2] = c->vx;
        st[3] = c->vy;
    };
    // checkpoint g: the state right after the sort of step g ckpt, that
    // sort's permutation and cell offsets
    auto save = [&](int g) -> int {
        state();
        for (int q = 0; q < 4; ++q) PIC_TRY(dcopy(ck.p + ((size_t)4 * g + q) * n, st[q], (size_t)n * 8));
        PIC_TRY(dcopy(perms.p + (size_t)g * n, c->perm2, (size_t)n * 4));
        PIC_TRY(dcopy(offs.p + noff * g, c->cell_off, noff * 8));
        return PIC_OK;
    };
    auto restore = [&](int g) -> int {
        state();
        for (int q = 0; q < 4; ++q) PIC_TRY(dcopy(st[q], ck.p + ((size_t)4 * g + q) * n, (size_t)n * 8));
        PIC_TRY(dcopy(c->cell_off, offs.p + noff * g, noff * 8));
        c->seg_valid = true;
        c->pending_perm = nullptr;
        return PIC_OK;
    };
    // ---------------------------------------------------------- forward
    for (int s = 0; s < nsteps; ++s) {
        if (sp.at(s)) {
            PIC_TRY(sort_by_cell(c, nullptr, true));
            if (s % ckpt == 0) PIC_TRY(save(s / ckpt));
        }
        hipLaunchKernelGGL(k_adj_drift, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx, c->vy,
                           h, c->dimx, c->dimy, (double *)nullptr, (double *)nullptr);
        PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
        PIC_TRY(launch_seg64(c, K, 3, 0.0, 0.0));
        PIC_TRY(poisson_solve_e(c, c->rho_acc, scale_rho, 0, -2));
        PIC_TRY(dcopy(Es.p + (size_t)2 * m * s, c->ex, (size_t)2 * m * 8));
        if (s == nsteps - 1) break;
        hipLaunchKernelGGL(k_adj_kick<K>, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx,
                           c->vy, (const double *)c->ex, (const double *)c->ey, nx, ny, dx, dy, dt,
                           h, c->dimx, c->dimy);
    }
    PIC_HIP(hipGetLastError());
    {
        double e2[2];
        PIC_HIP(hipMemcpyAsync(e2, c->energy_hist + 2 * (nsteps - 1), 16, hipMemcpyDeviceToHost,
                               c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        *J = e2[0] + e2[1];
    }
    // ---------------------------------------------------------- reverse
    // L[0..3] = lx, ly, lvx, lvy per slot; L[4] spare for the permutations
    double *L[5];
    for (int q = 0; q < 5; ++q) L[q] = lam.p + (size_t)q * n;
    hipLaunchKernelGGL(k_zero2, ga(n), dim3(AT), 0, c->stream, n, L[0], L[1]);
    hipLaunchKernelGGL(k_zero2, ga(n), dim3(AT), 0, c->stream, n, L[2], L[3]);
    // per-slot values into the slot order before the sort of permutation pg
    auto carry_back = [&](const uint32_t *pg) {
        for (int q = 0; q < 4; ++q) {
            hipLaunchKernelGGL(k_scatter_back, ga(n), dim3(AT), 0, c->stream, n, pg,
                               (const double *)L[q], L[4]);
            std::swap(L[q], L[4]);
        }
    };
    AdjSeg64 a;
    a.mux = mu.p;
    a.muy = mu.p + m;
    a.lrho = lrho.p;
    a.nx = nx;
    a.ny = ny;
    a.ncells = (int)m;
    a.dx = dx;
    a.dy = dy;
    a.dt = dt;
    a.h = h;
    a.wdd = c->w0 / dxdy;
    a.slow = c->slow_chunks;
    a.dlist = c->defer;
    a.dcount = c->defer_count;
    std::vector<int> sub_of((size_t)len);
    for (int g = nseg - 1; g >= 0; --g) {
        const int s0 = g * ckpt, s1 = std::min(nsteps, s0 + ckpt);
        PIC_TRY(restore(g));
        // recompute x'_s, s0 <= s < s1, each in the slot order of the sort in force
        int ns = 0;
        for (int s = s0; s < s1; ++s) {
            if (s > s0 && sp.at(s)) {
                PIC_TRY(sort_by_cell(c, nullptr, true));
                PIC_TRY(dcopy(subperm.p + (size_t)ns * n, c->perm2, (size_t)n * 4));
                PIC_TRY(dcopy(suboff.p + noff * ns, c->cell_off, noff * 8));
                ++ns;
            }
            sub_of[s - s0] = ns;
            double *xo = xs.p + (size_t)2 * n * (s - s0);
            hipLaunchKernelGGL(k_adj_drift, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx,
                               c->vy, h, c->dimx, c->dimy, xo, xo + n);
            if (s == nsteps - 1) break;
            const double *E = Es.p + (size_t)2 * m * s;
            hipLaunchKernelGGL(k_adj_kick<K>, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx,
                               c->vy, E, E + m, nx, ny, dx, dy, dt, h, c->dimx, c->dimy);
        }
        for (int s = s1 - 1; s >= s0; --s) {
            const int k = sub_of[s - s0];
            const double *xo = xs.p + (size_t)2 * n * (s - s0);
            const double *E = Es.p + (size_t)2 * m * s;
            a.xs = xo;
            a.ys = xo + n;
            a.lx = L[0];
            a.ly = L[1];
            a.lvx = L[2];
            a.lvy = L[3];
            a.ex = E;
            a.ey = E + m;
            a.off = k == 0 ? offs.p + noff * g : suboff.p + noff * (k - 1);
            if (s == nsteps - 1) {
                hipLaunchKernelGGL(k_adj_seed, ga(m), dim3(AT), 0, c->stream, m, E, E + m, dxdy,
                                   mu.p, mu.p + m);
            } else {
                hipLaunchKernelGGL(k_zero2, ga(m), dim3(AT), 0, c->stream, m, mu.p, mu.p + m);
                PIC_TRY((launch_adjseg64<K, true>(c, a)));
            }
            PIC_TRY(poisson_adjoint(c, mu.p, mu.p + m, lrho.p, scale));
            PIC_TRY((launch_adjseg64<K, false>(c, a)));
            // a sort inside the segment at step s: the slots into their order before it
            if (s > s0 && sp.at(s)) carry_back(subperm.p + (size_t)(k - 1) * n);
        }
        // into the slot order of the previous segment (or the caller's order)
        carry_back(perms.p + (size_t)g * n);
    }
    PIC_HIP(hipGetLastError());
    std::vector<double> h4((size_t)4 * n);
    for (int q = 0; q < 4; ++q)
        PIC_HIP(hipMemcpyAsync(h4.data() + (size_t)q * n, L[q], (size_t)n * 8,
                               hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        if (gx) {
            gx[2 * k] = h4[k];
            gx[2 * k + 1] = h4[n + k];
        }
        if (gv) {
            gv[2 * k] = h4[2 * n + k];
            gv[2 * k + 1] = h4[3 * n + k];
        }
    }
    // leave the particles as they were at entry: checkpoint 0 carried back
    // through its permutation into the caller's order
    PIC_TRY(restore(0));
    state();
    for (int q = 0; q < 4; ++q) {
        hipLaunchKernelGGL(k_scatter_back, ga(n), dim3(AT), 0, c->stream, n,
                           (const uint32_t *)perms.p, (const double *)st[q], L[4]);
        PIC_TRY(dcopy(st[q], L[4], (size_t)n * 8));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->seg_valid = false;
    c->sort_fresh = false;
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

extern "C" int pic_adjoint(pic_ctx *ctx, int kernel, double dt, int nsteps, int ckpt,
                           int sort_every, double *J, double *gx, double *gv) {
    PIC_CHECK(ctx && J, PIC_ERR_ARG, "null pointer");
    PIC_CHECK(kernel == PIC_CIC || kernel == PIC_M6, PIC_ERR_ARG, "kernel %d", kernel);
    PIC_CHECK(nsteps >= 1 && ckpt >= 1, PIC_ERR_ARG, "nsteps %d, ckpt %d", nsteps, ckpt);
    PIC_CHECK(sort_every >= 0, PIC_ERR_ARG, "sort_every %d", sort_every);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_CHECK(c->uniform_w, PIC_ERR_STATE, "the adjoint needs uniform particle weights");
    PIC_CHECK(c->bz_mode == PIC_BZ_NONE, PIC_ERR_STATE, "the adjoint supports bz = 0 only");
    PIC_CHECK(c->n > 0, PIC_ERR_STATE, "no particles");
    if (c->state == PIC_STATE_F32)
        return adjoint_f32(c, kernel, dt, nsteps, ckpt, sort_every, J, gx, gv);
    return kernel == PIC_CIC ? adjoint_impl<PIC_CIC>(c, dt, nsteps, ckpt, sort_every, J, gx, gv)
                             : adjoint_impl<PIC_M6>(c, dt, nsteps, ckpt, sort_every, J, gx, gv);
}

extern "C" int pic_set_adjoint_block(pic_ctx *ctx, int block) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(block >= 0, PIC_ERR_ARG, "block %d", block);
    ctx_of(ctx)->adj_block = block;
    return PIC_OK;
}

extern "C" int pic_get_adjoint_plan(pic_ctx *ctx, int *block, int64_t *bytes) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    const Ctx *c = ctx_of(ctx);
    if (block) *block = c->adj_block_used;
    if (bytes) *bytes = c->adj_bytes_used;
    return PIC_OK;
}

extern "C" int pic_get_adjoint_timing(pic_ctx *ctx, double *ms, int64_t *passes) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    const Ctx *c = ctx_of(ctx);
    for (int k = 0; k < 4; ++k) {
        if (ms) ms[k] = c->t_adj_ms[k];
        if (passes) passes[k] = c->t_adj_n[k];
    }
    return PIC_OK;
}
