// pic_adjoint32.hip -- the adjoint sweep of the FP32 cell-relative state.
//
// Same discrete adjoint as pic_adjoint.hip (the reference has none; its loop
// is examples/landau_damping_poisson_cic_m6.jl:70-83 with bz = 0), written
// for the FP32 loop, whose state is the mid-step position x'_s (an int32
// cell and an FP32 offset in it) and the velocity v_s:
//   x'_0 = x_0 + h v_0;   E_s = P D(x'_s)   (FP64 deposit sums and solve)
//   v_{s+1} = v_s + dt G(x'_s) E_s;   x'_{s+1} = x'_s + dt v_{s+1}.
// With A, B the adjoints of x'_{s+1}, v_{s+1} (zero after the last step):
//   s < last:  Bt = B + dt A;  A += dt (Bt . grad G(x'_s) E_s);
//              mu = dt G(x'_s)^T Bt                     (a scatter, FP64 sums)
//   s = last:  mu = 2 E_s dx dy
//   every s:   lrho = P^T mu (FP64 spectral, pic_poisson);
//              A += (w / dx dy) grad W(x'_s) . lrho      (a gather);  B = Bt
// and finally dJ/dx_0 = A, dJ/dv_0 = B + h A.  The particle adjoints (A, B)
// and all per-particle arithmetic are FP32; the field adjoints mu, lrho and
// the Poisson adjoint stay FP64, as the forward's charge sums and solve do.
//
// Forward: the time loop's own segment passes (drift + deposit, then kick +
// drift + deposit per step), a lane sort and a checkpoint (cell, offsets,
// velocities, sort permutation) every `ckpt` steps, and the FP32 field copy
// E_s of every step.  The reverse sweep recomputes each segment from its
// checkpoint with pic_push32.h's arithmetic and the stored E_s -- the code
// the forward passes run -- so the recomputed positions are bit-identical,
// then runs the adjoint passes backwards.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_internal.h"
#include "pic_push32.h"
#include "pic_segcore.h"

namespace pic {

int launch_seg32(Ctx *c, int kernel, int mode, double dt, double dtd);

namespace {

constexpr int AT = 256;

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + AT - 1) / AT)); }

template <int K>
struct St32 {
    static constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
};

// Shape weights and their derivatives (per cell length) of the stencil nodes
// at in-cell offset f.
template <int K>
__device__ __forceinline__ void shape32(float f, float *w, float *dw) {
    if (K == PIC_CIC) {
        w[0] = 1.0f - f;
        w[1] = f;
        dw[0] = -1.0f;
        dw[1] = 1.0f;
    } else {
        picd::m6_weights_f(f, w);
        m6_dweights_f(f, dw);
    }
}

struct Grid32 {
    int nx, ny, lognx;
};

// One step of the FP32 loop on every particle: a drift by h (KICK false,
// dtdx = h/dx) or the kick with E_s followed by a drift by dt.  The state is
// updated in place and the new position is also written to (pc, po).
template <int K, bool KICK>
__global__ __launch_bounds__(AT) void k_adj32_advance(int64_t n, int32_t *cell, float4 *ov,
                                                      const float2 *__restrict__ e32, Grid32 g,
                                                      float hdt, float dtdx, float dtdy,
                                                      int32_t *pc, float2 *po) {
    constexpr int HL = St32<K>::HL;
    const int64_t k = (int64_t)blockIdx.x * AT + threadIdx.x;
    if (k >= n) return;
    const int mx = g.nx - 1, my = g.ny - 1;
    const int c = cell[k];
    int ci = c & mx, cj = c >> g.lognx;
    const float4 o = ov[k];
    float fx = o.x, fy = o.y, ux = o.z, uy = o.w;
    if (KICK) {
        float e1, e2, b3 = 0.0f;
        interp32<K, PIC_BZ_NONE>(
            fx, fy, ci, cj,
            [&](int u, int v) {
                return e32[(int64_t)((ci - HL + u) & mx) + (int64_t)g.nx * ((cj - HL + v) & my)];
            },
            [&](int, int) { return 0.0f; }, e1, e2, b3);
        kick32<PIC_BZ_NONE>(hdt, 0.0f, e1, e2, b3, ux, uy);
    }
    float nfx, nfy;
    const int si = drift1s(fx, ux, dtdx, nfx);
    const int sj = drift1s(fy, uy, dtdy, nfy);
    ci = (ci + si) & mx;
    cj = (cj + sj) & my;
    const int cn = ci + g.nx * cj;
    cell[k] = cn;
    ov[k] = make_float4(nfx, nfy, ux, uy);
    pc[k] = cn;
    po[k] = make_float2(nfx, nfy);
}

__global__ __launch_bounds__(AT) void k_adj32_pos(int64_t n, const int32_t *__restrict__ cell,
                                                  const float4 *__restrict__ ov, int32_t *pc,
                                                  float2 *po) {
    const int64_t k = (int64_t)blockIdx.x * AT + threadIdx.x;
    if (k >= n) return;
    const float4 o = ov[k];
    pc[k] = cell[k];
    po[k] = make_float2(o.x, o.y);
}

// Adjoint of the gather and kick of step s < last at x'_s for one particle:
// lam = (A, B) -> (A + dt (Bt . grad G E_s), Bt) with Bt = B + dt A, and
// add_mu(u, v, qx, qy) receives dt Bt times the shape weight of node (u, v).
template <int K, class FE, class FA>
__device__ __forceinline__ float4 adj_gather_one(float fx, float fy, float4 L, float dt,
                                                 float dtdx, float dtdy, FE node_e, FA add_mu) {
    constexpr int S = St32<K>::S;
    float wx[6], wy[6], dwx[6], dwy[6];
    shape32<K>(fx, wx, dwx);
    shape32<K>(fy, wy, dwy);
    const float tx = __builtin_fmaf(dt, L.x, L.z), ty = __builtin_fmaf(dt, L.y, L.w);
    const float qx = dt * tx, qy = dt * ty;
    float gxx = 0.0f, gxy = 0.0f, gyx = 0.0f, gyy = 0.0f;  // d ex/dfx, d ex/dfy, d ey/dfx, ...
#pragma unroll
    for (int u = 0; u < S; ++u) {
        float sx1 = 0.0f, sx2 = 0.0f, sy1 = 0.0f, sy2 = 0.0f;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const float2 E = node_e(u, v);
            sx1 = __builtin_fmaf(wy[v], E.x, sx1);
            sx2 = __builtin_fmaf(wy[v], E.y, sx2);
            sy1 = __builtin_fmaf(dwy[v], E.x, sy1);
            sy2 = __builtin_fmaf(dwy[v], E.y, sy2);
            const float wab = wx[u] * wy[v];
            add_mu(u, v, qx * wab, qy * wab);
        }
        gxx = __builtin_fmaf(dwx[u], sx1, gxx);
        gyx = __builtin_fmaf(dwx[u], sx2, gyx);
        gxy = __builtin_fmaf(wx[u], sy1, gxy);
        gyy = __builtin_fmaf(wx[u], sy2, gyy);
    }
    const float ax = __builtin_fmaf(dtdx, __builtin_fmaf(tx, gxx, ty * gyx), L.x);
    const float ay = __builtin_fmaf(dtdy, __builtin_fmaf(tx, gxy, ty * gyy), L.y);
    return make_float4(ax, ay, tx, ty);
}

// Adjoint of the deposit at x'_s for one particle: A += (w / dx dy) grad W .
// lrho, with cx = w / (dx dy dx), cy = w / (dx dy dy); node_r(u, v) is lrho
// at node (u, v) rounded to FP32.
template <int K, class FR>
__device__ __forceinline__ float4 adj_deposit_one(float fx, float fy, float4 L, float cx,
                                                  float cy, FR node_r) {
    constexpr int S = St32<K>::S;
    float wx[6], wy[6], dwx[6], dwy[6];
    shape32<K>(fx, wx, dwx);
    shape32<K>(fy, wy, dwy);
    float gx = 0.0f, gy = 0.0f;
#pragma unroll
    for (int u = 0; u < S; ++u) {
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const float r = node_r(u, v);
            s1 = __builtin_fmaf(wy[v], r, s1);
            s2 = __builtin_fmaf(dwy[v], r, s2);
        }
        gx = __builtin_fmaf(dwx[u], s1, gx);
        gy = __builtin_fmaf(wx[u], s2, gy);
    }
    L.x = __builtin_fmaf(cx, gx, L.x);
    L.y = __builtin_fmaf(cy, gy, L.y);
    return L;
}

// Arguments of the reverse passes over the positions of one step, in the
// slot order of the lane sort whose offsets `off` are.
struct AdjSeg32 {
    const int32_t *pc;     // cell of x'_s per slot
    const float2 *po;      // in-cell offset of x'_s per slot
    float4 *lam;           // (A, B) per slot
    const float2 *e32;     // gather adjoint: E_s (FP32 copy)
    const double *lrho;    // deposit adjoint: P^T mu
    double *mux, *muy;     // gather adjoint: mu accumulators
    int nx, ny, lognx, ncells;
    float dt, dtdx, dtdy;  // gather adjoint: dt, dt/dx, dt/dy
    float cx, cy;          // deposit adjoint
    const int64_t *off;
    int bps, margin, cap;
    unsigned long long *slow;
    uint32_t *dlist;
    unsigned *dcount;
};

// Gather adjoint over one segment slice: E_s staged in an LDS window, mu
// summed in FP64 LDS tiles (ds_add_f64) and flushed once with FP64 global
// atomics; particles outside the window go to the deferred list.  Two copies
// of the mu tiles, lane l adding into copy l mod 2 (neighbouring lanes whose
// particles drifted into one cell then add to different words), in a 160 KiB
// window of a 1024-thread block: measured on MI355X, M6 on 512^2 with 5e8
// particles 23.9 -> 22.9 ms per pass, CIC on 128^2 with 2e6 35.8 -> 33.4 us
// (one copy in 80 KiB windows of 512-thread blocks before; two copies in
// those, a narrower window, took 36.7 ms).
constexpr int ADJ_GT = 1024;
template <int K>
__global__ __launch_bounds__(ADJ_GT) void k_adjseg_gather(AdjSeg32 a) {
    using St = SegStencil<K>;
    constexpr int S = St::S, HL = St::HL;
    extern __shared__ __align__(16) char smem_adj[];
    const int tid = threadIdx.x;
    const int nx = a.nx, ny = a.ny, mx = nx - 1, my = ny - 1;
    const int s = blockIdx.x / a.bps, bi = blockIdx.x - s * a.bps;
    const int ce = min(64 * s + 64, a.ncells);
    const int64_t s0 = a.off[64 * s], s1 = a.off[ce];
    const int64_t len = s1 - s0;
    const int64_t p0 = s0 + len * bi / a.bps, p1 = s0 + len * (bi + 1) / a.bps;
    const Win W = seg_window<S, HL>(s, nx, ny, a.ncells, a.margin, a.cap);
    const bool fits = W.tw * W.th <= a.cap;
    float2 *te = reinterpret_cast<float2 *>(smem_adj);
    double *tmx0 = reinterpret_cast<double *>(smem_adj + 8 * (size_t)a.cap), *tmy0 = tmx0 + a.cap;
    double *tmx1 = tmy0 + a.cap, *tmy1 = tmx1 + a.cap;
    double *tmx = (tid & 1) ? tmx1 : tmx0, *tmy = (tid & 1) ? tmy1 : tmy0;
    if (fits) {
        const int nt = W.tw * W.th;
        for (int t = tid; t < nt; t += ADJ_GT) {
            const int u = t % W.tw, v = t / W.tw;
            te[t] = a.e32[(int64_t)((W.x0 + u) & mx) + (int64_t)nx * ((W.y0 + v) & my)];
            tmx0[t] = 0.0;
            tmy0[t] = 0.0;
            tmx1[t] = 0.0;
            tmy1[t] = 0.0;
        }
    }
    __syncthreads();
    for (uint32_t p = (uint32_t)p0 + tid; p < (uint32_t)p1; p += ADJ_GT) {
        const int c = a.pc[p];
        const int ci = c & mx, cj = c >> a.lognx;
        const int lx = (ci - HL - W.x0) & mx, ly = (cj - HL - W.y0) & my;
        const bool in = fits && lx + S - 1 < W.tw && ly + S - 1 < W.th;
        defer_push(!in, p, a.dlist, a.dcount);
        if (!in) continue;
        const float2 f = a.po[p];
        a.lam[p] = adj_gather_one<K>(
            f.x, f.y, a.lam[p], a.dt, a.dtdx, a.dtdy,
            [&](int u, int v) { return te[(lx + u) + W.tw * (ly + v)]; },
            [&](int u, int v, float qa, float qb) {
                const int t = (lx + u) + W.tw * (ly + v);
                __hip_atomic_fetch_add(&tmx[t], (double)qa, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(&tmy[t], (double)qb, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            });
    }
    if (!fits) return;
    __syncthreads();
    const int nt = W.tw * W.th;
    for (int t = tid; t < nt; t += ADJ_GT) {
        const double qa = tmx0[t] + tmx1[t], qb = tmy0[t] + tmy1[t];
        if (qa != 0.0 || qb != 0.0) {
            const int u = t % W.tw, v = t / W.tw;
            const int64_t g = (int64_t)((W.x0 + u) & mx) + (int64_t)nx * ((W.y0 + v) & my);
            unsafeAtomicAdd(&a.mux[g], qa);
            unsafeAtomicAdd(&a.muy[g], qb);
        }
    }
}

// The deferred particles of k_adjseg_gather, from global memory.
template <int K>
__global__ __launch_bounds__(DEFER_T) void k_adjdefer_gather(AdjSeg32 a) {
    constexpr int HL = SegStencil<K>::HL;
    const unsigned cnt = __hip_atomic_load(a.dcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nx = a.nx, mx = a.nx - 1, my = a.ny - 1;
    for (unsigned i = blockIdx.x * DEFER_T + threadIdx.x; i < cnt; i += gridDim.x * DEFER_T) {
        const uint32_t p = a.dlist[i];
        const int c = a.pc[p];
        const int ci = c & mx, cj = c >> a.lognx;
        auto node = [&](int u, int v) {
            return (int64_t)((ci - HL + u) & mx) + (int64_t)nx * ((cj - HL + v) & my);
        };
        const float2 f = a.po[p];
        a.lam[p] = adj_gather_one<K>(
            f.x, f.y, a.lam[p], a.dt, a.dtdx, a.dtdy,
            [&](int u, int v) { return a.e32[node(u, v)]; },
            [&](int u, int v, float qa, float qb) {
                const int64_t g = node(u, v);
                unsafeAtomicAdd(&a.mux[g], (double)qa);
                unsafeAtomicAdd(&a.muy[g], (double)qb);
            });
    }
    defer_finish(cnt, a.dcount, a.slow);
}

// Deposit adjoint over one segment slice: lrho (rounded to FP32) staged in
// an LDS window.
template <int K>
__global__ __launch_bounds__(SEG_T) void k_adjseg_deposit(AdjSeg32 a) {
    using St = SegStencil<K>;
    constexpr int S = St::S, HL = St::HL;
    extern __shared__ __align__(16) char smem_adj[];
    const int tid = threadIdx.x;
    const int nx = a.nx, ny = a.ny, mx = nx - 1, my = ny - 1;
    const int s = blockIdx.x / a.bps, bi = blockIdx.x - s * a.bps;
    const int ce = min(64 * s + 64, a.ncells);
    const int64_t s0 = a.off[64 * s], s1 = a.off[ce];
    const int64_t len = s1 - s0;
    const int64_t p0 = s0 + len * bi / a.bps, p1 = s0 + len * (bi + 1) / a.bps;
    const Win W = seg_window<S, HL>(s, nx, ny, a.ncells, a.margin, a.cap);
    const bool fits = W.tw * W.th <= a.cap;
    float *tr = reinterpret_cast<float *>(smem_adj);
    if (fits) {
        const int nt = W.tw * W.th;
        for (int t = tid; t < nt; t += SEG_T) {
            const int u = t % W.tw, v = t / W.tw;
            tr[t] = (float)a.lrho[(int64_t)((W.x0 + u) & mx) + (int64_t)nx * ((W.y0 + v) & my)];
        }
    }
    __syncthreads();
    for (uint32_t p = (uint32_t)p0 + tid; p < (uint32_t)p1; p += SEG_T) {
        const int c = a.pc[p];
        const int ci = c & mx, cj = c >> a.lognx;
        const int lx = (ci - HL - W.x0) & mx, ly = (cj - HL - W.y0) & my;
        const bool in = fits && lx + S - 1 < W.tw && ly + S - 1 < W.th;
        defer_push(!in, p, a.dlist, a.dcount);
        if (!in) continue;
        const float2 f = a.po[p];
        a.lam[p] = adj_deposit_one<K>(f.x, f.y, a.lam[p], a.cx, a.cy, [&](int u, int v) {
            return tr[(lx + u) + W.tw * (ly + v)];
        });
    }
}

// The deferred particles of k_adjseg_deposit, from global memory.
template <int K>
__global__ __launch_bounds__(DEFER_T) void k_adjdefer_deposit(AdjSeg32 a) {
    constexpr int HL = SegStencil<K>::HL;
    const unsigned cnt = __hip_atomic_load(a.dcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nx = a.nx, mx = a.nx - 1, my = a.ny - 1;
    for (unsigned i = blockIdx.x * DEFER_T + threadIdx.x; i < cnt; i += gridDim.x * DEFER_T) {
        const uint32_t p = a.dlist[i];
        const int c = a.pc[p];
        const int ci = c & mx, cj = c >> a.lognx;
        const float2 f = a.po[p];
        a.lam[p] = adj_deposit_one<K>(f.x, f.y, a.lam[p], a.cx, a.cy, [&](int u, int v) {
            return (float)a.lrho[(int64_t)((ci - HL + u) & mx) +
                                 (int64_t)nx * ((cj - HL + v) & my)];
        });
    }
    defer_finish(cnt, a.dcount, a.slow);
}

// LDS per window node: FP32 field + two FP64 sums (gather adjoint), FP32
// lrho (deposit adjoint).  The gather adjoint's window takes 80 KiB (two
// blocks per CU): an M6 window of a 64-cell row segment then keeps a 14-cell
// drift margin instead of 7, and the 512^2 sweep (5e8 particles, ~2 cells of
// drift per step) can sort every 3 steps instead of 2 (measured: 5.36e9
// particle-steps/s against 5.0e9).  A window above 64 KiB needs the
// kernel's dynamic LDS limit raised; a failure there fails the launch, which
// hipGetLastError reports.
constexpr size_t ADJ_GATHER_NODE = 40, ADJ_DEPOSIT_NODE = 4;
constexpr size_t ADJ_GATHER_LDS_BYTES = 160 * 1024;

template <int K, bool GATHER>
int launch_adjseg(const Ctx *c, AdjSeg32 a) {
    const size_t per_node = GATHER ? ADJ_GATHER_NODE : ADJ_DEPOSIT_NODE;
    a.cap = (int)((GATHER ? ADJ_GATHER_LDS_BYTES : SEG_LDS_BYTES) / per_node) & ~15;
    a.margin = 32;
    const size_t sm = (size_t)a.cap * per_node;
    const void *k = GATHER ? reinterpret_cast<const void *>(k_adjseg_gather<K>)
                           : reinterpret_cast<const void *>(k_adjseg_deposit<K>);
    if (sm > 65536)
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    a.bps = seg_bps(c, k, sm, GATHER ? ADJ_GT : SEG_T);
    const dim3 grid((unsigned)(((c->nx * c->ny + 63) / 64) * a.bps));
    if (GATHER) {
        hipLaunchKernelGGL(k_adjseg_gather<K>, grid, dim3(ADJ_GT), sm, c->stream, a);
        hipLaunchKernelGGL(k_adjdefer_gather<K>, dim3(DEFER_BLOCKS), dim3(DEFER_T), 0, c->stream,
                           a);
    } else {
        hipLaunchKernelGGL(k_adjseg_deposit<K>, grid, dim3(SEG_T), sm, c->stream, a);
        hipLaunchKernelGGL(k_adjdefer_deposit<K>, dim3(DEFER_BLOCKS), dim3(DEFER_T), 0,
                           c->stream, a);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

__global__ void k_adj32_seed(int64_t m, const double *ex, const double *ey, double dxdy,
                             double *mux, double *muy) {
    const int64_t k = (int64_t)blockIdx.x * AT + threadIdx.x;
    if (k >= m) return;
    mux[k] = 2.0 * ex[k] * dxdy;
    muy[k] = 2.0 * ey[k] * dxdy;
}

// out[perm[r]] = in[r]: per-slot values into the order the permutation was
// taken from.
template <typename T>
__global__ void k_adj32_back(int64_t n, const uint32_t *__restrict__ perm,
                             const T *__restrict__ in, T *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * AT + threadIdx.x;
    if (r < n) out[perm[r]] = in[r];
}

// HIP events around the adjoint's particle passes when the context's timing
// is on: 0 forward passes, 1 gather adjoint, 2 deposit adjoint, 3 recompute.
struct PassTimer {
    Ctx *c;
    std::vector<hipEvent_t> ev[4];
    int64_t calls[4] = {0, 0, 0, 0};  // mark calls: begins and ends alternate
    bool on[4] = {false, false, false, false};
    explicit PassTimer(Ctx *cc) : c(cc) {}
    ~PassTimer() {
        for (auto &v : ev)
            for (hipEvent_t e : v) (void)hipEventDestroy(e);
    }
    // the passes of one kind are timed every timing_every-th time
    int mark(int k) {
        if (!c->timing) return PIC_OK;
        const bool begin = calls[k]++ % 2 == 0;
        if (begin) on[k] = (calls[k] / 2) % c->timing_every == 0;
        if (!on[k]) return PIC_OK;
        hipEvent_t e;
        PIC_HIP(hipEventCreate(&e));
        ev[k].push_back(e);
        PIC_HIP(hipEventRecord(e, c->stream));
        return PIC_OK;
    }
    // after the stream has drained: add the (begin, end) pairs to the context
    void collect() {
        for (int k = 0; k < 4; ++k)
            for (size_t i = 0; i + 1 < ev[k].size(); i += 2) {
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, ev[k][i], ev[k][i + 1]);
                c->t_adj_ms[k] += ms;
                c->t_adj_n[k] += 1;
            }
    }
};

// Steps at which the loop lane-sorts the particles: every checkpoint step and,
// with sort_every > 0, every multiple of it before the last step.  The
// forward pass and the reverse sweep's recomputation follow the same list,
// and the stable sort of a bit-identical state is the same permutation.
struct SortPlan {
    int ckpt, every, nsteps;
    bool at(int s) const {
        return s % ckpt == 0 || (every > 0 && s % every == 0 && s < nsteps - 1);
    }
};

// Memory plan of the reverse sweep.  Within a segment of `len` steps the
// positions of `q` consecutive steps are kept at once: the segment is
// recomputed once to store restart points (state, sort offsets and the
// permutation of a sort at that step) every q steps, then block by block
// from its restart point while the block's positions are stored and
// reversed.  q = len is the one-level sweep (one recomputation).
struct AdjPlan {
    int q = 0, nrestart = 0, nsub = 0;
    int64_t bytes = 0;
};

static AdjPlan plan_for(const SortPlan &sp, int nsteps, int ckpt, int q, int64_t n, int64_t m,
                        int64_t fixed) {
    AdjPlan p;
    p.q = q;
    const int len = std::min(ckpt, nsteps);
    p.nrestart = (len + q - 1) / q - 1;
    for (int s0 = 0; s0 < nsteps; s0 += ckpt) {
        const int s1 = std::min(nsteps, s0 + ckpt);
        for (int t0 = s0; t0 < s1; t0 += q) {
            int k = 0;
            for (int s = t0 + 1; s < std::min(s1, t0 + q); ++s) k += sp.at(s);
            p.nsub = std::max(p.nsub, k);
        }
    }
    const int64_t noff8 = (m + 1) * 8;
    p.bytes = fixed + (int64_t)p.nrestart * (24 * n + noff8) + (int64_t)q * 12 * n +
              (int64_t)p.nsub * (4 * n + noff8);
    return p;
}

template <int K>
int adjoint32_impl(Ctx *c, double dt, int nsteps, int ckpt, int sort_every, double *J,
                   double *gx, double *gv) {
    const int64_t n = c->n;
    const int nx = c->nx, ny = c->ny;
    const int64_t m = (int64_t)nx * ny;
    Grid32 gr;
    gr.nx = nx;
    gr.ny = ny;
    gr.lognx = 0;
    while ((1 << gr.lognx) < nx) ++gr.lognx;
    const double h = 0.5 * dt, dxdy = c->dx * c->dy;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    const int len = std::min(ckpt, nsteps);
    const SortPlan sp{ckpt, sort_every, nsteps};
    const double scale_rho = 1.0 / ((double)nx * (double)ny * dxdy);
    const double scale = 1.0 / ((double)nx * (double)ny);
    // the forward passes' FP32 constants (launch_seg32)
    const float hdt = (float)(0.5 * dt);
    const float dtdx_h = (float)(h / c->dx), dtdy_h = (float)(h / c->dy);
    const size_t noff = (size_t)m + 1;

    PIC_TRY(ensure_energy(c, nsteps));
    PIC_TRY(ensure_defer(c));
    if (c->pending_perm) PIC_TRY(settle_state(c));
    // the first sort also brings every sort and pass buffer into existence
    // before the free memory is measured
    PIC_TRY(sort_by_cell(c, nullptr, true));

    const int64_t fixed = (int64_t)m * nsteps * 8 + 40 * m + (int64_t)nseg * (24 * n + 8 * noff) +
                          36 * n;
    AdjPlan pl;
    if (c->adj_block > 0) {
        pl = plan_for(sp, nsteps, ckpt, std::min(c->adj_block, len), n, m, fixed);
    } else {
        size_t free_b = 0, total_b = 0;
        PIC_HIP(hipMemGetInfo(&free_b, &total_b));
        const int64_t budget = (int64_t)free_b - (int64_t)(512u << 20) - (int64_t)(total_b / 100);
        for (int q = len; q >= 1; --q) {
            pl = plan_for(sp, nsteps, ckpt, q, n, m, fixed);
            if (pl.bytes <= budget) break;
        }
        PIC_CHECK(pl.bytes <= budget, PIC_ERR_NOMEM,
                  "adjoint needs %lld bytes of device memory at one stored step, %lld free",
                  (long long)pl.bytes, (long long)budget);
    }
    c->adj_block_used = pl.q;
    c->adj_bytes_used = pl.bytes;
    const int q = pl.q;

    DevBuf<float2> Es, po;
    DevBuf<double> Elast, mu, lrho;
    DevBuf<int32_t> ckc, pc, tc, rsc;
    DevBuf<float4> cko, lam, t4, rso;
    DevBuf<uint32_t> perms, subperm, rsp;
    DevBuf<int64_t> offs, suboff, rsoff;
    PIC_TRY(Es.alloc((size_t)m * nsteps));
    PIC_TRY(Elast.alloc((size_t)2 * m));
    PIC_TRY(mu.alloc((size_t)2 * m));
    PIC_TRY(lrho.alloc((size_t)m));
    PIC_TRY(ckc.alloc((size_t)n * nseg));
    PIC_TRY(cko.alloc((size_t)n * nseg));
    PIC_TRY(perms.alloc((size_t)n * nseg));
    PIC_TRY(offs.alloc(noff * nseg));
    PIC_TRY(rsc.alloc((size_t)n * pl.nrestart));
    PIC_TRY(rso.alloc((size_t)n * pl.nrestart));
    PIC_TRY(rsp.alloc((size_t)n * pl.nrestart));
    PIC_TRY(rsoff.alloc(noff * pl.nrestart));
    PIC_TRY(subperm.alloc((size_t)n * pl.nsub));
    PIC_TRY(suboff.alloc(noff * pl.nsub));
    PIC_TRY(pc.alloc((size_t)n * q));
    PIC_TRY(po.alloc((size_t)n * q));
    PIC_TRY(lam.alloc((size_t)n));
    PIC_TRY(t4.alloc((size_t)n));
    PIC_TRY(tc.alloc((size_t)n));
    PIC_HIP(hipMemsetAsync(c->step_counter, 0, sizeof(int), c->stream));

    PassTimer tm(c);
    auto dcopy = [&](void *dst, const void *src, size_t bytes) -> int {
        PIC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    };
    // a restart point: state, sort offsets and permutation, in slot `g` of
    // the given buffers
    auto save_to = [&](int32_t *cc, float4 *co, uint32_t *cp, int64_t *cf, int g) -> int {
        PIC_TRY(dcopy(cc + (size_t)g * n, c->cell, (size_t)n * 4));
        PIC_TRY(dcopy(co + (size_t)g * n, c->ov32, (size_t)n * 16));
        PIC_TRY(dcopy(cp + (size_t)g * n, c->perm2, (size_t)n * 4));
        PIC_TRY(dcopy(cf + noff * g, c->cell_off, noff * 8));
        return PIC_OK;
    };
    // the state of a restart point, in its slot order, ready for segment passes
    auto restore_from = [&](const int32_t *cc, const float4 *co, const int64_t *cf,
                            int g) -> int {
        PIC_TRY(dcopy(c->cell, cc + (size_t)g * n, (size_t)n * 4));
        PIC_TRY(dcopy(c->ov32, co + (size_t)g * n, (size_t)n * 16));
        PIC_TRY(dcopy(c->cell_off, cf + noff * g, noff * 8));
        c->seg_valid = true;
        c->pending_perm = nullptr;
        return PIC_OK;
    };
    // The FP32 field the solve writes and the segment passes read is the
    // context's e32 array; pointed at slot s of Es, the forward solve stores
    // E_s there and the recomputation reads it there, with no copies.
    float *const e32_own = c->e32;
    auto field_at = [&](int s) { c->e32 = reinterpret_cast<float *>(Es.p + (size_t)m * s); };
    struct FieldBack {
        Ctx *c;
        float *p;
        ~FieldBack() { c->e32 = p; }
    } field_back{c, e32_own};
    // x'_s from x'_{s-1}: the loop's kick with E_{s-1} and drift (its segment
    // pass without the deposit), then the sort of step s if there is one
    auto recompute = [&](int s, bool *sorted) -> int {
        field_at(s - 1);
        PIC_TRY(tm.mark(3));
        PIC_TRY(launch_seg32(c, K, 2, dt, dt));
        PIC_TRY(tm.mark(3));
        *sorted = sp.at(s);
        if (*sorted) PIC_TRY(sort_by_cell(c, nullptr, true));
        return PIC_OK;
    };
    auto drift_half = [&](int32_t *pcq, float2 *poq) {
        hipLaunchKernelGGL((k_adj32_advance<K, false>), grid_for(n), dim3(AT), 0, c->stream, n,
                           c->cell, c->ov32, (const float2 *)nullptr, gr, hdt, dtdx_h, dtdy_h,
                           pcq, poq);
    };
    auto positions = [&](size_t slot) {
        hipLaunchKernelGGL(k_adj32_pos, grid_for(n), dim3(AT), 0, c->stream, n,
                           (const int32_t *)c->cell, (const float4 *)c->ov32, pc.p + slot * n,
                           po.p + slot * n);
    };

    // ---------------------------------------------------------- forward
    // checkpoint 0 holds x_0 (before the first half drift); checkpoint g > 0
    // holds (x'_s, v_s) at s = g ckpt.  Sorts between checkpoints leave their
    // permutation for the next pass to apply.
    PIC_TRY(save_to(ckc.p, cko.p, perms.p, offs.p, 0));
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
    PIC_TRY(tm.mark(0));
    PIC_TRY(launch_seg32(c, K, 0, dt, h));
    PIC_TRY(tm.mark(0));
    for (int s = 0; s < nsteps; ++s) {
        const bool last = s == nsteps - 1;
        if (s > 0 && sp.at(s)) {
            const bool ck = s % ckpt == 0;
            PIC_TRY(sort_by_cell(c, nullptr, true, ck));
            if (ck) PIC_TRY(save_to(ckc.p, cko.p, perms.p, offs.p, s / ckpt));
        }
        field_at(s);
        PIC_TRY(poisson_solve_e(c, c->rho_acc, scale_rho, last ? 1 : 3, -2));
        if (last) break;
        PIC_TRY(tm.mark(0));
        PIC_TRY(launch_seg32(c, K, 1, dt, dt));
        PIC_TRY(tm.mark(0));
    }
    PIC_TRY(dcopy(Elast.p, c->ex, (size_t)m * 8));
    PIC_TRY(dcopy(Elast.p + m, c->ey, (size_t)m * 8));
    {
        double e2[2];
        PIC_HIP(hipMemcpyAsync(e2, c->energy_hist + 2 * (nsteps - 1), 16, hipMemcpyDeviceToHost,
                               c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        *J = e2[0] + e2[1];
    }

    // ---------------------------------------------------------- reverse
    PIC_HIP(hipMemsetAsync(lam.p, 0, (size_t)n * 16, c->stream));
    AdjSeg32 a;
    a.lam = lam.p;
    a.mux = mu.p;
    a.muy = mu.p + m;
    a.lrho = lrho.p;
    a.nx = nx;
    a.ny = ny;
    a.lognx = gr.lognx;
    a.ncells = (int)m;
    a.dt = (float)dt;
    a.dtdx = (float)(dt / c->dx);
    a.dtdy = (float)(dt / c->dy);
    a.cx = (float)(c->w0 / (dxdy * c->dx));
    a.cy = (float)(c->w0 / (dxdy * c->dy));
    a.slow = c->slow_chunks;
    a.dlist = c->defer;
    a.dcount = c->defer_count;
    std::vector<int> sub_of((size_t)q);
    for (int g = nseg - 1; g >= 0; --g) {
        const int s0 = g * ckpt, s1 = std::min(nsteps, s0 + ckpt);
        const int nblk = (s1 - s0 + q - 1) / q;
        if (nblk > 1) {  // restart points every q steps through the segment
            PIC_TRY(restore_from(ckc.p, cko.p, offs.p, g));
            if (g == 0) drift_half(pc.p, po.p);
            for (int s = s0 + 1; s < s0 + (nblk - 1) * q + 1; ++s) {
                bool sorted;
                PIC_TRY(recompute(s, &sorted));
                if ((s - s0) % q == 0)
                    PIC_TRY(save_to(rsc.p, rso.p, rsp.p, rsoff.p, (s - s0) / q - 1));
            }
        }
        for (int j = nblk - 1; j >= 0; --j) {
            const int t0 = s0 + j * q, t1 = std::min(s1, t0 + q);
            // the restart point of the block: its offsets and permutation
            const int64_t *off0 = j == 0 ? offs.p + noff * g : rsoff.p + noff * (j - 1);
            const uint32_t *perm0 = j == 0 ? perms.p + (size_t)g * n : rsp.p + (size_t)(j - 1) * n;
            if (j == 0) {
                PIC_TRY(restore_from(ckc.p, cko.p, offs.p, g));
                if (g == 0)
                    drift_half(pc.p, po.p);
                else
                    positions(0);
            } else {
                PIC_TRY(restore_from(rsc.p, rso.p, rsoff.p, j - 1));
                positions(0);
            }
            // recompute x'_s, t0 <= s < t1, in the slot order of the sort in force
            int ns = 0;
            sub_of[0] = 0;
            for (int s = t0 + 1; s < t1; ++s) {
                bool sorted;
                PIC_TRY(recompute(s, &sorted));
                if (sorted) {
                    PIC_TRY(dcopy(subperm.p + (size_t)ns * n, c->perm2, (size_t)n * 4));
                    PIC_TRY(dcopy(suboff.p + noff * ns, c->cell_off, noff * 8));
                    ++ns;
                }
                positions((size_t)(s - t0));
                sub_of[s - t0] = ns;
            }
            for (int s = t1 - 1; s >= t0; --s) {
                const size_t qs = (size_t)(s - t0) * n;
                const int k = sub_of[s - t0];
                a.pc = pc.p + qs;
                a.po = po.p + qs;
                a.off = k == 0 ? off0 : suboff.p + noff * (k - 1);
                if (s == nsteps - 1) {
                    hipLaunchKernelGGL(k_adj32_seed, grid_for(m), dim3(AT), 0, c->stream, m,
                                       (const double *)Elast.p, (const double *)(Elast.p + m),
                                       dxdy, mu.p, mu.p + m);
                } else {
                    PIC_TRY(launch_fill(c, mu.p, 0.0, 2 * m));
                    a.e32 = Es.p + (size_t)m * s;
                    PIC_TRY(tm.mark(1));
                    PIC_TRY((launch_adjseg<K, true>(c, a)));
                    PIC_TRY(tm.mark(1));
                }
                PIC_TRY(poisson_adjoint(c, mu.p, mu.p + m, lrho.p, scale));
                PIC_TRY(tm.mark(2));
                PIC_TRY((launch_adjseg<K, false>(c, a)));
                PIC_TRY(tm.mark(2));
                // a sort at step s: the slots into their order before it
                if (sp.at(s)) {
                    const uint32_t *pg = k == 0 ? perm0 : subperm.p + (size_t)(k - 1) * n;
                    hipLaunchKernelGGL((k_adj32_back<float4>), grid_for(n), dim3(AT), 0,
                                       c->stream, n, pg, (const float4 *)lam.p, t4.p);
                    std::swap(lam.p, t4.p);
                    a.lam = lam.p;
                }
            }
        }
    }
    PIC_HIP(hipGetLastError());
    std::vector<float4> hl((size_t)n);
    PIC_HIP(hipMemcpyAsync(hl.data(), lam.p, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        const float4 L = hl[k];
        if (gx) {
            gx[2 * k] = L.x;
            gx[2 * k + 1] = L.y;
        }
        if (gv) {
            gv[2 * k] = (double)L.z + h * (double)L.x;
            gv[2 * k + 1] = (double)L.w + h * (double)L.y;
        }
    }
    // leave the particles as they were at entry: checkpoint 0 carried back
    // through its permutation into the caller's order
    PIC_TRY(restore_from(ckc.p, cko.p, offs.p, 0));
    hipLaunchKernelGGL((k_adj32_back<int32_t>), grid_for(n), dim3(AT), 0, c->stream, n,
                       (const uint32_t *)perms.p, (const int32_t *)c->cell, tc.p);
    hipLaunchKernelGGL((k_adj32_back<float4>), grid_for(n), dim3(AT), 0, c->stream, n,
                       (const uint32_t *)perms.p, (const float4 *)c->ov32, t4.p);
    PIC_TRY(dcopy(c->cell, tc.p, (size_t)n * 4));
    PIC_TRY(dcopy(c->ov32, t4.p, (size_t)n * 16));
    PIC_HIP(hipStreamSynchronize(c->stream));
    tm.collect();
    c->seg_valid = false;
    c->sort_fresh = false;
    return PIC_OK;
}

}  // namespace

int adjoint_f32(Ctx *c, int kernel, double dt, int nsteps, int ckpt, int sort_every, double *J,
                double *gx, double *gv) {
    return kernel == PIC_CIC
               ? adjoint32_impl<PIC_CIC>(c, dt, nsteps, ckpt, sort_every, J, gx, gv)
               : adjoint32_impl<PIC_M6>(c, dt, nsteps, ckpt, sort_every, J, gx, gv);
}

}  // namespace pic
