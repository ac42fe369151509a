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
            fx, fy,
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

// Adjoint of the gather and kick of step s < last at x'_s.  lam = (A, B).
template <int K>
__global__ __launch_bounds__(AT) void k_adj32_gather(int64_t n, const int32_t *__restrict__ pc,
                                                     const float2 *__restrict__ po, float4 *lam,
                                                     const float2 *__restrict__ e32, double *mux,
                                                     double *muy, Grid32 g, float dt, float dtdx,
                                                     float dtdy) {
    constexpr int S = St32<K>::S, HL = St32<K>::HL;
    const int64_t k = (int64_t)blockIdx.x * AT + threadIdx.x;
    if (k >= n) return;
    const int mx = g.nx - 1, my = g.ny - 1;
    const int c = pc[k];
    const int ci = c & mx, cj = c >> g.lognx;
    const float2 f = po[k];
    float wx[6], wy[6], dwx[6], dwy[6];
    shape32<K>(f.x, wx, dwx);
    shape32<K>(f.y, wy, dwy);
    const float4 L = lam[k];
    const float tx = __builtin_fmaf(dt, L.x, L.z), ty = __builtin_fmaf(dt, L.y, L.w);
    const float qx = dt * tx, qy = dt * ty;
    float gxx = 0.0f, gxy = 0.0f, gyx = 0.0f, gyy = 0.0f;  // d ex/dfx, d ex/dfy, d ey/dfx, ...
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int64_t iu = (ci - HL + u) & mx;
        float sx1 = 0.0f, sx2 = 0.0f, sy1 = 0.0f, sy2 = 0.0f;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const int64_t gi = iu + (int64_t)g.nx * ((cj - HL + v) & my);
            const float2 E = e32[gi];
            sx1 = __builtin_fmaf(wy[v], E.x, sx1);
            sx2 = __builtin_fmaf(wy[v], E.y, sx2);
            sy1 = __builtin_fmaf(dwy[v], E.x, sy1);
            sy2 = __builtin_fmaf(dwy[v], E.y, sy2);
            const float wab = wx[u] * wy[v];
            unsafeAtomicAdd(&mux[gi], (double)(qx * wab));
            unsafeAtomicAdd(&muy[gi], (double)(qy * wab));
        }
        gxx = __builtin_fmaf(dwx[u], sx1, gxx);
        gyx = __builtin_fmaf(dwx[u], sx2, gyx);
        gxy = __builtin_fmaf(wx[u], sy1, gxy);
        gyy = __builtin_fmaf(wx[u], sy2, gyy);
    }
    const float ax = __builtin_fmaf(dtdx, __builtin_fmaf(tx, gxx, ty * gyx), L.x);
    const float ay = __builtin_fmaf(dtdy, __builtin_fmaf(tx, gxy, ty * gyy), L.y);
    lam[k] = make_float4(ax, ay, tx, ty);
}

// Adjoint of the deposit at x'_s: A += (w / dx dy) grad W . lrho, with
// cx = w / (dx dy dx), cy = w / (dx dy dy).
template <int K>
__global__ __launch_bounds__(AT) void k_adj32_deposit(int64_t n, const int32_t *__restrict__ pc,
                                                      const float2 *__restrict__ po, float4 *lam,
                                                      const double *__restrict__ lrho, Grid32 g,
                                                      float cx, float cy) {
    constexpr int S = St32<K>::S, HL = St32<K>::HL;
    const int64_t k = (int64_t)blockIdx.x * AT + threadIdx.x;
    if (k >= n) return;
    const int mx = g.nx - 1, my = g.ny - 1;
    const int c = pc[k];
    const int ci = c & mx, cj = c >> g.lognx;
    const float2 f = po[k];
    float wx[6], wy[6], dwx[6], dwy[6];
    shape32<K>(f.x, wx, dwx);
    shape32<K>(f.y, wy, dwy);
    float gx = 0.0f, gy = 0.0f;
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int64_t iu = (ci - HL + u) & mx;
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const float r = (float)lrho[iu + (int64_t)g.nx * ((cj - HL + v) & my)];
            s1 = __builtin_fmaf(wy[v], r, s1);
            s2 = __builtin_fmaf(dwy[v], r, s2);
        }
        gx = __builtin_fmaf(dwx[u], s1, gx);
        gy = __builtin_fmaf(wx[u], s2, gy);
    }
    float4 L = lam[k];
    L.x = __builtin_fmaf(cx, gx, L.x);
    L.y = __builtin_fmaf(cy, gy, L.y);
    lam[k] = L;
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

template <int K>
int adjoint32_impl(Ctx *c, double dt, int nsteps, int ckpt, double *J, double *gx, double *gv) {
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
    const double scale_rho = 1.0 / ((double)nx * (double)ny * dxdy);
    const double scale = 1.0 / ((double)nx * (double)ny);
    // the forward passes' FP32 constants (launch_seg32)
    const float hdt = (float)(0.5 * dt);
    const float dtdx_h = (float)(h / c->dx), dtdy_h = (float)(h / c->dy);
    const float dtdx = (float)(dt / c->dx), dtdy = (float)(dt / c->dy);

    DevBuf<float2> Es, po;
    DevBuf<double> Elast, mu, lrho;
    DevBuf<int32_t> ckc, pc, tc;
    DevBuf<float4> cko, lam, t4;
    DevBuf<uint32_t> perms;
    PIC_TRY(Es.alloc((size_t)m * nsteps));
    PIC_TRY(Elast.alloc((size_t)2 * m));
    PIC_TRY(mu.alloc((size_t)2 * m));
    PIC_TRY(lrho.alloc((size_t)m));
    PIC_TRY(ckc.alloc((size_t)n * nseg));
    PIC_TRY(cko.alloc((size_t)n * nseg));
    PIC_TRY(perms.alloc((size_t)n * nseg));
    PIC_TRY(pc.alloc((size_t)n * len));
    PIC_TRY(po.alloc((size_t)n * len));
    PIC_TRY(lam.alloc((size_t)n));
    PIC_TRY(t4.alloc((size_t)n));
    PIC_TRY(tc.alloc((size_t)n));
    PIC_TRY(ensure_energy(c, nsteps));
    if (c->pending_perm) PIC_TRY(settle_state(c));
    PIC_HIP(hipMemsetAsync(c->step_counter, 0, sizeof(int), c->stream));

    auto save = [&](int g) -> int {
        PIC_HIP(hipMemcpyAsync(ckc.p + (size_t)g * n, c->cell, (size_t)n * 4,
                               hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(cko.p + (size_t)g * n, c->ov32, (size_t)n * 16,
                               hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(perms.p + (size_t)g * n, c->perm2, (size_t)n * 4,
                               hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    };
    auto restore = [&](int g) -> int {
        PIC_HIP(hipMemcpyAsync(c->cell, ckc.p + (size_t)g * n, (size_t)n * 4,
                               hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(c->ov32, cko.p + (size_t)g * n, (size_t)n * 16,
                               hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    };

    // ---------------------------------------------------------- forward
    // checkpoint 0 holds x_0 (before the first half drift); checkpoint g > 0
    // holds (x'_s, v_s) at s = g ckpt
    PIC_TRY(sort_by_cell(c, nullptr, true));
    PIC_TRY(save(0));
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
    PIC_TRY(launch_seg32(c, K, 0, dt, h));
    for (int s = 0; s < nsteps; ++s) {
        const bool last = s == nsteps - 1;
        if (s > 0 && s % ckpt == 0) {
            PIC_TRY(sort_by_cell(c, nullptr, true));
            PIC_TRY(save(s / ckpt));
        }
        PIC_TRY(poisson_solve_e(c, c->rho_acc, scale_rho, last ? 1 : 3, -2));
        PIC_HIP(hipMemcpyAsync(Es.p + (size_t)m * s, c->e32, (size_t)m * 8,
                               hipMemcpyDeviceToDevice, c->stream));
        if (last) break;
        PIC_TRY(launch_seg32(c, K, 1, dt, dt));
    }
    PIC_HIP(hipMemcpyAsync(Elast.p, c->ex, (size_t)m * 8, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(Elast.p + m, c->ey, (size_t)m * 8, hipMemcpyDeviceToDevice, c->stream));
    {
        double e2[2];
        PIC_HIP(hipMemcpyAsync(e2, c->energy_hist + 2 * (nsteps - 1), 16, hipMemcpyDeviceToHost,
                               c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        *J = e2[0] + e2[1];
    }

    // ---------------------------------------------------------- reverse
    PIC_HIP(hipMemsetAsync(lam.p, 0, (size_t)n * 16, c->stream));
    const float cx = (float)(c->w0 / (dxdy * c->dx)), cy = (float)(c->w0 / (dxdy * c->dy));
    const float dtf = (float)dt, dtdx_g = (float)(dt / c->dx), dtdy_g = (float)(dt / c->dy);
    for (int g = nseg - 1; g >= 0; --g) {
        const int s0 = g * ckpt, s1 = std::min(nsteps, s0 + ckpt);
        PIC_TRY(restore(g));
        if (g == 0)
            hipLaunchKernelGGL((k_adj32_advance<K, false>), grid_for(n), dim3(AT), 0, c->stream,
                               n, c->cell, c->ov32, (const float2 *)nullptr, gr, hdt, dtdx_h,
                               dtdy_h, pc.p, po.p);
        else
            hipLaunchKernelGGL(k_adj32_pos, grid_for(n), dim3(AT), 0, c->stream, n,
                               (const int32_t *)c->cell, (const float4 *)c->ov32, pc.p, po.p);
        for (int s = s0; s + 1 < s1; ++s) {
            const size_t q = (size_t)(s + 1 - s0) * n;
            hipLaunchKernelGGL((k_adj32_advance<K, true>), grid_for(n), dim3(AT), 0, c->stream,
                               n, c->cell, c->ov32, (const float2 *)(Es.p + (size_t)m * s), gr,
                               hdt, dtdx, dtdy, pc.p + q, po.p + q);
        }
        for (int s = s1 - 1; s >= s0; --s) {
            const size_t q = (size_t)(s - s0) * n;
            if (s == nsteps - 1) {
                hipLaunchKernelGGL(k_adj32_seed, grid_for(m), dim3(AT), 0, c->stream, m,
                                   (const double *)Elast.p, (const double *)(Elast.p + m), dxdy,
                                   mu.p, mu.p + m);
            } else {
                PIC_TRY(launch_fill(c, mu.p, 0.0, 2 * m));
                hipLaunchKernelGGL((k_adj32_gather<K>), grid_for(n), dim3(AT), 0, c->stream, n,
                                   (const int32_t *)(pc.p + q), (const float2 *)(po.p + q), lam.p,
                                   (const float2 *)(Es.p + (size_t)m * s), mu.p, mu.p + m, gr,
                                   dtf, dtdx_g, dtdy_g);
            }
            PIC_TRY(poisson_adjoint(c, mu.p, mu.p + m, lrho.p, scale));
            hipLaunchKernelGGL((k_adj32_deposit<K>), grid_for(n), dim3(AT), 0, c->stream, n,
                               (const int32_t *)(pc.p + q), (const float2 *)(po.p + q), lam.p,
                               (const double *)lrho.p, gr, cx, cy);
        }
        // into the slot order of the previous segment (or the caller's order)
        hipLaunchKernelGGL((k_adj32_back<float4>), grid_for(n), dim3(AT), 0, c->stream, n,
                           (const uint32_t *)(perms.p + (size_t)g * n), (const float4 *)lam.p,
                           t4.p);
        std::swap(lam.p, t4.p);
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
    PIC_TRY(restore(0));
    hipLaunchKernelGGL((k_adj32_back<int32_t>), grid_for(n), dim3(AT), 0, c->stream, n,
                       (const uint32_t *)perms.p, (const int32_t *)c->cell, tc.p);
    hipLaunchKernelGGL((k_adj32_back<float4>), grid_for(n), dim3(AT), 0, c->stream, n,
                       (const uint32_t *)perms.p, (const float4 *)c->ov32, t4.p);
    PIC_HIP(hipMemcpyAsync(c->cell, tc.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(c->ov32, t4.p, (size_t)n * 16, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->seg_valid = false;
    return PIC_OK;
}

}  // namespace

int adjoint_f32(Ctx *c, int kernel, double dt, int nsteps, int ckpt, double *J, double *gx,
                double *gv) {
    return kernel == PIC_CIC ? adjoint32_impl<PIC_CIC>(c, dt, nsteps, ckpt, J, gx, gv)
                             : adjoint32_impl<PIC_M6>(c, dt, nsteps, ckpt, J, gx, gv);
}

}  // namespace pic
