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
// copy), the deposit, the solve (E_s kept, 2 nx*ny doubles per step) and a
// kick + drift kernel.  Every `ckpt` steps the particles are lane-sorted and
// the state and the sort permutation are stored.  The reverse sweep
// recomputes each segment's x'_s from its checkpoint with the same kernels
// and the stored E_s, so the recomputed trajectory is bit-identical, then
// runs the adjoint passes backwards and carries lx, lv across segment
// boundaries (and finally into the caller's particle order) through the
// stored permutations.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_internal.h"

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

// Gather adjoint at x' (step s < last): lvt = lv + h lx;
// lx += dt (lvt . grad G E_s); mu += dt G^T lvt.  Writes lv <- lvt.
template <int K>
__global__ void k_adj_gather(int64_t n, const double *__restrict__ xs, const double *__restrict__ ys,
                             double *lx, double *ly, double *lvx, double *lvy,
                             const double *__restrict__ ex, const double *__restrict__ ey,
                             double *mux, double *muy, int nx, int ny, double dx, double dy,
                             double dt, double h) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
    double fx, fy;
    const int ci = ref_cell_a(xs[k], dx, nx, fx), cj = ref_cell_a(ys[k], dy, ny, fy);
    double wx[6], wy[6], dwx[6], dwy[6];
    shape_d<K>(fx, wx, dwx);
    shape_d<K>(fy, wy, dwy);
    const double ax = lx[k], ay = ly[k];
    const double tx = lvx[k] + h * ax, ty = lvy[k] + h * ay;
    double gxx = 0.0, gxy = 0.0, gyx = 0.0, gyy = 0.0;  // d ex/dx, d ex/dy, d ey/dx, d ey/dy
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int iu = picd::wrap2(ci - HL + u, nx);
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const int64_t g = iu + (int64_t)nx * picd::wrap2(cj - HL + v, ny);
            const double e1 = ex[g], e2 = ey[g];
            gxx += dwx[u] * wy[v] * e1;
            gxy += wx[u] * dwy[v] * e1;
            gyx += dwx[u] * wy[v] * e2;
            gyy += wx[u] * dwy[v] * e2;
            const double wab = wx[u] * wy[v];
            unsafeAtomicAdd(&mux[g], dt * tx * wab);
            unsafeAtomicAdd(&muy[g], dt * ty * wab);
        }
    }
    lx[k] = ax + dt * (tx * gxx + ty * gyx) / dx;
    ly[k] = ay + dt * (tx * gxy + ty * gyy) / dy;
    lvx[k] = tx;
    lvy[k] = ty;
}

// Deposit adjoint at x': lx += (w/dxdy) grad W . lrho; then lv += h lx.
template <int K>
__global__ void k_adj_deposit(int64_t n, const double *__restrict__ xs,
                              const double *__restrict__ ys, double *lx, double *ly, double *lvx,
                              double *lvy, const double *__restrict__ lrho, int nx, int ny,
                              double dx, double dy, double wdd, double h) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    constexpr int S = K == PIC_CIC ? 2 : 6, HL = K == PIC_CIC ? 0 : 2;
    double fx, fy;
    const int ci = ref_cell_a(xs[k], dx, nx, fx), cj = ref_cell_a(ys[k], dy, ny, fy);
    double wx[6], wy[6], dwx[6], dwy[6];
    shape_d<K>(fx, wx, dwx);
    shape_d<K>(fy, wy, dwy);
    double gx = 0.0, gy = 0.0;
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int iu = picd::wrap2(ci - HL + u, nx);
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const double r = lrho[iu + (int64_t)nx * picd::wrap2(cj - HL + v, ny)];
            gx += dwx[u] * wy[v] * r;
            gy += wx[u] * dwy[v] * r;
        }
    }
    const double ax = lx[k] + wdd * gx / dx, ay = ly[k] + wdd * gy / dy;
    lx[k] = ax;
    ly[k] = ay;
    lvx[k] = lvx[k] + h * ax;
    lvy[k] = lvy[k] + h * ay;
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

template <int K>
static int adjoint_impl(Ctx *c, double dt, int nsteps, int ckpt, double *J, double *gx,
                        double *gv) {
    const int64_t n = c->n;
    const int nx = c->nx, ny = c->ny;
    const int64_t m = (int64_t)nx * ny;
    const double h = 0.5 * dt, dx = c->dx, dy = c->dy, dxdy = c->dx * c->dy;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    const double scale_rho = 1.0 / ((double)nx * (double)ny * dxdy);
    const double scale = 1.0 / ((double)nx * (double)ny);
    DevBuf<double> Es, ck, xs, lam, mu, lrho, tmp;
    DevBuf<uint32_t> perms;
    PIC_TRY(Es.alloc((size_t)2 * m * nsteps));
    PIC_TRY(ck.alloc((size_t)4 * n * nseg));
    PIC_TRY(perms.alloc((size_t)n * nseg));
    PIC_TRY(xs.alloc((size_t)2 * n * ckpt));
    PIC_TRY(lam.alloc((size_t)4 * n));
    PIC_TRY(mu.alloc((size_t)2 * m));
    PIC_TRY(lrho.alloc((size_t)m));
    PIC_TRY(tmp.alloc((size_t)n));
    PIC_TRY(ensure_energy(c, nsteps));
    PIC_HIP(hipMemsetAsync(c->step_counter, 0, sizeof(int), c->stream));
    double *st[4] = {nullptr, nullptr, nullptr, nullptr};
    auto state = [&]() {
        st[0] = c->x;
        st[1] = c->y;
        st[2] = c->vx;
        st[3] = c->vy;
    };
    auto save = [&](int g) -> int {
        state();
        for (int q = 0; q < 4; ++q)
            PIC_HIP(hipMemcpyAsync(ck.p + ((size_t)4 * g + q) * n, st[q], (size_t)n * 8,
                                   hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(perms.p + (size_t)g * n, c->perm2, (size_t)n * 4,
                               hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    };
    auto restore = [&](int g) -> int {
        state();
        for (int q = 0; q < 4; ++q)
            PIC_HIP(hipMemcpyAsync(st[q], ck.p + ((size_t)4 * g + q) * n, (size_t)n * 8,
                                   hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    };
    // ---------------------------------------------------------- forward
    for (int s = 0; s < nsteps; ++s) {
        if (s % ckpt == 0) {
            PIC_TRY(sort_by_cell(c, nullptr, true));
            PIC_TRY(save(s / ckpt));
        }
        hipLaunchKernelGGL(k_adj_drift, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx, c->vy,
                           h, c->dimx, c->dimy, (double *)nullptr, (double *)nullptr);
        PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
        PIC_TRY(launch_seg64(c, K, 3, 0.0, 0.0));
        PIC_TRY(poisson_solve_e(c, c->rho_acc, scale_rho, 0, -2));
        PIC_HIP(hipMemcpyAsync(Es.p + (size_t)2 * m * s, c->ex, (size_t)2 * m * 8,
                               hipMemcpyDeviceToDevice, c->stream));
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
    double *lx = lam.p, *ly = lam.p + n, *lvx = lam.p + 2 * n, *lvy = lam.p + 3 * n;
    hipLaunchKernelGGL(k_zero2, ga(n), dim3(AT), 0, c->stream, n, lx, ly);
    hipLaunchKernelGGL(k_zero2, ga(n), dim3(AT), 0, c->stream, n, lvx, lvy);
    const double wdd = c->w0 / dxdy;
    for (int g = nseg - 1; g >= 0; --g) {
        const int s0 = g * ckpt, s1 = std::min(nsteps, s0 + ckpt);
        PIC_TRY(restore(g));
        for (int s = s0; s < s1; ++s) {
            double *xo = xs.p + (size_t)2 * n * (s - s0);
            hipLaunchKernelGGL(k_adj_drift, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx,
                               c->vy, h, c->dimx, c->dimy, xo, xo + n);
            if (s == nsteps - 1) break;
            const double *E = Es.p + (size_t)2 * m * s;
            hipLaunchKernelGGL(k_adj_kick<K>, ga(n), dim3(AT), 0, c->stream, n, c->x, c->y, c->vx,
                               c->vy, E, E + m, nx, ny, dx, dy, dt, h, c->dimx, c->dimy);
        }
        for (int s = s1 - 1; s >= s0; --s) {
            const double *xo = xs.p + (size_t)2 * n * (s - s0);
            const double *E = Es.p + (size_t)2 * m * s;
            if (s == nsteps - 1) {
                hipLaunchKernelGGL(k_adj_seed, ga(m), dim3(AT), 0, c->stream, m, E, E + m, dxdy,
                                   mu.p, mu.p + m);
            } else {
                hipLaunchKernelGGL(k_zero2, ga(m), dim3(AT), 0, c->stream, m, mu.p, mu.p + m);
                hipLaunchKernelGGL(k_adj_gather<K>, ga(n), dim3(AT), 0, c->stream, n, xo, xo + n,
                                   lx, ly, lvx, lvy, E, E + m, mu.p, mu.p + m, nx, ny, dx, dy, dt,
                                   h);
            }
            PIC_TRY(poisson_adjoint(c, mu.p, mu.p + m, lrho.p, scale));
            hipLaunchKernelGGL(k_adj_deposit<K>, ga(n), dim3(AT), 0, c->stream, n, xo, xo + n, lx,
                               ly, lvx, lvy, (const double *)lrho.p, nx, ny, dx, dy, wdd, h);
        }
        // into the slot order of the previous segment (or the caller's order)
        const uint32_t *pg = perms.p + (size_t)g * n;
        for (double *a : {lx, ly, lvx, lvy}) {
            hipLaunchKernelGGL(k_scatter_back, ga(n), dim3(AT), 0, c->stream, n, pg,
                               (const double *)a, tmp.p);
            PIC_HIP(hipMemcpyAsync(a, tmp.p, (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
        }
    }
    PIC_HIP(hipGetLastError());
    std::vector<double> h4((size_t)4 * n);
    PIC_HIP(hipMemcpyAsync(h4.data(), lam.p, (size_t)4 * n * 8, hipMemcpyDeviceToHost, c->stream));
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
                           (const uint32_t *)perms.p, (const double *)st[q], tmp.p);
        PIC_HIP(hipMemcpyAsync(st[q], tmp.p, (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->seg_valid = false;
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
    return kernel == PIC_CIC ? adjoint_impl<PIC_CIC>(c, dt, nsteps, ckpt, J, gx, gv)
                             : adjoint_impl<PIC_M6>(c, dt, nsteps, ckpt, J, gx, gv);
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
