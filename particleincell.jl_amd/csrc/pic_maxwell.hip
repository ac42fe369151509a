// pic_maxwell.hip -- the Vlasov-Maxwell loop: the FDTD solver of src/fdtd.jl
// on the Yee grid, the CIC current deposit of src/cloud_in_cell.jl:10-81 and
// the loop of examples/landau_damping_maxwell_fdtd.jl:58-83, on the device.
//
// Per step (pic_run_maxwell):
//   faraday(dt/2) from the second step        k_faraday
//   update_fields                             k_update_fields (nodes, periodic)
//   push_v!(dt), push_x!(dt/2), compute_current!, push_x!(dt/2)
//                                             k_vm_pass, one pass per particle
//   faraday(dt/2), ampere_maxwell!(dt)        k_faraday, k_ampere
//   compute_energy                            k_fdtd_energy
// The field kernels follow the reference's arithmetic term by term, so they
// agree with it bit for bit; the current is summed with FP64 atomics (in an
// LDS copy of the whole grid when it fits), which reorders its sums.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_internal.h"

namespace pic {

namespace {

constexpr int MT = 256;

__device__ __forceinline__ int m1z(int x, int n) {  // mod1(x, n) - 1 for x in [-n+1, 2n]
    int r = (x - 1) % n;
    return r < 0 ? r + n : r;
}

// faraday!: fdtd.jl:24-36, one thread per (i, j) of nx x ny.
__global__ void k_faraday(int nx, int ny, double dx, double dy, const double *__restrict__ ex,
                          const double *__restrict__ ey, double *__restrict__ bz, double dt) {
    const int t = blockIdx.x * MT + threadIdx.x;
    if (t >= nx * ny) return;
    const int i = t % nx + 1, j = t / nx + 1;  // 1-based
    const int ldx = nx, ldy = nx + 1;
    const double dex_dy = (ex[(i - 1) + ldx * m1z(j + 1, ny)] - ex[(i - 1) + ldx * (j - 1)]) / dy;
    const double dey_dx = (ey[m1z(i + 1, nx) + ldy * (j - 1)] - ey[(i - 1) + ldy * (j - 1)]) / dx;
    bz[(i - 1) + nx * (j - 1)] += dt * (dex_dy - dey_dx);
}

// ampere_maxwell!: fdtd.jl:40-56; threads [0, nx (ny+1)) update ex, the next
// (nx+1) ny update ey.  jx, jy are periodic nx x ny.
__global__ void k_ampere(int nx, int ny, double dx, double dy, double *__restrict__ ex,
                         double *__restrict__ ey, const double *__restrict__ bz,
                         const double *__restrict__ jx, const double *__restrict__ jy, double dt) {
    const int t = blockIdx.x * MT + threadIdx.x;
    const int nex = nx * (ny + 1);
    if (t < nex) {
        const int i = t % nx + 1, j = t / nx + 1;
        const double dbz_dy = (bz[(i - 1) + nx * m1z(j, ny)] - bz[(i - 1) + nx * m1z(j - 1, ny)]) / dy;
        const double ja = jx[(i - 1) + nx * m1z(j, ny)], jb = jx[m1z(i + 1, nx) + nx * m1z(j, ny)];
        ex[(i - 1) + nx * (j - 1)] += dt * dbz_dy - dt * 0.5 * (ja + jb);
        return;
    }
    const int u = t - nex;
    if (u >= (nx + 1) * ny) return;
    const int i = u % (nx + 1) + 1, j = u / (nx + 1) + 1;
    const double dbz_dx =
        (bz[m1z(i, nx) + nx * m1z(j, ny)] - bz[m1z(i - 1, nx) + nx * (j - 1)]) / dx;
    const double ja = jy[m1z(i, nx) + nx * (j - 1)], jb = jy[m1z(i, nx) + nx * m1z(j + 1, ny)];
    ey[(i - 1) + (nx + 1) * (j - 1)] -= dt * dbz_dx - dt * 0.5 * (ja + jb);
}

// update_fields!: fdtd.jl:60-77 on the nodes i <= nx, j <= ny (the ghost
// row and column of the reference are their periodic images).
__global__ void k_update_fields(int nx, int ny, const double *__restrict__ fex,
                                const double *__restrict__ fey, const double *__restrict__ fbz,
                                double *__restrict__ ex, double *__restrict__ ey,
                                double *__restrict__ bz) {
    const int t = blockIdx.x * MT + threadIdx.x;
    if (t >= nx * ny) return;
    const int i = t % nx + 1, j = t / nx + 1;
    ex[t] = 0.5 * (fex[m1z(i - 1, nx) + nx * (j - 1)] + fex[m1z(i, nx) + nx * (j - 1)]);
    ey[t] = 0.5 * (fey[(i - 1) + (nx + 1) * m1z(j - 1, ny)] + fey[(i - 1) + (nx + 1) * m1z(j, ny)]);
    bz[t] = 0.25 * (((fbz[m1z(i - 1, nx) + nx * m1z(j - 1, ny)] + fbz[m1z(i, nx) + nx * m1z(j - 1, ny)]) +
                     fbz[m1z(i - 1, nx) + nx * m1z(j, ny)]) +
                    fbz[m1z(i, nx) + nx * m1z(j, ny)]);
}

// The reference's CIC cell (floor) and offset of a coordinate.
__device__ __forceinline__ int cic_cell(double x, double d, int n, double &off) {
    const double xp = x / d;
    const long long i1 = (long long)floor(xp) + 1;
    off = (xp - (double)i1) + 1;
    long long r = (i1 - 1) % n;
    return (int)(r < 0 ? r + n : r);
}

struct VmArgs {
    int64_t n;
    double *x, *y, *vx, *vy;
    const double *w;  // per-particle weights, or null: w0
    double w0;
    const double *ex, *ey, *bz;  // node fields, nx x ny
    double *jx, *jy;              // periodic current, nx x ny
    int nx, ny;
    double dx, dy, dimx, dimy, dt, scaling;
    int tiled;                    // the whole current grid in LDS
};

// push_v!(dt) (pushers_cic.jl:3-53), push_x!(dt/2), the particle's current
// (cloud_in_cell.jl:31-60), push_x!(dt/2), in that order, per particle.
__global__ __launch_bounds__(MT) void k_vm_pass(VmArgs a) {
    extern __shared__ double jsh[];  // jx then jy, nx*ny each, when tiled
    const int nx = a.nx, ny = a.ny, m = nx * ny;
    if (a.tiled) {
        for (int t = threadIdx.x; t < 2 * m; t += MT) jsh[t] = 0.0;
        __syncthreads();
    }
    for (int64_t k = (int64_t)blockIdx.x * MT + threadIdx.x; k < a.n;
         k += (int64_t)gridDim.x * MT) {
        double px = a.x[k], py = a.y[k], v1 = a.vx[k], v2 = a.vy[k];
        const double wgt = a.w ? a.w[k] : a.w0;
        {  // push_v!
            double dxp, dyp;
            const int i = cic_cell(px, a.dx, nx, dxp), j = cic_cell(py, a.dy, ny, dyp);
            const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
            const double a1 = (1 - dxp) * (1 - dyp), a2 = dxp * (1 - dyp), a3 = dxp * dyp,
                         a4 = (1 - dxp) * dyp;
            const int c00 = i + nx * j, c10 = ip + nx * j, c11 = ip + nx * jp, c01 = i + nx * jp;
            const double e1 = a1 * a.ex[c00] + a2 * a.ex[c10] + a3 * a.ex[c11] + a4 * a.ex[c01];
            const double e2 = a1 * a.ey[c00] + a2 * a.ey[c10] + a3 * a.ey[c11] + a4 * a.ey[c01];
            const double b3 = a1 * a.bz[c00] + a2 * a.bz[c10] + a3 * a.bz[c11] + a4 * a.bz[c01];
            const double dt = a.dt;
            v1 += 0.5 * dt * e1;
            v2 += 0.5 * dt * e2;
            const double tantheta = 0.5 * dt * b3;
            const double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
            v1 += v2 * tantheta;
            v2 += -v1 * sintheta;
            v1 += v2 * tantheta;
            v1 = v1 + 0.5 * dt * e1;
            v2 = v2 + 0.5 * dt * e2;
        }
        const double h = 0.5 * a.dt;
        auto drift = [&]() {  // push_x!(dt/2): pushers_cic.jl:55-72
            double p1 = px + h * v1, p2 = py + h * v2;
            if (p1 > a.dimx) p1 -= a.dimx;
            if (p2 > a.dimy) p2 -= a.dimy;
            if (p1 < 0.0) p1 += a.dimx;
            if (p2 < 0.0) p2 += a.dimy;
            px = p1;
            py = p2;
        };
        drift();
        {  // compute_current!
            double dxp, dyp;
            const int i = cic_cell(px, a.dx, nx, dxp), j = cic_cell(py, a.dy, ny, dyp);
            const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
            const double a1 = (1 - dxp) * (1 - dyp), a2 = dxp * (1 - dyp), a3 = dxp * dyp,
                         a4 = (1 - dxp) * dyp;
            const double w1 = v1 * a.scaling * wgt, w2 = v2 * a.scaling * wgt;
            const int c[4] = {i + nx * j, ip + nx * j, ip + nx * jp, i + nx * jp};
            const double q[4] = {a1, a2, a3, a4};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (a.tiled) {
                    __hip_atomic_fetch_add(&jsh[c[r]], q[r] * w1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(&jsh[m + c[r]], q[r] * w2, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    unsafeAtomicAdd(&a.jx[c[r]], q[r] * w1);
                    unsafeAtomicAdd(&a.jy[c[r]], q[r] * w2);
                }
            }
        }
        drift();
        a.x[k] = px;
        a.y[k] = py;
        a.vx[k] = v1;
        a.vy[k] = v2;
    }
    if (!a.tiled) return;
    __syncthreads();
    for (int t = threadIdx.x; t < m; t += MT) {
        if (jsh[t] != 0.0) unsafeAtomicAdd(&a.jx[t], jsh[t]);
        if (jsh[m + t] != 0.0) unsafeAtomicAdd(&a.jy[t], jsh[m + t]);
    }
}

// The particles' current alone (compute_current! on the state as it is).
__global__ __launch_bounds__(MT) void k_current(VmArgs a) {
    for (int64_t k = (int64_t)blockIdx.x * MT + threadIdx.x; k < a.n;
         k += (int64_t)gridDim.x * MT) {
        double dxp, dyp;
        const int nx = a.nx;
        const int i = cic_cell(a.x[k], a.dx, nx, dxp), j = cic_cell(a.y[k], a.dy, a.ny, dyp);
        const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == a.ny ? 0 : j + 1;
        const double a1 = (1 - dxp) * (1 - dyp), a2 = dxp * (1 - dyp), a3 = dxp * dyp,
                     a4 = (1 - dxp) * dyp;
        const double wgt = a.w ? a.w[k] : a.w0;
        const double w1 = a.vx[k] * a.scaling * wgt, w2 = a.vy[k] * a.scaling * wgt;
        const int c[4] = {i + nx * j, ip + nx * j, ip + nx * jp, i + nx * jp};
        const double q[4] = {a1, a2, a3, a4};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            unsafeAtomicAdd(&a.jx[c[r]], q[r] * w1);
            unsafeAtomicAdd(&a.jy[c[r]], q[r] * w2);
        }
    }
}

// compute_energy = 0.5 log(sum(fdtd.ex .^ 2) dx dy) (fdtd.jl:80): one block,
// a fixed summation order; the result into out[slot] (slot < 0: *counter,
// then incremented).
__global__ __launch_bounds__(MT) void k_fdtd_energy(const double *__restrict__ fex, int64_t m,
                                                    double dxdy, double *out, int slot,
                                                    int *counter) {
    __shared__ double sh[MT];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < m; k += MT) s += fex[k] * fex[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = MT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int q = slot;
        if (slot < 0) {
            q = *counter;
            *counter = q + 1;
        }
        out[q] = 0.5 * log(sh[0] * dxdy);
    }
}

inline dim3 blocks_for(int64_t n) { return dim3((unsigned)((n + MT - 1) / MT)); }

int ensure_fdtd(Ctx *c) {
    const size_t nx = c->nx, ny = c->ny;
    if (!c->fd_ex) {
        PIC_HIP(hipMalloc(&c->fd_ex, nx * (ny + 1) * 8));
        PIC_HIP(hipMalloc(&c->fd_ey, (nx + 1) * ny * 8));
        PIC_HIP(hipMalloc(&c->fd_bz, nx * ny * 8));
        PIC_HIP(hipMalloc(&c->jx, nx * ny * 8));
        PIC_HIP(hipMalloc(&c->jy, nx * ny * 8));
        PIC_HIP(hipMemsetAsync(c->fd_ex, 0, nx * (ny + 1) * 8, c->stream));
        PIC_HIP(hipMemsetAsync(c->fd_ey, 0, (nx + 1) * ny * 8, c->stream));
        PIC_HIP(hipMemsetAsync(c->fd_bz, 0, nx * ny * 8, c->stream));
        PIC_HIP(hipMemsetAsync(c->jx, 0, nx * ny * 8, c->stream));
        PIC_HIP(hipMemsetAsync(c->jy, 0, nx * ny * 8, c->stream));
    }
    return PIC_OK;
}

int faraday(Ctx *c, double dt) {
    hipLaunchKernelGGL(k_faraday, blocks_for((int64_t)c->nx * c->ny), dim3(MT), 0, c->stream,
                       c->nx, c->ny, c->dx, c->dy, (const double *)c->fd_ex,
                       (const double *)c->fd_ey, c->fd_bz, dt);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int ampere(Ctx *c, double dt) {
    const int64_t t = (int64_t)c->nx * (c->ny + 1) + (int64_t)(c->nx + 1) * c->ny;
    hipLaunchKernelGGL(k_ampere, blocks_for(t), dim3(MT), 0, c->stream, c->nx, c->ny, c->dx,
                       c->dy, c->fd_ex, c->fd_ey, (const double *)c->fd_bz,
                       (const double *)c->jx, (const double *)c->jy, dt);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int update_fields(Ctx *c) {
    hipLaunchKernelGGL(k_update_fields, blocks_for((int64_t)c->nx * c->ny), dim3(MT), 0,
                       c->stream, c->nx, c->ny, (const double *)c->fd_ex,
                       (const double *)c->fd_ey, (const double *)c->fd_bz, c->ex, c->ey, c->bz);
    PIC_HIP(hipGetLastError());
    c->bz_mode = PIC_BZ_GRID;
    return PIC_OK;
}

VmArgs vm_args(Ctx *c, double dt) {
    VmArgs a;
    a.n = c->n;
    a.x = c->x; a.y = c->y; a.vx = c->vx; a.vy = c->vy;
    a.w = c->uniform_w ? nullptr : c->w;
    a.w0 = c->w0;
    a.ex = c->ex; a.ey = c->ey; a.bz = c->bz;
    a.jx = c->jx; a.jy = c->jy;
    a.nx = c->nx; a.ny = c->ny;
    a.dx = c->dx; a.dy = c->dy; a.dimx = c->dimx; a.dimy = c->dimy;
    a.dt = dt;
    a.scaling = 1 / (c->dx * c->dy);
    a.tiled = 0;
    return a;
}

// LDS bytes of the whole-grid current copy (jx and jy), if it fits.
constexpr size_t VM_LDS_MAX = 48 * 1024;

int vm_pass(Ctx *c, double dt) {
    if (c->n == 0) return PIC_OK;
    VmArgs a = vm_args(c, dt);
    const size_t grid_bytes = (size_t)2 * c->nx * c->ny * 8;
    a.tiled = grid_bytes <= VM_LDS_MAX;
    // tiled: enough particles per block that the flush stays small against
    // the particle work, and about two blocks per compute unit
    int64_t nb = (c->n + MT - 1) / MT;
    if (a.tiled) nb = std::min<int64_t>(nb, 2 * (int64_t)c->ncu);
    nb = std::max<int64_t>(1, std::min<int64_t>(nb, 65535));
    hipLaunchKernelGGL(k_vm_pass, dim3((unsigned)nb), dim3(MT), a.tiled ? grid_bytes : 0,
                       c->stream, a);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int need_vm(Ctx *c, int kernel) {
    PIC_CHECK(kernel == PIC_CIC, PIC_ERR_ARG,
              "the Vlasov-Maxwell current deposit is CloudInCell (cloud_in_cell.jl:10)");
    PIC_CHECK(c->state == PIC_STATE_F64, PIC_ERR_STATE, "the Vlasov-Maxwell loop runs in the FP64 state");
    PIC_TRY(ensure_fdtd(c));
    return PIC_OK;
}

}  // namespace

}  // namespace pic

using namespace pic;

extern "C" {

int pic_fdtd_set(pic_ctx *ctx, const double *ex, const double *ey, const double *bz) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    const size_t nx = c->nx, ny = c->ny;
    if (ex) PIC_HIP(hipMemcpyAsync(c->fd_ex, ex, nx * (ny + 1) * 8, hipMemcpyHostToDevice, c->stream));
    if (ey) PIC_HIP(hipMemcpyAsync(c->fd_ey, ey, (nx + 1) * ny * 8, hipMemcpyHostToDevice, c->stream));
    if (bz) PIC_HIP(hipMemcpyAsync(c->fd_bz, bz, nx * ny * 8, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_fdtd_get(pic_ctx *ctx, double *ex, double *ey, double *bz) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    const size_t nx = c->nx, ny = c->ny;
    if (ex) PIC_HIP(hipMemcpyAsync(ex, c->fd_ex, nx * (ny + 1) * 8, hipMemcpyDeviceToHost, c->stream));
    if (ey) PIC_HIP(hipMemcpyAsync(ey, c->fd_ey, (nx + 1) * ny * 8, hipMemcpyDeviceToHost, c->stream));
    if (bz) PIC_HIP(hipMemcpyAsync(bz, c->fd_bz, nx * ny * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_faraday(pic_ctx *ctx, double dt) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    PIC_TRY(faraday(c, dt));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_ampere_maxwell(pic_ctx *ctx, double dt) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    PIC_TRY(ampere(c, dt));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_update_fields(pic_ctx *ctx) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    PIC_TRY(update_fields(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_set_current(pic_ctx *ctx, const double *jx, const double *jy) {
    PIC_CHECK(ctx && jx && jy, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    const int nx = c->nx, ny = c->ny;
    std::vector<double> a((size_t)nx * ny), b((size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            a[(size_t)i + (size_t)nx * j] = jx[(size_t)i + (size_t)(nx + 1) * j];
            b[(size_t)i + (size_t)nx * j] = jy[(size_t)i + (size_t)(nx + 1) * j];
        }
    PIC_HIP(hipMemcpyAsync(c->jx, a.data(), a.size() * 8, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(c->jy, b.data(), b.size() * 8, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_compute_current(pic_ctx *ctx, int kernel, double *jx, double *jy) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_vm(c, kernel));
    const int nx = c->nx, ny = c->ny;
    const size_t m = (size_t)nx * ny;
    PIC_TRY(launch_fill(c, c->jx, 0.0, (int64_t)m));
    PIC_TRY(launch_fill(c, c->jy, 0.0, (int64_t)m));
    if (c->n > 0) {
        VmArgs a = vm_args(c, 0.0);
        const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((c->n + MT - 1) / MT, 65535));
        hipLaunchKernelGGL(k_current, dim3((unsigned)nb), dim3(MT), 0, c->stream, a);
        PIC_HIP(hipGetLastError());
    }
    if (jx || jy) {  // the reference's (nx+1) x (ny+1) arrays: ghosts are periodic images
        std::vector<double> a(m), b(m);
        PIC_HIP(hipMemcpyAsync(a.data(), c->jx, m * 8, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipMemcpyAsync(b.data(), c->jy, m * 8, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const size_t s = (size_t)(i % nx) + (size_t)nx * (j % ny);
                const size_t d = (size_t)i + (size_t)(nx + 1) * j;
                if (jx) jx[d] = a[s];
                if (jy) jy[d] = b[s];
            }
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_fdtd_energy(pic_ctx *ctx, double *out) {
    PIC_CHECK(ctx && out, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_fdtd(c));
    PIC_TRY(ensure_energy(c, 1));
    hipLaunchKernelGGL(k_fdtd_energy, dim3(1), dim3(MT), 0, c->stream, (const double *)c->fd_ex,
                       (int64_t)c->nx * (c->ny + 1), c->dx * c->dy, c->energy_hist, 0,
                       (int *)nullptr);
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipMemcpyAsync(out, c->energy_hist, 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_run_maxwell(pic_ctx *ctx, int kernel, double dt, int nsteps, double *energy) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(nsteps >= 0, PIC_ERR_ARG, "nsteps %d", nsteps);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(need_vm(c, kernel));
    if (nsteps == 0) return PIC_OK;
    PIC_TRY(ensure_energy(c, nsteps));
    PIC_HIP(hipMemsetAsync(c->step_counter, 0, sizeof(int), c->stream));
    const int64_t m = (int64_t)c->nx * c->ny;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (c->timing) {
        PIC_HIP(hipEventCreate(&ev[0]));
        PIC_HIP(hipEventCreate(&ev[1]));
    }
    int rc = PIC_OK;
    for (int s = 1; s <= nsteps && rc == PIC_OK; ++s) {
        if (s > 1) rc = faraday(c, 0.5 * dt);
        if (rc == PIC_OK) rc = update_fields(c);
        if (rc == PIC_OK) rc = launch_fill(c, c->jx, 0.0, m);
        if (rc == PIC_OK) rc = launch_fill(c, c->jy, 0.0, m);
        if (rc == PIC_OK && c->timing) rc = hipEventRecord(ev[0], c->stream) == hipSuccess ? PIC_OK : PIC_ERR_HIP;
        if (rc == PIC_OK) rc = vm_pass(c, dt);
        if (rc == PIC_OK && c->timing) {
            rc = hipEventRecord(ev[1], c->stream) == hipSuccess &&
                         hipEventSynchronize(ev[1]) == hipSuccess
                     ? PIC_OK
                     : PIC_ERR_HIP;
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
            c->t_pass_ms += ms;
            c->t_passes += 1;
        }
        if (rc == PIC_OK) rc = faraday(c, 0.5 * dt);
        if (rc == PIC_OK) rc = ampere(c, dt);
        if (rc == PIC_OK) {
            hipLaunchKernelGGL(k_fdtd_energy, dim3(1), dim3(MT), 0, c->stream,
                               (const double *)c->fd_ex, (int64_t)c->nx * (c->ny + 1),
                               c->dx * c->dy, c->energy_hist, -1, c->step_counter);
            if (hipGetLastError() != hipSuccess) rc = fail(PIC_ERR_HIP, "energy kernel launch");
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == PIC_OK && energy)
        rc = hipMemcpyAsync(energy, c->energy_hist, (size_t)nsteps * 8, hipMemcpyDeviceToHost,
                            c->stream) == hipSuccess
                 ? PIC_OK
                 : fail(PIC_ERR_HIP, "energy copy");
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == PIC_OK && se != hipSuccess) rc = fail(PIC_ERR_HIP, "pic_run_maxwell: %s", hipGetErrorString(se));
    c->seg_valid = false;
    return rc;
}

}  // extern "C"
