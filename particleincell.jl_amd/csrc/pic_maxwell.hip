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

// The reference's CIC cell (floor) and offset of a coordinate
// (pushers_cic.jl:18-24, cloud_in_cell.jl:36-42).
__device__ __forceinline__ int cic_cell(double x, double d, int n, double &off) {
    const double xp = x / d;
    return picd::cell_of(xp, floor(xp), n, off);
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
template <int BS>
__global__ __launch_bounds__(BS) void k_vm_pass(VmArgs a) {
    extern __shared__ double jsh[];  // jx then jy, nx*ny each, when tiled
    const int nx = a.nx, ny = a.ny, m = nx * ny;
    if (a.tiled) {
        for (int t = threadIdx.x; t < 2 * m; t += BS) jsh[t] = 0.0;
        __syncthreads();
    }
    for (int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x; k < a.n;
         k += (int64_t)gridDim.x * BS) {
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
    for (int t = threadIdx.x; t < m; t += BS) {
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

// (i, j) of t = i + ld j, walked by a fixed stride without divisions.
struct Walk {
    int i, j, di, dj, ld;
    __device__ Walk(int t, int ld_, int stride)
        : i(t % ld_), j(t / ld_), di(stride % ld_), dj(stride / ld_), ld(ld_) {}
    __device__ void next() {
        i += di;
        j += dj;
        if (i >= ld) {
            i -= ld;
            ++j;
        }
    }
};

constexpr int FT = 1024;  // threads of a block that works the whole grid in LDS
constexpr int VT = 1024;  // threads of a pass block that sums the current in LDS
constexpr size_t VM_LDS_BYTES_MAX = 120 * 1024;
constexpr size_t VM_STEP_LDS_MAX = 150 * 1024;

// ---- the FDTD updates on LDS copies of the fields (FT threads each)

// faraday!(dt): fdtd.jl:24-36, bz += dt curl e.
__device__ __forceinline__ void lds_faraday(int nx, int ny, double dx, double dy, double dt,
                                            const double *sex, const double *sey, double *sbz) {
    const int m = nx * ny, tid = threadIdx.x;
    Walk w(tid, nx, FT);
    for (int t = tid; t < m; t += FT, w.next()) {
        const int i = w.i, j = w.j;
        const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
        const double dex_dy = (sex[i + nx * jp] - sex[i + nx * j]) / dy;
        const double dey_dx = (sey[ip + (nx + 1) * j] - sey[i + (nx + 1) * j]) / dx;
        sbz[t] += dt * (dex_dy - dey_dx);
    }
}

// ampere_maxwell!(dt): fdtd.jl:40-54, with the periodic current sjx, sjy.
__device__ __forceinline__ void lds_ampere(int nx, int ny, double dx, double dy, double dt,
                                           double *sex, double *sey, const double *sbz,
                                           const double *sjx, const double *sjy) {
    const int mex = nx * (ny + 1), mey = (nx + 1) * ny, tid = threadIdx.x;
    Walk w(tid, nx, FT);
    for (int t = tid; t < mex; t += FT, w.next()) {
        const int i = w.i, j = w.j;  // j in [0, ny]
        const int jc = j == ny ? 0 : j, jm = j == 0 ? ny - 1 : j - 1;
        const int ip = i + 1 == nx ? 0 : i + 1;
        const double dbz_dy = (sbz[i + nx * jc] - sbz[i + nx * jm]) / dy;
        sex[t] += dt * dbz_dy - dt * 0.5 * (sjx[i + nx * jc] + sjx[ip + nx * jc]);
    }
    Walk v(tid, nx + 1, FT);
    for (int u = tid; u < mey; u += FT, v.next()) {
        const int i = v.i, j = v.j;  // i in [0, nx]
        const int ic = i == nx ? 0 : i, im = i == 0 ? nx - 1 : i - 1;
        const int jp = j + 1 == ny ? 0 : j + 1;
        const double dbz_dx = (sbz[ic + nx * j] - sbz[im + nx * j]) / dx;
        sey[u] -= dt * dbz_dx - dt * 0.5 * (sjy[ic + nx * j] + sjy[ic + nx * jp]);
    }
}

// compute_energy: fdtd.jl:79-80 in a fixed order (per-thread strided sums,
// a wave butterfly, the waves in order); every thread of the block calls
// it; thread 0 writes energy[*counter] and increments the counter.
__device__ __forceinline__ void lds_energy(const double *sex, int mex, double dxdy, double *red,
                                           double *energy, int *counter) {
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int t = tid; t < mex; t += FT) s += sex[t] * sex[t];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        double acc = 0.0;
        for (int w = 0; w < FT / 64; ++w) acc += red[w];
        const int q = *counter;
        *counter = q + 1;
        energy[q] = 0.5 * log(acc * dxdy);
    }
}

// update_fields!: fdtd.jl:60-77 on the nx x ny nodes, from the FDTD fields
// (global or LDS) into LDS node arrays; nodes [lo, hi) also to global memory.
template <int BS>
__device__ __forceinline__ void nodes_from_fdtd(int nx, int ny, const double *fex,
                                                const double *fey, const double *fbz,
                                                double *sex, double *sey, double *sbz, int lo,
                                                int hi, double *nex, double *ney, double *nbz) {
    const int m = nx * ny, tid = threadIdx.x;
    Walk w(tid, nx, BS);
    for (int t = tid; t < m; t += BS, w.next()) {
        const int i = w.i, j = w.j;
        const int im = i == 0 ? nx - 1 : i - 1, jm = j == 0 ? ny - 1 : j - 1;
        const double e1 = 0.5 * (fex[im + nx * j] + fex[i + nx * j]);
        const double e2 = 0.5 * (fey[i + (nx + 1) * jm] + fey[i + (nx + 1) * j]);
        const double b3 =
            0.25 * (((fbz[im + nx * jm] + fbz[i + nx * jm]) + fbz[im + nx * j]) + fbz[i + nx * j]);
        sex[t] = e1;
        sey[t] = e2;
        sbz[t] = b3;
        if (t >= lo && t < hi) {
            nex[t] = e1;
            ney[t] = e2;
            nbz[t] = b3;
        }
    }
}

// P particles per thread, loads first (kept in registers across calls).
template <int BS, int P>
struct VmBatch {
    double px[P], py[P], v1[P], v2[P], wg[P];
    __device__ void load(const VmArgs &a, int64_t base) {
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const int64_t k = base + r * BS + threadIdx.x;
            if (k < a.n) {
                px[r] = a.x[k];
                py[r] = a.y[k];
                v1[r] = a.vx[k];
                v2[r] = a.vy[k];
                wg[r] = a.w ? a.w[k] : a.w0;
            }
        }
    }
    // push_v!(dt) (pushers_cic.jl:3-53), push_x!(dt/2), compute_current!
    // (cloud_in_cell.jl:31-60), push_x!(dt/2), with the node fields and the
    // current accumulators in LDS.
    __device__ void push(const VmArgs &a, int64_t base, const double *sex, const double *sey,
                         const double *sbz, double *sjx, double *sjy) {
        const int nx = a.nx, ny = a.ny;
        const double dt = a.dt, h = 0.5 * dt;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const int64_t k = base + r * BS + threadIdx.x;
            if (k >= a.n) break;
            double x = px[r], y = py[r], u1 = v1[r], u2 = v2[r];
            {  // push_v!: pushers_cic.jl:13-51
                double dxp, dyp;
                const int i = cic_cell(x, a.dx, nx, dxp), j = cic_cell(y, a.dy, ny, dyp);
                const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
                const double a1 = (1 - dxp) * (1 - dyp), a2 = dxp * (1 - dyp), a3 = dxp * dyp,
                             a4 = (1 - dxp) * dyp;
                const int c00 = i + nx * j, c10 = ip + nx * j, c11 = ip + nx * jp, c01 = i + nx * jp;
                const double e1 = a1 * sex[c00] + a2 * sex[c10] + a3 * sex[c11] + a4 * sex[c01];
                const double e2 = a1 * sey[c00] + a2 * sey[c10] + a3 * sey[c11] + a4 * sey[c01];
                const double b3 = a1 * sbz[c00] + a2 * sbz[c10] + a3 * sbz[c11] + a4 * sbz[c01];
                u1 += 0.5 * dt * e1;
                u2 += 0.5 * dt * e2;
                const double tantheta = 0.5 * dt * b3;
                const double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
                u1 += u2 * tantheta;
                u2 += -u1 * sintheta;
                u1 += u2 * tantheta;
                u1 = u1 + 0.5 * dt * e1;
                u2 = u2 + 0.5 * dt * e2;
            }
            auto drift = [&]() {  // push_x!(dt/2): pushers_cic.jl:55-72
                double p1 = x + h * u1, p2 = y + h * u2;
                if (p1 > a.dimx) p1 -= a.dimx;
                if (p2 > a.dimy) p2 -= a.dimy;
                if (p1 < 0.0) p1 += a.dimx;
                if (p2 < 0.0) p2 += a.dimy;
                x = p1;
                y = p2;
            };
            drift();
            {  // compute_current!: cloud_in_cell.jl:31-60
                double dxp, dyp;
                const int i = cic_cell(x, a.dx, nx, dxp), j = cic_cell(y, a.dy, ny, dyp);
                const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
                const double a1 = (1 - dxp) * (1 - dyp), a2 = dxp * (1 - dyp), a3 = dxp * dyp,
                             a4 = (1 - dxp) * dyp;
                const double w1 = u1 * a.scaling * wg[r], w2 = u2 * a.scaling * wg[r];
                const int c[4] = {i + nx * j, ip + nx * j, ip + nx * jp, i + nx * jp};
                const double q[4] = {a1, a2, a3, a4};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    __hip_atomic_fetch_add(&sjx[c[e]], q[e] * w1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(&sjy[c[e]], q[e] * w2, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            drift();
            a.x[k] = x;
            a.y[k] = y;
            a.vx[k] = u1;
            a.vy[k] = u2;
        }
    }
};

// The block's LDS current added to the global one (non-zero nodes).
template <int BS>
__device__ __forceinline__ void flush_current(int m, const double *sjx, const double *sjy,
                                              double *jx, double *jy) {
    for (int t = threadIdx.x; t < m; t += BS) {
        if (sjx[t] != 0.0) unsafeAtomicAdd(&jx[t], sjx[t]);
        if (sjy[t] != 0.0) unsafeAtomicAdd(&jy[t], sjy[t]);
    }
}

// The loop's particle work for grids whose node fields and current fit in
// one workgroup's LDS.  Each block forms the node fields of update_fields!
// from the FDTD fields in LDS and writes its share of them to the context's
// node grids, then runs P particles per thread through the particle half of
// the iteration (loads of all P first), sums the current in LDS and adds it
// to jx, jy with global atomics.
template <int BS, int P>
__global__ __launch_bounds__(BS) void k_vm_pass_lds(VmArgs a, const double *__restrict__ fex,
                                                    const double *__restrict__ fey,
                                                    const double *__restrict__ fbz,
                                                    double *__restrict__ nex,
                                                    double *__restrict__ ney,
                                                    double *__restrict__ nbz) {
    extern __shared__ double lds[];
    const int nx = a.nx, ny = a.ny, m = nx * ny, tid = threadIdx.x;
    double *sex = lds, *sey = lds + m, *sbz = lds + 2 * m, *sjx = lds + 3 * m, *sjy = lds + 4 * m;
    const int64_t stride = (int64_t)gridDim.x * (BS * P);
    int64_t base = (int64_t)blockIdx.x * (BS * P);
    VmBatch<BS, P> B;
    B.load(a, base);
    const int lo = (int)((int64_t)m * blockIdx.x / gridDim.x);
    const int hi = (int)((int64_t)m * (blockIdx.x + 1) / gridDim.x);
    nodes_from_fdtd<BS>(nx, ny, fex, fey, fbz, sex, sey, sbz, lo, hi, nex, ney, nbz);
    for (int t = tid; t < m; t += BS) {
        sjx[t] = 0.0;
        sjy[t] = 0.0;
    }
    __syncthreads();
    while (base < a.n) {
        B.push(a, base, sex, sey, sbz, sjx, sjy);
        base += stride;
        if (base < a.n) B.load(a, base);
    }
    __syncthreads();
    flush_current<BS>(m, sjx, sjy, a.jx, a.jy);
}

// The field work between two such passes, in one workgroup with the FDTD
// fields and the current in LDS: one launch instead of five.  The same
// arithmetic as k_faraday, k_ampere and k_fdtd_energy, in the loop's order:
//   faraday!(dt/2), ampere_maxwell!(dt) with jx, jy, the energy into
//   energy[*counter] (then incremented), and with `next` (every iteration
//   but the last) jx = jy = 0 and the next iteration's faraday!(dt/2).
// fex/fey/fbz are read from (fin_*) and written to (fex...); the two may be
// the same arrays.
__global__ __launch_bounds__(FT) void k_vm_fields(int nx, int ny, double dx, double dy, double dt,
                                                  const double *fin_ex, const double *fin_ey,
                                                  const double *fin_bz, double *fex, double *fey,
                                                  double *fbz, double *jx, double *jy,
                                                  double *energy, int *counter, int next) {
    extern __shared__ double lds[];
    __shared__ double red[FT / 64];
    const int m = nx * ny, mex = nx * (ny + 1), mey = (nx + 1) * ny, tid = threadIdx.x;
    double *sex = lds, *sey = lds + mex, *sbz = sey + mey, *sjx = sbz + m, *sjy = sjx + m;
    for (int t = tid; t < mex; t += FT) sex[t] = fin_ex[t];
    for (int t = tid; t < mey; t += FT) sey[t] = fin_ey[t];
    for (int t = tid; t < m; t += FT) {
        sbz[t] = fin_bz[t];
        sjx[t] = jx[t];
        sjy[t] = jy[t];
    }
    __syncthreads();
    const double h = 0.5 * dt;
    lds_faraday(nx, ny, dx, dy, h, sex, sey, sbz);
    __syncthreads();
    lds_ampere(nx, ny, dx, dy, dt, sex, sey, sbz, sjx, sjy);
    __syncthreads();
    lds_energy(sex, mex, dx * dy, red, energy, counter);
    if (next) {  // the current is consumed: clear it for the next pass
        for (int t = tid; t < m; t += FT) {
            jx[t] = 0.0;
            jy[t] = 0.0;
        }
        lds_faraday(nx, ny, dx, dy, h, sex, sey, sbz);
        __syncthreads();
    }
    for (int t = tid; t < mex; t += FT) fex[t] = sex[t];
    for (int t = tid; t < mey; t += FT) fey[t] = sey[t];
    for (int t = tid; t < m; t += FT) fbz[t] = sbz[t];
}

// One whole iteration s of the loop in one launch, for grids small enough
// that every block holds the FDTD fields, the current and the node fields in
// LDS (up to 2048 cells).  Every block first finishes iteration s-1's field
// work on its own LDS copy, all with the same operations on the same inputs:
// faraday!(dt/2), ampere_maxwell!(dt) with the current j(s-1), then
// iteration s's leading faraday!(dt/2) and update_fields!; block 0 also
// records the energy of iteration s-1 and stores the fields for the next
// launch.  Then the particles as in k_vm_pass_lds, into j(s).  The FDTD
// fields alternate between two buffers and the current between three (the
// one read, the one summed into, the one cleared for the next launch), so no
// launch writes what another block of it reads.  For s = 1 the field work is
// only update_fields! of the initial fields.
struct VmStepIo {
    const double *bin_ex, *bin_ey, *bin_bz;  // the fields the field work starts from
    double *bout_ex, *bout_ey, *bout_bz;     // after it (block 0 writes)
    const double *jin_x, *jin_y;             // j(s-1)
    double *jacc_x, *jacc_y;                 // j(s), zero on entry
    double *jclr_x, *jclr_y;                 // cleared for the next launch
    double *nex, *ney, *nbz;                 // the context's node grids
    double *energy;
    int *counter;
    int first;
};

template <int P>
__global__ __launch_bounds__(VT) void k_vm_step(VmArgs a, VmStepIo io) {
    extern __shared__ double lds[];
    __shared__ double red[FT / 64];
    const int nx = a.nx, ny = a.ny, m = nx * ny, mex = nx * (ny + 1), mey = (nx + 1) * ny;
    const int tid = threadIdx.x;
    double *fex = lds, *fey = fex + mex, *fbz = fey + mey, *fjx = fbz + m, *fjy = fjx + m;
    double *sex = fjy + m, *sey = sex + m, *sbz = sey + m;  // node fields
    double *sjx = lds, *sjy = lds + m;  // the current, over the FDTD region once it is free
    const int64_t stride = (int64_t)gridDim.x * (VT * P);
    int64_t base = (int64_t)blockIdx.x * (VT * P);
    VmBatch<VT, P> B;
    B.load(a, base);
    for (int t = tid; t < mex; t += VT) fex[t] = io.bin_ex[t];
    for (int t = tid; t < mey; t += VT) fey[t] = io.bin_ey[t];
    for (int t = tid; t < m; t += VT) fbz[t] = io.bin_bz[t];
    if (!io.first)
        for (int t = tid; t < m; t += VT) {
            fjx[t] = io.jin_x[t];
            fjy[t] = io.jin_y[t];
        }
    __syncthreads();
    const double h = 0.5 * a.dt;
    if (!io.first) {
        lds_faraday(nx, ny, a.dx, a.dy, h, fex, fey, fbz);
        __syncthreads();
        lds_ampere(nx, ny, a.dx, a.dy, a.dt, fex, fey, fbz, fjx, fjy);
        __syncthreads();
        if (blockIdx.x == 0) lds_energy(fex, mex, a.dx * a.dy, red, io.energy, io.counter);
        lds_faraday(nx, ny, a.dx, a.dy, h, fex, fey, fbz);
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        for (int t = tid; t < mex; t += VT) io.bout_ex[t] = fex[t];
        for (int t = tid; t < mey; t += VT) io.bout_ey[t] = fey[t];
        for (int t = tid; t < m; t += VT) io.bout_bz[t] = fbz[t];
    }
    const int lo = (int)((int64_t)m * blockIdx.x / gridDim.x);
    const int hi = (int)((int64_t)m * (blockIdx.x + 1) / gridDim.x);
    nodes_from_fdtd<VT>(nx, ny, fex, fey, fbz, sex, sey, sbz, lo, hi, io.nex, io.ney, io.nbz);
    for (int t = lo + tid; t < hi; t += VT) {
        io.jclr_x[t] = 0.0;
        io.jclr_y[t] = 0.0;
    }
    __syncthreads();
    for (int t = tid; t < m; t += VT) {
        sjx[t] = 0.0;
        sjy[t] = 0.0;
    }
    __syncthreads();
    while (base < a.n) {
        B.push(a, base, sex, sey, sbz, sjx, sjy);
        base += stride;
        if (base < a.n) B.load(a, base);
    }
    __syncthreads();
    flush_current<VT>(m, sjx, sjy, io.jacc_x, io.jacc_y);
}

inline size_t vm_field_lds(const Ctx *c) {
    const size_t nx = c->nx, ny = c->ny;
    return (nx * (ny + 1) + (nx + 1) * ny + 3 * nx * ny) * 8;
}

inline size_t vm_step_lds(const Ctx *c) {
    const size_t nx = c->nx, ny = c->ny;
    return (nx * (ny + 1) + (nx + 1) * ny + 6 * nx * ny) * 8;
}

inline size_t vm_pass_lds(const Ctx *c) { return (size_t)5 * c->nx * c->ny * 8; }

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

int ensure_step_buffers(Ctx *c) {
    const size_t nx = c->nx, ny = c->ny;
    if (!c->fd2_ex) {
        PIC_HIP(hipMalloc(&c->fd2_ex, nx * (ny + 1) * 8));
        PIC_HIP(hipMalloc(&c->fd2_ey, (nx + 1) * ny * 8));
        PIC_HIP(hipMalloc(&c->fd2_bz, nx * ny * 8));
        for (double *&p : c->jb) PIC_HIP(hipMalloc(&p, nx * ny * 8));
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

int64_t vm_tiled_blocks(const Ctx *c) { return (int64_t)c->ncu / 2; }

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
constexpr size_t VM_LDS_MAX = 64 * 1024;

int vm_pass(Ctx *c, double dt) {
    if (c->n == 0) return PIC_OK;
    VmArgs a = vm_args(c, dt);
    const size_t grid_bytes = (size_t)2 * c->nx * c->ny * 8;
    a.tiled = grid_bytes <= VM_LDS_MAX;
    if (a.tiled) {
        // every block flushes its whole LDS grid with global atomics, which
        // all meet at the same 2 nx ny addresses: few, wide blocks
        int64_t nb = (c->n + VT - 1) / VT;
        nb = std::max<int64_t>(1, std::min<int64_t>(nb, vm_tiled_blocks(c)));
        hipLaunchKernelGGL(k_vm_pass<VT>, dim3((unsigned)nb), dim3(VT), grid_bytes, c->stream, a);
    } else {
        const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((c->n + MT - 1) / MT, 65535));
        hipLaunchKernelGGL(k_vm_pass<MT>, dim3((unsigned)nb), dim3(MT), 0, c->stream, a);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

constexpr size_t LDS_DEFAULT_MAX = 64 * 1024;

// Dynamic LDS above the default 64 KB needs the kernels' opt-in.
int vm_lds_attributes(size_t flds, size_t plds) {
    if (flds > LDS_DEFAULT_MAX)
        PIC_HIP(hipFuncSetAttribute((const void *)k_vm_fields,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)flds));
    if (plds > LDS_DEFAULT_MAX) {
        const void *fns[6] = {(const void *)k_vm_pass_lds<VT, 1>, (const void *)k_vm_pass_lds<VT, 2>,
                              (const void *)k_vm_pass_lds<VT, 4>, (const void *)k_vm_step<1>,
                              (const void *)k_vm_step<2>, (const void *)k_vm_step<4>};
        for (const void *f : fns)
            PIC_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds));
    }
    return PIC_OK;
}

// Particles per thread (1, 2 or 4) and blocks of a launch whose blocks need
// `lds` bytes each: the fewest particles per thread for which one round of
// resident blocks covers them all.
void vm_shape(const Ctx *c, size_t lds, int &P, int64_t &nb) {
    const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(2, (int64_t)(160 * 1024 / lds)));
    const int64_t resident = (int64_t)c->ncu * per_cu;
    P = 1;
    nb = (c->n + VT - 1) / VT;
    while (nb > resident && P < 4) {
        P *= 2;
        nb = (c->n + (int64_t)VT * P - 1) / ((int64_t)VT * P);
    }
    nb = std::max<int64_t>(1, std::min<int64_t>(nb, resident));
}

int vm_step_launch(Ctx *c, double dt, const VmStepIo &io, size_t lds) {
    const VmArgs a = vm_args(c, dt);
    int P;
    int64_t nb;
    vm_shape(c, lds, P, nb);
    if (P == 1)
        hipLaunchKernelGGL(k_vm_step<1>, dim3((unsigned)nb), dim3(VT), lds, c->stream, a, io);
    else if (P == 2)
        hipLaunchKernelGGL(k_vm_step<2>, dim3((unsigned)nb), dim3(VT), lds, c->stream, a, io);
    else
        hipLaunchKernelGGL(k_vm_step<4>, dim3((unsigned)nb), dim3(VT), lds, c->stream, a, io);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// k_vm_pass_lds over the particles: one or two blocks per compute unit (as
// the LDS allows), each thread taking the fewest particles (1, 2 or 4 per
// round) for which one round of resident blocks covers them all.
int vm_pass_lds_launch(Ctx *c, double dt) {
    const VmArgs a = vm_args(c, dt);
    const size_t lds = vm_pass_lds(c);
    int P;
    int64_t nb;
    vm_shape(c, lds, P, nb);
    const double *fx = c->fd_ex, *fy = c->fd_ey, *fb = c->fd_bz;
    if (P == 1)
        hipLaunchKernelGGL((k_vm_pass_lds<VT, 1>), dim3((unsigned)nb), dim3(VT), lds, c->stream, a,
                           fx, fy, fb, c->ex, c->ey, c->bz);
    else if (P == 2)
        hipLaunchKernelGGL((k_vm_pass_lds<VT, 2>), dim3((unsigned)nb), dim3(VT), lds, c->stream, a,
                           fx, fy, fb, c->ex, c->ey, c->bz);
    else
        hipLaunchKernelGGL((k_vm_pass_lds<VT, 4>), dim3((unsigned)nb), dim3(VT), lds, c->stream, a,
                           fx, fy, fb, c->ex, c->ey, c->bz);
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

int pic_set_maxwell_path(pic_ctx *ctx, int path) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(path >= 0 && path <= 3 && path != 1, PIC_ERR_ARG, "path %d", path);
    ctx_of(ctx)->vm_path = path;
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
    // timing: events around every particle pass, read after the loop
    std::vector<hipEvent_t> ev;
    auto mark = [&]() -> int {
        hipEvent_t e;
        PIC_HIP(hipEventCreate(&e));
        ev.push_back(e);
        PIC_HIP(hipEventRecord(e, c->stream));
        return PIC_OK;
    };
    int rc = PIC_OK;
    const size_t flds = vm_field_lds(c), plds = vm_pass_lds(c), slds = vm_step_lds(c);
    const bool one = slds <= VM_STEP_LDS_MAX && flds <= VM_STEP_LDS_MAX && c->vm_path != 2;
    const bool fused = !one && flds <= VM_LDS_BYTES_MAX && plds <= VM_LDS_BYTES_MAX &&
                       c->vm_path != 3;
    auto timed_pass = [&](bool lds) {
        int r = c->timing ? mark() : PIC_OK;
        if (r == PIC_OK) r = lds ? vm_pass_lds_launch(c, dt) : vm_pass(c, dt);
        if (r == PIC_OK && c->timing) r = mark();
        return r;
    };
    if (one) {
        // one launch per iteration (k_vm_step), then the last iteration's
        // field work; FDTD fields alternate between B0 (the context's) and B1,
        // the current among J0 (the context's), J1, J2
        rc = ensure_step_buffers(c);
        if (rc == PIC_OK) rc = vm_lds_attributes(flds, slds);
        double *B[2][3] = {{c->fd_ex, c->fd_ey, c->fd_bz}, {c->fd2_ex, c->fd2_ey, c->fd2_bz}};
        double *J[3][2] = {{c->jx, c->jy}, {c->jb[0], c->jb[1]}, {c->jb[2], c->jb[3]}};
        if (rc == PIC_OK) rc = launch_fill(c, J[1][0], 0.0, m);
        if (rc == PIC_OK) rc = launch_fill(c, J[1][1], 0.0, m);
        c->bz_mode = PIC_BZ_GRID;
        for (int s = 1; s <= nsteps && rc == PIC_OK; ++s) {
            VmStepIo io;
            double **bi = B[(s - 1) & 1], **bo = B[s & 1];
            io.bin_ex = bi[0]; io.bin_ey = bi[1]; io.bin_bz = bi[2];
            io.bout_ex = bo[0]; io.bout_ey = bo[1]; io.bout_bz = bo[2];
            io.jin_x = J[(s - 1) % 3][0]; io.jin_y = J[(s - 1) % 3][1];
            io.jacc_x = J[s % 3][0]; io.jacc_y = J[s % 3][1];
            io.jclr_x = J[(s + 1) % 3][0]; io.jclr_y = J[(s + 1) % 3][1];
            io.nex = c->ex; io.ney = c->ey; io.nbz = c->bz;
            io.energy = c->energy_hist;
            io.counter = c->step_counter;
            io.first = s == 1;
            const bool timed = c->timing && (s - 1) % c->timing_every == 0;
            if (timed) rc = mark();
            if (rc == PIC_OK) rc = vm_step_launch(c, dt, io, slds);
            if (rc == PIC_OK && timed) rc = mark();
        }
        if (rc == PIC_OK) {
            double **bi = B[nsteps & 1];
            double **jn = J[nsteps % 3];
            hipLaunchKernelGGL(k_vm_fields, dim3(1), dim3(FT), flds, c->stream, c->nx, c->ny,
                               c->dx, c->dy, dt, (const double *)bi[0], (const double *)bi[1],
                               (const double *)bi[2], c->fd_ex, c->fd_ey, c->fd_bz, jn[0], jn[1],
                               c->energy_hist, c->step_counter, 0);
            if (hipGetLastError() != hipSuccess) rc = fail(PIC_ERR_HIP, "field kernel launch");
        }
        if (rc == PIC_OK && nsteps % 3 != 0) {  // the context keeps the last current
            const size_t bytes = (size_t)m * 8;
            double **jn = J[nsteps % 3];
            if (hipMemcpyAsync(c->jx, jn[0], bytes, hipMemcpyDeviceToDevice, c->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(c->jy, jn[1], bytes, hipMemcpyDeviceToDevice, c->stream) !=
                    hipSuccess)
                rc = fail(PIC_ERR_HIP, "current copy");
        }
    }
    if (fused) {
        // iteration s: the pass (with update_fields! in every block), then
        // faraday!(dt/2), ampere_maxwell!, the energy and, but after the
        // last iteration, the next one's faraday!(dt/2) in one field launch
        rc = vm_lds_attributes(flds, plds);
        if (rc == PIC_OK) rc = launch_fill(c, c->jx, 0.0, m);
        if (rc == PIC_OK) rc = launch_fill(c, c->jy, 0.0, m);
        c->bz_mode = PIC_BZ_GRID;
        for (int s = 1; s <= nsteps && rc == PIC_OK; ++s) {
            rc = timed_pass(true);
            if (rc == PIC_OK) {
                hipLaunchKernelGGL(k_vm_fields, dim3(1), dim3(FT), flds, c->stream, c->nx, c->ny,
                                   c->dx, c->dy, dt, (const double *)c->fd_ex,
                                   (const double *)c->fd_ey, (const double *)c->fd_bz, c->fd_ex,
                                   c->fd_ey, c->fd_bz, c->jx, c->jy, c->energy_hist,
                                   c->step_counter, s < nsteps ? 1 : 0);
                if (hipGetLastError() != hipSuccess) rc = fail(PIC_ERR_HIP, "field kernel launch");
            }
        }
    }
    for (int s = 1; s <= nsteps && rc == PIC_OK && !fused && !one; ++s) {
        if (s > 1) rc = faraday(c, 0.5 * dt);
        if (rc == PIC_OK) rc = update_fields(c);
        if (rc == PIC_OK) rc = launch_fill(c, c->jx, 0.0, m);
        if (rc == PIC_OK) rc = launch_fill(c, c->jy, 0.0, m);
        if (rc == PIC_OK) rc = timed_pass(false);
        if (rc == PIC_OK) rc = faraday(c, 0.5 * dt);
        if (rc == PIC_OK) rc = ampere(c, dt);
        if (rc == PIC_OK) {
            hipLaunchKernelGGL(k_fdtd_energy, dim3(1), dim3(MT), 0, c->stream,
                               (const double *)c->fd_ex, (int64_t)c->nx * (c->ny + 1),
                               c->dx * c->dy, c->energy_hist, -1, c->step_counter);
            if (hipGetLastError() != hipSuccess) rc = fail(PIC_ERR_HIP, "energy kernel launch");
        }
    }
    if (rc == PIC_OK && energy)
        rc = hipMemcpyAsync(energy, c->energy_hist, (size_t)nsteps * 8, hipMemcpyDeviceToHost,
                            c->stream) == hipSuccess
                 ? PIC_OK
                 : fail(PIC_ERR_HIP, "energy copy");
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == PIC_OK && se != hipSuccess) rc = fail(PIC_ERR_HIP, "pic_run_maxwell: %s", hipGetErrorString(se));
    if (rc == PIC_OK && c->timing) {
        for (size_t k = 0; k + 1 < ev.size(); k += 2) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            c->t_pass_ms += ms;
            c->t_passes += 1;
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    c->seg_valid = false;
    c->sort_fresh = false;
    return rc;
}

}  // extern "C"
