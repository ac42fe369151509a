// pic_push64.h -- per-particle arithmetic of the FP64 step (the reference's
// operations), shared by the segment passes and the one-launch steps of
// small grids.
#pragma once

#include "pic_internal.h"

namespace pic {

struct SegF64 {
    double *x, *y, *vx, *vy;  // state written by the pass (slot p)
    // state read by the pass (see SegF32): the same arrays or the unsorted
    // ones read at perm[p]; weights w (null when uniform) are read at the
    // same place and, with a permutation, copied to w_out[p]
    const double *x_in, *y_in, *vx_in, *vy_in;
    const uint32_t *perm;
    const double *w;
    double *w_out;
    double w0;
    const double *ex, *ey, *bz;
    double bz0;
    int nx, ny, ncells;
    double dx, dy, dimx, dimy, dt, dtd;
    double *acc;
    const int64_t *off;
    int bps, margin, cap;
    unsigned long long *slow;
    uint32_t *dlist;  // deferred particles and their count
    unsigned *dcount;
};

// The reference's cell and in-cell offset (pushers_cic.jl:18-29): i =
// floor(xp)+1 (trunc in compute_rho CIC, pushers_cic.jl:86-87, equal for
// x >= 0), d = (xp - i) + 1; returns the 0-based wrapped cell.
__device__ __forceinline__ int ref_cell_s(double x, double dx, int n, bool trunc_, double &d) {
    const double xp = x / dx;
    return picd::cell_of(xp, trunc_ ? trunc(xp) : floor(xp), n, d);
}

// Field at a particle from the nodes (u, v) of its stencil, FP64, with the
// reference's operation order (pushers_cic.jl:31-44, pushers_m6.jl:63-83).
template <int K, int BZ, class FE, class FB>
__device__ __forceinline__ void interp64(double fx, double fy, int ci, int cj, FE node_e,
                                         FB node_b, double &e1, double &e2, double &b3) {
    constexpr int S = Shape<K>::S;
    if (K == PIC_CIC) {
        double a1 = (1 - fx) * (1 - fy);
        double a2 = fx * (1 - fy);
        double a3 = fx * fy;
        double a4 = (1 - fx) * fy;
        double2 E00 = node_e(0, 0), E10 = node_e(1, 0), E11 = node_e(1, 1),
                E01 = node_e(0, 1);
        e1 = a1 * E00.x + a2 * E10.x + a3 * E11.x + a4 * E01.x;
        e2 = a1 * E00.y + a2 * E10.y + a3 * E11.y + a4 * E01.y;
        if (BZ == PIC_BZ_GRID)
            b3 = a1 * node_b(0, 0) + a2 * node_b(1, 0) + a3 * node_b(1, 1) +
                 a4 * node_b(0, 1);
    } else {
        double wx[6], wy[6];
        gather_w<K>(fx, ci, wx);
        gather_w<K>(fy, cj, wy);
        e1 = 0.0;
        e2 = 0.0;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            double2 E0 = node_e(u, 0);
            double t1 = wy[0] * E0.x, t2 = wy[0] * E0.y;
#pragma unroll
            for (int v = 1; v < S; ++v) {
                double2 E = node_e(u, v);
                t1 = __builtin_fma(wy[v], E.x, t1);
                t2 = __builtin_fma(wy[v], E.y, t2);
            }
            e1 = __builtin_fma(wx[u], t1, e1);
            e2 = __builtin_fma(wx[u], t2, e2);
        }
        if (BZ == PIC_BZ_GRID) {
#pragma unroll
            for (int u = 0; u < S; ++u) {
                double t3 = wy[0] * node_b(u, 0);
#pragma unroll
                for (int v = 1; v < S; ++v) t3 = __builtin_fma(wy[v], node_b(u, v), t3);
                b3 = __builtin_fma(wx[u], t3, b3);
            }
        }
    }
}

// Half kick, rotation by bz, half kick with the reference's operations
// (pushers_cic.jl:38-49).
template <int BZ>
__device__ __forceinline__ void kick64(const SegF64 &a, double e1, double e2, double b3,
                                       double &pvx, double &pvy) {
    if (BZ == PIC_BZ_UNIFORM) b3 = a.bz0;
    const double dt = a.dt;
    double v1 = pvx, v2 = pvy;
    v1 += 0.5 * dt * e1;
    v2 += 0.5 * dt * e2;
    if (BZ != PIC_BZ_NONE) {
        double tantheta = 0.5 * dt * b3;
        double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
        v1 += v2 * tantheta;
        v2 += -v1 * sintheta;
        v1 += v2 * tantheta;
    }
    pvx = v1 + 0.5 * dt * e1;
    pvy = v2 + 0.5 * dt * e2;
}

// NDRIFT drifts by dtd with the single conditional wrap of push_x!
// (pushers_cic.jl:55-72).
template <int NDRIFT>
__device__ __forceinline__ void drift64(const SegF64 &a, double pvx, double pvy, double &px,
                                        double &py) {
#pragma unroll
    for (int r = 0; r < NDRIFT; ++r) {
        double p1 = px + a.dtd * pvx;
        double p2 = py + a.dtd * pvy;
        if (p1 > a.dimx) p1 -= a.dimx;
        if (p2 > a.dimy) p2 -= a.dimy;
        if (p1 < 0.0) p1 += a.dimx;
        if (p2 < 0.0) p2 += a.dimy;
        px = p1;
        py = p2;
    }
}

// Calls add(u, v, q) with the charge of stencil node (u, v)
// (pushers_cic.jl:91-94, pushers_m6.jl:20-27).
template <int K, class F>
__device__ __forceinline__ void charge64(double fx, double fy, int ci, int cj, double wgt,
                                         F &&add) {
    constexpr int S = Shape<K>::S;
    if (K == PIC_CIC) {
        double a1 = (1 - fx) * (1 - fy);
        double a2 = fx * (1 - fy);
        double a3 = fx * fy;
        double a4 = (1 - fx) * fy;
        add(0, 0, a1 * wgt);
        add(1, 0, a2 * wgt);
        add(1, 1, a3 * wgt);
        add(0, 1, a4 * wgt);
    } else {
        double wx[6], wy[6];
        deposit_w<K>(fx, ci, wx);
        deposit_w<K>(fy, cj, wy);
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int v = 0; v < S; ++v) add(u, v, wx[u] * wy[v] * wgt);
    }
}

}  // namespace pic
