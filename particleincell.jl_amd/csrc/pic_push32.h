// pic_push32.h -- per-particle arithmetic of the FP32 cell-relative step,
// shared by the time loop's passes and the adjoint's recomputation (the same
// inline code, so a recomputed trajectory is bit-identical to the forward).
#pragma once

#include "pic_internal.h"

namespace pic {

// Drift of an in-cell offset o in [0, 1) by v dtd cells: the new offset
// (in [0, 1)) and the whole number of cells crossed.
__device__ __forceinline__ int drift1s(float o, float v, float dtd, float &onew) {
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

// Field at a particle from the nodes (u, v) of its stencil, u, v in [0, S):
// CIC bilinear or M6 tensor-product weights, FP32, in a fixed operation order.
template <int K, int BZ, class FE, class FB>
__device__ __forceinline__ void interp32(float fx, float fy, int ci, int cj, FE node_e, FB node_b,
                                         float &e1, float &e2, float &b3) {
    constexpr int S = Shape<K>::S;
    if (K == PIC_CIC) {
        float sx = 1.0f - fx, sy = 1.0f - fy;
        float a0 = sx * sy, a1 = fx * sy, a2 = fx * fy, a3 = sx * fy;
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
        gather_w<K>(fx, ci, wx);
        gather_w<K>(fy, cj, wy);
        e1 = 0.0f;
        e2 = 0.0f;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            float2 E0 = node_e(u, 0);
            float t1 = wy[0] * E0.x, t2 = wy[0] * E0.y;
#pragma unroll
            for (int v = 1; v < S; ++v) {
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
            for (int u = 0; u < S; ++u) {
                float t3 = wy[0] * node_b(u, 0);
#pragma unroll
                for (int v = 1; v < S; ++v) t3 = __builtin_fmaf(wy[v], node_b(u, v), t3);
                b3 = u == 0 ? wx[0] * t3 : __builtin_fmaf(wx[u], t3, b3);
            }
        }
    }
}

// interp32 with the field read two nodes at a time: pair(u, v) returns the
// float4 (ex, ey) of nodes (u, v) and (u + 1, v), u even, which an LDS
// window laid out for it reads with one 16-byte ds_read_b128 (on MI355X
// twice the bytes per clock of the two 8-byte halves of a ds_read2_b64).
// The same operations in the same order as interp32, so the results are
// bit-identical; the S of every shape is even.
template <int K, int BZ, class FP, class FB>
__device__ __forceinline__ void interp32_pairs(float fx, float fy, int ci, int cj, FP pair,
                                               FB node_b, float &e1, float &e2, float &b3) {
    constexpr int S = Shape<K>::S;
    static_assert(S % 2 == 0, "pairs of nodes");
    if (K == PIC_CIC) {
        float sx = 1.0f - fx, sy = 1.0f - fy;
        float a0 = sx * sy, a1 = fx * sy, a2 = fx * fy, a3 = sx * fy;
        const float4 R0 = pair(0, 0), R1 = pair(0, 1);  // (E00, E10), (E01, E11)
        e1 = a0 * R0.x;
        e1 = __builtin_fmaf(a1, R0.z, e1);
        e1 = __builtin_fmaf(a2, R1.z, e1);
        e1 = __builtin_fmaf(a3, R1.x, e1);
        e2 = a0 * R0.y;
        e2 = __builtin_fmaf(a1, R0.w, e2);
        e2 = __builtin_fmaf(a2, R1.w, e2);
        e2 = __builtin_fmaf(a3, R1.y, e2);
        if (BZ == PIC_BZ_GRID) {
            b3 = a0 * node_b(0, 0);
            b3 = __builtin_fmaf(a1, node_b(1, 0), b3);
            b3 = __builtin_fmaf(a2, node_b(1, 1), b3);
            b3 = __builtin_fmaf(a3, node_b(0, 1), b3);
        }
    } else {
        float wx[6], wy[6];
        gather_w<K>(fx, ci, wx);
        gather_w<K>(fy, cj, wy);
        e1 = 0.0f;
        e2 = 0.0f;
#pragma unroll
        for (int u = 0; u < S; u += 2) {
            const float4 P0 = pair(u, 0);
            float t1 = wy[0] * P0.x, t2 = wy[0] * P0.y;
            float q1 = wy[0] * P0.z, q2 = wy[0] * P0.w;
#pragma unroll
            for (int v = 1; v < S; ++v) {
                const float4 Pv = pair(u, v);
                t1 = __builtin_fmaf(wy[v], Pv.x, t1);
                t2 = __builtin_fmaf(wy[v], Pv.y, t2);
                q1 = __builtin_fmaf(wy[v], Pv.z, q1);
                q2 = __builtin_fmaf(wy[v], Pv.w, q2);
            }
            if (u == 0) {
                e1 = wx[0] * t1;
                e2 = wx[0] * t2;
            } else {
                e1 = __builtin_fmaf(wx[u], t1, e1);
                e2 = __builtin_fmaf(wx[u], t2, e2);
            }
            e1 = __builtin_fmaf(wx[u + 1], q1, e1);
            e2 = __builtin_fmaf(wx[u + 1], q2, e2);
        }
        if (BZ == PIC_BZ_GRID) {
#pragma unroll
            for (int u = 0; u < S; ++u) {
                float t3 = wy[0] * node_b(u, 0);
#pragma unroll
                for (int v = 1; v < S; ++v) t3 = __builtin_fmaf(wy[v], node_b(u, v), t3);
                b3 = u == 0 ? wx[0] * t3 : __builtin_fmaf(wx[u], t3, b3);
            }
        }
    }
}

// Velocity update of the FP32 step from the gathered field: half kick,
// rotation by bz, half kick (pushers_cic.jl:38-49 in FP32 with FMAs).
template <int BZ>
__device__ __forceinline__ void kick32(float hdt, float bz0, float e1, float e2, float b3,
                                       float &ux, float &uy) {
    if (BZ == PIC_BZ_UNIFORM) b3 = bz0;
    float v1 = __builtin_fmaf(hdt, e1, ux);
    float v2 = __builtin_fmaf(hdt, e2, uy);
    if (BZ != PIC_BZ_NONE) {
        float tn = hdt * b3;
        float sn = (2.0f * tn) / __builtin_fmaf(tn, tn, 1.0f);
        v1 = __builtin_fmaf(v2, tn, v1);
        v2 = __builtin_fmaf(-v1, sn, v2);
        v1 = __builtin_fmaf(v2, tn, v1);
    }
    ux = __builtin_fmaf(hdt, e1, v1);
    uy = __builtin_fmaf(hdt, e2, v2);
}

// Derivatives d w_k / dt of the M6 weights of m6_weights_f, from the same
// even/odd split in z = t - 1/2: w = E(z^2) +- z Q(z^2) gives
// dw/dt = E'(z) +- (Q + z Q'), with E' = z P(z^2) and Q + z Q' = R(z^2).
__device__ __forceinline__ void m6_dweights_f(float t, float d[6]) {
    const float z = t - 0.5f, z2 = z * z;
    float e, r;
    e = z * __builtin_fmaf(PIC_F(1, 12), z2, PIC_F(1, 48));
    r = __builtin_fmaf(__builtin_fmaf(PIC_F(1, 24), z2, PIC_F(1, 16)), z2, PIC_F(1, 384));
    d[5] = e + r;
    d[0] = e - r;
    e = z * __builtin_fmaf(PIC_F(-1, 4), z2, PIC_F(7, 16));
    r = __builtin_fmaf(__builtin_fmaf(PIC_F(-5, 24), z2, PIC_F(3, 16)), z2, PIC_F(25, 128));
    d[4] = e + r;
    d[1] = e - r;
    e = z * __builtin_fmaf(PIC_F(1, 6), z2, PIC_F(-11, 24));
    r = __builtin_fmaf(__builtin_fmaf(PIC_F(5, 12), z2, PIC_F(-7, 8)), z2, PIC_F(77, 192));
    d[3] = e + r;
    d[2] = e - r;
}

}  // namespace pic
