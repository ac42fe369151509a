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

// Fixed-point tiles: exponent e of the quantum u = 2^e with which a block of
// `count` particles accumulates its charge as int32 multiples of u.  Every
// contribution |q| <= qmax |w| (qmax = the largest product of two shape
// weights: 1 for CIC, (66/120)^2 < 0.31 for M6) is rounded to the nearest
// multiple of u (error <= u/2); u is large enough that |q| < 2^22 u (so the
// rounding below is exact integer arithmetic) and that a node's sum stays
// below 2^31 u (sum |q|/u <= 2^30, plus at most count/2 from the roundings).
__host__ __device__ inline int fixed_exponent(float w, double qmax, int64_t count) {
    const double aw = qmax * fabs((double)w) + 1e-300;
    int e1, e2;
    const double m1 = frexp(aw * (double)(count + 5), &e1);
    const double m2 = frexp(aw, &e2);
    e1 = (m1 == 0.5 ? e1 - 1 : e1) - 30;  // ceil(log2(.)) - 30
    e2 = (m2 == 0.5 ? e2 - 1 : e2) - 22;
    return e1 > e2 ? e1 : e2;
}

// Largest block for which the fixed-point tiles conserve charge particle by
// particle (charge32_fixed): the node that takes a particle's remainder
// receives at most its own share plus the S^2 / 2 roundings of the others
// and the 1e-6 by which the FP32 weights may miss a sum of 1 (at most ~50
// units with w/u < 2^24), so a node's sum stays below 2^30 u + 50 count,
// within int32 for count <= 2^24.  Larger blocks round every share.
constexpr int64_t CONSERVE_MAX_COUNT = (int64_t)1 << 24;

// Calls add(u, v, q) with the FP32 charge q of stencil node (u, v) of a
// particle at offset (fx, fy), in a fixed order.
template <int K, class F>
__device__ __forceinline__ void charge32(float fx, float fy, int ci, int cj, float w, F &&add) {
    constexpr int S = Shape<K>::S;
    if (K == PIC_CIC) {
        float sx = 1.0f - fx, sy = 1.0f - fy;
        add(0, 0, (sx * sy) * w);
        add(1, 0, (fx * sy) * w);
        add(1, 1, (fx * fy) * w);
        add(0, 1, (sx * fy) * w);
    } else {
        float wx[6], wy[6];
        deposit_w<K>(fx, ci, wx);
        deposit_w<K>(fy, cj, wy);
#pragma unroll
        for (int v = 0; v < S; ++v) wy[v] = wy[v] * w;
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int v = 0; v < S; ++v) add(u, v, wx[u] * wy[v]);
    }
}

// The int32 fixed-point charges of a particle's stencil nodes, in units of
// the quantum u = 2^qe: add(u, v, n) with n = the exact product of the two
// shape factors times w/u, rounded to the nearest integer by one fused
// multiply-add onto 1.5 * 2^23 (the float grid there is the integers; the
// products stay below 2^22 by the choice of u).
//
// With `conserve` the node (HL | 1, HL) (the cell's own node for TSC and M4,
// the next one for CIC, NGP and M6: always an odd column of the stencil, so
// that in the packed tiles below it is the high half of its pair) takes R
// minus the sum of the other nodes' rounded shares, R = round(|w|/u) the same
// for every particle of the block, so that every particle adds exactly R
// units and the flush's scale w/R gives the block exactly its count times w.  Rounding every share to nearest instead drops
// the many shares below u/2 at a stencil's far nodes (M6: products of two
// weights of ~1e-3): measured on 5e8 particles on 512^2, ~2e4 units per
// particle, the deposit lost 2.2e-5 of the charge.  The remainder node
// absorbs those roundings (at most S^2/2 units) and the ~1e-6 by which the
// FP32 weights miss a sum of 1.
template <int K, class F>
__device__ __forceinline__ void charge32_fixed(float fx, float fy, int ci, int cj, float wu, int R,
                                               bool conserve, F &&add) {
    constexpr int S = Shape<K>::S, HL = Shape<K>::HL, UR = HL | 1, VR = HL;
    constexpr float C = 12582912.0f;  // 1.5 * 2^23
    auto n = [&](float a, float b) {
        return __float_as_int(__builtin_fmaf(a, b, C)) - __float_as_int(C);
    };
    if (K == PIC_CIC) {
        const float sx = 1.0f - fx, sy = (1.0f - fy) * wu, ty = fy * wu;
        const int n00 = n(sx, sy), n11 = n(fx, ty), n01 = n(sx, ty);
        add(0, 0, n00);
        add(1, 1, n11);
        add(0, 1, n01);
        add(1, 0, conserve ? R - (n00 + n11 + n01) : n(fx, sy));
    } else {
        float wx[6], wy[6];
        deposit_w<K>(fx, ci, wx);
        deposit_w<K>(fy, cj, wy);
#pragma unroll
        for (int v = 0; v < S; ++v) wy[v] = wy[v] * wu;
        int s = 0;
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int v = 0; v < S; ++v) {
                if (u == UR && v == VR) continue;
                const int q = n(wx[u], wy[v]);
                s += q;
                add(u, v, q);
            }
        add(UR, VR, conserve ? R - s : n(wx[UR], wy[VR]));
    }
}

}  // namespace pic
