// pic_fftlanes.h -- radix-2 transforms across the lanes of a wave: partners
// exchanged by DPP moves, swizzles and lane permutes instead of LDS round
// trips.  Used by the one-launch small-grid step (pic_small.hip) and by the
// three-kernel E solve of pic_poisson.hip for 64 <= nx, ny <= 512.
#pragma once

#include <hip/hip_runtime.h>

namespace pic {

__device__ __forceinline__ int brev_bits(int i, int logn) {
    return (int)(__brev((unsigned)i) >> (32 - logn));
}

// Value of lane i ^ H of a wave, for H = 1, 2, 8 by DPP (quad permutes,
// a rotation by 8 within rows of 16 lanes: VALU moves), H = 4, 16 by
// ds_swizzle (bitmask mode: xor within 32 lanes, no LDS memory access),
// H = 32 by ds_bpermute.
template <int H>
__device__ __forceinline__ int xor_lane(int v) {
    if constexpr (H == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    else if constexpr (H == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    else if constexpr (H == 8) return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);
    else if constexpr (H == 4 || H == 16) return __builtin_amdgcn_ds_swizzle(v, (H << 10) | 0x1F);
    else return __shfl_xor(v, H);
}

template <int H>
__device__ __forceinline__ double xor_lane(double v) {
    const int lo = xor_lane<H>(__double2loint(v)), hi = xor_lane<H>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Stage 2^LH of the transforms below (skipped when 2^LH >= N): positions
// h = 2^LH apart pair up; the lower takes a + b, the upper (a - b) W^k,
// k = (i mod h) N / (2h).  Branch-free: t = y + s z (s = -1 for the upper,
// exact), then t times w (w = tw[0] = 1 for the lower, no product in the
// last stage).
template <int LH, int R>
__device__ __forceinline__ void fft_stage(double2 (&z)[R], int i, int logN, const double2 *tw,
                                          bool inv) {
    constexpr int h = 1 << LH;
    if (LH + 1 > logN) return;
    double yx[R], yy[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        yx[r] = xor_lane<h>(z[r].x);
        yy[r] = xor_lane<h>(z[r].y);
    }
    const bool up = (i & h) != 0;
    const double sg = up ? -1.0 : 1.0;
    if (LH > 0) {
        const double2 w = tw[up ? (i & (h - 1)) << (logN - 1 - LH) : 0];
        const double wy = inv ? -w.y : w.y;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double tx = __builtin_fma(sg, z[r].x, yx[r]);
            const double ty = __builtin_fma(sg, z[r].y, yy[r]);
            z[r] = make_double2(__builtin_fma(tx, w.x, -(ty * wy)), __builtin_fma(tx, wy, ty * w.x));
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
            z[r] = make_double2(__builtin_fma(sg, z[r].x, yx[r]), __builtin_fma(sg, z[r].y, yy[r]));
    }
}

// R independent transforms of length N = 2^logN <= 64, each held by N
// consecutive lanes (lane position i holds element i of every one),
// interleaved so that their exchanges overlap.  On return position i holds
// element brev(i) of X[k] = sum_j x[j] exp(-+2 pi i jk/N) (- forward, + inv),
// with tw[k] = exp(-2 pi i k/N), k < N/2 (conjugated for inv): radix-2
// decimation in frequency.
template <int R>
__device__ __forceinline__ void fft_lanes(double2 (&z)[R], int i, int logN, const double2 *tw,
                                          bool inv) {
    fft_stage<5, R>(z, i, logN, tw, inv);
    fft_stage<4, R>(z, i, logN, tw, inv);
    fft_stage<3, R>(z, i, logN, tw, inv);
    fft_stage<2, R>(z, i, logN, tw, inv);
    fft_stage<1, R>(z, i, logN, tw, inv);
    fft_stage<0, R>(z, i, logN, tw, inv);
}

// One transform of length N = 64 P (P = 1, 2, 4, 8) per wave: lane l holds
// elements l + 64 r in z[r].  The first log2(P) stages of the decimation in
// frequency pair points held by one lane (distance 64 H, H = P/2 .. 1); the
// remaining six are fft_lanes' stages on the P sub-transforms of 64 lanes,
// whose twiddle indices (l mod h) N/(2h) are those of the full length.  On
// return position l + 64 r holds element brev(l + 64 r) of
// X[k] = sum_j x[j] exp(-+2 pi i jk/N), with tw[k] = exp(-2 pi i k/N),
// k < N/2 (conjugated for inv).
template <int P>
__device__ __forceinline__ void fft_wave(double2 (&z)[P], int l, int logN, const double2 *tw,
                                         bool inv) {
#pragma unroll
    for (int H = P / 2; H >= 1; H /= 2) {
#pragma unroll
        for (int r = 0; r < P; ++r) {
            if (r & H) continue;
            const double2 a = z[r], b = z[r + H];
            const double2 w = tw[(l + 64 * (r % H)) * (P / (2 * H))];
            const double wy = inv ? -w.y : w.y;
            const double dx = a.x - b.x, dy = a.y - b.y;
            z[r] = make_double2(a.x + b.x, a.y + b.y);
            z[r + H] = make_double2(__builtin_fma(dx, w.x, -(dy * wy)), __builtin_fma(dx, wy, dy * w.x));
        }
    }
    fft_lanes<P>(z, l, logN, tw, inv);
}

}  // namespace pic
