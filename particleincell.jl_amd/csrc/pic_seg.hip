// pic_seg.hip -- segment-persistent particle passes of the device time loop.
//
// After a lane sort (pic_sort_lanes) segment s of 64 consecutive cells owns
// the slots [off[64 s], off[min(64 s + 64, ncells)]) until the next sort, and
// its particles stay within a few cells of the segment's home cells.  Each
// block takes a contiguous slice of one segment and:
//   1. stages the field window (home cells + stencil halo + drift margin,
//      periodic) into LDS and clears the charge window;
//   2. streams its particles (coalesced, 64 distinct cells per wave in the
//      lane-spread order): gather from the LDS window, kick, drift, deposit
//      into the LDS window;
//   3. flushes the non-zero charge nodes with FP64 global atomics, once.
// No block-wide synchronisation inside the particle loop.  A particle whose
// stencil leaves the window (drifted further than the margin since the sort)
// is appended to a deferred list instead, whole or for its deposit only; a
// small kernel launched after the pass completes those from global memory
// with the same arithmetic and counts them.  Keeping the global-memory path
// out of the pass kernel keeps its register count, and so its occupancy, at
// what the LDS path needs.
//
// Arithmetic per particle is the same as in pic_step.hip's kernels (FP64:
// the reference's operations; FP32: the cell-relative step), so results are
// bit-identical to them up to the order of the charge sums.
//
// FP32 charge accumulation, three modes (Ctx::f32_acc):
//  2 fixed-point int32 LDS tiles: every contribution is rounded to a multiple
//    of a power-of-two quantum u (per block, 2^-30 of the largest sum a node
//    of the block can reach) and summed exactly with ds_add_u32; the flush
//    converts to FP64.  Each particle adds exactly R = round(w/u) units: the
//    node of its own cell takes R minus the rounded shares of the others, and
//    the flush scales units by w/R, so a block deposits exactly its count
//    times w (see charge32_fixed).  The passes that drift before they
//    deposit keep 2 or 4 copies of the tile (Ctx::f32_tiles), lane l adding
//    into copy l mod copies: a ds_add_u32 whose lanes address one word
//    serialises
//    them, and once particles drift neighbouring lanes often sit in one
//    cell.  Measured on MI355X with the 6x6 stencil's address pattern, a
//    64-lane ds_add_u32 takes ~4.1 clocks per CU with the lanes in 64
//    adjacent cells, ~8.3 after a drift of half a cell (sigma), ~6.3 with
//    two copies, ~5.6 with four; bank conflicts alone (rows apart) cost
//    little (~4.9).
//  1 FP64 LDS tiles: every FP32 contribution is added exactly converted to
//    FP64 (ds_add_f64);
//  0 compensated (hi, lo) FP32 tile pairs.  Every contribution q is split as
//    q_hi = (q + C) - C, q_lo = q - q_hi, with C = 1.5 * 2^23 * u for a
//    power-of-two quantum u: q_hi is a multiple of u, so the hi tile sums
//    exactly while below 2^24 u (u is chosen from the particle count per
//    block), and q_lo, below u/2, carries the rest.  Both adds are
//    non-returning LDS atomics (ds_add_f32).  hi + lo is flushed in FP64.
// Measured on MI355X, ds_add_f32 retires about 0.33 lanes per clock per CU
// against 7.3 for ds_add_f64 and 13 for ds_add_u32.
#include <climits>
#include <cmath>

#include "pic_internal.h"
#include "pic_push32.h"
#include "pic_push64.h"
#include "pic_segcore.h"

namespace pic {

// ============================================================ FP32 state

struct SegF32 {
    int32_t *cell;  // state written by the pass (slot p)
    float4 *ov;     // (ox, oy, vx, vy)
    // state read by the pass: the same arrays, or, right after a lane sort
    // whose permutation has not been applied, the unsorted arrays read at
    // perm[p] (the pass then moves every particle into its sorted slot)
    const int32_t *cell_in;
    const float4 *ov_in;
    const uint32_t *perm;
    float w;
    const float2 *e32;
    const float *bz;
    float bz0;
    int nx, ny, lognx, ncells;
    float hdt, dtdx, dtdy;
    double *acc;
    const int64_t *off;
    int bps, margin, cap;
    unsigned long long *slow;
    uint32_t *dlist;  // deferred particles and their count
    unsigned *dcount;
    int tiles;  // fixed-point charge tile copies (1, 2 or 4, see k_seg_f32)
};

// C = 1.5 * 2^23 * u of the (hi, lo) split for a block of `count`
// particles: the hi tile of a node sums at most count * max(w * w_x * w_y)
// <= 0.31 w count, which stays below 2^24 u for the power-of-two quantum
// u = 2^(ceil(log2(0.31 w (count + 5))) - 24).
__device__ __forceinline__ float split_constant(float w, int64_t count) {
    const double bound = 0.31 * fabs((double)w) * (double)(count + 5) + 1e-300;
    int ex;
    const double m = frexp(bound, &ex);  // bound = m 2^ex, m in [0.5, 1)
    const int e = (m == 0.5 ? ex - 1 : ex) - 24;
    return (float)ldexp(1.5, e + 23);
}

__device__ __forceinline__ int64_t tile_sum(const int *iacc, int t, int stride, int tiles) {
    int64_t q = 0;
    for (int k = 0; k < tiles; ++k) q += iacc[t + k * stride];
    return q;
}

// The same shares two nodes at a time: add(u, v, q) for u even with the
// 64-bit q = n(u, v) + 2^32 n(u + 1, v), one ds_add_u64 into a packed tile
// whose 8-byte words hold the int32 sums of nodes (u, u + 1) (on MI355X one
// 64-lane ds_add_u64 takes 6.8 clocks per CU against 5.3 for a ds_add_u32,
// both with lanes on distinct words).  The low halves are never negative
// (|w| units; the remainder is a high half), so the word's 64-bit sum
// carries nothing between its halves that the flush cannot undo: with L and
// H the sums of the low and high halves, both below 2^31 in magnitude, the
// word holds L + 2^32 H modulo 2^64.
//
// The share n of a node is read from the bits of fma(a, b, C), which are
// bits(C) + n; the two raw words of a pair are packed first and bits(C) taken
// off both halves by one 64-bit subtraction (a single v_lshl_add_u64 on
// MI355X, where two 32-bit ones cost two instructions).  The low half never
// borrows (n >= 0), so each half is exactly its share, as with two 32-bit
// subtractions.  The conserving remainder is formed from the raw sum the
// same way: sum of the raw words minus bits(C) per share.
template <int K, class F>
__device__ __forceinline__ void charge32_packed(float fx, float fy, int ci, int cj, float wu, int R,
                                                bool conserve, F &&add) {
    constexpr int S = Shape<K>::S, HL = Shape<K>::HL, UR = HL | 1, VR = HL;
    constexpr float C = 12582912.0f;  // 1.5 * 2^23
    constexpr unsigned CB = 0x4B400000u;  // bits of C
    constexpr unsigned long long CB2 = (unsigned long long)CB << 32 | CB;
    auto n = [&](float a, float b) { return __float_as_uint(__builtin_fmaf(a, b, C)); };
    auto pack = [](unsigned lo, unsigned hi) {
        return ((unsigned long long)lo | (unsigned long long)hi << 32) - CB2;
    };
    float wx[6], wy[6];
    if (K == PIC_CIC) {
        wx[0] = 1.0f - fx;
        wx[1] = fx;
        wy[0] = (1.0f - fy) * wu;
        wy[1] = fy * wu;
    } else {
        deposit_w<K>(fx, ci, wx);
        deposit_w<K>(fy, cj, wy);
#pragma unroll
        for (int v = 0; v < S; ++v) wy[v] = wy[v] * wu;
    }
    unsigned sr = 0;  // raw words of the S * S - 1 shares before the last
#pragma unroll
    for (int v = 0; v < S; ++v)
#pragma unroll
        for (int u = 0; u < S; u += 2) {
            const unsigned a = n(wx[u], wy[v]);
            sr += a;
            if (u + 1 == UR && v == VR) continue;
            const unsigned b = n(wx[u + 1], wy[v]);
            sr += b;
            add(u, v, pack(a, b));
        }
    const int s = (int)(sr - (unsigned)(S * S - 1) * CB);
    const unsigned a = n(wx[UR - 1], wy[VR]);
    add(UR - 1, VR, pack(a, conserve ? (unsigned)(R - s) + CB : n(wx[UR], wy[VR])));
}

// Signed int32 sum of the half at position pos of a packed tile.
__device__ __forceinline__ int64_t packed_half(const unsigned long long *tile, int pos) {
    const unsigned long long w = tile[pos >> 1];
    const int lo = (int)(unsigned)w;
    if (!(pos & 1)) return lo;
    return (int)(unsigned)((w - (unsigned long long)(long long)lo) >> 32);
}

// The particles of one segment slice.  A particle whose stencil lies outside
// the block's LDS window is not touched here: it goes to the deferred list
// whole (gather outside) or for its deposit only (drifted outside), and
// k_defer_f32 completes it from global memory after this kernel.
//
// Register budget: the passes with 40 KiB windows are held to 64 VGPRs
// (eight waves per SIMD), so that four blocks share a CU instead of the three
// that the 66 registers the compiler picks unbounded allow (the M6 pass with
// one charge tile ran ~1.5 % faster so).  The multi-tile fixed-point drift-
// and-deposit passes are bounded at four waves per SIMD (seg32_waves below).
// Shapes whose fixed-point drift-and-deposit passes keep several copies of
// the charge tile (packed pairs of int32 sums, 160 KiB windows, 1024-thread
// blocks): every shape.  Measured on MI355X with four copies: M6 +4 %, TSC
// +3 %, M4 +5 % on 1e7 particles, 128^2; NGP, whose particles put all their
// charge on one node and pile onto shared words as they drift, 140 -> 97 us
// per pass in the loop (1e7 particles, 128^2, a sort every 8 steps); CIC
// 962 -> 910 us (1e8 particles, 256^2, bz = 20).
template <int K>
constexpr bool multi_tile_shape() {
    return K == PIC_M6 || K == PIC_TSC || K == PIC_M4 || K == PIC_NGP || K == PIC_CIC;
}
inline bool multi_tile_shape(int kernel) {
    return valid_shape(kernel);
}

// The multi-tile fixed-point drift-and-deposit passes run one 1024-thread
// block per CU with the default 160 KiB window, four waves per SIMD, so they
// are bounded at four waves per SIMD (128 VGPRs) and load each thread's next
// particle while processing the current one.  Measured on MI355X (Landau M6,
// 1e7 particles, 128^2): 166.7 us per pass bounded at eight waves per SIMD,
// 164.2 at four, 161.2 at four with the early load (at eight with the early
// load the pass spilled: 165.7).
template <int K, bool DRIFT, bool DEPOSIT, int ACC>
constexpr bool seg32_wide() {
    return multi_tile_shape<K>() && ACC >= 2 && DRIFT && DEPOSIT;
}
template <int K, bool DRIFT, bool DEPOSIT, int ACC, bool PERM>
constexpr int seg32_waves() {
    return seg32_wide<K, DRIFT, DEPOSIT, ACC>() ? 4 : 8;
}

// Block size: the M6 fixed-point drift-and-deposit passes run 1024 threads
// per block, so that their two 80 KiB blocks per CU hold 32 waves instead of
// 16 (measured with two tile copies: 1.5 % faster at 1e7 particles on 128^2,
// 4 % at 5e8 on 512^2; the pass after a sort, at 74 VGPRs, then runs one
// such block per CU); the others SEG_T.
template <int K, bool DRIFT, bool DEPOSIT, int ACC>
constexpr int seg32_threads() {
    return multi_tile_shape<K>() && ACC >= 2 && DRIFT && DEPOSIT ? 1024 : SEG_T;
}

template <int K, bool GATHER, bool DRIFT, bool DEPOSIT, int BZ, int ACC, bool PERM>
__global__ __launch_bounds__((seg32_threads<K, DRIFT, DEPOSIT, ACC>()),
                             (seg32_waves<K, DRIFT, DEPOSIT, ACC, PERM>())) void k_seg_f32(SegF32 a) {
    constexpr int BT = seg32_threads<K, DRIFT, DEPOSIT, ACC>();
    using St = SegStencil<K>;
    constexpr int S = St::S, HL = St::HL;
    extern __shared__ __align__(16) char smem_s32[];
    const int tid = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const int mx = nx - 1, my = ny - 1;  // nx, ny are powers of two
    // XCD-aware block order (the pass ran ~1 % faster on the M6 workload,
    // the pass that applies a sort's permutation ~7 %)
    const int lb = seg_block_order();
    const int s = lb / a.bps, bi = lb - s * a.bps;
    const int ce = min(64 * s + 64, a.ncells);
    const int64_t s0 = a.off[64 * s], s1 = a.off[ce];
    const int64_t len = s1 - s0;
    const int64_t p0 = s0 + len * bi / a.bps, p1 = s0 + len * (bi + 1) / a.bps;
    // The field window is held twice, as float2 (ex, ey) rows of pitch P
    // (even): copy A with node u at position u, copy B with node u at u + 1.
    // A particle whose stencil starts at an even column reads A, one that
    // starts at an odd column B, both at an even position, so every pair of
    // nodes (u, u + 1) it needs is one aligned 16-byte read (interp32_pairs).
    // ACC 3, the packed fixed-point tiles, keeps its charge tiles in the same
    // two shifted layouts as the field (tile pairs A, B): pairs of stencil
    // nodes are aligned 8-byte words in one of them.
    constexpr bool FIX = ACC == 2 || ACC == 3, PACK = ACC == 3 && DEPOSIT;
    constexpr int PAD = GATHER || PACK ? 2 : 0;
    const Win W = seg_window<S, HL>(s, nx, ny, a.ncells, a.margin, a.cap, PAD);
    const bool fits = (W.tw + PAD) * W.th <= a.cap;
    const int cap = a.cap;
    const int P = GATHER || PACK ? (W.tw + 2) & ~1 : W.tw;
    // window position of a stencil whose reference cell is (ci, cj):
    // ((ci - ox) & mx, (cj - oy) & my); the products of window rows below are
    // 24-bit (full-rate v_mul_u32_u24: a window holds fewer than 2^24 nodes)
    const int ox = W.x0 + HL, oy = W.y0 + HL;
    // ACC 0: the (hi, lo) split constant; ACC 2, 3: the fixed-point quantum
    // 2^qe and |w| in units of it (exact: a power-of-two scaling)
    const float cq = ACC == 0 ? split_constant(a.w, p1 - p0) : 0.0f;
    const int qe = FIX ? fixed_exponent(a.w, Shape<K>::QMAX, p1 - p0) : 0;
    const float wu = FIX ? ldexpf(fabsf(a.w), -qe) : 0.0f;
    // charge-conserving fixed point: R units per particle, worth w/R each
    const bool conserve = FIX && p1 - p0 <= CONSERVE_MAX_COUNT;
    const int R = conserve ? (int)rintf(wu) : 0;
    const double unit = !FIX                      ? 0.0
                        : (conserve && R != 0) ? (double)a.w / R
                                               : copysign(ldexp(1.0, qe), (double)a.w);
    float2 *teA = reinterpret_cast<float2 *>(smem_s32);
    float2 *teB = teA + cap;  // cap is a multiple of 16: 16-byte aligned
    float *tb = reinterpret_cast<float *>(smem_s32 + 16 * (size_t)cap);
    char *dep = smem_s32 + (GATHER ? (BZ == PIC_BZ_GRID ? 20 : 16) * (size_t)cap : 0);
    float *hi = reinterpret_cast<float *>(dep), *lo = hi + cap;
    double *dacc = reinterpret_cast<double *>(dep);
    int *iacc = reinterpret_cast<int *>(dep);
    // fixed-point tiles: lane l adds into copy l mod a.tiles,
    // iacc[k cap, (k + 1) cap), so that neighbouring lanes whose particles
    // drifted into one cell do not add to one LDS word in the same
    // instruction
    int *itile = iacc + (tid & (a.tiles - 1)) * cap;
    if (PACK && fits) {
        const int np = P * W.th;
        for (int t = tid; t < np; t += BT)
            for (int k = 0; k < a.tiles; ++k) iacc[t + k * cap] = 0;
    }
    if (fits) {
        const int nt = W.tw * W.th;
        for (int t = tid; t < nt; t += BT) {
            const int u = t % W.tw, v = t / W.tw;
            const int64_t g = (int64_t)((W.x0 + u) & mx) + (int64_t)nx * ((W.y0 + v) & my);
            if (GATHER) {
                const float2 ev = a.e32[g];
                teA[u + P * v] = ev;
                teB[u + 1 + P * v] = ev;
                if (BZ == PIC_BZ_GRID) tb[t] = a.bz[g];
            }
            if (DEPOSIT) {
                if (ACC == 0) {
                    hi[t] = 0.0f;
                    lo[t] = 0.0f;
                } else if (ACC == 1) {
                    dacc[t] = 0.0;
                } else if (ACC == 2) {
                    iacc[t] = 0;
                    for (int k = 1; k < a.tiles; ++k) iacc[t + k * cap] = 0;
                }
            }
        }
    }
    __syncthreads();

    // the next particle's state is loaded while this one is processed
    int c_next = 0;
    float4 o_next = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    constexpr bool PREFETCH = seg32_wide<K, DRIFT, DEPOSIT, ACC>();
    if (PREFETCH && (uint32_t)p0 + tid < (uint32_t)p1) {
        const uint32_t src = PERM ? a.perm[(uint32_t)p0 + tid] : (uint32_t)p0 + tid;
        c_next = a.cell_in[src];
        o_next = a.ov_in[src];
    }
    for (uint32_t p = (uint32_t)p0 + tid; p < (uint32_t)p1; p += BT) {
        int c;
        float4 o;
        if (PREFETCH) {
            c = c_next;
            o = o_next;
            // unconditional: past the end of the slice it re-reads the last
            // particle (p < p1 here), so no copy of the previous value is
            // kept for a skipped load (five register moves per particle)
            const uint32_t q = min(p + (uint32_t)BT, (uint32_t)p1 - 1u);
            const uint32_t src = PERM ? a.perm[q] : q;
            c_next = a.cell_in[src];
            o_next = a.ov_in[src];
        } else {
            const uint32_t src = PERM ? a.perm[p] : p;
            c = a.cell_in[src];
            o = a.ov_in[src];
        }
        int ci = c & mx, cj = c >> a.lognx;
        float fx = o.x, fy = o.y, ux = o.z, uy = o.w;
        if (GATHER) {
            const int lx = (ci - ox) & mx, ly = (cj - oy) & my;
            const bool in = fits && lx + S - 1 < W.tw && ly + S - 1 < W.th;
            defer_push(!in, (uint32_t)p, a.dlist, a.dcount);
            if (!in) continue;
            float e1, e2, b3 = 0.0f;
            const int odd = lx & 1;
            const float4 *row =
                reinterpret_cast<const float4 *>((odd ? teB : teA) + (lx + odd) + __umul24(P, ly));
            const int P2 = P >> 1;
            interp32_pairs<K, BZ>(
                fx, fy, ci, cj, [&](int u, int v) { return row[(u >> 1) + P2 * v]; },
                [&](int u, int v) { return tb[(lx + u) + W.tw * (ly + v)]; }, e1, e2, b3);
            kick32<BZ>(a.hdt, a.bz0, e1, e2, b3, ux, uy);
        }
        if (DRIFT) {
            float nfx, nfy;
            const int si = drift1s(fx, ux, a.dtdx, nfx);
            const int sj = drift1s(fy, uy, a.dtdy, nfy);
            ci = (ci + si) & mx;
            cj = (cj + sj) & my;
            fx = nfx;
            fy = nfy;
        }
        if (DRIFT || PERM) a.cell[p] = ci + (cj << a.lognx);
        if (GATHER || DRIFT || PERM) a.ov[p] = make_float4(fx, fy, ux, uy);
        if (DEPOSIT) {
            const int lx = (ci - ox) & mx, ly = (cj - oy) & my;
            const bool in = fits && lx + S - 1 < W.tw && ly + S - 1 < W.th;
            defer_push(!in, (uint32_t)p | DEFER_DEPOSIT_ONLY, a.dlist, a.dcount);
            if (!in) continue;
            if (ACC == 2) {
                charge32_fixed<K>(fx, fy, ci, cj, wu, R, conserve, [&](int u, int v, int qi) {
                    __hip_atomic_fetch_add(&itile[(lx + u) + W.tw * (ly + v)], qi,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                });
                continue;
            }
            if (ACC == 3) {
                // tile pair (A, B) of lane copy tid mod (tiles / 2); B for a
                // stencil that starts at an odd column
                const int odd = lx & 1;
                // word (pair) index of stencil node (0, 0): (lx + odd + P ly) / 2,
                // P even; node (u, v), u even, is word w0 + u / 2 + v P / 2
                unsigned long long *w0 =
                    reinterpret_cast<unsigned long long *>(
                        iacc + __umul24((a.tiles >= 4 ? 2 * (tid & 1) : 0) + odd, cap)) +
                    ((lx + odd) >> 1) + __umul24(P >> 1, ly);
                const int P2 = P >> 1;
                charge32_packed<K>(fx, fy, ci, cj, wu, R, conserve,
                                   [&](int u, int v, unsigned long long q) {
                                       __hip_atomic_fetch_add(&w0[(u >> 1) + P2 * v], q,
                                                              __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_WORKGROUP);
                                   });
                continue;
            }
            charge32<K>(fx, fy, ci, cj, a.w, [&](int u, int v, float q) {
                const int t = (lx + u) + W.tw * (ly + v);
                if (ACC == 0) {
                    const float qh = (q + cq) - cq;
                    const float ql = q - qh;
                    __hip_atomic_fetch_add(&hi[t], qh, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(&lo[t], ql, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    __hip_atomic_fetch_add(&dacc[t], (double)q, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            });
        }
    }
    if (!DEPOSIT || !fits) return;
    __syncthreads();
    const int nt = W.tw * W.th;
    for (int t = tid; t < nt; t += BT) {
        const int u = t % W.tw, v = t / W.tw;
        double q;
        if (ACC == 0) {
            q = (double)hi[t] + (double)lo[t];
        } else if (ACC == 1) {
            q = dacc[t];
        } else if (ACC == 2) {
            q = (double)tile_sum(iacc, t, cap, a.tiles) * unit;
        } else {
            // node u of the window: position u of tile A, u + 1 of tile B
            int64_t qi = 0;
            for (int k = 0; k + 1 < a.tiles; k += 2) {
                const auto *A = reinterpret_cast<const unsigned long long *>(iacc + k * cap);
                const auto *B = reinterpret_cast<const unsigned long long *>(iacc + (k + 1) * cap);
                qi += packed_half(A, u + P * v) + packed_half(B, u + 1 + P * v);
            }
            q = (double)qi * unit;
        }
        if (q != 0.0) {
            unsafeAtomicAdd(&a.acc[(int64_t)((W.x0 + u) & mx) + (int64_t)nx * ((W.y0 + v) & my)],
                            q);
        }
    }
}

// Completes the particles k_seg_f32 deferred, from global memory (same
// arithmetic), adds their number to the slow-path counter, and the last
// block to finish resets the list for the next pass.
template <int K, bool GATHER, bool DRIFT, bool DEPOSIT, int BZ, bool PERM>
__global__ __launch_bounds__(DEFER_T) void k_defer_f32(SegF32 a) {
    constexpr int HL = SegStencil<K>::HL;
    const unsigned cnt = __hip_atomic_load(a.dcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nx = a.nx, mx = a.nx - 1, my = a.ny - 1;
    for (unsigned i = blockIdx.x * DEFER_T + threadIdx.x; i < cnt; i += gridDim.x * DEFER_T) {
        const uint32_t e = a.dlist[i];
        const int64_t p = e & ~DEFER_DEPOSIT_ONLY;
        const bool whole = (e & DEFER_DEPOSIT_ONLY) == 0;
        // a whole particle has not been touched: read it where the pass
        // would have; a deposit-only one is already in its slot
        const uint32_t src = (PERM && whole) ? a.perm[p] : (uint32_t)p;
        const int c = whole ? a.cell_in[src] : a.cell[p];
        int ci = c & mx, cj = c >> a.lognx;
        const float4 o = whole ? a.ov_in[src] : a.ov[p];
        float fx = o.x, fy = o.y;
        if (whole) {
            float ux = o.z, uy = o.w;
            if (GATHER) {
                float e1, e2, b3 = 0.0f;
                interp32<K, BZ>(
                    fx, fy, ci, cj,
                    [&](int u, int v) {
                        return a.e32[(int64_t)((ci - HL + u) & mx) +
                                     (int64_t)nx * ((cj - HL + v) & my)];
                    },
                    [&](int u, int v) {
                        return a.bz[(int64_t)((ci - HL + u) & mx) +
                                    (int64_t)nx * ((cj - HL + v) & my)];
                    },
                    e1, e2, b3);
                kick32<BZ>(a.hdt, a.bz0, e1, e2, b3, ux, uy);
            }
            if (DRIFT) {
                float nfx, nfy;
                const int si = drift1s(fx, ux, a.dtdx, nfx);
                const int sj = drift1s(fy, uy, a.dtdy, nfy);
                ci = (ci + si) & mx;
                cj = (cj + sj) & my;
                fx = nfx;
                fy = nfy;
            }
            if (DRIFT || PERM) a.cell[p] = ci + nx * cj;
            if (GATHER || DRIFT || PERM) a.ov[p] = make_float4(fx, fy, ux, uy);
        }
        if (DEPOSIT)
            charge32<K>(fx, fy, ci, cj, a.w, [&](int u, int v, float q) {
                unsafeAtomicAdd(
                    &a.acc[(int64_t)((ci - HL + u) & mx) + (int64_t)nx * ((cj - HL + v) & my)],
                    (double)q);
            });
    }
    defer_finish(cnt, a.dcount, a.slow);
}

// ============================================================ FP64 state

// FP64 counterpart of k_seg_f32; out-of-window particles go to k_defer_f64.
// The FP64 passes run one 1024-thread block per CU with a 160 KiB window
// (16 B of field per node, 8 B more with a bz grid, and two copies of the
// FP64 charge tile, 16 B): an M6 window of a 64-cell row segment keeps a drift
// margin of 20 cells, so the state can be sorted as seldom as the FP32 state.
// Measured on MI355X, FP64 state, against 512-thread blocks in 40 KiB windows
// (a 7-cell margin, one tile): Landau M6, 1e7 particles on 128^2, 1.86e10
// particle-steps/s sorting every 4 steps -> 2.19e10 every 12 (2.09e10 with
// one tile copy); 5e8 particles on 512^2, 1.12e10 sorting every step ->
// 1.81e10 every 4.
constexpr int SEG64_T = 1024;
constexpr size_t SEG64_LDS_BYTES = 160 * 1024;

template <int K, bool GATHER, int NDRIFT, bool DEPOSIT, int BZ, bool PERM>
__global__ __launch_bounds__(SEG64_T) void k_seg_f64(SegF64 a) {
    using St = SegStencil<K>;
    constexpr int S = St::S, HL = St::HL;
    extern __shared__ __align__(16) char smem_s64[];
    const int tid = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const int s = blockIdx.x / a.bps, bi = blockIdx.x - s * a.bps;
    const int ce = min(64 * s + 64, a.ncells);
    const int64_t s0 = a.off[64 * s], s1 = a.off[ce];
    const int64_t len = s1 - s0;
    const int64_t p0 = s0 + len * bi / a.bps, p1 = s0 + len * (bi + 1) / a.bps;
    const Win W = seg_window<S, HL>(s, nx, ny, a.ncells, a.margin, a.cap);
    const bool fits = W.tw * W.th <= a.cap;
    const int cap = a.cap;
    double2 *te = reinterpret_cast<double2 *>(smem_s64);
    double *tb = reinterpret_cast<double *>(smem_s64 + 16 * (size_t)cap);
    // two copies of the charge tile, lane l adding into copy l mod 2
    double *dacc = reinterpret_cast<double *>(
        smem_s64 + (GATHER ? (BZ == PIC_BZ_GRID ? 24 : 16) * (size_t)cap : 0));
    double *dtile = dacc + (tid & 1) * cap;
    if (fits) {
        const int nt = W.tw * W.th;
        for (int t = tid; t < nt; t += SEG64_T) {
            const int u = t % W.tw, v = t / W.tw;
            const int64_t g = (int64_t)picd::wrap2(W.x0 + u, nx) +
                              (int64_t)nx * picd::wrap2(W.y0 + v, ny);
            if (GATHER) {
                te[t] = make_double2(a.ex[g], a.ey[g]);
                if (BZ == PIC_BZ_GRID) tb[t] = a.bz[g];
            }
            if (DEPOSIT) {
                dacc[t] = 0.0;
                dacc[t + cap] = 0.0;
            }
        }
    }
    __syncthreads();

    for (uint32_t p = (uint32_t)p0 + tid; p < (uint32_t)p1; p += SEG64_T) {
        const uint32_t src = PERM ? a.perm[p] : p;
        double px = a.x_in[src], py = a.y_in[src];
        double pvx = 0.0, pvy = 0.0;
        if (GATHER) {
            double fx, fy;
            const int ci = ref_cell_s(px, a.dx, nx, false, fx);
            const int cj = ref_cell_s(py, a.dy, ny, false, fy);
            const int lx = picd::wrap2(ci - HL - W.x0, nx), ly = picd::wrap2(cj - HL - W.y0, ny);
            const bool in = fits && lx + S - 1 < W.tw && ly + S - 1 < W.th;
            defer_push(!in, (uint32_t)p, a.dlist, a.dcount);
            if (!in) continue;
            pvx = a.vx_in[src];
            pvy = a.vy_in[src];
            double e1, e2, b3 = 0.0;
            interp64<K, BZ>(
                fx, fy, ci, cj, [&](int u, int v) { return te[(lx + u) + W.tw * (ly + v)]; },
                [&](int u, int v) { return tb[(lx + u) + W.tw * (ly + v)]; }, e1, e2, b3);
            kick64<BZ>(a, e1, e2, b3, pvx, pvy);
        } else if (NDRIFT > 0 || PERM) {
            pvx = a.vx_in[src];
            pvy = a.vy_in[src];
        }
        if (NDRIFT > 0) drift64<NDRIFT>(a, pvx, pvy, px, py);
        if (GATHER || PERM) {
            a.vx[p] = pvx;
            a.vy[p] = pvy;
        }
        if (NDRIFT > 0 || PERM) {
            a.x[p] = px;
            a.y[p] = py;
        }
        if (PERM && a.w) a.w_out[p] = a.w[src];
        if (DEPOSIT) {
            double fx, fy;
            const int ci = ref_cell_s(px, a.dx, nx, K == PIC_CIC, fx);
            const int cj = ref_cell_s(py, a.dy, ny, K == PIC_CIC, fy);
            const int lx = picd::wrap2(ci - HL - W.x0, nx), ly = picd::wrap2(cj - HL - W.y0, ny);
            const bool in = fits && lx + S - 1 < W.tw && ly + S - 1 < W.th;
            defer_push(!in, (uint32_t)p | DEFER_DEPOSIT_ONLY, a.dlist, a.dcount);
            if (!in) continue;
            charge64<K>(fx, fy, ci, cj, a.w ? a.w[src] : a.w0, [&](int u, int v, double q) {
                __hip_atomic_fetch_add(&dtile[(lx + u) + W.tw * (ly + v)], q, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            });
        }
    }
    if (!DEPOSIT || !fits) return;
    __syncthreads();
    const int nt = W.tw * W.th;
    for (int t = tid; t < nt; t += SEG64_T) {
        const double q = dacc[t] + dacc[t + cap];
        if (q != 0.0) {
            const int u = t % W.tw, v = t / W.tw;
            unsafeAtomicAdd(&a.acc[(int64_t)picd::wrap2(W.x0 + u, nx) +
                                   (int64_t)nx * picd::wrap2(W.y0 + v, ny)],
                            q);
        }
    }
}

// Completes the particles k_seg_f64 deferred, from global memory.
template <int K, bool GATHER, int NDRIFT, bool DEPOSIT, int BZ, bool PERM>
__global__ __launch_bounds__(DEFER_T) void k_defer_f64(SegF64 a) {
    constexpr int HL = SegStencil<K>::HL;
    const unsigned cnt = __hip_atomic_load(a.dcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nx = a.nx, ny = a.ny;
    for (unsigned i = blockIdx.x * DEFER_T + threadIdx.x; i < cnt; i += gridDim.x * DEFER_T) {
        const uint32_t e = a.dlist[i];
        const int64_t p = e & ~DEFER_DEPOSIT_ONLY;
        const bool whole = (e & DEFER_DEPOSIT_ONLY) == 0;
        // a whole particle is read where the pass would have read it; a
        // deposit-only one is already in its slot
        const uint32_t src = (PERM && whole) ? a.perm[p] : (uint32_t)p;
        double px = whole ? a.x_in[src] : a.x[p], py = whole ? a.y_in[src] : a.y[p];
        const double wgt = !a.w ? a.w0 : whole ? a.w[src] : PERM ? a.w_out[p] : a.w[p];
        if (whole) {
            double pvx = 0.0, pvy = 0.0;
            if (GATHER || NDRIFT > 0 || PERM) {
                pvx = a.vx_in[src];
                pvy = a.vy_in[src];
            }
            if (GATHER) {
                double fx, fy;
                const int ci = ref_cell_s(px, a.dx, nx, false, fx);
                const int cj = ref_cell_s(py, a.dy, ny, false, fy);
                double e1, e2, b3 = 0.0;
                interp64<K, BZ>(
                    fx, fy, ci, cj,
                    [&](int u, int v) {
                        const int64_t g = (int64_t)picd::wrap2(ci - HL + u, nx) +
                                          (int64_t)nx * picd::wrap2(cj - HL + v, ny);
                        return make_double2(a.ex[g], a.ey[g]);
                    },
                    [&](int u, int v) {
                        return a.bz[(int64_t)picd::wrap2(ci - HL + u, nx) +
                                    (int64_t)nx * picd::wrap2(cj - HL + v, ny)];
                    },
                    e1, e2, b3);
                kick64<BZ>(a, e1, e2, b3, pvx, pvy);
            }
            if (NDRIFT > 0) drift64<NDRIFT>(a, pvx, pvy, px, py);
            if (GATHER || PERM) {
                a.vx[p] = pvx;
                a.vy[p] = pvy;
            }
            if (NDRIFT > 0 || PERM) {
                a.x[p] = px;
                a.y[p] = py;
            }
            if (PERM && a.w) a.w_out[p] = a.w[src];
        }
        if (DEPOSIT) {
            double fx, fy;
            const int ci = ref_cell_s(px, a.dx, nx, K == PIC_CIC, fx);
            const int cj = ref_cell_s(py, a.dy, ny, K == PIC_CIC, fy);
            charge64<K>(fx, fy, ci, cj, wgt, [&](int u, int v, double q) {
                unsafeAtomicAdd(&a.acc[(int64_t)picd::wrap2(ci - HL + u, nx) +
                                       (int64_t)nx * picd::wrap2(cj - HL + v, ny)],
                                q);
            });
        }
    }
    defer_finish(cnt, a.dcount, a.slow);
}

// ============================================================== dispatch

// Launches the pass kernel `k` over every segment, then its deferred-list
// kernel `kd`.
template <class Kern, class Defer>
static void seg32_launch(const Ctx *c, SegF32 a, Kern k, Defer kd, size_t sm, int bt) {
    // windows above 64 KiB need the kernel's dynamic LDS limit raised; a
    // failure here fails the launch, which hipGetLastError reports
    if (sm > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    a.bps = seg_bps(c, reinterpret_cast<const void *>(k), sm, bt);
    dim3 grid((unsigned)(((c->nx * c->ny + 63) / 64) * a.bps));
    hipLaunchKernelGGL(k, grid, dim3(bt), sm, c->stream, a);
    hipLaunchKernelGGL(kd, dim3(DEFER_BLOCKS), dim3(DEFER_T), 0, c->stream, a);
}

template <int K, bool G, bool DR, bool D, int ACC, bool P>
static void seg32_bz(const Ctx *c, const SegF32 &a, int bz, size_t sm) {
    constexpr int bt = seg32_threads<K, DR, D, ACC>();
    switch (bz) {
        case PIC_BZ_NONE:
            seg32_launch(c, a, k_seg_f32<K, G, DR, D, PIC_BZ_NONE, ACC, P>,
                         k_defer_f32<K, G, DR, D, PIC_BZ_NONE, P>, sm, bt);
            break;
        case PIC_BZ_UNIFORM:
            seg32_launch(c, a, k_seg_f32<K, G, DR, D, PIC_BZ_UNIFORM, ACC, P>,
                         k_defer_f32<K, G, DR, D, PIC_BZ_UNIFORM, P>, sm, bt);
            break;
        default:
            seg32_launch(c, a, k_seg_f32<K, G, DR, D, PIC_BZ_GRID, ACC, P>,
                         k_defer_f32<K, G, DR, D, PIC_BZ_GRID, P>, sm, bt);
    }
}

template <int K, int ACC, bool P>
static int seg32_mode(const Ctx *c, const SegF32 &a, int mode, int bz, size_t sm) {
    switch (mode) {
        case 0: seg32_bz<K, false, true, true, ACC, P>(c, a, PIC_BZ_NONE, sm); break;
        case 1: seg32_bz<K, true, true, true, ACC, P>(c, a, bz, sm); break;
        case 2: seg32_bz<K, true, true, false, ACC, P>(c, a, bz, sm); break;
        case 3: seg32_bz<K, false, false, true, ACC, P>(c, a, PIC_BZ_NONE, sm); break;
        default: return fail(PIC_ERR_ARG, "pass mode %d", mode);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

template <int K, int ACC>
static int seg32_perm(const Ctx *c, const SegF32 &a, int mode, int bz, size_t sm) {
    return a.perm ? seg32_mode<K, ACC, true>(c, a, mode, bz, sm)
                  : seg32_mode<K, ACC, false>(c, a, mode, bz, sm);
}

// Fixed-point passes: the packed tiles (ACC 3) for the shapes that keep
// several tile copies, in the passes that drift and deposit; else ACC 2.
template <int K>
static int seg32_fixed(const Ctx *c, const SegF32 &a, int mode, int bz, size_t sm, bool packed) {
    if constexpr (multi_tile_shape<K>()) {
        if (packed && (mode == 0 || mode == 1)) {
            if (a.perm) {
                if (mode == 0) seg32_bz<K, false, true, true, 3, true>(c, a, PIC_BZ_NONE, sm);
                else seg32_bz<K, true, true, true, 3, true>(c, a, bz, sm);
            } else {
                if (mode == 0) seg32_bz<K, false, true, true, 3, false>(c, a, PIC_BZ_NONE, sm);
                else seg32_bz<K, true, true, true, 3, false>(c, a, bz, sm);
            }
            PIC_HIP(hipGetLastError());
            return PIC_OK;
        }
    }
    return seg32_perm<K, 2>(c, a, mode, bz, sm);
}

// One segment-persistent pass in FP32 state.  The particles are in the lane
// order of the last sort_by_cell(..., lanes = true), or that sort left its
// permutation pending (c->pending_perm): then this pass reads every particle
// through it, writes it into its sorted slot of the alternate arrays, and
// swaps those in.
int launch_seg32(Ctx *c, int kernel, int mode, double dt, double dtd) {
    if (c->n == 0) return PIC_OK;
    if (mode != 3) c->sort_fresh = false;  // the pass moves the particles
    PIC_CHECK(c->seg_valid, PIC_ERR_STATE, "segment passes need a lane sort first");
    SegF32 a;
    a.cell = c->cell; a.ov = c->ov32;
    a.cell_in = c->cell; a.ov_in = c->ov32;
    a.perm = c->pending_perm;
    if (a.perm) PIC_TRY(alt_f32(c, &a.cell, &a.ov));
    a.w = (float)c->w0;
    a.e32 = reinterpret_cast<const float2 *>(c->e32);
    a.bz = c->bz32;
    a.bz0 = (float)c->bz0;
    a.nx = c->nx; a.ny = c->ny;
    a.lognx = 0;
    while ((1 << a.lognx) < c->nx) ++a.lognx;
    a.ncells = c->nx * c->ny;
    a.hdt = (float)(0.5 * dt);
    a.dtdx = (float)(dtd / c->dx);
    a.dtdy = (float)(dtd / c->dy);
    a.acc = c->rho_acc;
    a.off = c->cell_off;
    a.bps = 1;
    a.margin = 32;
    const bool gather = mode == 1 || mode == 2;
    const int acc = c->f32_acc;
    // copies of the fixed-point tile (pic_set_f32_tiles, default 4) for the
    // passes that drift and then deposit.  Measured with two: the M6 fused
    // pass 216 -> 208 us (1e7 particles, 128^2) and 12.8 -> 11.6 ms (5e8,
    // 512^2); four: 200 -> 190 us at 1e7; the deposit-only pass ~10 % slower
    // with two (its particles are in their sorted lanes after a sort).  CIC
    // and NGP were 4-8 % slower with two copies in 53 KiB windows of 512-thread
    // blocks (on freshly sorted particles); with four packed copies, 160 KiB
    // windows and 1024-thread blocks they gain 5 % (CIC) and 30 % (NGP) in
    // the loop, as TSC and M4 gain 3-5 %.
    a.tiles = acc == 2 && multi_tile_shape(kernel) && (mode == 0 || mode == 1) ? c->f32_tiles : 1;
    // field window twice (16 B per node), bz (4 B), charge tiles
    const size_t per_node = (gather ? (c->bz_mode == PIC_BZ_GRID ? 20 : 16) : 0) +
                            (mode == 2 ? 0 : acc == 2 ? 4 * a.tiles : 8);
    // the drift-and-deposit passes with fixed-point tiles run 1024-thread
    // blocks with the context's window (pic_set_f32_window, 160 KiB by
    // default: one block per CU); the others 40 KiB
    const bool wide = acc == 2 && multi_tile_shape(kernel) && (mode == 0 || mode == 1);
    const size_t win = wide ? (size_t)c->f32_window_kib * 1024 : SEG_LDS_BYTES;
    a.cap = (int)(win / (per_node ? per_node : 8)) & ~15;
    a.slow = c->slow_chunks;
    PIC_TRY(ensure_defer(c));
    a.dlist = c->defer;
    a.dcount = c->defer_count;
    const size_t sm = (size_t)a.cap * (per_node ? per_node : 8);
    int rc;
    // two or more tile copies: packed pairs of int32 sums (ds_add_u64)
    const bool packed = acc == 2 && a.tiles >= 2;
    PIC_SHAPE_SWITCH(kernel, rc = acc == 0   ? seg32_perm<KK, 0>(c, a, mode, c->bz_mode, sm)
                                  : acc == 1 ? seg32_perm<KK, 1>(c, a, mode, c->bz_mode, sm)
                                             : seg32_fixed<KK>(c, a, mode, c->bz_mode, sm,
                                                               packed));
    if (rc == PIC_OK && a.perm) {
        swap_alt_f32(c);
        c->pending_perm = nullptr;
    }
    return rc;
}

template <class Kern, class Defer>
static void seg64_launch(const Ctx *c, SegF64 a, Kern k, Defer kd, size_t sm) {
    if (sm > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    a.bps = seg_bps(c, reinterpret_cast<const void *>(k), sm, SEG64_T);
    dim3 grid((unsigned)(((c->nx * c->ny + 63) / 64) * a.bps));
    hipLaunchKernelGGL(k, grid, dim3(SEG64_T), sm, c->stream, a);
    hipLaunchKernelGGL(kd, dim3(DEFER_BLOCKS), dim3(DEFER_T), 0, c->stream, a);
}

template <int K, bool G, int ND, bool D, bool P>
static void seg64_bz(const Ctx *c, const SegF64 &a, int bz, size_t sm) {
    switch (bz) {
        case PIC_BZ_NONE:
            seg64_launch(c, a, k_seg_f64<K, G, ND, D, PIC_BZ_NONE, P>,
                         k_defer_f64<K, G, ND, D, PIC_BZ_NONE, P>, sm);
            break;
        case PIC_BZ_UNIFORM:
            seg64_launch(c, a, k_seg_f64<K, G, ND, D, PIC_BZ_UNIFORM, P>,
                         k_defer_f64<K, G, ND, D, PIC_BZ_UNIFORM, P>, sm);
            break;
        default:
            seg64_launch(c, a, k_seg_f64<K, G, ND, D, PIC_BZ_GRID, P>,
                         k_defer_f64<K, G, ND, D, PIC_BZ_GRID, P>, sm);
    }
}

template <int K, bool P>
static int seg64_mode(const Ctx *c, const SegF64 &a, int mode, int bz, size_t sm) {
    switch (mode) {
        case 0: seg64_bz<K, false, 1, true, P>(c, a, PIC_BZ_NONE, sm); break;
        case 1: seg64_bz<K, true, 2, true, P>(c, a, bz, sm); break;
        case 2: seg64_bz<K, true, 1, false, P>(c, a, bz, sm); break;
        case 3: seg64_bz<K, false, 0, true, P>(c, a, PIC_BZ_NONE, sm); break;
        default: return fail(PIC_ERR_ARG, "pass mode %d", mode);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// The FP64 counterpart of launch_seg32 (same pending-permutation rule).
int launch_seg64(Ctx *c, int kernel, int mode, double dt, double dtd) {
    if (c->n == 0) return PIC_OK;
    if (mode != 3) c->sort_fresh = false;  // the pass moves the particles
    PIC_CHECK(c->seg_valid, PIC_ERR_STATE, "segment passes need a lane sort first");
    SegF64 a;
    a.x = c->x; a.y = c->y; a.vx = c->vx; a.vy = c->vy;
    a.x_in = c->x; a.y_in = c->y; a.vx_in = c->vx; a.vy_in = c->vy;
    a.w = c->uniform_w ? nullptr : c->w;
    a.w_out = nullptr;
    a.perm = c->pending_perm;
    if (a.perm) PIC_TRY(alt_f64(c, &a.x, &a.y, &a.vx, &a.vy, a.w ? &a.w_out : nullptr));
    a.w0 = c->w0;
    a.ex = c->ex; a.ey = c->ey; a.bz = c->bz; a.bz0 = c->bz0;
    a.nx = c->nx; a.ny = c->ny; a.ncells = c->nx * c->ny;
    a.dx = c->dx; a.dy = c->dy; a.dimx = c->dimx; a.dimy = c->dimy;
    a.dt = dt; a.dtd = dtd;
    a.acc = c->rho_acc;
    a.off = c->cell_off;
    a.bps = 1;
    a.margin = 32;
    const bool gather = mode == 1 || mode == 2;
    const size_t per_node =
        (gather ? (c->bz_mode == PIC_BZ_GRID ? 24 : 16) : 0) + (mode == 2 ? 0 : 16);
    a.cap = (int)(SEG64_LDS_BYTES / (per_node ? per_node : 8)) & ~15;
    a.slow = c->slow_chunks;
    PIC_TRY(ensure_defer(c));
    a.dlist = c->defer;
    a.dcount = c->defer_count;
    const size_t sm = (size_t)a.cap * (per_node ? per_node : 8);
    int rc;
    PIC_SHAPE_SWITCH(kernel, rc = a.perm ? seg64_mode<KK, true>(c, a, mode, c->bz_mode, sm)
                                         : seg64_mode<KK, false>(c, a, mode, c->bz_mode, sm));
    if (rc == PIC_OK && a.perm) {
        swap_alt_f64(c);
        c->pending_perm = nullptr;
    }
    return rc;
}

}  // namespace pic
