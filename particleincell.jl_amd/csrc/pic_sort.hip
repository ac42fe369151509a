// pic_sort.hip -- stable sort of the particles by cell.
//
// LSD radix sort of (cell, index) pairs: ceil(B/9) passes over digits of at
// most 9 bits, B = bits of the largest cell index (2 passes up to 512 x 512
// cells).  Every pass is a stable counting sort of one digit:
//   k_digit_hist   per tile of 4096 keys, a histogram of the digit
//                  (LDS counters, one atomic per run of equal digits)
//   k_scan_*       exclusive scan of the histograms in digit-major order,
//                  giving each (digit, tile) its first output slot
//   k_digit_scatter  per tile, each wave ranking its own quarter of the
//                  tile: the rank of each key among the equal digits before it
//                  (wave ballots within a round of 64, a wave-private count
//                  per digit across rounds, one prefix over the waves), which
//                  places it in the tile's digit-ordered copy in LDS; the copy
//                  is then written out, each digit's run to consecutive slots
// The first pass reads the keys where they are (the FP32 state's cell array)
// and takes the item index as the value.
// Equal keys keep their input order in every pass, so the result is the
// permutation std::stable_sort by cell produces.  The permutation is then
// applied to the whole particle state by one gather kernel into a second set
// of arrays whose pointers are swapped in.
#include <algorithm>
#include <climits>
#include <vector>

#include "pic_internal.h"

namespace pic {

constexpr int SORT_T = 256;          // threads per tile
constexpr int SORT_R = 16;           // rounds (items per thread)
constexpr int SORT_TILE = SORT_T * SORT_R;
constexpr int SORT_MAXB = 9;         // widest digit
constexpr int SORT_MAXR = 1 << SORT_MAXB;

// keys from the current state: cell index i + nx*j (FP64 state: the cell of
// floor(x/dx), floor(y/dy), wrapped).
__global__ void k_keys_f64(int64_t n, const double *x, const double *y, double dx, double dy,
                           int nx, int ny, uint32_t *key) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = picd::wrapi((int)floor(x[k] / dx), nx);
    int j = picd::wrapi((int)floor(y[k] / dy), ny);
    key[k] = (uint32_t)(i + nx * j);
}

__global__ __launch_bounds__(SORT_T) void k_digit_hist(const uint32_t *__restrict__ key, int64_t n,
                                                       int shift, int bits, int ntiles,
                                                       uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[SORT_MAXR];
    const int R = 1 << bits;
    const int tid = threadIdx.x;
    for (int d = tid; d < R; d += SORT_T) cnt[d] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    // every key of this thread first, so the loads are in flight together
    uint32_t kr[SORT_R];
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        const int64_t p = base + (int64_t)r * SORT_T + tid;
        kr[r] = p < n ? key[p] : 0u;
    }
    uint32_t cur = 0xffffffffu, run = 0;
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        const int64_t p = base + (int64_t)r * SORT_T + tid;
        if (p >= n) break;
        uint32_t d = (kr[r] >> shift) & (uint32_t)(R - 1);
        if (d == cur) {
            ++run;
        } else {
            if (run) atomicAdd(&cnt[cur], run);
            cur = d;
            run = 1;
        }
    }
    if (run) atomicAdd(&cnt[cur], run);
    __syncthreads();
    for (int d = tid; d < R; d += SORT_T) hist[(int64_t)d * ntiles + blockIdx.x] = cnt[d];
}

// --- exclusive scan of uint32 (3 phases, 2048 items per block)
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_B = SCAN_T * SCAN_I;

// Exclusive scan of v over the SCAN_T threads of the block, and the total:
// an inclusive scan within each wave by lane shuffles, the wave totals
// through LDS (two barriers in all).
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t &total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN_T / 64; ++k) {
        const uint32_t t = sh[k];
        off += k < w ? t : 0u;
        tot += t;
    }
    total = tot;
    __syncthreads();  // sh is reused by the next call
    return off + x - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const uint32_t *in, int64_t n,
                                                        uint32_t *bsum) {
    __shared__ uint32_t sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_B;
    uint32_t s = 0;
    for (int i = 0; i < SCAN_I; ++i) {
        int64_t p = base + (int64_t)threadIdx.x * SCAN_I + i;
        if (p < n) s += in[p];
    }
    uint32_t tot;
    block_excl_scan(s, sh, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb values in place, 16 consecutive
// values per thread (one block scan per 4096 values)
__global__ __launch_bounds__(SCAN_T) void k_scan_top(uint32_t *bsum, int nb) {
    constexpr int PER = 16;
    __shared__ uint32_t sh[SCAN_T];
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nb; b0 += SCAN_T * PER) {
        const int i0 = b0 + threadIdx.x * PER;
        uint32_t v[PER], s = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            v[k] = i0 + k < nb ? bsum[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t e = carry + block_excl_scan(s, sh, tot);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (i0 + k < nb) bsum[i0 + k] = e;
            e += v[k];
        }
        carry += tot;
    }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_down(uint32_t *inout, int64_t n,
                                                      const uint32_t *bsum) {
    __shared__ uint32_t sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_B;
    uint32_t v[SCAN_I], s = 0;
    for (int i = 0; i < SCAN_I; ++i) {
        int64_t p = base + (int64_t)threadIdx.x * SCAN_I + i;
        v[i] = p < n ? inout[p] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t e = block_excl_scan(s, sh, tot) + bsum[blockIdx.x];
    for (int i = 0; i < SCAN_I; ++i) {
        int64_t p = base + (int64_t)threadIdx.x * SCAN_I + i;
        if (p < n) inout[p] = e;
        e += v[i];
    }
}

template <bool IDX>
__global__ __launch_bounds__(SORT_T) void k_digit_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, int64_t n, int shift,
    int bits, int ntiles, const uint32_t *__restrict__ offs, uint32_t *__restrict__ kout,
    uint32_t *__restrict__ vout) {
    constexpr int NW = SORT_T / 64, WITEMS = SORT_TILE / NW;
    __shared__ uint32_t gbase[SORT_MAXR];   // first output slot of each digit of this tile
    __shared__ uint32_t dstart[SORT_MAXR];  // its first position in the tile's digit order
    // per wave: its items of each digit, then (after the prefix over waves)
    // the position of its first item of each digit in the tile's digit order
    __shared__ uint32_t wcnt[NW][SORT_MAXR];
    __shared__ uint32_t lk[SORT_TILE], lv[SORT_TILE];
    __shared__ uint32_t sh[SCAN_T];
    const int R = 1 << bits;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    const int cnt_tile = (int)min((int64_t)SORT_TILE, n - base);
    // this tile's count of digits 2 tid and 2 tid + 1, from the scanned
    // histogram (the next entry in digit-major order ends the run; the last
    // entry ends at n)
    const int64_t total = (int64_t)R * ntiles;
    uint32_t cd[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int d = 2 * tid + q;
        cd[q] = 0;
        if (d < R) {
            const int64_t e = (int64_t)d * ntiles + blockIdx.x;
            const uint32_t g0 = offs[e];
            const uint32_t g1 = e + 1 < total ? offs[e + 1] : (uint32_t)n;
            gbase[d] = g0;
            cd[q] = g1 - g0;
            for (int w = 0; w < NW; ++w) wcnt[w][d] = 0;
        }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan(cd[0] + cd[1], sh, tot);
    if (2 * tid < R) dstart[2 * tid] = ex;
    if (2 * tid + 1 < R) dstart[2 * tid + 1] = ex + cd[0];
    // wave wv owns the tile's items [wv WITEMS, (wv + 1) WITEMS), 64 per
    // round; every (key, value) of this thread first, so the loads are in
    // flight together
    const int64_t wbase = base + (int64_t)wv * WITEMS + lane;
    uint32_t kr[SORT_R], vr[SORT_R], lpos[SORT_R];
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        const int64_t p = wbase + 64 * r;
        kr[r] = p < n ? kin[p] : 0u;
        vr[r] = IDX ? (uint32_t)p : (p < n ? vin[p] : 0u);
    }
    __syncthreads();
    // rank of every item among the equal digits before it in its wave: the
    // wave's earlier rounds (a wave-private counter per digit, advanced by
    // the lowest lane of each digit group) plus the lower lanes of its round
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        const bool valid = wbase + 64 * r < n;
        const uint32_t d = (kr[r] >> shift) & (uint32_t)(R - 1);
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const bool leader = valid && (peers & lt) == 0;
        uint32_t before = 0;
        if (leader) before = atomicAdd(&wcnt[wv][d], (uint32_t)__popcll(peers));
        before = (uint32_t)__shfl((int)before, peers ? __ffsll((long long)peers) - 1 : 0);
        lpos[r] = before + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    // prefix over the waves, from the digit's first position in the tile
    for (int d = tid; d < R; d += SORT_T) {
        uint32_t a = dstart[d];
        for (int w = 0; w < NW; ++w) {
            const uint32_t t = wcnt[w][d];
            wcnt[w][d] = a;
            a += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        if (wbase + 64 * r < n) {
            const uint32_t d = (kr[r] >> shift) & (uint32_t)(R - 1);
            const uint32_t lp = wcnt[wv][d] + lpos[r];
            lk[lp] = kr[r];
            lv[lp] = vr[r];
        }
    }
    __syncthreads();
    // the tile in digit order: each digit's items go to consecutive slots
    for (int i = tid; i < cnt_tile; i += SORT_T) {
        const uint32_t kk = lk[i];
        const uint32_t d = (kk >> shift) & (uint32_t)(R - 1);
        const uint32_t g = gbase[d] + ((uint32_t)i - dstart[d]);
        kout[g] = kk;
        vout[g] = lv[i];
    }
}

// Exclusive offsets of every cell's run in the sorted keys (ncells + 1).
__global__ void k_cell_offsets(const uint32_t *__restrict__ key, int64_t n, int ncells,
                               int64_t *__restrict__ off) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    int lo = r == 0 ? 0 : (int)key[r - 1] + 1;
    int hi = r == n ? ncells : (int)key[r];
    for (int c = lo; c <= hi; ++c) off[c] = r;
}

// ---------------------------------------------------------- lane spread
//
// After the stable sort by cell, the particles of every segment of 64
// consecutive cells (s = cell / 64, l = cell % 64) are re-ordered by
// (rank k within the cell, l): round k holds the k-th particle of each cell of
// the segment that has more than k, in cell order.  A wave of 64 consecutive
// particles then spans 64 different, adjacent cells.  With A_s(k) =
// sum_l min(count_l, k) and M_s(k) = {l : count_l > k}, the particle of rank
// k in cell l goes to
//     off[64 s] + A_s(k) + popcount(M_s(k) & ((1 << l) - 1)).

// One wave per segment: the largest cell count of the segment.
__global__ void k_seg_max(const int64_t *__restrict__ off, int ncells, int nseg,
                          uint32_t *__restrict__ kmax) {
    const int s = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    if (s >= nseg) return;
    const int c = 64 * s + l;
    int cnt = c < ncells ? (int)(off[c + 1] - off[c]) : 0;
    for (int o = 32; o > 0; o >>= 1) cnt = max(cnt, __shfl_xor(cnt, o));
    if (l == 0) kmax[s] = (uint32_t)cnt;
}

// Tables A_s(k), M_s(k) at tbl[s] + k, in chunks of 64 entries: block
// (s, y) of SEGT_W waves fills the chunks y SEGT_W + w, advancing by
// gridDim.y SEGT_W, of segment s.  A chunk starts from
// A_s(k0) = sum_l min(count_l, k0) (a wave sum) and steps through its 64
// ranks with one ballot each.
constexpr int SEGT_W = 4, SEGT_Y = 8;
__global__ __launch_bounds__(64 * SEGT_W) void k_seg_tables(
    const int64_t *__restrict__ off, int ncells, const uint32_t *__restrict__ kmax,
    const uint32_t *__restrict__ tbl, uint32_t *__restrict__ A, unsigned long long *__restrict__ M) {
    const int s = blockIdx.x;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = 64 * s + l;
    const uint32_t cnt = c < ncells ? (uint32_t)(off[c + 1] - off[c]) : 0u;
    const uint32_t K = kmax[s], base = tbl[s];
    for (uint32_t k0 = 64u * (blockIdx.y * SEGT_W + w); k0 < K; k0 += 64u * SEGT_W * gridDim.y) {
        uint32_t acc = min(cnt, k0);
        for (int o = 32; o > 0; o >>= 1) acc += (uint32_t)__shfl_xor((int)acc, o);
        unsigned long long mine = 0;
        uint32_t acc_mine = 0;
        const uint32_t kend = min(K - k0, 64u);
        for (uint32_t j = 0; j < kend; ++j) {
            const unsigned long long m = __ballot(cnt > k0 + j);
            if ((uint32_t)l == j) {
                mine = m;
                acc_mine = acc;
            }
            acc += (uint32_t)__popcll(m);
        }
        if ((uint32_t)l < kend) {
            A[base + k0 + l] = acc_mine;
            M[base + k0 + l] = mine;
        }
    }
}

// The lane-spread permutation as a transpose, segment by segment: in the
// sorted order a segment's values are cell-major (cell l, rank k); in the
// lane-spread order they are rank-major.  A block moves LANE_KC ranks of all
// 64 cells at a time through LDS: it reads each cell's run of ranks
// (consecutive sorted slots) and writes each rank's cells (consecutive
// destination slots).  Blocks are (segment, b) with b = 0 .. gridDim.y-1
// taking the rank chunks b, b + gridDim.y, ... of the segment.
constexpr int LANE_KC = 32, LANE_T = 256;

__global__ __launch_bounds__(LANE_T) void k_lane_transpose(
    const uint32_t *__restrict__ val, const int64_t *__restrict__ off, int ncells,
    const uint32_t *__restrict__ kmax, const uint32_t *__restrict__ tbl,
    const uint32_t *__restrict__ A, const unsigned long long *__restrict__ M,
    uint32_t *__restrict__ perm2) {
    __shared__ uint32_t tile[LANE_KC][65];
    __shared__ int64_t coff[65];
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int c0 = 64 * s;
    const int nc = min(64, ncells - c0);
    if (tid <= 64) coff[tid] = off[min(c0 + tid, ncells)];
    __syncthreads();
    const uint32_t K = kmax[s], tb = tbl[s];
    const int64_t sbase = coff[0];
    for (uint32_t k0 = (uint32_t)blockIdx.y * LANE_KC; k0 < K; k0 += gridDim.y * LANE_KC) {
        // read: cell l's ranks k0 .. k0 + LANE_KC - 1 (those it has)
        for (int i = tid; i < 64 * LANE_KC; i += LANE_T) {
            const int l = i / LANE_KC, kk = i % LANE_KC;
            if (l < nc) {
                const int64_t r = coff[l] + k0 + kk;
                if (r < coff[l + 1]) tile[kk][l] = val[r];
            }
        }
        __syncthreads();
        // write: rank k's cells, in cell order, to consecutive slots
        for (int i = tid; i < 64 * LANE_KC; i += LANE_T) {
            const int kk = i / 64, l = i % 64;
            const uint32_t k = k0 + kk;
            if (k < K) {
                const unsigned long long m = M[tb + k];
                if ((m >> l) & 1ull) {
                    const unsigned long long below = l ? (m & (~0ull >> (64 - l))) : 0ull;
                    perm2[sbase + A[tb + k] + __popcll(below)] = tile[kk][l];
                }
            }
        }
        __syncthreads();
    }
}

int ensure_sort(Ctx *c) {
    if (c->sort_cap < c->cap || !c->sort_keys[0]) {
        for (int k = 0; k < 2; ++k) {
            if (c->sort_keys[k]) (void)hipFree(c->sort_keys[k]);
            if (c->sort_vals[k]) (void)hipFree(c->sort_vals[k]);
            c->sort_keys[k] = c->sort_vals[k] = nullptr;
        }
        int64_t cap = c->cap > 0 ? c->cap : 1;
        for (int k = 0; k < 2; ++k) {
            PIC_HIP(hipMalloc(&c->sort_keys[k], cap * 4));
            PIC_HIP(hipMalloc(&c->sort_vals[k], cap * 4));
        }
        c->sort_cap = cap;
    }
    const int64_t ntiles = (c->cap + SORT_TILE - 1) / SORT_TILE + 1;
    // the histograms, and room for the scans of the digit histograms, of the
    // per-cell counts (ncells) and of the segment tables (block sums in
    // sort_blk, one per SCAN_B items)
    const int64_t hcap = std::max<int64_t>((int64_t)SORT_MAXR * ntiles, (int64_t)c->nx * c->ny);
    if (hcap > c->sort_hist_cap) {
        if (c->sort_hist) (void)hipFree(c->sort_hist);
        if (c->sort_blk) (void)hipFree(c->sort_blk);
        PIC_HIP(hipMalloc(&c->sort_hist, hcap * 4));
        PIC_HIP(hipMalloc(&c->sort_blk, ((hcap + SCAN_B - 1) / SCAN_B + 1) * 4));
        c->sort_hist_cap = hcap;
    }
    if (!c->cell_off) PIC_HIP(hipMalloc(&c->cell_off, ((size_t)c->nx * c->ny + 1) * 8));
    return PIC_OK;
}

static int scan(Ctx *c, uint32_t *a, int64_t n) {
    if (n <= 2 * SCAN_B) {  // short arrays (segment tables): one block, one launch
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, c->stream, a, (int)n);
        PIC_HIP(hipGetLastError());
        return PIC_OK;
    }
    const int nb = (int)((n + SCAN_B - 1) / SCAN_B);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, c->stream, a, n, c->sort_blk);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, c->stream, c->sort_blk, nb);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_T), 0, c->stream, a, n, c->sort_blk);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// Stable sort by key of the first n items of `keys0` (read, not written; the
// values are the item indices): the sorted keys and values end in
// (sort_keys[0], sort_vals[0]).  `keys0` may be sort_keys[0] itself.
static int radix_sort_pairs(Ctx *c, const uint32_t *keys0, int64_t n, int keybits) {
    const int passes = (keybits + SORT_MAXB - 1) / SORT_MAXB;
    const int ntiles = (int)((n + SORT_TILE - 1) / SORT_TILE);
    const uint32_t *kin = keys0, *vin = nullptr;
    int out = 1, shift = 0;
    for (int ps = 0; ps < passes; ++ps) {
        const int bits = (keybits - shift + (passes - ps) - 1) / (passes - ps);
        hipLaunchKernelGGL(k_digit_hist, dim3(ntiles), dim3(SORT_T), 0, c->stream, kin, n, shift,
                           bits, ntiles, c->sort_hist);
        PIC_TRY(scan(c, c->sort_hist, (int64_t)(1 << bits) * ntiles));
        if (ps == 0)
            hipLaunchKernelGGL(k_digit_scatter<true>, dim3(ntiles), dim3(SORT_T), 0, c->stream,
                               kin, vin, n, shift, bits, ntiles, c->sort_hist,
                               c->sort_keys[out], c->sort_vals[out]);
        else
            hipLaunchKernelGGL(k_digit_scatter<false>, dim3(ntiles), dim3(SORT_T), 0, c->stream,
                               kin, vin, n, shift, bits, ntiles, c->sort_hist,
                               c->sort_keys[out], c->sort_vals[out]);
        PIC_HIP(hipGetLastError());
        kin = c->sort_keys[out];
        vin = c->sort_vals[out];
        out = 1 - out;
        shift += bits;
    }
    if (kin != c->sort_keys[0]) {  // the result is in slot 1
        std::swap(c->sort_keys[0], c->sort_keys[1]);
        std::swap(c->sort_vals[0], c->sort_vals[1]);
    }
    return PIC_OK;
}

// The alternate array of state array `arr` (allocated once, same capacity,
// freed with the particles): the sort gathers into it and then swaps the two
// pointers.
template <typename T>
static int alt_slot(Ctx *c, T *&arr, int *slot_out) {
    void **key = reinterpret_cast<void **>(&arr);
    int slot = -1;
    for (int k = 0; k < Ctx::NALT; ++k)
        if (c->alt_of[k] == key) slot = k;
    if (slot < 0) {
        for (int k = 0; k < Ctx::NALT && slot < 0; ++k)
            if (!c->alt_of[k]) slot = k;
        PIC_CHECK(slot >= 0, PIC_ERR_STATE, "no free alternate array slot");
        void *q = nullptr;
        PIC_HIP(hipMalloc(&q, (size_t)(c->cap > 0 ? c->cap : 1) * sizeof(T)));
        c->alt_of[slot] = key;
        c->alt[slot] = q;
    }
    *slot_out = slot;
    return PIC_OK;
}

template <typename T>
static void alt_swap(Ctx *c, T *&arr, int slot) {
    T *alt = static_cast<T *>(c->alt[slot]);
    c->alt[slot] = arr;
    arr = alt;
}

// out_q[r] = in_q[perm[r]] for the state arrays of one particle
// representation, in one pass (the permutation is read once).  The FP32
// state is the cell index and one 16-byte (ox, oy, vx, vy) record; the FP64
// state N8 arrays of 8-byte words (a null input is skipped).
struct StateGather {
    const uint64_t *in8[5];
    uint64_t *out8[5];
};

__global__ void k_gather_f32state(int64_t n, const uint32_t *__restrict__ perm,
                                  const int32_t *__restrict__ cell_in, int32_t *__restrict__ cell_out,
                                  const float4 *__restrict__ ov_in, float4 *__restrict__ ov_out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t src = perm[r];
    cell_out[r] = cell_in[src];
    ov_out[r] = ov_in[src];
}

template <int N8>
__global__ void k_gather_state(int64_t n, const uint32_t *__restrict__ perm, StateGather g) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t src = perm[r];
#pragma unroll
    for (int q = 0; q < N8; ++q)
        if (g.in8[q]) g.out8[q][r] = g.in8[q][src];
}

// The alternate state arrays (allocated on first use) and their swap with
// the current ones, for a pass that moves the particles into sorted slots.
int alt_f32(Ctx *c, int32_t **cell, float4 **ov) {
    int s0, s1;
    PIC_TRY(alt_slot(c, c->cell, &s0));
    PIC_TRY(alt_slot(c, c->ov32, &s1));
    *cell = static_cast<int32_t *>(c->alt[s0]);
    *ov = static_cast<float4 *>(c->alt[s1]);
    return PIC_OK;
}

void swap_alt_f32(Ctx *c) {
    int s0 = 0, s1 = 0;
    (void)alt_slot(c, c->cell, &s0);
    (void)alt_slot(c, c->ov32, &s1);
    alt_swap(c, c->cell, s0);
    alt_swap(c, c->ov32, s1);
}

int alt_f64(Ctx *c, double **x, double **y, double **vx, double **vy, double **w) {
    int s[5];
    PIC_TRY(alt_slot(c, c->x, &s[0]));
    PIC_TRY(alt_slot(c, c->y, &s[1]));
    PIC_TRY(alt_slot(c, c->vx, &s[2]));
    PIC_TRY(alt_slot(c, c->vy, &s[3]));
    *x = static_cast<double *>(c->alt[s[0]]);
    *y = static_cast<double *>(c->alt[s[1]]);
    *vx = static_cast<double *>(c->alt[s[2]]);
    *vy = static_cast<double *>(c->alt[s[3]]);
    if (w) {
        PIC_TRY(alt_slot(c, c->w, &s[4]));
        *w = static_cast<double *>(c->alt[s[4]]);
    }
    return PIC_OK;
}

void swap_alt_f64(Ctx *c) {
    int s[5] = {0, 0, 0, 0, 0};
    (void)alt_slot(c, c->x, &s[0]);
    (void)alt_slot(c, c->y, &s[1]);
    (void)alt_slot(c, c->vx, &s[2]);
    (void)alt_slot(c, c->vy, &s[3]);
    alt_swap(c, c->x, s[0]);
    alt_swap(c, c->y, s[1]);
    alt_swap(c, c->vx, s[2]);
    alt_swap(c, c->vy, s[3]);
    if (!c->uniform_w) {
        (void)alt_slot(c, c->w, &s[4]);
        alt_swap(c, c->w, s[4]);
    }
}

// Applies `perm` to the current state (FP32: cell, ox, oy, vx, vy; FP64: x,
// y, vx, vy and w when weights differ).
static int permute_state(Ctx *c, int64_t n, const uint32_t *perm) {
    StateGather g = {};
    dim3 grid((unsigned)((n + 255) / 256));
    if (c->state == PIC_STATE_F32) {
        int sl[2];
        PIC_TRY(alt_slot(c, c->cell, &sl[0]));
        PIC_TRY(alt_slot(c, c->ov32, &sl[1]));
        hipLaunchKernelGGL(k_gather_f32state, grid, dim3(256), 0, c->stream, n, perm, c->cell,
                           static_cast<int32_t *>(c->alt[sl[0]]), c->ov32,
                           static_cast<float4 *>(c->alt[sl[1]]));
        PIC_HIP(hipGetLastError());
        alt_swap(c, c->cell, sl[0]);
        alt_swap(c, c->ov32, sl[1]);
    } else {
        const bool w = !c->uniform_w;
        int sl[5] = {0, 0, 0, 0, 0};
        PIC_TRY(alt_slot(c, c->x, &sl[0]));
        PIC_TRY(alt_slot(c, c->y, &sl[1]));
        PIC_TRY(alt_slot(c, c->vx, &sl[2]));
        PIC_TRY(alt_slot(c, c->vy, &sl[3]));
        if (w) PIC_TRY(alt_slot(c, c->w, &sl[4]));
        double *src[5] = {c->x, c->y, c->vx, c->vy, w ? c->w : nullptr};
        for (int q = 0; q < 5; ++q) {
            g.in8[q] = reinterpret_cast<const uint64_t *>(src[q]);
            g.out8[q] = src[q] ? static_cast<uint64_t *>(c->alt[sl[q]]) : nullptr;
        }
        hipLaunchKernelGGL((k_gather_state<5>), grid, dim3(256), 0, c->stream, n, perm, g);
        PIC_HIP(hipGetLastError());
        alt_swap(c, c->x, sl[0]);
        alt_swap(c, c->y, sl[1]);
        alt_swap(c, c->vx, sl[2]);
        alt_swap(c, c->vy, sl[3]);
        if (w) alt_swap(c, c->w, sl[4]);
    }
    return PIC_OK;
}

// Lane-spread tables (per-segment A_s(k), M_s(k)) from the cell offsets.  No host
// synchronisation: the tables hold sum_s max_l count_l <= n entries, so
// capacity n + 64 always suffices.
static int lane_spread(Ctx *c) {
    const int64_t n = c->n;
    const int ncells = c->nx * c->ny;
    const int nseg = (ncells + 63) / 64;
    if (!c->seg_k) {
        PIC_HIP(hipMalloc(&c->seg_k, (size_t)nseg * 4));
        PIC_HIP(hipMalloc(&c->seg_tbl, (size_t)(nseg + 1) * 4));
    }
    if (!c->perm2) PIC_HIP(hipMalloc(&c->perm2, (size_t)(c->cap > 0 ? c->cap : 1) * 4));
    const int64_t need = (c->cap > 0 ? c->cap : 1) + 64;
    if (need > c->seg_cap) {
        if (c->seg_A) (void)hipFree(c->seg_A);
        if (c->seg_M) (void)hipFree(c->seg_M);
        c->seg_A = nullptr;
        c->seg_M = nullptr;
        PIC_HIP(hipMalloc(&c->seg_A, (size_t)need * 4));
        PIC_HIP(hipMalloc(&c->seg_M, (size_t)need * 8));
        c->seg_cap = need;
    }
    const int wpb = 4;  // waves (segments) per block
    dim3 gs((unsigned)((nseg + wpb - 1) / wpb)), bs(64 * wpb);
    hipLaunchKernelGGL(k_seg_max, gs, bs, 0, c->stream, c->cell_off, ncells, nseg, c->seg_k);
    PIC_HIP(hipMemcpyAsync(c->seg_tbl, c->seg_k, (size_t)nseg * 4, hipMemcpyDeviceToDevice,
                           c->stream));
    PIC_TRY(scan(c, c->seg_tbl, nseg));
    hipLaunchKernelGGL(k_seg_tables, dim3((unsigned)nseg, SEGT_Y), dim3(64 * SEGT_W), 0,
                       c->stream, c->cell_off, ncells, c->seg_k, c->seg_tbl, c->seg_A, c->seg_M);
    PIC_HIP(hipGetLastError());
    (void)n;
    return PIC_OK;
}

// Applies a permutation a lane sort left pending (no pass consumed it).
int settle_state(Ctx *c) {
    if (!c->pending_perm) return PIC_OK;
    const uint32_t *perm = c->pending_perm;
    c->pending_perm = nullptr;
    return permute_state(c, c->n, perm);
}

int sort_by_cell(Ctx *c, int32_t *perm_host, bool lanes, bool apply) {
    const int64_t n = c->n;
    PIC_TRY(settle_state(c));
    c->lanes_order = lanes;
    c->seg_valid = false;
    c->sort_fresh = false;
    if (n == 0) return PIC_OK;
    PIC_TRY(ensure_sort(c));
    const int ncells = c->nx * c->ny;
    int keybits = 1;
    while ((1 << keybits) < ncells) ++keybits;
    // keys: the FP32 state's cell array as it is, or the FP64 state's cells
    const uint32_t *keys0 = reinterpret_cast<const uint32_t *>(c->cell);
    if (c->state != PIC_STATE_F32) {
        hipLaunchKernelGGL(k_keys_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           n, c->x, c->y, c->dx, c->dy, c->nx, c->ny, c->sort_keys[0]);
        PIC_HIP(hipGetLastError());
        keys0 = c->sort_keys[0];
    }
    PIC_TRY(radix_sort_pairs(c, keys0, n, keybits));
    hipLaunchKernelGGL(k_cell_offsets, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0,
                       c->stream, c->sort_keys[0], n, ncells, c->cell_off);
    PIC_HIP(hipGetLastError());
    const uint32_t *perm = c->sort_vals[0];
    if (lanes) {
        PIC_TRY(lane_spread(c));
        // rank chunks per segment: enough blocks to fill the device several
        // times over at the mean cell count (blocks past a segment's largest
        // count exit at once)
        const int nseg = (ncells + 63) / 64;
        const int64_t per_cell = n / ncells + 1;
        const int64_t chunks = (per_cell + LANE_KC - 1) / LANE_KC + 1;
        const int64_t want = std::max<int64_t>(1, (int64_t)c->ncu * 16 / nseg);
        const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(chunks, want));
        hipLaunchKernelGGL(k_lane_transpose, dim3((unsigned)nseg, gy), dim3(LANE_T), 0,
                           c->stream, (const uint32_t *)c->sort_vals[0],
                           (const int64_t *)c->cell_off, ncells, (const uint32_t *)c->seg_k,
                           (const uint32_t *)c->seg_tbl, (const uint32_t *)c->seg_A,
                           (const unsigned long long *)c->seg_M, c->perm2);
        PIC_HIP(hipGetLastError());
        perm = c->perm2;
        c->seg_valid = true;
    }
    if (apply || !lanes || perm_host)
        PIC_TRY(permute_state(c, n, perm));
    else
        c->pending_perm = perm;  // the next segment pass moves the particles
    if (perm_host) {
        PIC_HIP(hipMemcpyAsync(perm_host, perm, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
    }
    c->sort_fresh = lanes;
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_sort(pic_ctx *ctx, int32_t *perm) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(sort_by_cell(c, perm, false));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_sort_lanes(pic_ctx *ctx, int32_t *perm) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(sort_by_cell(c, perm, true));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_cell_offsets(pic_ctx *ctx, int64_t *offsets) {
    PIC_CHECK(ctx && offsets, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    PIC_CHECK(c->cell_off, PIC_ERR_STATE, "no sort has run on this context");
    PIC_HIP(hipMemcpyAsync(offsets, c->cell_off, ((size_t)c->nx * c->ny + 1) * 8,
                           hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // extern "C"
