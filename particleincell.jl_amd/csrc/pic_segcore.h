// pic_segcore.h -- the segment-persistent pass machinery shared by the time
// loop's passes (pic_seg.hip) and the adjoint's reverse passes: segment
// windows, the deferred-particle list, the launch shape.
//
// After a lane sort segment s of 64 consecutive cells owns the slots
// [off[64 s], off[min(64 s + 64, ncells)]) until the next sort; a block takes
// a slice of one segment, stages a window of the grid around the segment's
// home cells in LDS and hands particles whose stencil leaves it to a
// deferred list that a global-memory kernel completes after the pass.
#pragma once

#include "pic_internal.h"

namespace pic {

constexpr int SEG_T = 512;  // threads per block

// Stencil of shape K (S nodes per direction, HL left of the cell).
template <int K>
using SegStencil = Shape<K>;

struct Win {
    int x0, y0, tw, th;  // global index of window node (0,0) (may be negative), size
};

// Window of segment s: its home cells widened by the stencil and a drift
// margin, the margin reduced until the window, with `pad` spare columns per
// row, fits `cap` nodes.  A window as wide as the periodic grid plus the
// stencil covers every cell of that direction (x0 = -HL then), duplicating
// S-1 columns.
template <int S, int HL>
__device__ Win seg_window(int s, int nx, int ny, int ncells, int margin, int cap, int pad = 0) {
    const int cf = 64 * s, cl = min(64 * s + 63, ncells - 1);
    const int jh0 = cf / nx, jh1 = cl / nx;
    int ih0, ih1;
    if (jh0 == jh1) {
        ih0 = cf - jh0 * nx;
        ih1 = cl - jh0 * nx;
    } else {
        ih0 = 0;
        ih1 = nx - 1;
    }
    const int hw = ih1 - ih0 + 1, hh = jh1 - jh0 + 1;
    int M = margin;
    int tw = 0, th = 0;
    for (;; --M) {
        tw = min(hw + S - 1 + 2 * M, nx + S - 1);
        th = min(hh + S - 1 + 2 * M, ny + S - 1);
        if ((tw + pad) * th <= cap || M == 0) break;
    }
    Win w;
    w.tw = tw;
    w.th = th;
    w.x0 = tw == nx + S - 1 ? -HL : ih0 - HL - M;
    w.y0 = th == ny + S - 1 ? -HL : jh0 - HL - M;
    return w;
}

// XCD-aware block order of the segment passes: workgroups are dealt to the
// 8 XCDs round-robin, so block b takes slice (b % 8) nb/8 + b/8 of the
// nb = gridDim.x slices; each XCD then works on one contiguous band of
// segments, whose overlapping windows and source slots share its L2.
__device__ __forceinline__ int seg_block_order() {
    const int nbk = (int)gridDim.x;
    return nbk % 8 == 0 ? (int)(blockIdx.x % 8) * (nbk / 8) + (int)(blockIdx.x / 8)
                        : (int)blockIdx.x;
}

// End of a deferred-list kernel: block 0 adds the list length to the
// slow-path counter; the last block to arrive (every block has read the
// count by then) clears count and arrivals for the next pass.
__device__ __forceinline__ void defer_finish(unsigned cnt, unsigned *dcount,
                                             unsigned long long *slow) {
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (blockIdx.x == 0 && cnt) atomicAdd(slow, (unsigned long long)cnt);
    const unsigned arrived = atomicAdd(&dcount[1], 1u);
    if (arrived == gridDim.x - 1) {
        __hip_atomic_store(&dcount[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&dcount[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Appends `entry` for every lane with `want` set (one atomic per wave).  The
// leader's slot is broadcast with v_readlane, not a shuffle: a shuffle is an
// LDS instruction, and waiting for its result would also wait for every LDS
// atomic the wave still has in flight.
__device__ __forceinline__ void defer_push(bool want, uint32_t entry, uint32_t *list,
                                           unsigned *count) {
    const unsigned long long b = __ballot(want);
    if (!b) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)b) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(b));
    base = (unsigned)__builtin_amdgcn_readlane((int)base, leader);
    if (want) list[base + (unsigned)__popcll(b & ((1ull << lane) - 1ull))] = entry;
}

constexpr uint32_t DEFER_DEPOSIT_ONLY = 0x80000000u;

constexpr int DEFER_T = 256, DEFER_BLOCKS = 256;

// LDS per block for the window of the FP32 passes other than the
// fixed-point drift-and-deposit ones (and of the adjoint's deposit pass):
// 8 B field (+4 B bz) + 8 B charge (pair or FP64).  40 KiB lets four
// 512-thread blocks share a CU's 160 KiB; an M6 window of a 64-cell row
// segment then keeps a drift margin of 11 cells.  The FP64 passes and the
// FP32 fixed-point drift-and-deposit passes run one 160 KiB block per CU.
constexpr size_t SEG_LDS_BYTES = 40 * 1024;

// Blocks per segment: as many as keep every CU busy in one round of
// resident blocks (the occupancy of this kernel instantiation at this LDS
// size), at least one.  A larger grid than one round leaves a partial last
// round; a grid with more segments than resident blocks takes several full
// rounds.
inline int seg_bps(const Ctx *c, const void *kernel, size_t sm, int threads = SEG_T) {
    const int nseg = (c->nx * c->ny + 63) / 64;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, sm) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    int bps = c->ncu * per_cu / nseg;
    if (bps < 1) bps = 1;
    if (bps > 64) bps = 64;
    return bps;
}

}  // namespace pic
