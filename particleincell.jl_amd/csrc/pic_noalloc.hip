// pic_noalloc.hip -- compute_rho on caller-owned device memory, allocating
// nothing: for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// pic_compute_rho deposits into the context's FP64 accumulator and reduces
// its total in the context's partial sums.  This variant takes the particles
// (the reference's p.array, 5 x n doubles, in device memory) and the output
// grid from the caller and keeps the accumulator, the partial sums and the
// reduction's arrival counter in one file-scope device array, `scratch`,
// sized for grids up to NOALLOC_MAX_CELLS nodes (2 MiB + 2 KiB, held by every
// GPU the library runs on until the process ends).  Host code never passes
// the array's name to a kernel: it resolves the array's address on each
// device once (hipGetSymbolAddress) and passes that.  Every call clears what
// must start at zero with a fill kernel (no hipMemsetAsync, which graph
// capture records poorly), enqueues its kernels on the caller's stream and
// returns without synchronising.  Calls that overlap on one GPU share the
// array and race.
#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>

#include "pic_internal.h"

namespace pic {

namespace {

// Largest grid of the allocation-free compute_rho: 512 x 512 nodes, the
// largest mesh of the reference's configurations.
constexpr int64_t NOALLOC_MAX_CELLS = 512 * 512;
// accumulator, RED_BLOCKS partial sums, the total, the arrival counter
constexpr int64_t NOALLOC_DOUBLES = NOALLOC_MAX_CELLS + RED_BLOCKS + 2;

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

// The address of `scratch` on the current device, resolved on the first call
// there.
int scratch_on(int device, double **out) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        PIC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PIC_CHECK(p, PIC_ERR_HIP, "no device address for the scratch array on device %d", device);
        scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = scratch_ptr[device];
    return PIC_OK;
}

// Zeroes the accumulator's first m nodes and the arrival counter.
__global__ void k_na_clear(double *acc, int64_t m, unsigned *count) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * blockDim.x)
        acc[k] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0u;
}

// The raw deposit of compute_rho (pushers_cic.jl:84-105 for CIC with its
// trunc index, pushers_m6.jl:246-338 and the other shapes with floor), one
// particle per thread, read from p.array (x, y at rows 1-2, w at row 5),
// added to the accumulator with FP64 atomics.  The same shares, in the same
// operations, as pic_compute_rho's deposit.
template <int K>
__global__ void k_na_deposit(const double *__restrict__ p, int64_t n, int nx, int ny, double dx,
                             double dy, double *acc) {
    using St = Shape<K>;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *q = p + 5 * k;
    const double xp = q[0] / dx, yp = q[1] / dy, wgt = q[4];
    double fx, fy;
    const int ci = picd::cell_of(xp, K == PIC_CIC ? trunc(xp) : floor(xp), nx, fx);
    const int cj = picd::cell_of(yp, K == PIC_CIC ? trunc(yp) : floor(yp), ny, fy);
    auto add = [&](int u, int v, double c) {
        const int64_t g = (int64_t)picd::wrap2(ci - St::HL + u, nx) +
                          (int64_t)nx * picd::wrap2(cj - St::HL + v, ny);
        unsafeAtomicAdd(&acc[g], c);
    };
    if (K == PIC_CIC) {
        const double a1 = (1 - fx) * (1 - fy);
        const double a2 = fx * (1 - fy);
        const double a3 = fx * fy;
        const double a4 = (1 - fx) * fy;
        add(0, 0, a1 * wgt);
        add(1, 0, a2 * wgt);
        add(1, 1, a3 * wgt);
        add(0, 1, a4 * wgt);
    } else {
        double wx[6], wy[6];
        deposit_w<K>(fx, ci, wx);
        deposit_w<K>(fy, cj, wy);
#pragma unroll
        for (int u = 0; u < St::S; ++u)
#pragma unroll
            for (int v = 0; v < St::S; ++v) add(u, v, wx[u] * wy[v] * wgt);
    }
}

}  // namespace

}  // namespace pic

using namespace pic;

extern "C" {

int pic_compute_rho_noalloc(int kernel, const double *d_p, int64_t n, int nx, int ny, double xmin,
                            double xmax, double ymin, double ymax, double *d_rho, void *stream) {
    PIC_CHECK(valid_shape(kernel), PIC_ERR_ARG, "kernel %d", kernel);
    PIC_CHECK(nx >= 4 && ny >= 4, PIC_ERR_ARG, "nx=%d, ny=%d: need at least 4 cells each", nx, ny);
    PIC_CHECK((int64_t)nx * ny <= NOALLOC_MAX_CELLS, PIC_ERR_ARG,
              "grid %dx%d exceeds the %lld nodes of the allocation-free scratch", nx, ny,
              (long long)NOALLOC_MAX_CELLS);
    PIC_CHECK(xmax > xmin && ymax > ymin, PIC_ERR_ARG, "empty domain");
    PIC_CHECK(n >= 0 && n <= INT32_MAX, PIC_ERR_ARG, "particle count %lld", (long long)n);
    PIC_CHECK(d_rho && (n == 0 || d_p), PIC_ERR_ARG, "null device pointer");
    int device = 0;
    PIC_HIP(hipGetDevice(&device));
    PIC_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, PIC_ERR_ARG, "device %d", device);
    double *s = nullptr;
    PIC_TRY(scratch_on(device, &s));
    const int64_t m = (int64_t)nx * ny;
    double *acc = s, *part = s + NOALLOC_MAX_CELLS, *total = part + RED_BLOCKS;
    unsigned *count = reinterpret_cast<unsigned *>(total + 1);
    const double dx = (xmax - xmin) / nx, dy = (ymax - ymin) / ny;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_na_clear, dim3((unsigned)std::min<int64_t>((m + 255) / 256, 1024)),
                       dim3(256), 0, st, acc, m, count);
    PIC_HIP(hipGetLastError());
    if (n) {
        const dim3 grid((unsigned)((n + 255) / 256));
        PIC_SHAPE_SWITCH(kernel, hipLaunchKernelGGL(k_na_deposit<KK>, grid, dim3(256), 0, st, d_p,
                                                    n, nx, ny, dx, dy, acc));
        PIC_HIP(hipGetLastError());
    }
    int blocks = (int)((m + RED_THREADS - 1) / RED_THREADS);
    if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
    hipLaunchKernelGGL(k_sum, dim3(blocks), dim3(RED_THREADS), 0, st, acc, m, part, count, total);
    hipLaunchKernelGGL(k_finalize_rho, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, acc,
                       d_rho, m, dx * dy, total, kernel, dx, dy, xmax - xmin, ymax - ymin);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

}  // extern "C"
