// pic_small.hip -- pic_run's time loop with one launch per iteration on small
// grids (power-of-two nx, ny <= 64).
//
// On a 64 x 64 grid the loop of separate kernels (three transform kernels, the
// particle pass and its deferred-list kernel, a lane sort every few steps)
// is bound by launch latency: each kernel runs a few microseconds with most
// of the chip idle.  Here every block of one launch holds the whole grid in
// LDS and does:
//   1. the E solve of the previous deposit, redundantly on its own LDS copy
//      (the same operations on the same inputs, so every block holds the
//      same field): transforms along x and y of the raw charge, the spectral
//      multiplier of the E solve, the inverse transforms along y and x,
//      (ex, ey) = (Re, Im).  Block 0 writes the field energy, and at the last
//      iteration ex, ey and their FP32 copy;
//   2. the particles in any order (no sort): gather from the LDS field, kick,
//      drift, deposit into an LDS charge grid;
//   3. the charge grid added to the global accumulator with global atomics.
// The raw charge rotates through three buffers (the one read, the one summed
// into, the one cleared for the next launch), so no launch writes what
// another block of the same launch reads.
//
// Transforms: a 1-D transform of length N <= 64 runs in N lanes of one wave
// (radix-2 decimation in frequency, partners exchanged between lanes by DPP
// moves and swizzles), each lane working on up to four rows or columns at
// once; no block-wide barrier between the stages, four between the passes
// over rows and columns.  Same transforms (unnormalised, exp(-+2 pi i jk/N))
// and the same spectral arithmetic as the three-kernel solve of
// pic_poisson.hip; results agree to rounding.
//
// Charge: FP32 state: int32 fixed-point LDS grid (charge32_fixed) with one
// quantum for every block, flushed as int64 sums, so the raw charge is exact
// and the same on every run; FP64 state: FP64 LDS atomics and FP64 global
// atomics, as in the segment passes.
//
// Per-particle arithmetic is the segment passes' (pic_push32.h,
// pic_push64.h): the same operations on the same field values.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_fftlanes.h"
#include "pic_internal.h"
#include "pic_push32.h"
#include "pic_push64.h"

namespace pic {

constexpr int ES_T = 1024;                 // threads per block
constexpr int ES_WAVES = ES_T / 64;
constexpr int ES_MAX_SIDE = 64;            // nx, ny <= 64: one transform per wave
// rows (or columns) per wave and pass: nx ny / (64 ES_WAVES) <= 4
constexpr int ES_MAX_ROUNDS = ES_MAX_SIDE * ES_MAX_SIDE / (64 * ES_WAVES);
// Particles per block.  Every block adds its whole charge grid to the global
// one, so fewer, fuller blocks flush fewer atomics (each node receives one
// per block); measured on MI355X, Landau CIC 64 x 64 with 204,800 particles:
// 128 blocks (1.6 particles per thread) 25.5 us per launch in the FP32
// state and 27.9 in the FP64 state, 200 blocks 27.5 and 29.5, 100 blocks
// 25.4 and 28.5.
constexpr int64_t ES_BLOCK_PARTICLES = 1600;

struct EsStep {
    int nx, ny, lognx, logny;
    int mode;     // 0: drift dt/2 + deposit; 1: kick + drift dt + deposit; 2: kick + drift dt/2
    int solve;    // E from rin first (modes 1, 2)
    int write_e;  // block 0 stores ex, ey, e32
    int slot;     // energy history entry (-1: none)
    int64_t n;
    SegF64 s64;   // FP64 state: arrays and the step's constants (kick64, drift64)
    int32_t *cell;  // FP32 state
    float4 *ov;
    float hdt, dtdx, dtdy, bz0;
    float wu;     // |w| in quanta
    int R, conserve;
    double unit;  // charge of one quantum
    const void *rin;  // raw charge read (FP64, or int64 quanta in the FP32 state)
    void *racc;       // summed into
    void *rclr;       // cleared
    const double *kxk2, *kyk2;
    double scale, dxdy;
    double *ex, *ey;
    float2 *e32;
    double *hist;
    const double2 *tw;  // twiddles: nx/2 along x, then ny/2 along y
    int nbp;            // particle blocks; block nbp (if launched) the field block
};

// -i m v scale for complex v.
__device__ __forceinline__ double2 es_mul_mi(double m, double2 v, double scale) {
    return make_double2(m * v.y * scale, -(m * v.x) * scale);
}

// (Fx, Fy) = -i (kx, ky)/k2 R at half-spectrum index (ih, jh), R[0,0] -> 0.
__device__ __forceinline__ void es_half(const double2 *S, int PX, int nx, int ih, int jh,
                                        const double *__restrict__ kxk2,
                                        const double *__restrict__ kyk2, double scale,
                                        double2 &f, double2 &g) {
    const int q = ih + (nx / 2 + 1) * jh;
    if (q == 0) {
        f = g = make_double2(0.0, 0.0);
        return;
    }
    const double2 R = S[jh * PX + ih];
    f = es_mul_mi(kxk2[q], R, scale);
    g = es_mul_mi(kyk2[q], R, scale);
}

// Full-grid value (ik, jk) of the spectrum whose inverse transform is
// ex + i ey, from the forward transform S of the charge: the half-spectrum
// values, the self-conjugate columns (ih = 0, nx/2) projected, columns past
// nx/2 the conjugates of their mirrors.
__device__ __forceinline__ double2 es_spectral(const double2 *S, int PX, int nx, int ny, int ik,
                                               int jk, const double *kxk2, const double *kyk2,
                                               double scale) {
    const bool mirror = ik > nx / 2;
    const int ih = mirror ? nx - ik : ik;
    const int jh = mirror ? (ny - jk) & (ny - 1) : jk;
    double2 f, g;
    es_half(S, PX, nx, ih, jh, kxk2, kyk2, scale, f, g);
    if (ih == 0 || 2 * ih == nx) {
        double2 fm, gm;
        es_half(S, PX, nx, ih, (ny - jh) & (ny - 1), kxk2, kyk2, scale, fm, gm);
        f = make_double2(0.5 * (f.x + fm.x), 0.5 * (f.y - fm.y));
        g = make_double2(0.5 * (g.x + gm.x), 0.5 * (g.y - gm.y));
    }
    if (mirror) {
        f.y = -f.y;
        g.y = -g.y;
    }
    return make_double2(f.x - g.y, f.y + g.x);
}

template <int K, bool F32, int BZ>
__global__ __launch_bounds__(ES_T) void k_es_step(EsStep a) {
    constexpr int HL = Shape<K>::HL;
    constexpr int NR = ES_MAX_ROUNDS;
    extern __shared__ __align__(16) char es_smem[];
    __shared__ double red[2 * ES_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nx = a.nx, ny = a.ny, mx = nx - 1, my = ny - 1, m = nx * ny;
    const int PX = nx + 1;  // odd row pitch: a column's lanes spread over the banks
    const int nq = (nx / 2 + 1) * ny;
    // LDS: the grid (complex), twiddles (nx/2 along x, then ny/2 along y),
    // the spectral tables, the charge grid
    double2 *X = reinterpret_cast<double2 *>(es_smem);
    double2 *tw = X + ny * PX;
    double *tkx = reinterpret_cast<double *>(tw + (nx / 2 + ny / 2)), *tky = tkx + nq;
    char *tiles = reinterpret_cast<char *>(tky + nq);
    int *itile = reinterpret_cast<int *>(tiles);
    double *dtile = reinterpret_cast<double *>(tiles);
    const bool gather = a.mode != 0, deposit = a.mode != 2;
    const SegF64 &s = a.s64;

    // every global load of the launch's first phase is issued here, so that
    // their latencies overlap: this thread's first particle (used after the
    // solve), its rows of the charge, the spectral tables
    // blocks [0, nbp) move the particles; block nbp, launched with the solve,
    // writes the field energy (and the last field) instead
    const bool field = (int)blockIdx.x == a.nbp;
    int64_t p = field ? a.n : (int64_t)blockIdx.x * ES_T + tid;
    const int64_t stride = (int64_t)a.nbp * ES_T;
    int pc = 0;
    float4 po = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    double px = 0.0, py = 0.0, pvx = 0.0, pvy = 0.0, pw = 0.0;
    if (p < a.n) {
        if (F32) {
            pc = a.cell[p];
            po = a.ov[p];
        } else {
            px = s.x[p];
            py = s.y[p];
            pvx = s.vx[p];
            pvy = s.vy[p];
            pw = s.w ? s.w[p] : s.w0;
        }
    }
    // rows: 64 / nx per wave and round, lane position i along x; columns:
    // 64 / ny per wave and round, lane position jp along y; NR rounds
    const int rw = 64 >> a.lognx, i = lane & mx, ib = brev_bits(i, a.lognx);
    const int j0 = wave * rw + (lane >> a.lognx), jstep = ES_WAVES * rw;
    const int cw = 64 >> a.logny, jp = lane & my, jb = brev_bits(jp, a.logny);
    const int c0 = wave * cw + (lane >> a.logny), cstep = ES_WAVES * cw;
    double2 zr[NR / 2];
    if (a.solve) {
        // real rows j and j + ny/2 packed as one complex sequence
#pragma unroll
        for (int r = 0; r < NR / 2; ++r) {
            const int j = j0 + r * jstep;
            double u = 0.0, v = 0.0;
            if (j < ny / 2) {
                const int64_t g = (int64_t)j * nx + i, g2 = g + (int64_t)(ny / 2) * nx;
                if (F32) {
                    const long long *q = reinterpret_cast<const long long *>(a.rin);
                    u = (double)q[g] * a.unit;
                    v = (double)q[g2] * a.unit;
                } else {
                    u = reinterpret_cast<const double *>(a.rin)[g];
                    v = reinterpret_cast<const double *>(a.rin)[g2];
                }
            }
            zr[r] = make_double2(u, v);
        }
        for (int k = tid; k < nq; k += ES_T) {
            tkx[k] = a.kxk2[k];
            tky[k] = a.kyk2[k];
        }
        for (int k = tid; k < nx / 2 + ny / 2; k += ES_T) tw[k] = a.tw[k];
    }
    if (deposit && !field)
        for (int t = tid; t < m; t += ES_T) {
            if (F32) itile[t] = 0;
            else dtile[t] = 0.0;
        }
    if (a.rclr) {
        const int lo = (int)((int64_t)m * blockIdx.x / gridDim.x);
        const int hi = (int)((int64_t)m * (blockIdx.x + 1) / gridDim.x);
        for (int t = lo + tid; t < hi; t += ES_T)
            reinterpret_cast<unsigned long long *>(a.rclr)[t] = 0ull;
    }
    __syncthreads();

    if (a.solve) {
        const double2 *twx = tw, *twy = tw + nx / 2;
        // forward along x of the packed rows; split into the transforms of
        // rows j, j + ny/2: A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / 2i
        fft_lanes<NR / 2>(zr, i, a.lognx, twx, false);
#pragma unroll
        for (int r = 0; r < NR / 2; ++r) {
            const int j = j0 + r * jstep;
            if (j >= ny / 2) continue;
            double2 *row = X + j * PX;
            row[ib] = zr[r];
            const double2 zk = row[i], zm = row[(nx - i) & mx];
            row[i] = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
            X[(j + ny / 2) * PX + i] = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
        }
        __syncthreads();
        // forward along y of the half spectrum, columns 0 .. nx/2
        constexpr int NC = NR - 1;  // (nx/2 + 1) ny / (64 ES_WAVES) <= 3
        double2 zc[NC];
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            const int col = c0 + r * cstep;
            zc[r] = make_double2(0.0, 0.0);
            if (col <= nx / 2) zc[r] = X[jp * PX + col];
        }
        fft_lanes<NC>(zc, jp, a.logny, twy, false);
#pragma unroll
        for (int r = 0; r < NC; ++r)
            if (c0 + r * cstep <= nx / 2) X[jb * PX + c0 + r * cstep] = zc[r];
        __syncthreads();
        // spectral values, then inverse along y
        double2 z[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int col = c0 + r * cstep;
            z[r] = make_double2(0.0, 0.0);
            if (col < nx) z[r] = es_spectral(X, PX, nx, ny, col, jp, tkx, tky, a.scale);
        }
        __syncthreads();
        fft_lanes<NR>(z, jp, a.logny, twy, true);
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (c0 + r * cstep < nx) X[jb * PX + c0 + r * cstep] = z[r];
        __syncthreads();
        // inverse along x: (ex, ey) = (Re, Im); energy sums
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int j = j0 + r * jstep;
            z[r] = make_double2(0.0, 0.0);
            if (j < ny) z[r] = X[j * PX + i];
        }
        fft_lanes<NR>(z, i, a.lognx, twx, true);
        double sx = 0.0, sy = 0.0;
        const bool store = a.write_e && field;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int j = j0 + r * jstep;
            if (j >= ny) continue;
            X[j * PX + ib] = z[r];
            sx += z[r].x * z[r].x;
            sy += z[r].y * z[r].y;
            if (store) {
                const int64_t g = (int64_t)j * nx + ib;
                a.ex[g] = z[r].x;
                a.ey[g] = z[r].y;
                a.e32[g] = make_float2((float)z[r].x, (float)z[r].y);
            }
        }
        const bool energy = field && a.slot >= 0;
        if (energy) {  // fixed-order sums: wave trees, then the waves in order
            for (int o = 32; o > 0; o >>= 1) {
                sx += __shfl_xor(sx, o);
                sy += __shfl_xor(sy, o);
            }
            if (lane == 0) {
                red[wave] = sx;
                red[ES_WAVES + wave] = sy;
            }
        }
        __syncthreads();
        if (energy && tid == 0) {
            double tx = 0.0, ty = 0.0;
            for (int k = 0; k < ES_WAVES; ++k) {
                tx += red[k];
                ty += red[ES_WAVES + k];
            }
            a.hist[2 * a.slot] = tx * a.dxdy;
            a.hist[2 * a.slot + 1] = ty * a.dxdy;
        }
    }

    while (p < a.n) {
        if (F32) {
            int ci = pc & mx, cj = pc >> a.lognx;
            float fx = po.x, fy = po.y, ux = po.z, uy = po.w;
            if (gather) {
                float e1, e2, b3 = 0.0f;
                interp32<K, BZ>(
                    fx, fy, ci, cj,
                    [&](int u, int v) {
                        const double2 e = X[((cj - HL + v) & my) * PX + ((ci - HL + u) & mx)];
                        return make_float2((float)e.x, (float)e.y);
                    },
                    [&](int, int) { return 0.0f; }, e1, e2, b3);
                kick32<BZ>(a.hdt, a.bz0, e1, e2, b3, ux, uy);
            }
            float nfx, nfy;
            const int si = drift1s(fx, ux, a.dtdx, nfx);
            const int sj = drift1s(fy, uy, a.dtdy, nfy);
            ci = (ci + si) & mx;
            cj = (cj + sj) & my;
            fx = nfx;
            fy = nfy;
            a.cell[p] = ci + nx * cj;
            a.ov[p] = make_float4(fx, fy, ux, uy);
            if (deposit)
                charge32_fixed<K>(fx, fy, ci, cj, a.wu, a.R, a.conserve != 0,
                                  [&](int u, int v, int q) {
                                      __hip_atomic_fetch_add(
                                          &itile[((cj - HL + v) & my) * nx + ((ci - HL + u) & mx)],
                                          q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                  });
        } else {
            if (gather) {
                double fx, fy;
                const int ci = ref_cell_s(px, s.dx, nx, false, fx);
                const int cj = ref_cell_s(py, s.dy, ny, false, fy);
                double e1, e2, b3 = 0.0;
                interp64<K, BZ>(
                    fx, fy, ci, cj,
                    [&](int u, int v) {
                        return X[((cj - HL + v) & my) * PX + ((ci - HL + u) & mx)];
                    },
                    [&](int, int) { return 0.0; }, e1, e2, b3);
                kick64<BZ>(s, e1, e2, b3, pvx, pvy);
                s.vx[p] = pvx;
                s.vy[p] = pvy;
            }
            if (a.mode == 1) drift64<2>(s, pvx, pvy, px, py);
            else drift64<1>(s, pvx, pvy, px, py);
            s.x[p] = px;
            s.y[p] = py;
            if (deposit) {
                double fx, fy;
                const int ci = ref_cell_s(px, s.dx, nx, K == PIC_CIC, fx);
                const int cj = ref_cell_s(py, s.dy, ny, K == PIC_CIC, fy);
                charge64<K>(fx, fy, ci, cj, pw, [&](int u, int v, double q) {
                    __hip_atomic_fetch_add(
                        &dtile[((cj - HL + v) & my) * nx + ((ci - HL + u) & mx)], q,
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                });
            }
        }
        p += stride;
        if (p < a.n) {
            if (F32) {
                pc = a.cell[p];
                po = a.ov[p];
            } else {
                px = s.x[p];
                py = s.y[p];
                pvx = s.vx[p];
                pvy = s.vy[p];
                pw = s.w ? s.w[p] : s.w0;
            }
        }
    }
    if (!deposit || field) return;
    __syncthreads();
    // each block starts at its own node, so that the blocks' atomics spread
    // over the grid instead of meeting at the same addresses (measured: 0.7
    // us less per launch on 64 x 64)
    const int off = (int)((int64_t)blockIdx.x * m / a.nbp);
    for (int t0 = tid; t0 < m; t0 += ES_T) {
        const int t = (t0 + off) & (m - 1);
        if (F32) {
            const int q = itile[t];
            if (q) atomicAdd(reinterpret_cast<unsigned long long *>(a.racc) + t,
                             (unsigned long long)(long long)q);
        } else {
            const double q = dtile[t];
            if (q != 0.0) unsafeAtomicAdd(reinterpret_cast<double *>(a.racc) + t, q);
        }
    }
}

// Twiddles exp(-2 pi i k / N) of the transforms along x (k < nx/2), then y.
__global__ void k_es_twiddles(int nx, int ny, int lognx, int logny, double2 *tw) {
    const int k = threadIdx.x;
    if (k >= nx / 2 + ny / 2) return;
    const bool along_x = k < nx / 2;
    const int kk = along_x ? k : k - nx / 2;
    double sn, cs;
    sincospi(ldexp((double)(-2 * kk), -(along_x ? lognx : logny)), &sn, &cs);
    tw[k] = make_double2(cs, sn);
}

// The raw charge of an accumulator buffer as FP64 (FP32 state: quanta times
// the unit).
__global__ void k_es_raw(const void *__restrict__ buf, int quanta, double unit, int64_t m,
                         double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    out[k] = quanta ? (double)reinterpret_cast<const long long *>(buf)[k] * unit
                    : reinterpret_cast<const double *>(buf)[k];
}

static bool pow2_within(int n, int lo, int hi) {
    return n >= lo && n <= hi && (n & (n - 1)) == 0;
}

static int log2_of(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

bool es_step_ok(const Ctx *c) {
    return c->fused_step && pow2_within(c->nx, 4, ES_MAX_SIDE) &&
           pow2_within(c->ny, 4, ES_MAX_SIDE) && c->bz_mode != PIC_BZ_GRID &&
           (c->state == PIC_STATE_F64 || c->f32_acc == 2);
}

static size_t es_lds(const Ctx *c) {
    const size_t nx = c->nx, ny = c->ny;
    return (ny * (nx + 1) + nx / 2 + ny / 2) * sizeof(double2) +
           2 * (nx / 2 + 1) * ny * sizeof(double) + nx * ny * (c->state == PIC_STATE_F32 ? 4 : 8);
}

template <int K, bool F32>
static int es_launch(const Ctx *c, const EsStep &a, size_t sm) {
    const int nb = a.nbp + (a.solve ? 1 : 0);
    auto go = [&](auto kern) -> int {
        PIC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(ES_T), sm, c->stream, a);
        PIC_HIP(hipGetLastError());
        return PIC_OK;
    };
    if (c->bz_mode == PIC_BZ_UNIFORM && a.mode != 0) return go(k_es_step<K, F32, PIC_BZ_UNIFORM>);
    return go(k_es_step<K, F32, PIC_BZ_NONE>);
}

// pic_run's loop (pic_run.hip) as one launch per iteration: a prologue
// launch (drift dt/2 + deposit), then launch s solves the deposit of launch
// s - 1, records energy s and moves the particles on (the last one without
// deposit).  Leaves rho of the last solve's deposit (finalize_rho), ex, ey
// and e32 of the last solve.
int es_run(Ctx *c, int kernel, double dt, int nsteps) {
    const int64_t m = (int64_t)c->nx * c->ny;
    // three raw charge buffers, then the twiddles
    const size_t ntw = (size_t)(c->nx / 2 + c->ny / 2);
    if (!c->es_buf) PIC_HIP(hipMalloc(&c->es_buf, (size_t)(3 * m) * 8 + ntw * sizeof(double2)));
    if (c->pending_perm) PIC_TRY(settle_state(c));
    const bool f32 = c->state == PIC_STATE_F32;
    // particle blocks: ~ES_BLOCK_PARTICLES each, one per CU at most, one CU
    // left to the field block
    int nb = (int)std::min<int64_t>(c->ncu - 1, (c->n + ES_BLOCK_PARTICLES - 1) / ES_BLOCK_PARTICLES);
    if (nb < 1) nb = 1;
    EsStep a = {};
    a.nx = c->nx;
    a.ny = c->ny;
    a.lognx = log2_of(c->nx);
    a.logny = log2_of(c->ny);
    a.nbp = nb;
    double2 *tw = reinterpret_cast<double2 *>(c->es_buf + 3 * m);
    hipLaunchKernelGGL(k_es_twiddles, dim3(1), dim3(64), 0, c->stream, a.nx, a.ny, a.lognx,
                       a.logny, tw);
    PIC_HIP(hipGetLastError());
    a.tw = tw;
    a.n = c->n;
    SegF64 &s = a.s64;
    s.x = c->x; s.y = c->y; s.vx = c->vx; s.vy = c->vy;
    s.w = c->uniform_w ? nullptr : c->w;
    s.w0 = c->w0;
    s.bz0 = c->bz0;
    s.nx = c->nx; s.ny = c->ny; s.ncells = (int)m;
    s.dx = c->dx; s.dy = c->dy; s.dimx = c->dimx; s.dimy = c->dimy;
    s.dt = dt;
    s.dtd = 0.5 * dt;
    a.cell = c->cell;
    a.ov = c->ov32;
    a.hdt = (float)(0.5 * dt);
    a.bz0 = (float)c->bz0;
    if (f32) {
        // one quantum for every block: the largest particle count of a block
        const int64_t per_round = (int64_t)nb * ES_T;
        const int64_t count = (c->n + per_round - 1) / per_round * ES_T;
        const float w = (float)c->w0;
        double qmax = 1.0;
        PIC_SHAPE_SWITCH(kernel, qmax = Shape<KK>::QMAX);
        const int qe = fixed_exponent(w, qmax, count);
        a.wu = ldexpf(fabsf(w), -qe);
        a.conserve = count <= CONSERVE_MAX_COUNT;
        a.R = a.conserve ? (int)rintf(a.wu) : 0;
        a.unit = (a.conserve && a.R != 0) ? (double)w / a.R : copysign(ldexp(1.0, qe), (double)w);
    }
    a.kxk2 = c->kx_k2;
    a.kyk2 = c->ky_k2;
    a.scale = 1.0 / ((double)c->nx * (double)c->ny * c->dx * c->dy);
    a.dxdy = c->dx * c->dy;
    a.ex = c->ex;
    a.ey = c->ey;
    a.e32 = reinterpret_cast<float2 *>(c->e32);
    a.hist = c->energy_hist;
    const size_t sm = es_lds(c);
    double *buf[3] = {c->es_buf, c->es_buf + m, c->es_buf + 2 * m};
    auto launch = [&](const EsStep &e) -> int {
        int rc = PIC_OK;
        PIC_SHAPE_SWITCH(kernel, rc = f32 ? es_launch<KK, true>(c, e, sm)
                                          : es_launch<KK, false>(c, e, sm));
        return rc;
    };
    PIC_TRY(launch_fill(c, buf[0], 0.0, m));
    c->sort_fresh = false;  // the launches move the particles
    // prologue: drift dt/2 and deposit into buffer 0, buffer 1 cleared
    EsStep e = a;
    e.mode = 0;
    e.solve = 0;
    e.slot = -1;
    e.dtdx = (float)(0.5 * dt / c->dx);
    e.dtdy = (float)(0.5 * dt / c->dy);
    e.racc = buf[0];
    e.rclr = buf[1];
    PIC_TRY(launch(e));
    std::vector<hipEvent_t> ev;
    int rc = PIC_OK;
    for (int st = 0; st < nsteps && rc == PIC_OK; ++st) {
        const bool last = st == nsteps - 1;
        e = a;
        e.mode = last ? 2 : 1;
        e.solve = 1;
        e.slot = st;
        e.write_e = last;
        // FP32: one drift of dt (mode 1) or dt/2; FP64 drifts by dt/2 once or twice
        const double dtd = last ? 0.5 * dt : dt;
        e.dtdx = (float)(dtd / c->dx);
        e.dtdy = (float)(dtd / c->dy);
        e.rin = buf[st % 3];
        e.racc = buf[(st + 1) % 3];
        e.rclr = buf[(st + 2) % 3];
        // events around every timing_every-th launch
        const bool timed = c->timing && st % c->timing_every == 0;
        auto mark = [&]() -> int {
            hipEvent_t h;
            if (hipEventCreate(&h) != hipSuccess) return fail(PIC_ERR_HIP, "event create failed");
            ev.push_back(h);
            (void)hipEventRecord(h, c->stream);
            return PIC_OK;
        };
        if (timed) rc = mark();
        if (rc == PIC_OK) rc = launch(e);
        if (rc == PIC_OK && timed) rc = mark();
    }
    if (c->timing && rc == PIC_OK && !ev.empty() &&
        hipEventSynchronize(ev.back()) == hipSuccess) {
        for (size_t k = 0; k + 1 < ev.size(); k += 2) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            c->t_pass_ms += ms;
            c->t_passes += 1;
        }
    }
    for (hipEvent_t h : ev) (void)hipEventDestroy(h);
    PIC_TRY(rc);
    // the charge the last solve read, as the raw accumulator, then rho
    const void *last_in = nsteps > 0 ? buf[(nsteps - 1) % 3] : buf[0];
    hipLaunchKernelGGL(k_es_raw, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                       last_in, f32 ? 1 : 0, a.unit, m, c->rho_acc);
    PIC_HIP(hipGetLastError());
    return finalize_rho(c, kernel);
}

}  // namespace pic

using namespace pic;

extern "C" int pic_set_fused_step(pic_ctx *ctx, int fused) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    ctx_of(ctx)->fused_step = fused != 0;
    return PIC_OK;
}
