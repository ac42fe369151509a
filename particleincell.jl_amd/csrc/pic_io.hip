// pic_io.hip -- particle initial conditions made on the device and particle
// files: the Landau quiet start of sample!(pg, mesh, LandauDamping(alpha, kx))
// (src/landau_damping.jl:77-97) as one kernel, a seeded perturbed
// Maxwellian (the strong-field workload's start), and read_particles
// (src/read_particles.jl:3-35) with its "ix iy dpx dpy vx vy" line format,
// which maps onto the FP32 cell-relative state one to one.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pic_internal.h"

namespace pic {

namespace {

// Direction numbers of the first three Sobol dimensions, 30 bits: the van
// der Corput sequence, then the primitive polynomials x + 1 (m = 1) and
// x^2 + x + 1 (m = 1, 3) of Joe and Kuo's table, by their recurrence
// m_k = 2 a_1 m_{k-1} ^ ... ^ 2^s m_{k-s} ^ m_{k-s}.
constexpr int SOBOL_BITS = 30;

struct SobolDirs {
    uint32_t v[3][SOBOL_BITS];
};

SobolDirs sobol_dirs() {
    SobolDirs d;
    uint32_t m1[SOBOL_BITS], m2[SOBOL_BITS];
    m1[0] = 1;
    for (int k = 1; k < SOBOL_BITS; ++k) m1[k] = (2u * m1[k - 1]) ^ m1[k - 1];
    m2[0] = 1;
    m2[1] = 3;
    for (int k = 2; k < SOBOL_BITS; ++k) m2[k] = (2u * m2[k - 1]) ^ (4u * m2[k - 2]) ^ m2[k - 2];
    for (int k = 0; k < SOBOL_BITS; ++k) {
        const int sh = SOBOL_BITS - 1 - k;
        d.v[0][k] = 1u << sh;
        d.v[1][k] = m1[k] << sh;
        d.v[2][k] = m2[k] << sh;
    }
    return d;
}

// sample!(pg, mesh, LandauDamping(alpha, kx)): landau_damping.jl:77-97 with
// the Sobol point i (1-based; the sequence's all-zero point 0 is not used),
// newton() of :19-28 (the interleaved pair update, at most 200 iterations).
__global__ void k_sample_landau(int64_t n, SobolDirs d, double alpha, double kx, double ymin,
                                double dimy, double *x, double *y, double *vx, double *vy) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t i = (uint64_t)k + 1;
    uint64_t g = i ^ (i >> 1);
    uint32_t s0 = 0, s1 = 0, s2 = 0;
    for (int b = 0; g != 0 && b < SOBOL_BITS; ++b, g >>= 1)
        if (g & 1) {
            s0 ^= d.v[0][b];
            s1 ^= d.v[1][b];
            s2 ^= d.v[2][b];
        }
    const double scale = 1.0 / (double)(1u << SOBOL_BITS);
    const double r1 = s0 * scale, r2 = s1 * scale, r3 = s2 * scale;
    const double r = r1 * (2 * M_PI / kx);
    double x0 = 0.0, x1 = 1.0;
    for (int it = 0; it < 200 && fabs(x1 - x0) > 1e-12; ++it) {
        const double p = x0 + alpha * sin(kx * x0) / kx;
        const double f = 1 + alpha * cos(kx * x0);
        const double nx1 = x0 - (p - r) / f;
        x0 = x1;
        x1 = nx1;
    }
    const double v = sqrt(-2 * log(((double)i - 0.5) / (double)n));
    const double theta = r3 * 2 * M_PI;
    x[k] = x1;
    y[k] = ymin + r2 * dimy;
    vx[k] = v * cos(theta);
    vy[k] = v * sin(theta);
}

// ------------------------------------------- perturbed Maxwellian, seeded

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32 x 32 -> 64-bit products
// on a 128-bit counter under a 64-bit key bumped by the Weyl constants.
// Counter-based: particle k draws from counters (k, 0, j, 0), so the sample
// does not depend on the launch shape.
struct U4 {
    uint32_t a, b, c, d;
};

__device__ __forceinline__ U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * ctr.a;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr.c;
        ctr = U4{(uint32_t)(p1 >> 32) ^ ctr.b ^ k0, (uint32_t)p1,
                 (uint32_t)(p0 >> 32) ^ ctr.d ^ k1, (uint32_t)p0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ctr;
}

// A double in [0, 1) from 53 bits of two words.
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    return (double)(((uint64_t)(hi >> 5) << 26) | (lo >> 6)) * 0x1.0p-53;
}

// n particles of density 1 + alpha cos(kx (x - xmin)) cos(ky (y - ymin)) on
// the mesh and Maxwellian velocities of width vth.  Integrated over whole
// periods in y the density is uniform in x, so x is uniform; y then follows
// its conditional density 1 + a cos(ky y'), a = alpha cos(kx (x - xmin)),
// whose distribution function y' + a sin(ky y') / ky = u dimy is inverted
// by Newton's method (|a| < 1 keeps it monotone).  Velocities by the
// Box-Muller transform.  Particle k takes the words of counters (k, 0, 0, 0)
// and (k, 0, 1, 0): x, y from the first, the velocity from the second.
__global__ void k_sample_maxwellian(int64_t n, uint32_t k0, uint32_t k1, double alpha, double kx,
                                    double ky, double xmin, double dimx, double ymin, double dimy,
                                    double vth, double *x, double *y, double *vx, double *vy) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t lo = (uint32_t)k, hi = (uint32_t)((uint64_t)k >> 32);
    const U4 r = philox4x32_10(U4{lo, hi, 0u, 0u}, k0, k1);
    const U4 s = philox4x32_10(U4{lo, hi, 1u, 0u}, k0, k1);
    const double xr = u53(r.a, r.b) * dimx;
    const double a = alpha * cos(kx * xr);
    const double target = u53(r.c, r.d) * dimy;
    double yr = target;
    for (int it = 0; it < 60; ++it) {
        const double f = yr + a * sin(ky * yr) / ky - target;
        const double step = f / (1 + a * cos(ky * yr));
        yr -= step;
        if (fabs(step) <= 1e-15 * dimy) break;
    }
    yr = fmin(fmax(yr, 0.0), dimy);
    if (yr >= dimy) yr = 0.0;
    const double rad = vth * sqrt(-2 * log(1.0 - u53(s.a, s.b)));  // 1 - u in (0, 1]
    const double th = 2 * M_PI * u53(s.c, s.d);
    x[k] = xmin + xr;
    y[k] = ymin + yr;
    vx[k] = rad * cos(th);
    vy[k] = rad * sin(th);
}

// ------------------------------------------------------------ text files

struct Line {
    int32_t ix, iy;
    double dpx, dpy, vx, vy;
};

// Parses the lines of [b, e) (blank lines skipped); returns the 1-based
// number, within the chunk, of the first malformed line, or 0.
int64_t parse_chunk(const char *b, const char *e, std::vector<Line> &out) {
    int64_t ln = 0;
    while (b < e) {
        const char *eol = static_cast<const char *>(memchr(b, '\n', (size_t)(e - b)));
        if (!eol) eol = e;
        ++ln;
        std::string s(b, eol);
        b = eol + 1;
        if (s.find_first_not_of(" \t\r") == std::string::npos) continue;
        const char *p = s.c_str();
        char *q = nullptr;
        Line L;
        errno = 0;
        const long ix = strtol(p, &q, 10);
        if (q == p) return ln;
        p = q;
        const long iy = strtol(p, &q, 10);
        if (q == p || errno) return ln;
        p = q;
        double vals[4];
        for (double &v : vals) {
            v = strtod(p, &q);
            if (q == p) return ln;
            p = q;
        }
        while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
        if (*p != '\0') return ln;
        if (ix < INT32_MIN || ix > INT32_MAX || iy < INT32_MIN || iy > INT32_MAX) return ln;
        L.ix = (int32_t)ix;
        L.iy = (int32_t)iy;
        L.dpx = vals[0];
        L.dpy = vals[1];
        L.vx = vals[2];
        L.vy = vals[3];
        out.push_back(L);
    }
    return 0;
}

int read_lines(const char *path, std::vector<Line> &all) {
    FILE *f = fopen(path, "rb");
    PIC_CHECK(f, PIC_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    std::string buf;
    char tmp[1 << 16];
    size_t got;
    while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, got);
    const bool bad = ferror(f) != 0;
    fclose(f);
    PIC_CHECK(!bad, PIC_ERR_ARG, "read error on %s", path);
    // chunks at line boundaries, parsed in parallel
    const size_t nth = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
    const size_t want = std::max<size_t>(1, std::min(nth, buf.size() / (1 << 20) + 1));
    std::vector<size_t> cut{0};
    for (size_t t = 1; t < want; ++t) {
        size_t pos = buf.size() * t / want;
        const size_t nl = buf.find('\n', std::max(pos, cut.back()));
        if (nl == std::string::npos) break;
        cut.push_back(nl + 1);
    }
    cut.push_back(buf.size());
    const size_t nc = cut.size() - 1;
    std::vector<std::vector<Line>> parts(nc);
    std::vector<int64_t> bad_line(nc, 0), lines_in(nc, 0);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nc; ++t)
        th.emplace_back([&, t]() {
            const char *b = buf.data() + cut[t], *e = buf.data() + cut[t + 1];
            bad_line[t] = parse_chunk(b, e, parts[t]);
            lines_in[t] = std::count(b, e, '\n');
        });
    for (auto &x : th) x.join();
    int64_t before = 0;
    for (size_t t = 0; t < nc; ++t) {
        PIC_CHECK(bad_line[t] == 0, PIC_ERR_ARG, "%s:%lld: expected \"ix iy dpx dpy vx vy\"", path,
                  (long long)(before + bad_line[t]));
        before += lines_in[t];
    }
    size_t total = 0;
    for (auto &p : parts) total += p.size();
    all.clear();
    all.reserve(total);
    for (auto &p : parts) all.insert(all.end(), p.begin(), p.end());
    return PIC_OK;
}

}  // namespace

}  // namespace pic

using namespace pic;

extern "C" {

int pic_sample_landau(pic_ctx *ctx, int64_t n, double alpha, double kx, int state) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(n >= 0 && n < (int64_t)1 << SOBOL_BITS, PIC_ERR_ARG, "particle count %lld",
              (long long)n);
    PIC_CHECK(kx > 0.0, PIC_ERR_ARG, "kx must be positive");
    PIC_CHECK(state == PIC_STATE_F64 || state == PIC_STATE_F32, PIC_ERR_ARG, "state %d", state);
    Ctx *c = ctx_of(ctx);
    PIC_CHECK(fabs(c->dimx - 2 * M_PI / kx) <= 1e-12 * c->dimx, PIC_ERR_ARG,
              "the Landau sampler needs dimx == 2pi/kx (landau_damping.jl:81)");
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_particles(c, n));
    c->n = n;
    c->state = PIC_STATE_F64;
    c->seg_valid = false;
    c->sort_fresh = false;
    c->uniform_w = true;
    c->w0 = n ? (c->dimx * c->dimy) / (double)n : 0.0;
    if (n) {
        hipLaunchKernelGGL(k_sample_landau, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           c->stream, n, sobol_dirs(), alpha, kx, c->ymin, c->dimy, c->x, c->y,
                           c->vx, c->vy);
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(pic_set_state(ctx, state));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_sample_maxwellian(pic_ctx *ctx, int64_t n, double alpha, double kx, double ky,
                          double vth, uint64_t seed, int state) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(n >= 0 && n <= INT32_MAX, PIC_ERR_ARG, "particle count %lld", (long long)n);
    PIC_CHECK(fabs(alpha) < 1.0, PIC_ERR_ARG, "|alpha| must be below 1");
    PIC_CHECK(kx > 0.0 && ky > 0.0 && vth >= 0.0, PIC_ERR_ARG, "kx, ky > 0 and vth >= 0");
    PIC_CHECK(state == PIC_STATE_F64 || state == PIC_STATE_F32, PIC_ERR_ARG, "state %d", state);
    Ctx *c = ctx_of(ctx);
    // the conditional inversion needs whole periods of cos(ky y) in the mesh
    const double periods = ky * c->dimy / (2 * M_PI);
    PIC_CHECK(periods >= 0.5 && fabs(periods - nearbyint(periods)) <= 1e-9 * periods,
              PIC_ERR_ARG, "the mesh's dimy must hold whole periods 2pi/ky");
    PIC_HIP(hipSetDevice(c->device));
    PIC_TRY(ensure_particles(c, n));
    c->n = n;
    c->state = PIC_STATE_F64;
    c->seg_valid = false;
    c->sort_fresh = false;
    c->uniform_w = true;
    c->w0 = n ? (c->dimx * c->dimy) / (double)n : 0.0;
    if (n) {
        hipLaunchKernelGGL(k_sample_maxwellian, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           c->stream, n, (uint32_t)seed, (uint32_t)(seed >> 32), alpha, kx, ky,
                           c->xmin, c->dimx, c->ymin, c->dimy, vth, c->x, c->y, c->vx, c->vy);
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(pic_set_state(ctx, state));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_read_particles(pic_ctx *ctx, const char *path, int state, int64_t *n_out) {
    PIC_CHECK(ctx && path, PIC_ERR_ARG, "null pointer");
    PIC_CHECK(state == PIC_STATE_F64 || state == PIC_STATE_F32, PIC_ERR_ARG, "state %d", state);
    Ctx *c = ctx_of(ctx);
    std::vector<Line> L;
    PIC_TRY(read_lines(path, L));
    const int64_t n = (int64_t)L.size();
    const double w = n ? (c->dimx * c->dimy) / (double)n : 0.0;
    if (n_out) *n_out = n;
    if (state == PIC_STATE_F32) {
        std::vector<int32_t> cell(n);
        std::vector<float> ox(n), oy(n), vx(n), vy(n);
        for (int64_t k = 0; k < n; ++k) {
            PIC_CHECK(L[k].ix >= 0 && L[k].ix < c->nx && L[k].iy >= 0 && L[k].iy < c->ny,
                      PIC_ERR_ARG, "%s: particle %lld in cell (%d, %d) outside the %d x %d mesh",
                      path, (long long)k + 1, L[k].ix, L[k].iy, c->nx, c->ny);
            // an offset just below 1 that rounds to 1.0f starts the next cell
            int ci = L[k].ix, cj = L[k].iy;
            ox[k] = (float)L[k].dpx;
            oy[k] = (float)L[k].dpy;
            if (ox[k] == 1.0f) {
                ox[k] = 0.0f;
                ci = ci + 1 == c->nx ? 0 : ci + 1;
            }
            if (oy[k] == 1.0f) {
                oy[k] = 0.0f;
                cj = cj + 1 == c->ny ? 0 : cj + 1;
            }
            cell[k] = ci + c->nx * cj;
            vx[k] = (float)L[k].vx;
            vy[k] = (float)L[k].vy;
        }
        return pic_set_state_f32(ctx, cell.data(), ox.data(), oy.data(), vx.data(), vy.data(), w,
                                 n);
    }
    // read_particles.jl:28-31
    std::vector<double> x((size_t)2 * n), v((size_t)2 * n);
    for (int64_t k = 0; k < n; ++k) {
        x[2 * k] = (L[k].dpx + L[k].ix) * c->dx;
        x[2 * k + 1] = (L[k].dpy + L[k].iy) * c->dy;
        v[2 * k] = L[k].vx;
        v[2 * k + 1] = L[k].vy;
    }
    return pic_set_particles_xv(ctx, x.data(), v.data(), w, n);
}

int pic_write_particles(pic_ctx *ctx, const char *path) {
    PIC_CHECK(ctx && path, PIC_ERR_ARG, "null pointer");
    Ctx *c = ctx_of(ctx);
    const int64_t n = c->n;
    std::string out;
    out.reserve((size_t)n * 96);
    char line[256];
    if (c->state == PIC_STATE_F32) {
        std::vector<int32_t> cell(n);
        std::vector<float> ox(n), oy(n), vx(n), vy(n);
        if (n)
            PIC_TRY(pic_get_state_f32(ctx, cell.data(), ox.data(), oy.data(), vx.data(),
                                      vy.data()));
        for (int64_t k = 0; k < n; ++k) {
            const int m = snprintf(line, sizeof line, "%d %d %.9g %.9g %.9g %.9g\n",
                                   cell[k] % c->nx, cell[k] / c->nx, ox[k], oy[k], vx[k], vy[k]);
            out.append(line, (size_t)m);
        }
    } else {
        std::vector<double> x((size_t)2 * n), v((size_t)2 * n);
        if (n) PIC_TRY(pic_get_particles_xv(ctx, x.data(), v.data()));
        for (int64_t k = 0; k < n; ++k) {
            const double px = x[2 * k] / c->dx, py = x[2 * k + 1] / c->dy;
            const double fi = floor(px), fj = floor(py);
            const int m = snprintf(line, sizeof line, "%d %d %.17g %.17g %.17g %.17g\n", (int)fi,
                                   (int)fj, px - fi, py - fj, v[2 * k], v[2 * k + 1]);
            out.append(line, (size_t)m);
        }
    }
    FILE *f = fopen(path, "wb");
    PIC_CHECK(f, PIC_ERR_ARG, "cannot create %s: %s", path, strerror(errno));
    const size_t put = fwrite(out.data(), 1, out.size(), f);
    const int rc = fclose(f);
    PIC_CHECK(put == out.size() && rc == 0, PIC_ERR_ARG, "write error on %s", path);
    return PIC_OK;
}

}  // extern "C"
