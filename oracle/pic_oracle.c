/*
 * pic_oracle.c -- CPU restatement of ParticleInCell.jl's 2D particle-mesh step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only to check or time it.  The product path never calls it.
 *
 * Conventions
 *   - Every array is column-major as in Julia: element (i, j) (1-based in the
 *     reference) lives at C index (i-1) + ld*(j-1).
 *   - A particle array `p` is the reference's ParticleGroup{2,2}.array, 5 x n
 *     column-major: p[5*k + 0..4] = x, y, vx, vy, w of particle k.
 *   - Operation order follows the cited reference lines so that IEEE results
 *     agree bit for bit where the reference fixes them.  Build with
 *     -ffp-contract=off (no FMA contraction), as Julia and the reference's
 *     Fortran twin (x86-64 baseline) do.
 *
 * Deviation from the reference, applied identically on the GPU: a particle
 * sitting exactly on x == dimx (push_x! keeps it, pushers_cic.jl:64-67 uses a
 * strict `>`) gets cell index nx+1 in the reference, which is an out-of-bounds
 * access (silent in compute_rho's @inbounds loop, a BoundsError in push_v!).
 * Here that index is wrapped periodically to 1, which is what the Fortran twin
 * does implicitly through its ghost row/column (pic.f90:20-29).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IX(i, j, ld) ((size_t)(i) + (size_t)(ld) * (size_t)(j))

/* Julia mod1(x, n) = mod(x - 1, n) + 1 for integers. */
static inline int64_t mod1(int64_t x, int64_t n)
{
    int64_t r = (x - 1) % n;
    if (r < 0) r += n;
    return r + 1;
}

/* Julia mod(x, n) for integers (result in [0, n)). */
static inline int64_t imod(int64_t x, int64_t n)
{
    int64_t r = x % n;
    if (r < 0) r += n;
    return r;
}

/* Julia mod(x, y) for Float64 (base/float.jl): fmod, then shift into the sign
 * of y.  r + y can round up to y itself for tiny negative r. */
static inline double jl_fmod(double x, double y)
{
    double r = fmod(x, y);
    if (r == 0.0) return copysign(r, y);
    if ((r > 0) != (y > 0)) return r + y;
    return r;
}

/* x^n for Float64 and integer n >= 0, restating Julia's Base.Math.pow_body
 * (compensated power by squaring, used by `x^5` in compute_rho_m6.jl:14-18).
 * muladd is taken as a fused multiply-add, which is what Julia emits on a host
 * with FMA hardware.  The last-bit agreement with Julia is unverified here
 * (no Julia in this environment). */
static double jl_pow_int(double x, int64_t n)
{
    if (n == 0) return 1.0;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            double err = fma(y, xnlo, x * ynlo);
            double hi = x * y;
            double lo = fma(x, y, -hi);
            y = hi;
            ynlo = lo + err;
        }
        double err = x * 2 * xnlo;
        double hi = x * x;
        double lo = fma(x, x, -hi);
        x = hi;
        xnlo = lo + err;
        n >>= 1;
    }
    double err = fma(y, xnlo, x * ynlo);
    if (isfinite(x) && isfinite(err)) return fma(x, y, err);
    return x * y;
}

/* f_m6: compute_rho_m6.jl:10-25 (Monaghan M6 = centred quintic B-spline). */
double or_f_m6(double q)
{
    double f;
    if (q < 1.0)
        f = (jl_pow_int(3 - q, 5) - 6 * jl_pow_int(2 - q, 5)) + 15 * jl_pow_int(1 - q, 5);
    else if (q >= 1.0 && q < 2.0)
        f = jl_pow_int(3 - q, 5) - 6 * jl_pow_int(2 - q, 5);
    else if (q >= 2 && q < 3)
        f = jl_pow_int(3 - q, 5);
    else
        f = 0.0;
    return f / 120;
}

/* ------------------------------------------------------------------ CIC */

/* push_v!(p, ::CloudInCell, m, ex, ey, bz, dt): pushers_cic.jl:3-53.
 * Fields are nx x ny (ld = nx) as in the Julia method, or (nx+1) x (ny+1)
 * (ld = nx+1) as in the Fortran twin pic.f90:20-77; the +1 neighbours wrap
 * with mod1 as in pushers_cic.jl:31-32.  bz may be NULL (field identically 0).
 */
void or_push_v_cic(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                   const double *ex, const double *ey, const double *bz, int ld)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double *q = p + 5 * k;
        double v1 = q[2];
        double v2 = q[3];
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)floor(xp) + 1;
        int64_t j = (int64_t)floor(yp) + 1;
        double dxp = (xp - (double)i) + 1;
        double dyp = (yp - (double)j) + 1;
        double a1 = (1 - dxp) * (1 - dyp);
        double a2 = dxp * (1 - dyp);
        double a3 = dxp * dyp;
        double a4 = (1 - dxp) * dyp;
        int64_t ip1 = mod1(i + 1, nx);
        int64_t jp1 = mod1(j + 1, ny);
        i = mod1(i, nx);
        j = mod1(j, ny);
        size_t c00 = IX(i - 1, j - 1, ld), c10 = IX(ip1 - 1, j - 1, ld);
        size_t c11 = IX(ip1 - 1, jp1 - 1, ld), c01 = IX(i - 1, jp1 - 1, ld);
        double e1 = a1 * ex[c00] + a2 * ex[c10] + a3 * ex[c11] + a4 * ex[c01];
        double e2 = a1 * ey[c00] + a2 * ey[c10] + a3 * ey[c11] + a4 * ey[c01];
        double b3 = bz ? a1 * bz[c00] + a2 * bz[c10] + a3 * bz[c11] + a4 * bz[c01] : 0.0;
        v1 += 0.5 * dt * e1;
        v2 += 0.5 * dt * e2;
        double tantheta = 0.5 * dt * b3;
        double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
        v1 += v2 * tantheta;
        v2 += -v1 * sintheta;
        v1 += v2 * tantheta;
        q[2] = v1 + 0.5 * dt * e1;
        q[3] = v2 + 0.5 * dt * e2;
    }
}

/* push_x!(p, mesh, dt): pushers_cic.jl:55-72 (single conditional wrap, strict
 * comparisons, domain assumed to start at 0). */
void or_push_x(int64_t n, double *p, double dt, double dimx, double dimy)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double *q = p + 5 * k;
        double p1 = q[0] + dt * q[2];
        double p2 = q[1] + dt * q[3];
        if (p1 > dimx) p1 -= dimx;
        if (p2 > dimy) p2 -= dimy;
        if (p1 < 0.0) p1 += dimx;
        if (p2 < 0.0) p2 += dimy;
        q[0] = p1;
        q[1] = p2;
    }
}

/* Sum of an array the way the HIP library's FP64 reductions define it is not
 * the point here: Julia's `sum` is a pairwise reduction whose leaves are SIMD
 * loops, which no portable code reproduces bit for bit.  A pairwise sum keeps
 * the error at the same O(log n) level; parity on means is tolerance-only. */
static double pairwise_sum(const double *a, size_t n)
{
    if (n <= 128) {
        double s = 0.0;
        for (size_t k = 0; k < n; ++k) s += a[k];
        return s;
    }
    size_t h = n / 2;
    return pairwise_sum(a, h) + pairwise_sum(a + h, n - h);
}

/* compute_rho(p, ::CloudInCell, m): pushers_cic.jl:75-114.  Writes the
 * returned array rho .- mean(rho) into rho (nx x ny). */
void or_compute_rho_cic(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                        double *rho)
{
    memset(rho, 0, sizeof(double) * (size_t)nx * ny);
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p + 5 * k;
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)xp + 1; /* trunc, :86-87 */
        int64_t j = (int64_t)yp + 1;
        double dxp = (xp - (double)i) + 1;
        double dyp = (yp - (double)j) + 1;
        double a1 = (1 - dxp) * (1 - dyp);
        double a2 = dxp * (1 - dyp);
        double a3 = dxp * dyp;
        double a4 = (1 - dxp) * dyp;
        int64_t ip1 = mod1(i + 1, nx);
        int64_t jp1 = mod1(j + 1, ny);
        i = mod1(i, nx);
        j = mod1(j, ny);
        double w = q[4];
        rho[IX(i - 1, j - 1, nx)] += a1 * w;
        rho[IX(ip1 - 1, j - 1, nx)] += a2 * w;
        rho[IX(ip1 - 1, jp1 - 1, nx)] += a3 * w;
        rho[IX(i - 1, jp1 - 1, nx)] += a4 * w;
    }
    size_t m = (size_t)nx * ny;
    for (size_t k = 0; k < m; ++k) rho[k] /= (dx * dy);
    double mean = pairwise_sum(rho, m) / (double)m;
    for (size_t k = 0; k < m; ++k) rho[k] -= mean;
}

/* interpolate_cic!(epx, epy, mesh, p, ex, ey): interpolate_cic.jl:3-33
 * (trunc index, mod1 on +1 neighbours). ex, ey are nx x ny. */
void or_interpolate_cic(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                        const double *ex, const double *ey, double *epx, double *epy)
{
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p + 5 * k;
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)xp + 1;
        int64_t j = (int64_t)yp + 1;
        double dxp = (xp - (double)i) + 1;
        double dyp = (yp - (double)j) + 1;
        double a1 = (1 - dxp) * (1 - dyp);
        double a2 = dxp * (1 - dyp);
        double a3 = dxp * dyp;
        double a4 = (1 - dxp) * dyp;
        int64_t ip1 = mod1(i + 1, nx);
        int64_t jp1 = mod1(j + 1, ny);
        i = mod1(i, nx);
        j = mod1(j, ny);
        size_t c00 = IX(i - 1, j - 1, nx), c10 = IX(ip1 - 1, j - 1, nx);
        size_t c11 = IX(ip1 - 1, jp1 - 1, nx), c01 = IX(i - 1, jp1 - 1, nx);
        epx[k] = a1 * ex[c00] + a2 * ex[c10] + a3 * ex[c11] + a4 * ex[c01];
        epy[k] = a1 * ey[c00] + a2 * ey[c10] + a3 * ey[c11] + a4 * ey[c01];
    }
}

/* ------------------------------------------------------------------- M6 */

/* The 7 x 7 stencil of pushers_m6.jl: offsets -3..+3 around the node i with
 * weights f_m6(3+d), f_m6(2+d), f_m6(1+d), f_m6(d), f_m6(1-d), f_m6(2-d),
 * f_m6(3-d) (pushers_m6.jl:42-55).  The first weight is always 0 for
 * d in [0, 1); it is kept so the sums run over the reference's 49 terms. */
static inline void m6_weights(double d, double c[7])
{
    c[0] = or_f_m6(3 + d);
    c[1] = or_f_m6(2 + d);
    c[2] = or_f_m6(1 + d);
    c[3] = or_f_m6(d);
    c[4] = or_f_m6(1 - d);
    c[5] = or_f_m6(2 - d);
    c[6] = or_f_m6(3 - d);
}

/* push_v!(p, ::M6, m, ex, ey, bz, dt): pushers_m6.jl:5-224.  The 49-term sums
 * run x-offset outer, y-offset inner, each term (cx*cy)*field, as in :56-207.
 * Fields are nx x ny (ld = nx).  bz may be NULL. */
void or_push_v_m6(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                  const double *ex, const double *ey, const double *bz, int ld)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double *q = p + 5 * k;
        double v1 = q[2];
        double v2 = q[3];
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)floor(xp) + 1;
        int64_t j = (int64_t)floor(yp) + 1;
        double dpx = (xp - (double)i) + 1;
        double dpy = (yp - (double)j) + 1;
        int64_t ii[7], jj[7];
        for (int a = 0; a < 7; ++a) {
            ii[a] = mod1(i - 3 + a, nx);
            jj[a] = mod1(j - 3 + a, ny);
        }
        double cx[7], cy[7];
        m6_weights(dpx, cx);
        m6_weights(dpy, cy);
        double e1 = 0.0, e2 = 0.0, b3 = 0.0;
        for (int a = 0; a < 7; ++a)
            for (int b = 0; b < 7; ++b)
                e1 += cx[a] * cy[b] * ex[IX(ii[a] - 1, jj[b] - 1, ld)];
        for (int a = 0; a < 7; ++a)
            for (int b = 0; b < 7; ++b)
                e2 += cx[a] * cy[b] * ey[IX(ii[a] - 1, jj[b] - 1, ld)];
        if (bz)
            for (int a = 0; a < 7; ++a)
                for (int b = 0; b < 7; ++b)
                    b3 += cx[a] * cy[b] * bz[IX(ii[a] - 1, jj[b] - 1, ld)];
        v1 += 0.5 * dt * e1;
        v2 += 0.5 * dt * e2;
        double tantheta = 0.5 * dt * b3;
        double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
        v1 += v2 * tantheta;
        v2 += -v1 * sintheta;
        v1 += v2 * tantheta;
        q[2] = v1 + 0.5 * dt * e1;
        q[3] = v2 + 0.5 * dt * e2;
    }
}

/* compute_rho(p, ::M6, mesh): pushers_m6.jl:226-348.  Scatter order :284-338,
 * normalisation :342, mean removal rho_total/dimx/dimy :344-346. */
void or_compute_rho_m6(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                       double dimx, double dimy, double *rho)
{
    memset(rho, 0, sizeof(double) * (size_t)nx * ny);
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p + 5 * k;
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)floor(xp) + 1;
        int64_t j = (int64_t)floor(yp) + 1;
        double dpx = (xp - (double)i) + 1;
        double dpy = (yp - (double)j) + 1;
        double weight = q[4];
        int64_t ii[7], jj[7];
        for (int a = 0; a < 7; ++a) {
            ii[a] = mod1(i - 3 + a, nx);
            jj[a] = mod1(j - 3 + a, ny);
        }
        double cx[7], cy[7];
        m6_weights(dpx, cx);
        m6_weights(dpy, cy);
        for (int a = 0; a < 7; ++a)
            for (int b = 0; b < 7; ++b)
                rho[IX(ii[a] - 1, jj[b] - 1, nx)] += cx[a] * cy[b] * weight;
    }
    size_t m = (size_t)nx * ny;
    for (size_t k = 0; k < m; ++k) rho[k] /= (dx * dy);
    double rho_total = pairwise_sum(rho, m) * dx * dy;
    double shift = rho_total / dimx / dimy;
    for (size_t k = 0; k < m; ++k) rho[k] -= shift;
}

/* compute_rho_m6!(fields, particles): compute_rho_m6.jl:181-316, the
 * MeshFields API.  x is Particles.x (2 x n), rewritten in place; rho is
 * fields.rho, (nx+1) x (ny+1).
 *
 * compat = 1 restates the reference line for line, including two quirks:
 *   - xmin/xmax (and ymin/ymax) are swapped (:190-191), so the wrap is taken
 *     relative to xmax and the rewritten positions land in [xmax, xmax+dimx)
 *     (:197-201);
 *   - the neighbours use mod1(i+k, nx) on the 0-based trunc index (:216-227),
 *     which is the 1-based slot of node i+k-1: every off-centre weight is
 *     deposited one node to the left/below of where compute_rho(p, M6(), m)
 *     (pushers_m6.jl:256-267) puts it, and node i-1 receives nothing.
 * compat = 0 uses the corrected neighbours mod(i+k, nx)+1 (the form of the
 * UA variant :70-81 and of interpol_eb_m6!, interpolation_m6.jl:155-166) and
 * the unswapped xmin; the result then equals compute_rho(p, M6(), m) up to
 * rounding plus the ghost copy.
 * The println of :311 is not restated. */
void or_compute_rho_m6_fields(int64_t n, double *x, double w, int nx, int ny, double xmin,
                              double xmax, double ymin, double ymax, int compat, double *rho)
{
    int ld = nx + 1;
    memset(rho, 0, sizeof(double) * (size_t)(nx + 1) * (ny + 1));
    double dx = (xmax - xmin) / nx, dy = (ymax - ymin) / ny;
    double x0 = compat ? xmax : xmin;
    double y0 = compat ? ymax : ymin;
    double dimx = xmax - xmin, dimy = ymax - ymin;
    for (int64_t k = 0; k < n; ++k) {
        double xt1 = jl_fmod(x[2 * k] - x0, dimx);
        double xt2 = jl_fmod(x[2 * k + 1] - y0, dimy);
        x[2 * k] = xt1 + x0;
        x[2 * k + 1] = xt2 + y0;
        double px = xt1 / dx;
        double py = xt2 / dy;
        int32_t i = (int32_t)px;
        int32_t j = (int32_t)py;
        double dpx = px - (double)i;
        double dpy = py - (double)j;
        int64_t ii[7], jj[7];
        for (int a = 0; a < 7; ++a) {
            if (compat) {
                ii[a] = mod1((int64_t)i - 3 + a, nx);
                jj[a] = mod1((int64_t)j - 3 + a, ny);
            } else {
                ii[a] = imod((int64_t)i - 3 + a, nx) + 1;
                jj[a] = imod((int64_t)j - 3 + a, ny) + 1;
            }
        }
        /* the centre node is i+1 (1-based), unwrapped (:229-230) */
        ii[3] = (int64_t)i + 1;
        jj[3] = (int64_t)j + 1;
        double cx[7], cy[7];
        m6_weights(dpx, cx);
        m6_weights(dpy, cy);
        for (int a = 0; a < 7; ++a)
            for (int b = 0; b < 7; ++b)
                rho[IX(ii[a] - 1, jj[b] - 1, ld)] += cx[a] * cy[b] * w;
    }
    /* ghost copies :303-305 */
    for (int i = 0; i < nx; ++i) rho[IX(i, ny, ld)] = rho[IX(i, 0, ld)];
    for (int j = 0; j < ny; ++j) rho[IX(nx, j, ld)] = rho[IX(0, j, ld)];
    rho[IX(nx, ny, ld)] = rho[IX(0, 0, ld)];
    size_t m = (size_t)(nx + 1) * (ny + 1);
    for (size_t k = 0; k < m; ++k) rho[k] /= (dx * dy);
    double s = 0.0;
    for (int j = 0; j < ny; ++j) s += pairwise_sum(rho + IX(0, j, ld), (size_t)nx);
    double rho_total = s * dx * dy;
    double shift = rho_total / dimx / dimy;
    for (size_t k = 0; k < m; ++k) rho[k] -= shift;
}

/* interpol_eb_m6!(particles, fields): interpolation_m6.jl:125-247.  x is
 * Particles.x (2 x n, rewritten into [xmin, xmax)), e_f is fields.e
 * (2 x (nx+1) x (ny+1), component fastest), e_p is Particles.e (2 x n). */
void or_interpol_eb_m6(int64_t n, double *x, int nx, int ny, double xmin, double xmax,
                       double ymin, double ymax, const double *e_f, double *e_p)
{
    int ld = nx + 1;
    double dx = (xmax - xmin) / nx, dy = (ymax - ymin) / ny;
    double dimx = xmax - xmin, dimy = ymax - ymin;
    for (int64_t k = 0; k < n; ++k) {
        double xn = jl_fmod(x[2 * k] - xmin, dimx);
        double yn = jl_fmod(x[2 * k + 1] - ymin, dimy);
        double px = xn / dx;
        double py = yn / dy;
        int32_t i = (int32_t)px;
        int32_t j = (int32_t)py;
        double dpx = px - (double)i;
        double dpy = py - (double)j;
        x[2 * k] = xmin + xn;
        x[2 * k + 1] = ymin + yn;
        int64_t ii[7], jj[7];
        for (int a = 0; a < 7; ++a) {
            ii[a] = imod((int64_t)i - 3 + a, nx) + 1;
            jj[a] = imod((int64_t)j - 3 + a, ny) + 1;
        }
        ii[3] = (int64_t)i + 1;
        jj[3] = (int64_t)j + 1;
        double cx[7], cy[7];
        m6_weights(dpx, cx);
        m6_weights(dpy, cy);
        for (int l = 0; l < 2; ++l) {
            double e = 0.0;
            for (int a = 0; a < 7; ++a)
                for (int b = 0; b < 7; ++b)
                    e += cx[a] * cy[b] * e_f[l + 2 * IX(ii[a] - 1, jj[b] - 1, ld)];
            e_p[2 * k + l] = e;
        }
    }
}

/* -------------------------------------------------------------- Poisson */

/* In-place radix-2 complex FFT of length n (power of two) over elements
 * a[0], a[s], ..., a[(n-1)s].  sign = -1 forward (FFTW/Julia convention
 * exp(-2 pi i jk/n)), +1 backward (unnormalised). */
static void fft_strided(double complex *a, int n, size_t s, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double complex t = a[(size_t)i * s];
            a[(size_t)i * s] = a[(size_t)j * s];
            a[(size_t)j * s] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int h = len / 2;
        for (int k = 0; k < h; ++k) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            double complex w = cos(ang) + I * sin(ang);
            for (int i = 0; i < n; i += len) {
                double complex u = a[(size_t)(i + k) * s];
                double complex v = a[(size_t)(i + k + h) * s] * w;
                a[(size_t)(i + k) * s] = u + v;
                a[(size_t)(i + k + h) * s] = u - v;
            }
        }
    }
}

/* The same transform over a contiguous array, for the other oracle files and
 * the tests: sign -1 forward, +1 backward (unnormalised). */
void or_fft_c(double complex *a, int n, int sign) { fft_strided(a, n, 1, sign); }

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* rfft of an nx x ny real array (x fastest): output (nx/2+1) x ny, as Julia's
 * rfft (FFTW r2c) lays it out. Returns 0, or -1 for non power-of-two sizes. */
int or_rfft2(int nx, int ny, const double *in, int ld_in, double complex *out)
{
    if (!is_pow2(nx) || !is_pow2(ny)) return -1;
    double complex *f = malloc(sizeof(double complex) * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) f[IX(i, j, nx)] = in[IX(i, j, ld_in)];
    for (int j = 0; j < ny; ++j) fft_strided(f + IX(0, j, nx), nx, 1, -1);
    for (int i = 0; i <= nx / 2; ++i) fft_strided(f + i, ny, (size_t)nx, -1);
    int h = nx / 2 + 1;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < h; ++i) out[IX(i, j, h)] = f[IX(i, j, nx)];
    free(f);
    return 0;
}

/* irfft(Y, nx) of a (nx/2+1) x ny half spectrum, normalised by 1/(nx*ny).
 * The missing half is the Hermitian image of the given one; the self-
 * conjugate columns ik = 0 and ik = nx/2 contribute their real part only,
 * which is what FFTW's c2r (a c2c pass along y followed by a c2r along x that
 * reads the real parts of the x-DC and x-Nyquist terms) returns for input
 * that is not Hermitian in those columns. */
int or_irfft2(int nx, int ny, const double complex *in, double *out, int ld_out)
{
    if (!is_pow2(nx) || !is_pow2(ny)) return -1;
    int h = nx / 2 + 1;
    double complex *f = malloc(sizeof(double complex) * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            if (i < h)
                f[IX(i, j, nx)] = in[IX(i, j, h)];
            else
                f[IX(i, j, nx)] = conj(in[IX(nx - i, (ny - j) % ny, h)]);
        }
    for (int i = 0; i < nx; ++i) fft_strided(f + i, ny, (size_t)nx, +1);
    for (int j = 0; j < ny; ++j) fft_strided(f + IX(0, j, nx), nx, 1, +1);
    double scale = 1.0 / ((double)nx * (double)ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) out[IX(i, j, ld_out)] = creal(f[IX(i, j, nx)]) * scale;
    free(f);
    return 0;
}

/* TwoDPoissonPeriodic(grid): poisson_2d.jl:28-61.  Tables are (nx/2+1) x ny:
 * kx/k2, ky/k2 and k2, with ky negative from jk = ny/2+1 on (Nyquist
 * included, :48-51) and kx[1,1] = 1 (:54). */
void or_poisson_tables(int nx, int ny, double dimx, double dimy, double *kx, double *ky,
                       double *k2)
{
    int h = nx / 2 + 1;
    double kx0 = 2 * M_PI / dimx;
    double ky0 = 2 * M_PI / dimy;
    for (int ik = 1; ik <= h; ++ik) {
        double kx1 = (ik - 1) * kx0;
        for (int jk = 1; jk <= ny / 2; ++jk) {
            kx[IX(ik - 1, jk - 1, h)] = kx1;
            ky[IX(ik - 1, jk - 1, h)] = (jk - 1) * ky0;
        }
        for (int jk = ny / 2 + 1; jk <= ny; ++jk) {
            kx[IX(ik - 1, jk - 1, h)] = kx1;
            ky[IX(ik - 1, jk - 1, h)] = (jk - 1 - ny) * ky0;
        }
    }
    kx[0] = 1.0;
    size_t m = (size_t)h * ny;
    for (size_t k = 0; k < m; ++k) k2[k] = kx[k] * kx[k] + ky[k] * ky[k];
    for (size_t k = 0; k < m; ++k) {
        kx[k] = kx[k] / k2[k];
        ky[k] = ky[k] / k2[k];
    }
}

/* solve!(phi, poisson, rho): poisson_2d.jl:78-84, phi = irfft(rfft(rho)./k2).
 * No DC zeroing (k2[1,1] = 1 passes the mean of rho through). */
int or_solve_phi(int nx, int ny, double dimx, double dimy, const double *rho, double *phi)
{
    int h = nx / 2 + 1;
    size_t m = (size_t)h * ny;
    double *kx = malloc(8 * m), *ky = malloc(8 * m), *k2 = malloc(8 * m);
    double complex *r = malloc(sizeof(double complex) * m);
    or_poisson_tables(nx, ny, dimx, dimy, kx, ky, k2);
    int rc = or_rfft2(nx, ny, rho, nx, r);
    if (rc == 0) {
        for (size_t k = 0; k < m; ++k) r[k] = r[k] / k2[k];
        rc = or_irfft2(nx, ny, r, phi, nx);
    }
    free(kx); free(ky); free(k2); free(r);
    return rc;
}

/* solve!(ex, ey, poisson, rho): poisson_2d.jl:97-113.  rht = rfft(rho);
 * rht[1,1] = 0; ex = irfft(-i kx/k2 .* rht); the same for ey (the reference
 * recomputes the identical rfft before the second half). */
int or_solve_e(int nx, int ny, double dimx, double dimy, const double *rho, double *ex,
               double *ey)
{
    int h = nx / 2 + 1;
    size_t m = (size_t)h * ny;
    double *kx = malloc(8 * m), *ky = malloc(8 * m), *k2 = malloc(8 * m);
    double complex *r = malloc(sizeof(double complex) * m);
    double complex *t = malloc(sizeof(double complex) * m);
    or_poisson_tables(nx, ny, dimx, dimy, kx, ky, k2);
    int rc = or_rfft2(nx, ny, rho, nx, r);
    if (rc == 0) {
        r[0] = 0.0;
        for (size_t k = 0; k < m; ++k) t[k] = (-I * kx[k]) * r[k];
        rc = or_irfft2(nx, ny, t, ex, nx);
        for (size_t k = 0; k < m; ++k) t[k] = (-I * ky[k]) * r[k];
        if (rc == 0) rc = or_irfft2(nx, ny, t, ey, nx);
    }
    free(kx); free(ky); free(k2); free(r); free(t);
    return rc;
}

/* Poisson2D(mesh) + solve!(fields, p): meshfields.jl:27-108.  rho is
 * fields.rho ((nx+1) x (ny+1)), e is fields.e (2 x (nx+1) x (ny+1)).  The
 * spectral E is taken on rho[1:nx, 1:ny] without DC zeroing (kx[1,1] = 1,
 * :58-60), ghosts are copied (:100-103) and the return value is
 * sum(ex^2 + ey^2) * dx * dy over all (nx+1)(ny+1) nodes (:105-106). */
int or_solve_fields(int nx, int ny, double xmin, double xmax, double ymin, double ymax,
                    const double *rho, double *e, double *energy)
{
    int h = nx / 2 + 1, ld = nx + 1;
    size_t m = (size_t)h * ny;
    double dimx = xmax - xmin, dimy = ymax - ymin;
    double dx = dimx / nx, dy = dimy / ny;
    double *kx = malloc(8 * m), *ky = malloc(8 * m), *k2 = malloc(8 * m);
    double complex *r = malloc(sizeof(double complex) * m);
    double complex *t = malloc(sizeof(double complex) * m);
    double *tmp = malloc(sizeof(double) * (size_t)nx * ny);
    or_poisson_tables(nx, ny, dimx, dimy, kx, ky, k2);
    int rc = or_rfft2(nx, ny, rho, ld, r);
    for (int l = 0; l < 2 && rc == 0; ++l) {
        const double *kk = l == 0 ? kx : ky;
        for (size_t k = 0; k < m; ++k) t[k] = (-I * kk[k]) * r[k];
        rc = or_irfft2(nx, ny, t, tmp, nx);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) e[l + 2 * IX(i, j, ld)] = tmp[IX(i, j, nx)];
        for (int j = 0; j < ny; ++j) e[l + 2 * IX(nx, j, ld)] = e[l + 2 * IX(0, j, ld)];
        for (int i = 0; i <= nx; ++i) e[l + 2 * IX(i, ny, ld)] = e[l + 2 * IX(i, 0, ld)];
    }
    if (rc == 0 && energy) {
        size_t nn = (size_t)(nx + 1) * (ny + 1);
        double *sq = malloc(sizeof(double) * nn);
        for (size_t k = 0; k < nn; ++k) sq[k] = e[2 * k] * e[2 * k] + e[2 * k + 1] * e[2 * k + 1];
        *energy = pairwise_sum(sq, nn) * dx * dy;
        free(sq);
    }
    free(kx); free(ky); free(k2); free(r); free(t); free(tmp);
    return rc;
}

/* Electric energy of the hot-path driver, sqrt(sum(ex.^2) * dx * dy)
 * (examples/landau_damping_poisson_cic_m6.jl:65).  Also returns the two raw
 * sums sum(ex^2)*dx*dy and sum(ey^2)*dx*dy when `parts` is not NULL. */
double or_energy(int nx, int ny, double dx, double dy, const double *ex, const double *ey,
                 double *parts)
{
    size_t m = (size_t)nx * ny;
    double *sq = malloc(sizeof(double) * m);
    for (size_t k = 0; k < m; ++k) sq[k] = ex[k] * ex[k];
    double sx = pairwise_sum(sq, m) * dx * dy;
    double sy = 0.0;
    if (ey) {
        for (size_t k = 0; k < m; ++k) sq[k] = ey[k] * ey[k];
        sy = pairwise_sum(sq, m) * dx * dy;
    }
    free(sq);
    if (parts) {
        parts[0] = sx;
        parts[1] = sy;
    }
    return sqrt(sx);
}

/* ----------------------------------------------------------- sort by cell */

/* Stable counting sort by integer key in [0, nkeys): perm[r] = index of the
 * particle placed at rank r, i.e. the permutation std::stable_sort by key
 * produces.  Also returns the exclusive bucket offsets (nkeys + 1 entries). */
void or_counting_sort(int64_t n, const int32_t *key, int32_t nkeys, int32_t *perm,
                      int64_t *offsets)
{
    int64_t *cnt = calloc((size_t)nkeys + 1, sizeof(int64_t));
    for (int64_t k = 0; k < n; ++k) cnt[key[k] + 1]++;
    for (int32_t c = 0; c < nkeys; ++c) cnt[c + 1] += cnt[c];
    if (offsets) memcpy(offsets, cnt, sizeof(int64_t) * ((size_t)nkeys + 1));
    for (int64_t k = 0; k < n; ++k) perm[cnt[key[k]]++] = (int32_t)k;
    free(cnt);
}

/* ------------------------------------------ NGP, TSC and M4 (shapes) */

/* f_tsc(x): pushers_tsc.jl:3-12 (x^2 is x*x in Julia). */
double or_f_tsc(double x)
{
    if (x <= 0.5)
        return 0.75 - x * x;
    else if (0.5 < x && x <= 1.5)
        return 0.5 * ((1.5 - x) * (1.5 - x));
    return 0.0;
}

/* f_m4(x): pushers_m4.jl:8-19 (x^3 is (x*x)*x), times 2/3 as (f*2)/3. */
double or_f_m4(double x)
{
    double f = 0.0;
    if (x > 2)
        f = 0.;
    else if (x >= 1 && x <= 2) {
        double t = 2 - x;
        f = 0.25 * ((t * t) * t);
    } else if (x <= 1)
        f = (1 - 1.5 * (x * x)) + 0.75 * ((x * x) * x);
    return f * 2 / 3;
}

/* Weights of the offsets -1, 0, +1, +2 (and +3 for M4's deposit, whose
 * weight f_m4(3-d) is always 0), in the order of the reference's
 * cm1, c, cp1, cp2(, cp3). */
static inline void shape4_weights(double (*f)(double), double d, double c[5])
{
    c[0] = f(1 + d);
    c[1] = f(d);
    c[2] = f(1 - d);
    c[3] = f(2 - d);
    c[4] = f(3 - d);
}

/* push_v! of NearestGridPoint (pushers_ngp.jl:5-44): the node of
 * round(Int, xp) (ties to even) + 1, mod1. */
void or_push_v_ngp(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                   const double *ex, const double *ey, const double *bz, int ld)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double *q = p + 5 * k;
        double v1 = q[2];
        double v2 = q[3];
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = mod1((int64_t)nearbyint(xp) + 1, nx);
        int64_t j = mod1((int64_t)nearbyint(yp) + 1, ny);
        double e1 = ex[IX(i - 1, j - 1, ld)];
        double e2 = ey[IX(i - 1, j - 1, ld)];
        double b3 = bz ? bz[IX(i - 1, j - 1, ld)] : 0.0;
        v1 += 0.5 * dt * e1;
        v2 += 0.5 * dt * e2;
        double tantheta = 0.5 * dt * b3;
        double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
        v1 += v2 * tantheta;
        v2 += -v1 * sintheta;
        v1 += v2 * tantheta;
        q[2] = v1 + 0.5 * dt * e1;
        q[3] = v2 + 0.5 * dt * e2;
    }
}

/* push_v! of TriangularShapeCloud (pushers_tsc.jl:16-127) and of M4
 * (pushers_m4.jl:22-131): both gather with the M4 weights f_m4 on the 4 x 4
 * nodes i-1 .. i+2 (the TSC method calls f_m4, pushers_tsc.jl:46-53), x
 * offset outer, y offset inner, each term (cx*cy)*field. */
void or_push_v_m4(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                  const double *ex, const double *ey, const double *bz, int ld)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double *q = p + 5 * k;
        double v1 = q[2];
        double v2 = q[3];
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)floor(xp) + 1;
        int64_t j = (int64_t)floor(yp) + 1;
        double dpx = (xp - (double)i) + 1;
        double dpy = (yp - (double)j) + 1;
        int64_t ii[4], jj[4];
        for (int a = 0; a < 4; ++a) {
            ii[a] = mod1(i - 1 + a, nx);
            jj[a] = mod1(j - 1 + a, ny);
        }
        double cx[5], cy[5];
        shape4_weights(or_f_m4, dpx, cx);
        shape4_weights(or_f_m4, dpy, cy);
        double e1 = 0.0, e2 = 0.0, b3 = 0.0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                e1 += cx[a] * cy[b] * ex[IX(ii[a] - 1, jj[b] - 1, ld)];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                e2 += cx[a] * cy[b] * ey[IX(ii[a] - 1, jj[b] - 1, ld)];
        if (bz)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    b3 += cx[a] * cy[b] * bz[IX(ii[a] - 1, jj[b] - 1, ld)];
        v1 += 0.5 * dt * e1;
        v2 += 0.5 * dt * e2;
        double tantheta = 0.5 * dt * b3;
        double sintheta = 2 * tantheta / (1 + tantheta * tantheta);
        v1 += v2 * tantheta;
        v2 += -v1 * sintheta;
        v1 += v2 * tantheta;
        q[2] = v1 + 0.5 * dt * e1;
        q[3] = v2 + 0.5 * dt * e2;
    }
}

/* The tail shared by the NGP, TSC and M4 compute_rho: rho ./= dx*dy, then
 * rho .- (sum(rho) dx dy)/dimx/dimy (pushers_ngp.jl:74-80). */
static void shape_rho_finish(double *rho, int nx, int ny, double dx, double dy, double dimx,
                             double dimy)
{
    size_t m = (size_t)nx * ny;
    for (size_t k = 0; k < m; ++k) rho[k] /= (dx * dy);
    double rho_total = pairwise_sum(rho, m) * dx * dy;
    double shift = rho_total / dimx / dimy;
    for (size_t k = 0; k < m; ++k) rho[k] -= shift;
}

/* compute_rho(p, ::NearestGridPoint, mesh): pushers_ngp.jl:46-82. */
void or_compute_rho_ngp(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                        double dimx, double dimy, double *rho)
{
    memset(rho, 0, sizeof(double) * (size_t)nx * ny);
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p + 5 * k;
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = mod1((int64_t)nearbyint(xp) + 1, nx);
        int64_t j = mod1((int64_t)nearbyint(yp) + 1, ny);
        rho[IX(i - 1, j - 1, nx)] += q[4];
    }
    shape_rho_finish(rho, nx, ny, dx, dy, dimx, dimy);
}

/* compute_rho(p, ::TriangularShapeCloud, mesh) (pushers_tsc.jl:129-204, the
 * 4 x 4 nodes i-1 .. i+2 with f_tsc) and compute_rho(p, ::M4, mesh)
 * (pushers_m4.jl:135-228, the 5 x 5 nodes i-1 .. i+3 with f_m4, the last row
 * and column of weight 0).  Scatter order x offset outer, y offset inner,
 * each charge (cx*cy)*weight.  m4 selects the M4 method. */
void or_compute_rho_shape4(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                           double dimx, double dimy, int m4, double *rho)
{
    memset(rho, 0, sizeof(double) * (size_t)nx * ny);
    const int S = m4 ? 5 : 4;
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p + 5 * k;
        double xp = q[0] / dx;
        double yp = q[1] / dy;
        int64_t i = (int64_t)floor(xp) + 1;
        int64_t j = (int64_t)floor(yp) + 1;
        double dpx = (xp - (double)i) + 1;
        double dpy = (yp - (double)j) + 1;
        double weight = q[4];
        int64_t ii[5], jj[5];
        for (int a = 0; a < 5; ++a) {
            ii[a] = mod1(i - 1 + a, nx);
            jj[a] = mod1(j - 1 + a, ny);
        }
        double cx[5], cy[5];
        shape4_weights(m4 ? or_f_m4 : or_f_tsc, dpx, cx);
        shape4_weights(m4 ? or_f_m4 : or_f_tsc, dpy, cy);
        for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b)
                rho[IX(ii[a] - 1, jj[b] - 1, nx)] += cx[a] * cy[b] * weight;
    }
    shape_rho_finish(rho, nx, ny, dx, dy, dimx, dimy);
}

/* ------------------------------------------------------------- time loop */

/* The hot-path driver's time loop (examples/landau_damping_poisson_cic_m6.jl
 * :70-83): drift 1/2, deposit, Poisson solve, energy, kick, drift 1/2.
 * kernel 0 = CloudInCell, 1 = M6, 2 = NearestGridPoint, 3 =
 * TriangularShapeCloud, 4 = M4 (examples/landau_damping_poisson_cic_m6.jl
 * :88-103 runs the loop with each).  push_v!/push_x! run with OpenMP threads
 * (as @threads), the deposit is serial (as compute_rho).  energy[s] receives
 * sqrt(sum(ex^2) dx dy) after the solve of step s.  The domain is [0,dimx) x
 * [0,dimy) and bz is identically `bz0`. */
int or_landau_loop(int64_t n, double *p, int nx, int ny, double dimx, double dimy, double dt,
                   int nsteps, int kernel, double bz0, double *energy)
{
    double dx = dimx / nx, dy = dimy / ny;
    size_t m = (size_t)nx * ny;
    double *rho = malloc(8 * m), *ex = malloc(8 * m), *ey = malloc(8 * m), *bz = NULL;
    if (bz0 != 0.0) {
        bz = malloc(8 * m);
        for (size_t k = 0; k < m; ++k) bz[k] = bz0;
    }
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        or_push_x(n, p, 0.5 * dt, dimx, dimy);
        if (kernel == 0)
            or_compute_rho_cic(n, p, nx, ny, dx, dy, rho);
        else if (kernel == 1)
            or_compute_rho_m6(n, p, nx, ny, dx, dy, dimx, dimy, rho);
        else if (kernel == 2)
            or_compute_rho_ngp(n, p, nx, ny, dx, dy, dimx, dimy, rho);
        else
            or_compute_rho_shape4(n, p, nx, ny, dx, dy, dimx, dimy, kernel == 4, rho);
        rc = or_solve_e(nx, ny, dimx, dimy, rho, ex, ey);
        if (energy) energy[s] = or_energy(nx, ny, dx, dy, ex, NULL, NULL);
        if (kernel == 0)
            or_push_v_cic(n, p, dt, nx, ny, dx, dy, ex, ey, bz, nx);
        else if (kernel == 1)
            or_push_v_m6(n, p, dt, nx, ny, dx, dy, ex, ey, bz, nx);
        else if (kernel == 2)
            or_push_v_ngp(n, p, dt, nx, ny, dx, dy, ex, ey, bz, nx);
        else
            or_push_v_m4(n, p, dt, nx, ny, dx, dy, ex, ey, bz, nx);
        or_push_x(n, p, 0.5 * dt, dimx, dimy);
    }
    free(rho); free(ex); free(ey); free(bz);
    return rc;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
