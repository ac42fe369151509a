/*
 * oracle_f32.c -- CPU restatement of the FP32 cell-relative particle step.
 *
 * TEST INFRASTRUCTURE ONLY (see pic_oracle.c).
 *
 * The reference has no FP32 mode; this state is a build addition.  It is the
 * reference's step (pushers_cic.jl:3-72, pushers_m6.jl:5-224 and the loop of
 * examples/landau_damping_poisson_cic_m6.jl:70-83) carried out on
 *     cell  int32  i + nx*j          (0-based)
 *     ox,oy float  in-cell offsets in units of the cell, in [0, 1)
 *     vx,vy float  velocities
 * with uniform particle weight w.  It is the layout of the reference's
 * text particle reader (read_particles.jl:19-28, `ix iy dpx dpy vx vy`).
 *
 * Every FP32 operation below is written out (fmaf where a fused multiply-add
 * is meant, plain * and + otherwise, build with -ffp-contract=off) so that the
 * HIP kernel, which spells the same operations, produces identical cells,
 * offsets and velocities bit for bit from identical inputs.  Charge is
 * accumulated here in FP64 from the FP32 per-node contributions; the GPU's
 * compensated FP32 tile sums agree with it to ~1e-12 relative.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Quintic B-spline (M6) weights of the 6 nodes i-2..i+3 for offset t in
 * [0,1).  Expanded pieces of f_m6 (compute_rho_m6.jl:10-25):
 *   120 w0 = (1-t)^5                     = f(2+t)
 *   120 w1 = P4(1-t)                     = f(1+t)
 *   120 w2 = P3(1-t)                     = f(t)
 *   120 w3 = P3(t) = 26+50t+20t^2-20t^3-20t^4+10t^5 = f(1-t)
 *   120 w4 = P4(t) = 1+5t+10t^2+10t^3+5t^4-5t^5     = f(2-t)
 *   120 w5 = t^5                         = f(3-t)
 * Written in z = t - 1/2, each mirrored pair is
 *   w5, w0 = E0(z^2) +/- z Q0(z^2),  E0 = 1/3840 + z^2/96 + z^4/48,
 *                                    Q0 = 1/384 + z^2/48 + z^4/120,
 *   w4, w1 = E1 +/- z Q1,            E1 = 79/1280 + 7/32 z^2 - z^4/16,
 *                                    Q1 = 25/128 + z^2/16 - z^4/24,
 *   w3, w2 = E2 +/- z Q2,            E2 = 841/1920 - 11/48 z^2 + z^4/24,
 *                                    Q2 = 77/192 - 7/24 z^2 + z^4/12,
 * (the binomial expansion of the pieces about t = 1/2), coefficients rounded
 * to float, evaluated with fused multiply-adds in this order. */
#define QF(num, den) ((float)((double)(num) / (double)(den)))

void orf_m6_weights(float t, float w[6])
{
    const float z = t - 0.5f, z2 = z * z, mz = -z;
    float e, q;
    e = fmaf(fmaf(QF(1, 48), z2, QF(1, 96)), z2, QF(1, 3840));
    q = fmaf(fmaf(QF(1, 120), z2, QF(1, 48)), z2, QF(1, 384));
    w[5] = fmaf(z, q, e);
    w[0] = fmaf(mz, q, e);
    e = fmaf(fmaf(QF(-1, 16), z2, QF(7, 32)), z2, QF(79, 1280));
    q = fmaf(fmaf(QF(-1, 24), z2, QF(1, 16)), z2, QF(25, 128));
    w[4] = fmaf(z, q, e);
    w[1] = fmaf(mz, q, e);
    e = fmaf(fmaf(QF(1, 24), z2, QF(-11, 48)), z2, QF(841, 1920));
    q = fmaf(fmaf(QF(1, 12), z2, QF(-7, 24)), z2, QF(77, 192));
    w[3] = fmaf(z, q, e);
    w[2] = fmaf(mz, q, e);
}

static inline void cic_weights(float ox, float oy, float a[4])
{
    float sx = 1.0f - ox, sy = 1.0f - oy;
    a[0] = sx * sy;
    a[1] = ox * sy;
    a[2] = ox * oy;
    a[3] = sx * oy;
}

static inline int wrapi(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

/* FP32 weights of the other separable shapes, build additions beside
 * the reference's FP64 f_tsc (pushers_tsc.jl:3-12) and f_m4
 * (pushers_m4.jl:8-19):
 *   f_tsc(x) = 3/4 - x^2 (x <= 1/2), (3/2 - x)^2 / 2 (x <= 3/2), else 0;
 *   f_m4(x)  = (1 - 3/2 x^2 + 3/4 x^3) 2/3 = 2/3 + x^2 (x/2 - 1) (x < 1),
 *              (2 - x)^3 / 6 (1 <= x <= 2), else 0,
 * the second written with one fused multiply-add in that expanded form. */
static inline float orf_f_tsc(float x)
{
    if (x <= 0.5f) return 0.75f - x * x;
    if (x <= 1.5f) return 0.5f * ((1.5f - x) * (1.5f - x));
    return 0.0f;
}

static inline float orf_f_m4(float x)
{
    if (x > 2.0f) return 0.0f;
    if (x >= 1.0f) {
        const float t = 2.0f - x;
        return (t * t) * t * QF(1, 6);
    }
    return fmaf(x * x, fmaf(0.5f, x, -1.0f), QF(2, 3));
}

/* Stencil of shape `kernel` (1 M6, 2 NGP, 3 TSC, 4 M4): S nodes per
 * direction starting HL nodes before the particle's cell. */
static inline int shape_s(int kernel) { return kernel == 1 ? 6 : kernel == 2 ? 2 : 4; }
static inline int shape_hl(int kernel) { return kernel == 1 ? 2 : kernel == 2 ? 0 : 1; }

/* Weights of the S nodes at in-cell offset t of a particle in (0-based)
 * cell `cell` of one direction.  gather = 1: push_v!'s weights, 0:
 * compute_rho's.  TSC gathers with f_m4 and deposits with f_tsc, as the
 * reference's pushers_tsc.jl:46-57 and :165-176 do.  NGP puts the whole
 * weight on the node of round(Int, xp) (ties to even, pushers_ngp.jl:22-23),
 * xp = cell + t: the lower node unless t > 1/2, or t == 1/2 in an odd cell. */
void orf_shape_weights(int kernel, int gather, float t, int cell, float w[6])
{
    if (kernel == 1) {
        orf_m6_weights(t, w);
    } else if (kernel == 2) {
        const int up = t > 0.5f || (t == 0.5f && (cell & 1));
        w[0] = up ? 0.0f : 1.0f;
        w[1] = up ? 1.0f : 0.0f;
    } else if (kernel == 3 && !gather) {
        w[0] = orf_f_tsc(1.0f + t);
        w[1] = orf_f_tsc(t);
        w[2] = orf_f_tsc(1.0f - t);
        w[3] = orf_f_tsc(2.0f - t);
    } else {
        w[0] = orf_f_m4(1.0f + t);
        w[1] = orf_f_m4(t);
        w[2] = orf_f_m4(1.0f - t);
        w[3] = orf_f_m4(2.0f - t);
    }
}

/* FP32 field of a separable shape at a particle of cell (i, j) and offsets
 * (fx, fy): x offset outer, y offset inner, each inner sum started with a
 * product and continued with fused multiply-adds, the outer likewise. */
static void sep_gather(int kernel, int i, int j, float fx, float fy, int nx, int ny,
                       const float *e32, int bz_mode, const float *bz32, float *e1o, float *e2o,
                       float *b3o)
{
    const int S = shape_s(kernel), HL = shape_hl(kernel);
    float wx[6], wy[6];
    orf_shape_weights(kernel, 1, fx, i, wx);
    orf_shape_weights(kernel, 1, fy, j, wy);
    int ii[6], jj[6];
    for (int a = 0; a < S; ++a) {
        ii[a] = wrapi(i - HL + a, nx);
        jj[a] = wrapi(j - HL + a, ny);
    }
    float e1 = 0.0f, e2 = 0.0f, b = 0.0f;
    for (int a = 0; a < S; ++a) {
        float t1 = wy[0] * e32[2 * (ii[a] + nx * jj[0])];
        float t2 = wy[0] * e32[2 * (ii[a] + nx * jj[0]) + 1];
        for (int bb = 1; bb < S; ++bb) {
            int node = ii[a] + nx * jj[bb];
            t1 = fmaf(wy[bb], e32[2 * node], t1);
            t2 = fmaf(wy[bb], e32[2 * node + 1], t2);
        }
        if (a == 0) {
            e1 = wx[0] * t1;
            e2 = wx[0] * t2;
        } else {
            e1 = fmaf(wx[a], t1, e1);
            e2 = fmaf(wx[a], t2, e2);
        }
    }
    if (bz_mode == 2) {
        for (int a = 0; a < S; ++a) {
            float t3 = wy[0] * bz32[ii[a] + nx * jj[0]];
            for (int bb = 1; bb < S; ++bb) t3 = fmaf(wy[bb], bz32[ii[a] + nx * jj[bb]], t3);
            b = a == 0 ? wx[0] * t3 : fmaf(wx[a], t3, b);
        }
    }
    *e1o = e1;
    *e2o = e2;
    *b3o = b;
}

/* FP32 charges of a separable shape, (w_x[u] * (w_y[v] * w)), added in FP64
 * to the S x S nodes from (i - HL, j - HL). */
static void sep_deposit(int kernel, int i, int j, float fx, float fy, int nx, int ny, float w,
                        double *rho_acc)
{
    const int S = shape_s(kernel), HL = shape_hl(kernel);
    float wx[6], wy[6];
    orf_shape_weights(kernel, 0, fx, i, wx);
    orf_shape_weights(kernel, 0, fy, j, wy);
    for (int bb = 0; bb < S; ++bb) wy[bb] = wy[bb] * w;
    for (int a = 0; a < S; ++a) {
        int ia = wrapi(i - HL + a, nx);
        for (int bb = 0; bb < S; ++bb)
            rho_acc[ia + nx * wrapi(j - HL + bb, ny)] += (double)(wx[a] * wy[bb]);
    }
}


/* One offset component drifted by u = v * dt/d (in cells): returns the integer
 * cell shift and the new offset in [0,1).  An offset that rounds up to 1.0
 * (t in [-2^-25, 0) after the floor) is moved to 0 of the next cell. */
static inline int drift1(float o, float v, float dtd, float *onew)
{
    float t = fmaf(v, dtd, o);
    float s = floorf(t);
    float r = t - s;
    if (r >= 1.0f) {
        r = 0.0f;
        s = s + 1.0f;
    }
    *onew = r;
    return (int)s;
}

/*
 * One fused FP32 step for n particles, in this order per particle:
 *   gather E (and Bz) at the current position,
 *   kick (half E, rotation, half E: pushers_cic.jl:38-49),
 *   drift by dt_drift (the two half drifts of consecutive loop iterations,
 *   pushers_cic.jl:55-72, taken as one),
 *   deposit w at the new position into rho_acc (FP64, nx*ny, ADDED to).
 * kernel 0 = CIC, 1 = M6, 2 = NGP, 3 = TSC, 4 = M4 (the library's shape codes).
 * e32 holds (ex, ey) interleaved per node, nx*ny.
 * bz_mode 0 = no magnetic field, 1 = uniform bz0, 2 = grid bz32 (nx*ny).
 * hdt = (float)(dt/2), dtdx = (float)(dt_drift/dx), dtdy likewise, all
 * rounded once on the host.  deposit = 0 skips the deposit.
 */
void orf_step(int64_t n, int32_t *cell, float *ox, float *oy, float *vx, float *vy, int nx,
              int ny, int kernel, const float *e32, int bz_mode, float bz0, const float *bz32,
              float hdt, float dtdx, float dtdy, float w, int deposit, double *rho_acc)
{
    for (int64_t k = 0; k < n; ++k) {
        int c = cell[k];
        int i = c % nx, j = c / nx;
        float fx = ox[k], fy = oy[k];
        float e1, e2, b = 0.0f;
        if (kernel == 0) {
            float a[4];
            cic_weights(fx, fy, a);
            int ip = i + 1 == nx ? 0 : i + 1;
            int jp = j + 1 == ny ? 0 : j + 1;
            int n00 = i + nx * j, n10 = ip + nx * j, n11 = ip + nx * jp, n01 = i + nx * jp;
            e1 = a[0] * e32[2 * n00];
            e1 = fmaf(a[1], e32[2 * n10], e1);
            e1 = fmaf(a[2], e32[2 * n11], e1);
            e1 = fmaf(a[3], e32[2 * n01], e1);
            e2 = a[0] * e32[2 * n00 + 1];
            e2 = fmaf(a[1], e32[2 * n10 + 1], e2);
            e2 = fmaf(a[2], e32[2 * n11 + 1], e2);
            e2 = fmaf(a[3], e32[2 * n01 + 1], e2);
            if (bz_mode == 2) {
                b = a[0] * bz32[n00];
                b = fmaf(a[1], bz32[n10], b);
                b = fmaf(a[2], bz32[n11], b);
                b = fmaf(a[3], bz32[n01], b);
            }
        } else {
            sep_gather(kernel, i, j, fx, fy, nx, ny, e32, bz_mode, bz32, &e1, &e2, &b);
        }
        if (bz_mode == 1) b = bz0;
        float v1 = fmaf(hdt, e1, vx[k]);
        float v2 = fmaf(hdt, e2, vy[k]);
        if (bz_mode != 0) {
            float tn = hdt * b;
            float sn = (2.0f * tn) / fmaf(tn, tn, 1.0f);
            v1 = fmaf(v2, tn, v1);
            v2 = fmaf(-v1, sn, v2);
            v1 = fmaf(v2, tn, v1);
        }
        v1 = fmaf(hdt, e1, v1);
        v2 = fmaf(hdt, e2, v2);
        vx[k] = v1;
        vy[k] = v2;
        float nfx, nfy;
        int si = drift1(fx, v1, dtdx, &nfx);
        int sj = drift1(fy, v2, dtdy, &nfy);
        i = wrapi(i + si, nx);
        j = wrapi(j + sj, ny);
        cell[k] = i + nx * j;
        ox[k] = nfx;
        oy[k] = nfy;
        if (!deposit) continue;
        if (kernel == 0) {
            float a[4];
            cic_weights(nfx, nfy, a);
            int ip = i + 1 == nx ? 0 : i + 1;
            int jp = j + 1 == ny ? 0 : j + 1;
            rho_acc[i + nx * j] += (double)(a[0] * w);
            rho_acc[ip + nx * j] += (double)(a[1] * w);
            rho_acc[ip + nx * jp] += (double)(a[2] * w);
            rho_acc[i + nx * jp] += (double)(a[3] * w);
        } else {
            sep_deposit(kernel, i, j, nfx, nfy, nx, ny, w, rho_acc);
        }
    }
}

/* Deposit only (no push), FP32 contributions into an FP64 accumulator: the
 * charge of the current cell-relative state, as orf_step deposits it. */
void orf_deposit(int64_t n, const int32_t *cell, const float *ox, const float *oy, int nx,
                 int ny, int kernel, float w, double *rho_acc)
{
    for (int64_t k = 0; k < n; ++k) {
        int c = cell[k];
        int i = c % nx, j = c / nx;
        if (kernel == 0) {
            float a[4];
            cic_weights(ox[k], oy[k], a);
            int ip = i + 1 == nx ? 0 : i + 1;
            int jp = j + 1 == ny ? 0 : j + 1;
            rho_acc[i + nx * j] += (double)(a[0] * w);
            rho_acc[ip + nx * j] += (double)(a[1] * w);
            rho_acc[ip + nx * jp] += (double)(a[2] * w);
            rho_acc[i + nx * jp] += (double)(a[3] * w);
        } else {
            sep_deposit(kernel, i, j, ox[k], oy[k], nx, ny, w, rho_acc);
        }
    }
}

/* FP64 (x, y, vx, vy) -> FP32 cell-relative state, and back.  x in [0, dimx):
 * xp = x/dx, i = floor(xp), ox = (float)(xp - i), an offset that rounds to
 * 1.0f moves to 0 of the next cell, i wraps periodically. */
void orf_from_f64(int64_t n, const double *p5, int nx, int ny, double dx, double dy,
                  int32_t *cell, float *ox, float *oy, float *vx, float *vy)
{
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p5 + 5 * k;
        double xp = q[0] / dx, yp = q[1] / dy;
        double fi = floor(xp), fj = floor(yp);
        float fx = (float)(xp - fi), fy = (float)(yp - fj);
        int i = (int)fi, j = (int)fj;
        if (fx >= 1.0f) { fx = 0.0f; i += 1; }
        if (fy >= 1.0f) { fy = 0.0f; j += 1; }
        i = wrapi(i, nx);
        j = wrapi(j, ny);
        cell[k] = i + nx * j;
        ox[k] = fx;
        oy[k] = fy;
        vx[k] = (float)q[2];
        vy[k] = (float)q[3];
    }
}

void orf_to_f64(int64_t n, const int32_t *cell, const float *ox, const float *oy,
                const float *vx, const float *vy, int nx, double dx, double dy, double w,
                double *p5)
{
    for (int64_t k = 0; k < n; ++k) {
        int i = cell[k] % nx, j = cell[k] / nx;
        double *q = p5 + 5 * k;
        q[0] = ((double)i + (double)ox[k]) * dx;
        q[1] = ((double)j + (double)oy[k]) * dy;
        q[2] = (double)vx[k];
        q[3] = (double)vy[k];
        q[4] = w;
    }
}
