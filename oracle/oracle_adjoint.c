/*
 * oracle_adjoint.c -- CPU restatement of the adjoint (reverse-time) sweep of
 * the particle-mesh time loop.
 *
 * TEST INFRASTRUCTURE ONLY (see pic_oracle.c).
 *
 * The reference has no adjoint; this is the exact discrete adjoint of its
 * time loop (examples/landau_damping_poisson_cic_m6.jl:70-83) with bz = 0:
 *   step s:  x' = x + h v            push_x!(dt/2)       (h = dt/2)
 *            E_s = P D(x')           compute_rho; solve!(ex, ey)
 *            v  <- v + dt G(x') E_s  push_v! (the two half kicks)
 *            x  <- x' + h v          push_x!(dt/2)
 *   J = sum(ex^2 + ey^2) dx dy of E_{S-1}, the last energy the loop records.
 * D is the deposit (w times the shape weights over dx dy, mean removed), G
 * the gather (interpolation of a grid at the particle), P the spectral
 * E = -grad(lap^-1 rho).  The reverse sweep, from lx = lv = 0:
 *   s = S-1:    mu = dJ/dE = 2 E dx dy;  lx' = 0;  lvt = 0
 *   s < S-1:    lvt = lv + h lx;  lx' = lx
 *               mu = dt G(x')^T lvt                  (scatter: deposit-like)
 *               lx' += dt (lvt . grad_x G(x') E_s)    (gradient gather of E)
 *   all s:      lrho = P^T mu = -(P_x mu_x + P_y mu_y)
 *               lx' += D_x(x')^T lrho                 (gradient gather of lrho)
 *               lx = lx';  lv = lvt + h lx'
 * giving dJ/dx0 = lx and dJ/dv0 = lv.  P^T = -P because P is a convolution
 * with a real kernel whose symbol -i k/|k|^2 is odd and purely imaginary.
 *
 * Checkpointing: the forward pass keeps every E_s (nx*ny*2 doubles per step)
 * and the particle state every `ckpt` steps; the reverse sweep recomputes
 * each segment's positions x'_s from its checkpoint with the stored E_s, so
 * the recomputed trajectory is bit-identical to the forward one.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void or_push_v_cic(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                   const double *ex, const double *ey, const double *bz, int ld);
void or_push_v_m6(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                  const double *ex, const double *ey, const double *bz, int ld);
void or_push_x(int64_t n, double *p, double dt, double dimx, double dimy);
void or_compute_rho_cic(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                        double *rho);
void or_compute_rho_m6(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                       double dimx, double dimy, double *rho);
int or_solve_e(int nx, int ny, double dimx, double dimy, const double *rho, double *ex,
               double *ey);
double or_f_m6(double q);

static inline int64_t wrapl(int64_t i, int64_t n)
{
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

/* Shape weights and their derivatives in units of the cell: nodes
 * i0 + o .. i0 + o + S - 1 with o = 0 (CIC, S = 2) or -2 (M6, S = 6). */
static inline void shape1d(int kernel, double f, double *w, double *dw)
{
    if (kernel == 0) {
        w[0] = 1 - f; w[1] = f;
        dw[0] = -1; dw[1] = 1;
    } else {
        /* f_m6 at distances 2+f, 1+f, f, 1-f, 2-f, 3-f (compute_rho_m6.jl:10-25);
         * derivative of the quintic pieces (1/120) d/df of
         * (1-f)^5, (2-f)^5-6(1-f)^5, (3-f)^5-6(2-f)^5+15(1-f)^5 at their mirror
         * images, written through the symmetric closed forms below */
        double t = f, s = 1 - f;
        w[0] = or_f_m6(2 + t);
        w[1] = or_f_m6(1 + t);
        w[2] = or_f_m6(t);
        w[3] = or_f_m6(1 - t);
        w[4] = or_f_m6(2 - t);
        w[5] = or_f_m6(3 - t);
        /* 120 w3 = 26+50t+20t^2-20t^3-20t^4+10t^5, 120 w4 = 1+5t+10t^2+10t^3+5t^4-5t^5,
         * 120 w5 = t^5, and w_k(t) = w_{5-k}(1-t) */
        double d3 = (50 + t * (40 + t * (-60 + t * (-80 + t * 50)))) / 120;
        double d4 = (5 + t * (20 + t * (30 + t * (20 - t * 25)))) / 120;
        double d5 = (5 * t * t * t * t) / 120;
        double e3 = (50 + s * (40 + s * (-60 + s * (-80 + s * 50)))) / 120;
        double e4 = (5 + s * (20 + s * (30 + s * (20 - s * 25)))) / 120;
        double e5 = (5 * s * s * s * s) / 120;
        dw[0] = -e5; dw[1] = -e4; dw[2] = -e3;
        dw[3] = d3; dw[4] = d4; dw[5] = d5;
    }
}

/* Cell and in-cell offset as the reference computes them (floor; x >= 0). */
static inline void cell_of(double x, double d, int n, int64_t *i, double *f)
{
    double xp = x / d;
    int64_t i1 = (int64_t)floor(xp) + 1;
    *f = (xp - (double)i1) + 1;
    *i = wrapl(i1 - 1, n);
}

/* mu += scale * G(x)^T (ux, uy): scatter per-particle vectors to the grid. */
static void scatter(int kernel, int64_t n, const double *p, int nx, int ny, double dx, double dy,
                    const double *ux, const double *uy, double scale, double *mux, double *muy)
{
    const int S = kernel == 0 ? 2 : 6, o = kernel == 0 ? 0 : -2;
    for (int64_t k = 0; k < n; ++k) {
        int64_t i, j;
        double fx, fy, wx[6], wy[6], dwx[6], dwy[6];
        cell_of(p[5 * k], dx, nx, &i, &fx);
        cell_of(p[5 * k + 1], dy, ny, &j, &fy);
        shape1d(kernel, fx, wx, dwx);
        shape1d(kernel, fy, wy, dwy);
        for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b) {
                size_t g = (size_t)wrapl(i + o + a, nx) + (size_t)nx * wrapl(j + o + b, ny);
                double wab = wx[a] * wy[b];
                mux[g] += scale * ux[k] * wab;
                muy[g] += scale * uy[k] * wab;
            }
    }
}

/* g += scale * (sum_nodes grad_x W(x - X) * field) per particle, for the
 * scalar field f (gx, gy receive the x and y components). */
static void grad_gather(int kernel, int64_t n, const double *p, int nx, int ny, double dx,
                        double dy, const double *f, const double *sx, const double *sy,
                        double scale, double *gx, double *gy)
{
    const int S = kernel == 0 ? 2 : 6, o = kernel == 0 ? 0 : -2;
    for (int64_t k = 0; k < n; ++k) {
        int64_t i, j;
        double fx, fy, wx[6], wy[6], dwx[6], dwy[6];
        cell_of(p[5 * k], dx, nx, &i, &fx);
        cell_of(p[5 * k + 1], dy, ny, &j, &fy);
        shape1d(kernel, fx, wx, dwx);
        shape1d(kernel, fy, wy, dwy);
        double ax = 0, ay = 0;
        for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b) {
                size_t g = (size_t)wrapl(i + o + a, nx) + (size_t)nx * wrapl(j + o + b, ny);
                ax += dwx[a] * wy[b] * f[g];
                ay += wx[a] * dwy[b] * f[g];
            }
        double s = scale * (sx ? sx[k] : 1.0);
        double t = scale * (sy ? sy[k] : 1.0);
        gx[k] += s * ax / dx;
        gy[k] += t * ay / dy;
    }
}

/* lrho = P^T (mux, muy) = -(P_x mux + P_y muy). */
int or_poisson_adjoint(int nx, int ny, double dimx, double dimy, const double *mux,
                       const double *muy, double *lrho)
{
    size_t m = (size_t)nx * ny;
    double *a = malloc(8 * m), *b = malloc(8 * m), *c = malloc(8 * m), *d = malloc(8 * m);
    int rc = or_solve_e(nx, ny, dimx, dimy, mux, a, b);
    if (rc == 0) rc = or_solve_e(nx, ny, dimx, dimy, muy, c, d);
    if (rc == 0)
        for (size_t k = 0; k < m; ++k) lrho[k] = -(a[k] + d[k]);
    free(a); free(b); free(c); free(d);
    return rc;
}

static void deposit(int kernel, int64_t n, const double *p, int nx, int ny, double dx, double dy,
                    double dimx, double dimy, double *rho)
{
    if (kernel == 0)
        or_compute_rho_cic(n, p, nx, ny, dx, dy, rho);
    else
        or_compute_rho_m6(n, p, nx, ny, dx, dy, dimx, dimy, rho);
}

static void kick(int kernel, int64_t n, double *p, double dt, int nx, int ny, double dx,
                 double dy, const double *ex, const double *ey)
{
    if (kernel == 0)
        or_push_v_cic(n, p, dt, nx, ny, dx, dy, ex, ey, NULL, nx);
    else
        or_push_v_m6(n, p, dt, nx, ny, dx, dy, ex, ey, NULL, nx);
}

/* Forward loop + reverse sweep.  p0 is the initial 5 x n particle array
 * (unchanged).  Outputs: *J, gx (2 x n: dJ/dx0, dJ/dy0), gv (2 x n:
 * dJ/dvx0, dJ/dvy0), and the energy history e_hist (S entries of
 * sum(ex^2+ey^2) dx dy; may be NULL).  ckpt >= 1. */
int or_adjoint(int64_t n, const double *p0, int nx, int ny, double dimx, double dimy, double dt,
               int nsteps, int kernel, int ckpt, double *J, double *gx, double *gv,
               double *e_hist)
{
    if (nsteps < 1 || ckpt < 1) return -1;
    const double dx = dimx / nx, dy = dimy / ny, h = 0.5 * dt;
    const size_t m = (size_t)nx * ny, pn = (size_t)5 * n;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    double *E = malloc(sizeof(double) * 2 * m * (size_t)nsteps);
    double *ck = malloc(sizeof(double) * pn * (size_t)nseg);
    double *p = malloc(sizeof(double) * pn);
    double *rho = malloc(8 * m);
    if (!E || !ck || !p || !rho) return -2;
    memcpy(p, p0, sizeof(double) * pn);
    int rc = 0;
    /* forward */
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        if (s % ckpt == 0) memcpy(ck + pn * (size_t)(s / ckpt), p, sizeof(double) * pn);
        or_push_x(n, p, h, dimx, dimy);
        deposit(kernel, n, p, nx, ny, dx, dy, dimx, dimy, rho);
        double *ex = E + 2 * m * (size_t)s, *ey = ex + m;
        rc = or_solve_e(nx, ny, dimx, dimy, rho, ex, ey);
        if (e_hist) {
            double t = 0;
            for (size_t k = 0; k < m; ++k) t += ex[k] * ex[k] + ey[k] * ey[k];
            e_hist[s] = t * dx * dy;
        }
        kick(kernel, n, p, dt, nx, ny, dx, dy, ex, ey);
        or_push_x(n, p, h, dimx, dimy);
    }
    {
        const double *ex = E + 2 * m * (size_t)(nsteps - 1), *ey = ex + m;
        double t = 0;
        for (size_t k = 0; k < m; ++k) t += ex[k] * ex[k] + ey[k] * ey[k];
        *J = t * dx * dy;
    }
    /* reverse */
    double *lx = calloc(n, 8), *ly = calloc(n, 8), *lvx = calloc(n, 8), *lvy = calloc(n, 8);
    double *tx = malloc(8 * n), *ty = malloc(8 * n);
    double *mux = malloc(8 * m), *muy = malloc(8 * m), *lrho = malloc(8 * m);
    double *xs = malloc(sizeof(double) * pn * (size_t)ckpt);  /* x'_s of a segment */
    for (int g = nseg - 1; g >= 0 && rc == 0; --g) {
        const int s0 = g * ckpt, s1 = s0 + ckpt < nsteps ? s0 + ckpt : nsteps;
        memcpy(p, ck + pn * (size_t)g, sizeof(double) * pn);
        for (int s = s0; s < s1; ++s) {
            const double *ex = E + 2 * m * (size_t)s, *ey = ex + m;
            or_push_x(n, p, h, dimx, dimy);
            memcpy(xs + pn * (size_t)(s - s0), p, sizeof(double) * pn);
            kick(kernel, n, p, dt, nx, ny, dx, dy, ex, ey);
            or_push_x(n, p, h, dimx, dimy);
        }
        for (int s = s1 - 1; s >= s0 && rc == 0; --s) {
            const double *xp = xs + pn * (size_t)(s - s0);
            const double *ex = E + 2 * m * (size_t)s, *ey = ex + m;
            memset(mux, 0, 8 * m);
            memset(muy, 0, 8 * m);
            if (s == nsteps - 1) {
                for (size_t k = 0; k < m; ++k) {
                    mux[k] = 2 * ex[k] * dx * dy;
                    muy[k] = 2 * ey[k] * dx * dy;
                }
                for (int64_t k = 0; k < n; ++k) lx[k] = ly[k] = lvx[k] = lvy[k] = 0.0;
            } else {
                for (int64_t k = 0; k < n; ++k) {  /* lvt = lv + h lx; lx' = lx */
                    lvx[k] += h * lx[k];
                    lvy[k] += h * ly[k];
                }
                scatter(kernel, n, xp, nx, ny, dx, dy, lvx, lvy, dt, mux, muy);
                /* lx' += dt (lvx d(ex)/dx + lvy d(ey)/dx), y likewise */
                grad_gather(kernel, n, xp, nx, ny, dx, dy, ex, lvx, lvx, dt, lx, ly);
                grad_gather(kernel, n, xp, nx, ny, dx, dy, ey, lvy, lvy, dt, lx, ly);
            }
            rc = or_poisson_adjoint(nx, ny, dimx, dimy, mux, muy, lrho);
            /* lx' += D_x^T lrho: per particle (w / dxdy) sum lrho grad W */
            for (int64_t k = 0; k < n; ++k) tx[k] = ty[k] = xp[5 * k + 4] / (dx * dy);
            grad_gather(kernel, n, xp, nx, ny, dx, dy, lrho, tx, ty, 1.0, lx, ly);
            for (int64_t k = 0; k < n; ++k) {  /* lv = lvt + h lx' */
                lvx[k] += h * lx[k];
                lvy[k] += h * ly[k];
            }
        }
    }
    for (int64_t k = 0; k < n; ++k) {
        gx[2 * k] = lx[k];
        gx[2 * k + 1] = ly[k];
        gv[2 * k] = lvx[k];
        gv[2 * k + 1] = lvy[k];
    }
    free(E); free(ck); free(p); free(rho); free(lx); free(ly); free(lvx); free(lvy);
    free(tx); free(ty); free(mux); free(muy); free(lrho); free(xs);
    return rc;
}

/* J = sum(ex^2+ey^2) dx dy of the last solve of the loop, forward only. */
int or_energy_objective(int64_t n, const double *p0, int nx, int ny, double dimx, double dimy,
                        double dt, int nsteps, int kernel, double *J)
{
    const double dx = dimx / nx, dy = dimy / ny, h = 0.5 * dt;
    const size_t m = (size_t)nx * ny, pn = (size_t)5 * n;
    double *p = malloc(8 * pn), *rho = malloc(8 * m), *ex = malloc(8 * m), *ey = malloc(8 * m);
    memcpy(p, p0, 8 * pn);
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        or_push_x(n, p, h, dimx, dimy);
        deposit(kernel, n, p, nx, ny, dx, dy, dimx, dimy, rho);
        rc = or_solve_e(nx, ny, dimx, dimy, rho, ex, ey);
        if (s == nsteps - 1) {
            double t = 0;
            for (size_t k = 0; k < m; ++k) t += ex[k] * ex[k] + ey[k] * ey[k];
            *J = t * dx * dy;
        }
        kick(kernel, n, p, dt, nx, ny, dx, dy, ex, ey);
        or_push_x(n, p, h, dimx, dimy);
    }
    free(p); free(rho); free(ex); free(ey);
    return rc;
}
