/*
 * oracle_ua.c -- CPU restatement of ParticleInCell.jl's uniformly accurate
 * (UA) PIC step: the UA type (src/ua_type.jl), its steps (src/ua_steps.jl),
 * the UA deposit (src/compute_rho_m6.jl:29-179) and gather
 * (src/interpolation_m6.jl:3-123), and the driver loop of test/bupdate.jl.
 *
 * TEST INFRASTRUCTURE ONLY, like pic_oracle.c: loaded by tests/ and by
 * bench.py's cpu_baseline leg as the checker, never by the product path.
 *
 * Layouts are the reference's, column-major: xt, yt, fx, ... are
 * (ntau, 2, n) ComplexF64, element (m, c, k) (0-based) at m + ntau (c + 2 k);
 * et is (ntau, 2, n) Float64; pl, ql are (ntau, n); Particles.x, v, e are
 * (2, n); b, t are n-vectors.  The tau FFTs are the radix-2 transform of
 * pic_oracle.c (FFTW's sign convention; ifft divides by ntau).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void or_fft_c(double complex *a, int n, int sign);
void or_compute_rho_m6_fields(int64_t n, double *x, double w, int nx, int ny, double xmin,
                              double xmax, double ymin, double ymax, int compat, double *rho);
void or_interpol_eb_m6(int64_t n, double *x, int nx, int ny, double xmin, double xmax,
                       double ymin, double ymax, const double *e_f, double *e_p);
int or_solve_fields(int nx, int ny, double xmin, double xmax, double ymin, double ymax,
                    const double *rho, double *e, double *energy);

#define UX(m, c, k, ntau) ((size_t)(m) + (size_t)(ntau) * ((size_t)(c) + 2 * (size_t)(k)))

/* UA(ntau, eps, nbpart): ua_type.jl:18-24. */
static void ua_tables(int ntau, double *tau, double *ltau)
{
    double dtau = 2 * M_PI / ntau;
    for (int m = 0; m < ntau; ++m) {
        tau[m] = (double)m * dtau;
        ltau[m] = m < ntau / 2 ? (double)m : (double)(m - ntau);
    }
}

/* exp(s * 1im * l * t / eps) for s = -1 or +1, as Julia evaluates it:
 * ((s 1im * l) * t) / eps has real part +-0 and imaginary part
 * ((s l) t) / eps, whose exp is (cos, sin) of that. */
static inline double complex ua_cis(int s, double l, double t, double eps)
{
    double th = (((double)s * l) * t) / eps;
    return cos(th) + I * sin(th);
}

/* preparation!(ua, dt, particles, xt, yt): ua_steps.jl:3-83. */
void or_ua_preparation(int ntau, double eps, double dt, int64_t n, const double *x,
                       const double *v, const double *e, double *b, double *t,
                       double complex *pl, double complex *ql, double complex *xt,
                       double complex *yt)
{
    double *tau = malloc(8 * ntau), *ltau = malloc(8 * ntau);
    double complex *r = malloc(sizeof(double complex) * 2 * ntau);
    ua_tables(ntau, tau, ltau);
    for (int64_t k = 0; k < n; ++k) {
        double x1 = x[2 * k], x2 = x[2 * k + 1];
        double bk = 1 + 0.5 * sin(x1) * sin(x2);
        double tk = dt * bk;
        b[k] = bk;
        t[k] = tk;
        pl[(size_t)ntau * k] = tk;
        ql[(size_t)ntau * k] = tk * tk / 2;
        for (int m = 1; m < ntau; ++m) {
            double complex elt = ua_cis(-1, ltau[m], tk, eps);
            double l = ltau[m];
            /* eps * 1im * (elt - 1) / l and
             * eps * (eps * (1 - elt) - 1im * l * t) / l^2 */
            pl[m + (size_t)ntau * k] = (eps * I) * (elt - 1) / l;
            ql[m + (size_t)ntau * k] = eps * (eps * (1 - elt) - I * l * tk) / (l * l);
        }
        double ex = e[2 * k], ey = e[2 * k + 1];
        double vx = v[2 * k], vy = v[2 * k + 1];
        double vxb = vx / bk, vyb = vy / bk;
        for (int m = 0; m < ntau; ++m) {
            double tm = tau[m];
            double h1 = eps * (sin(tm) * vxb - cos(tm) * vyb);
            double h2 = eps * (sin(tm) * vyb + cos(tm) * vxb);
            double xt1 = x1 + h1 + eps * vyb;
            double xt2 = x2 + h2 - eps * vxb;
            xt[UX(m, 0, k, ntau)] = xt1;
            xt[UX(m, 1, k, ntau)] = xt2;
            double interv = (1 + 0.5 * sin(xt1) * sin(xt2) - bk) / eps;
            double exb = ((cos(tm) * vy - sin(tm) * vx) * interv + ex) / bk;
            double eyb = ((-cos(tm) * vx - sin(tm) * vy) * interv + ey) / bk;
            r[m] = cos(tm) * exb - sin(tm) * eyb;
            r[ntau + m] = sin(tm) * exb + cos(tm) * eyb;
        }
        for (int c = 0; c < 2; ++c) {
            double complex *rc = r + (size_t)ntau * c;
            or_fft_c(rc, ntau, -1);
            for (int m = 1; m < ntau; ++m) rc[m] = -I * rc[m] / ltau[m];
            or_fft_c(rc, ntau, +1);
            for (int m = 0; m < ntau; ++m) rc[m] /= ntau;
        }
        for (int m = 0; m < ntau; ++m) {
            yt[UX(m, 0, k, ntau)] = vx + (r[m] - r[0]) * eps;
            yt[UX(m, 1, k, ntau)] = vy + (r[ntau + m] - r[ntau]) * eps;
        }
    }
    free(tau); free(ltau); free(r);
}

/* update_particles_e!(particles, et, fields, ua, xt) = interpol_eb_m6!(et,
 * fields, xt, nbpart, ntau): interpolation_m6.jl:3-123, the gather of
 * interpol_eb_m6!(particles, fields) at every real(xt[m, :, k]). */
void or_ua_interpol(int ntau, int64_t n, const double complex *xt, int nx, int ny, double xmin,
                    double xmax, double ymin, double ymax, const double *e_f, double *et)
{
    double pos[2], ep[2];
    for (int64_t k = 0; k < n; ++k)
        for (int m = 0; m < ntau; ++m) {
            pos[0] = creal(xt[UX(m, 0, k, ntau)]);
            pos[1] = creal(xt[UX(m, 1, k, ntau)]);
            or_interpol_eb_m6(1, pos, nx, ny, xmin, xmax, ymin, ymax, e_f, ep);
            et[UX(m, 0, k, ntau)] = ep[0];
            et[UX(m, 1, k, ntau)] = ep[1];
        }
}

/* compute_f!(fx, fy, ua, particles, xt, yt, et): ua_steps.jl:107-148,
 * including the closing fft!(fx, 1), fft!(fy, 1). */
void or_ua_compute_f(int ntau, double eps, int64_t n, const double *b,
                     const double complex *xt, const double complex *yt, const double *et,
                     double complex *fx, double complex *fy)
{
    double *tau = malloc(8 * ntau), *ltau = malloc(8 * ntau);
    ua_tables(ntau, tau, ltau);
    for (int64_t k = 0; k < n; ++k) {
        double bk = b[k];
        for (int m = 0; m < ntau; ++m) {
            double xt1 = creal(xt[UX(m, 0, k, ntau)]);
            double xt2 = creal(xt[UX(m, 1, k, ntau)]);
            double complex yt1 = yt[UX(m, 0, k, ntau)];
            double complex yt2 = yt[UX(m, 1, k, ntau)];
            double tm = tau[m];
            fx[UX(m, 0, k, ntau)] = (cos(tm) * yt1 + sin(tm) * yt2) / bk;
            fx[UX(m, 1, k, ntau)] = (-sin(tm) * yt1 + cos(tm) * yt2) / bk;
            double interv = (1 + 0.5 * sin(xt1) * sin(xt2) - bk) / eps;
            double complex tmp1 = et[UX(m, 0, k, ntau)] + (cos(tm) * yt2 - sin(tm) * yt1) * interv;
            double complex tmp2 = et[UX(m, 1, k, ntau)] + (-cos(tm) * yt1 - sin(tm) * yt2) * interv;
            fy[UX(m, 0, k, ntau)] = (cos(tm) * tmp1 - sin(tm) * tmp2) / bk;
            fy[UX(m, 1, k, ntau)] = (sin(tm) * tmp1 + cos(tm) * tmp2) / bk;
        }
        for (int c = 0; c < 2; ++c) {
            or_fft_c(fx + UX(0, c, k, ntau), ntau, -1);
            or_fft_c(fy + UX(0, c, k, ntau), ntau, -1);
        }
    }
    free(tau); free(ltau);
}

/* ua_step!(xt, x~t, ua, particles, fx) (ua_steps.jl:152-172) and, with gx,
 * ua_step!(xt, x~t, ua, particles, fx, gx) (:174-202). */
void or_ua_step(int ntau, double eps, int64_t n, const double *t, const double complex *pl,
                const double complex *ql, const double complex *xtilde,
                const double complex *fx, const double complex *gx, double complex *xt)
{
    double *tau = malloc(8 * ntau), *ltau = malloc(8 * ntau);
    ua_tables(ntau, tau, ltau);
    for (int64_t k = 0; k < n; ++k) {
        double tk = t[k];
        for (int m = 0; m < ntau; ++m) {
            double complex elt = ua_cis(-1, ltau[m], tk, eps);
            double complex p = pl[m + (size_t)ntau * k];
            for (int c = 0; c < 2; ++c) {
                size_t u = UX(m, c, k, ntau);
                double complex f = fx[u];
                double complex z = elt * xtilde[u] + p * f;
                if (gx) z += ql[m + (size_t)ntau * k] * (gx[u] - f) / tk;
                xt[u] = z;
            }
        }
    }
    free(tau); free(ltau);
}

/* The reconstruction of compute_rho_m6!(fields, particles, xt, ua)
 * (compute_rho_m6.jl:43-66): x(t) = real(sum(fft(xt[:, c, k]) .*
 * exp(1im l t / eps) / ntau)), brought into the domain, written to
 * Particles.x; then the deposit of :68-178, which is the centred M6
 * deposit (compute_rho_m6_fields with compat = 0) at those positions. */
void or_ua_compute_rho(int ntau, double eps, int64_t n, const double *t,
                       const double complex *xt, double w, int nx, int ny, double xmin,
                       double xmax, double ymin, double ymax, double *x, double *rho)
{
    double *tau = malloc(8 * ntau), *ltau = malloc(8 * ntau);
    double complex *f = malloc(sizeof(double complex) * ntau);
    ua_tables(ntau, tau, ltau);
    for (int64_t k = 0; k < n; ++k) {
        for (int c = 0; c < 2; ++c) {
            memcpy(f, xt + UX(0, c, k, ntau), sizeof(double complex) * ntau);
            or_fft_c(f, ntau, -1);
            double complex s = 0;
            for (int m = 0; m < ntau; ++m) {
                f[m] *= ua_cis(+1, ltau[m], t[k], eps) / ntau;
                s += f[m];
            }
            x[2 * k + c] = creal(s);
        }
    }
    or_compute_rho_m6_fields(n, x, w, nx, ny, xmin, xmax, ymin, ymax, 0, rho);
    free(tau); free(ltau); free(f);
}

/* compute_v!(yt, particles, ua): ua_steps.jl:206-224. */
void or_ua_compute_v(int ntau, double eps, int64_t n, const double *t, const double complex *yt,
                     double *v)
{
    double *tau = malloc(8 * ntau), *ltau = malloc(8 * ntau);
    ua_tables(ntau, tau, ltau);
    for (int64_t k = 0; k < n; ++k) {
        double tk = t[k];
        double complex px = 0, py = 0;
        for (int m = 0; m < ntau; ++m) {
            double complex elt = ua_cis(+1, ltau[m], tk, eps);
            px += yt[UX(m, 0, k, ntau)] / ntau * elt;
            py += yt[UX(m, 1, k, ntau)] / ntau * elt;
        }
        v[2 * k] = creal(cos(tk / eps) * px + sin(tk / eps) * py);
        v[2 * k + 1] = creal(cos(tk / eps) * py - sin(tk / eps) * px);
    }
    free(tau); free(ltau);
}

/* nsteps iterations of the loop of test/bupdate.jl:62-116 on Particles
 * (x, v, e: 2 x n, uniform weight w) and MeshFields2D fields.e (2 x (nx+1) x
 * (ny+1)), which hold the prologue's state (bupdate.jl:56-60).  energy
 * receives the two solve! energies of every iteration (prediction,
 * correction), 2 x nsteps.  The loop's last update_particles_e! (:111) is not
 * restated: its et is overwritten before any use. */
int or_ua_run(int ntau, double eps, double dt, int nsteps, int64_t n, double *x, double *v,
              const double *e, double w, int nx, int ny, double xmin, double xmax, double ymin,
              double ymax, double *e_f, double *energy)
{
    size_t nt = (size_t)ntau * 2 * (size_t)n, nr = (size_t)(nx + 1) * (ny + 1);
    double complex *xt = malloc(sizeof(double complex) * nt), *xtt = malloc(sizeof(double complex) * nt);
    double complex *yt = malloc(sizeof(double complex) * nt), *ytt = malloc(sizeof(double complex) * nt);
    double complex *fx = malloc(sizeof(double complex) * nt), *fy = malloc(sizeof(double complex) * nt);
    double complex *gx = malloc(sizeof(double complex) * nt), *gy = malloc(sizeof(double complex) * nt);
    double complex *pl = malloc(sizeof(double complex) * (size_t)ntau * n);
    double complex *ql = malloc(sizeof(double complex) * (size_t)ntau * n);
    double *et = malloc(8 * nt), *b = malloc(8 * (size_t)n), *t = malloc(8 * (size_t)n);
    double *rho = malloc(8 * nr);
    int rc = 0;
    if (!xt || !xtt || !yt || !ytt || !fx || !fy || !gx || !gy || !pl || !ql || !et || !b || !t ||
        !rho)
        rc = -1;
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        or_ua_preparation(ntau, eps, dt, n, x, v, e, b, t, pl, ql, xt, yt);
        or_ua_interpol(ntau, n, xt, nx, ny, xmin, xmax, ymin, ymax, e_f, et);
        or_ua_compute_f(ntau, eps, n, b, xt, yt, et, fx, fy);
        /* mul!(x~t, ftau, xt); ua_step!; the same for yt; ifft!(xt, 1), ifft!(yt, 1) */
        memcpy(xtt, xt, sizeof(double complex) * nt);
        memcpy(ytt, yt, sizeof(double complex) * nt);
        for (size_t q = 0; q < nt; q += ntau) {
            or_fft_c(xtt + q, ntau, -1);
            or_fft_c(ytt + q, ntau, -1);
        }
        or_ua_step(ntau, eps, n, t, pl, NULL, xtt, fx, NULL, xt);
        or_ua_step(ntau, eps, n, t, pl, NULL, ytt, fy, NULL, yt);
        for (size_t q = 0; q < nt; q += ntau) {
            or_fft_c(xt + q, ntau, +1);
            or_fft_c(yt + q, ntau, +1);
            for (int m = 0; m < ntau; ++m) {
                xt[q + m] /= ntau;
                yt[q + m] /= ntau;
            }
        }
        or_ua_compute_rho(ntau, eps, n, t, xt, w, nx, ny, xmin, xmax, ymin, ymax, x, rho);
        rc = or_solve_fields(nx, ny, xmin, xmax, ymin, ymax, rho, e_f, energy + 2 * s);
        if (rc) break;
        or_ua_interpol(ntau, n, xt, nx, ny, xmin, xmax, ymin, ymax, e_f, et);
        or_ua_compute_f(ntau, eps, n, b, xt, yt, et, gx, gy);
        or_ua_step(ntau, eps, n, t, pl, ql, xtt, fx, gx, xt);
        or_ua_step(ntau, eps, n, t, pl, ql, ytt, fy, gy, yt);
        for (size_t q = 0; q < nt; q += ntau) {
            or_fft_c(xt + q, ntau, +1);
            for (int m = 0; m < ntau; ++m) xt[q + m] /= ntau;
        }
        or_ua_compute_rho(ntau, eps, n, t, xt, w, nx, ny, xmin, xmax, ymin, ymax, x, rho);
        rc = or_solve_fields(nx, ny, xmin, xmax, ymin, ymax, rho, e_f, energy + 2 * s + 1);
        if (rc) break;
        or_ua_compute_v(ntau, eps, n, t, yt, v);
    }
    free(xt); free(xtt); free(yt); free(ytt); free(fx); free(fy); free(gx); free(gy);
    free(pl); free(ql); free(et); free(b); free(t); free(rho);
    return rc;
}
