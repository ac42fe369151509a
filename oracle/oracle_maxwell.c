/*
 * oracle_maxwell.c -- CPU restatement of ParticleInCell.jl's Vlasov-Maxwell
 * loop: the FDTD solver (src/fdtd.jl), the CIC current deposit
 * (src/cloud_in_cell.jl:10-81) and the loop of
 * examples/landau_damping_maxwell_fdtd.jl:58-83.
 *
 * TEST INFRASTRUCTURE ONLY, like pic_oracle.c: loaded by tests/ and by
 * bench.py's cpu_baseline leg as the checker, never by the product path.
 *
 * Layouts are the reference's, column-major, 1-based in the comments:
 *   fdtd.ex  nx x (ny+1)    fdtd.ey  (nx+1) x ny    fdtd.bz  nx x ny
 *   ex, ey, bz, jx, jy on the nodes, (nx+1) x (ny+1).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define AT(a, ld, i, j) (a)[(size_t)((i) - 1) + (size_t)(ld) * (size_t)((j) - 1)]

static inline int64_t m1(int64_t x, int64_t n)
{
    int64_t r = (x - 1) % n;
    if (r < 0) r += n;
    return r + 1;
}

/* faraday!(fdtd, m, dt): fdtd.jl:24-36. */
void or_faraday(int nx, int ny, double dx, double dy, const double *ex, const double *ey,
                double *bz, double dt)
{
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            double dex_dy = (AT(ex, nx, i, m1(j + 1, ny)) - AT(ex, nx, i, j)) / dy;
            double dey_dx = (AT(ey, nx + 1, m1(i + 1, nx), j) - AT(ey, nx + 1, i, j)) / dx;
            AT(bz, nx, i, j) += dt * (dex_dy - dey_dx);
        }
}

/* ampere_maxwell!(fdtd, m, jx, jy, dt): fdtd.jl:40-56. */
void or_ampere_maxwell(int nx, int ny, double dx, double dy, double *ex, double *ey,
                       const double *bz, const double *jx, const double *jy, double dt)
{
    const int ldj = nx + 1;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny + 1; ++j) {
            double dbz_dy = (AT(bz, nx, i, m1(j, ny)) - AT(bz, nx, i, m1(j - 1, ny))) / dy;
            AT(ex, nx, i, j) +=
                dt * dbz_dy - dt * 0.5 * (AT(jx, ldj, i, j) + AT(jx, ldj, m1(i + 1, nx), j));
        }
    for (int i = 1; i <= nx + 1; ++i)
        for (int j = 1; j <= ny; ++j) {
            double dbz_dx =
                (AT(bz, nx, m1(i, nx), m1(j, ny)) - AT(bz, nx, m1(i - 1, nx), j)) / dx;
            AT(ey, nx + 1, i, j) -=
                dt * dbz_dx - dt * 0.5 * (AT(jy, ldj, i, j) + AT(jy, ldj, i, m1(j + 1, ny)));
        }
}

/* update_fields!(ex, ey, bz, m, fdtd): fdtd.jl:60-77. */
void or_update_fields(int nx, int ny, const double *fex, const double *fey, const double *fbz,
                      double *ex, double *ey, double *bz)
{
    const int ld = nx + 1;
    for (int i = 1; i <= nx + 1; ++i)
        for (int j = 1; j <= ny + 1; ++j) {
            AT(ex, ld, i, j) = 0.5 * (AT(fex, nx, m1(i - 1, nx), j) + AT(fex, nx, m1(i, nx), j));
            AT(ey, ld, i, j) =
                0.5 * (AT(fey, nx + 1, i, m1(j - 1, ny)) + AT(fey, nx + 1, i, m1(j, ny)));
            AT(bz, ld, i, j) =
                0.25 * (((AT(fbz, nx, m1(i - 1, nx), m1(j - 1, ny)) +
                          AT(fbz, nx, m1(i, nx), m1(j - 1, ny))) +
                         AT(fbz, nx, m1(i - 1, nx), m1(j, ny))) +
                        AT(fbz, nx, m1(i, nx), m1(j, ny)));
        }
}

/* compute_current!(jx, jy, m, ::CloudInCell, p): cloud_in_cell.jl:10-81 with
 * one thread (the per-thread partial arrays of :23-25 then hold everything),
 * then the ghost folding of :66-79.  jx, jy are (nx+1) x (ny+1). */
void or_compute_current(int64_t n, const double *p, int nx, int ny, double dx, double dy,
                        double *jx, double *jy)
{
    const int ld = nx + 1;
    memset(jx, 0, sizeof(double) * (size_t)ld * (ny + 1));
    memset(jy, 0, sizeof(double) * (size_t)ld * (ny + 1));
    const double scaling = 1 / (dx * dy);
    for (int64_t k = 0; k < n; ++k) {
        const double *q = p + 5 * k;
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
        double w = q[4];
        double w1 = q[2] * scaling * w;
        double w2 = q[3] * scaling * w;
        AT(jx, ld, i, j) += a1 * w1;
        AT(jy, ld, i, j) += a1 * w2;
        AT(jx, ld, i + 1, j) += a2 * w1;
        AT(jy, ld, i + 1, j) += a2 * w2;
        AT(jx, ld, i + 1, j + 1) += a3 * w1;
        AT(jy, ld, i + 1, j + 1) += a3 * w2;
        AT(jx, ld, i, j + 1) += a4 * w1;
        AT(jy, ld, i, j + 1) += a4 * w2;
    }
    for (int i = 1; i <= nx + 1; ++i) {
        AT(jx, ld, i, 1) += AT(jx, ld, i, ny + 1);
        AT(jx, ld, i, ny + 1) = AT(jx, ld, i, 1);
        AT(jy, ld, i, 1) += AT(jy, ld, i, ny + 1);
        AT(jy, ld, i, ny + 1) = AT(jy, ld, i, 1);
    }
    for (int j = 1; j <= ny + 1; ++j) {
        AT(jx, ld, 1, j) += AT(jx, ld, nx + 1, j);
        AT(jx, ld, nx + 1, j) = AT(jx, ld, 1, j);
        AT(jy, ld, 1, j) += AT(jy, ld, nx + 1, j);
        AT(jy, ld, nx + 1, j) = AT(jy, ld, 1, j);
    }
}

/* Julia's sum (pairwise) over a contiguous array. */
static double psum(const double *a, size_t n)
{
    if (n <= 1024) {
        double s = 0.0;
        for (size_t k = 0; k < n; ++k) s += a[k];
        return s;
    }
    size_t h = n / 2;
    return psum(a, h) + psum(a + h, n - h);
}

/* compute_energy(fdtd, m) = 0.5 log(sum(fdtd.ex .^ 2) dx dy): fdtd.jl:80. */
double or_fdtd_energy(int nx, int ny, double dx, double dy, const double *fex)
{
    size_t m = (size_t)nx * (ny + 1);
    double *sq = malloc(sizeof(double) * m);
    for (size_t k = 0; k < m; ++k) sq[k] = fex[k] * fex[k];
    double s = psum(sq, m);
    free(sq);
    return 0.5 * log(s * dx * dy);
}

void or_push_v_cic(int64_t n, double *p, double dt, int nx, int ny, double dx, double dy,
                   const double *ex, const double *ey, const double *bz, int ld);
void or_push_x(int64_t n, double *p, double dt, double dimx, double dimy);

/* The Vlasov-Maxwell loop of examples/landau_damping_maxwell_fdtd.jl:58-83
 * (and test/test_landau_damping.jl:50-82): faraday(dt/2) from the second
 * step, update_fields, push_v!(dt), push_x!(dt/2), compute_current!,
 * push_x!(dt/2), faraday(dt/2), ampere_maxwell!(dt); energy[s] =
 * compute_energy after the step (examples/vm2d.jl:42-56).  fex, fey, fbz are
 * the FDTD fields, updated in place. */
int or_maxwell_loop(int64_t n, double *p, int nx, int ny, double dimx, double dimy, double dt,
                    int nsteps, double *fex, double *fey, double *fbz, double *energy)
{
    const double dx = dimx / nx, dy = dimy / ny;
    const size_t mn = (size_t)(nx + 1) * (ny + 1);
    double *ex = calloc(mn, 8), *ey = calloc(mn, 8), *bz = calloc(mn, 8);
    double *jx = calloc(mn, 8), *jy = calloc(mn, 8);
    if (!ex || !ey || !bz || !jx || !jy) return -1;
    for (int s = 1; s <= nsteps; ++s) {
        if (s > 1) or_faraday(nx, ny, dx, dy, fex, fey, fbz, 0.5 * dt);
        or_update_fields(nx, ny, fex, fey, fbz, ex, ey, bz);
        or_push_v_cic(n, p, dt, nx, ny, dx, dy, ex, ey, bz, nx + 1);
        or_push_x(n, p, 0.5 * dt, dimx, dimy);
        or_compute_current(n, p, nx, ny, dx, dy, jx, jy);
        or_push_x(n, p, 0.5 * dt, dimx, dimy);
        or_faraday(nx, ny, dx, dy, fex, fey, fbz, 0.5 * dt);
        or_ampere_maxwell(nx, ny, dx, dy, fex, fey, fbz, jx, jy, dt);
        if (energy) energy[s - 1] = or_fdtd_energy(nx, ny, dx, dy, fex);
    }
    free(ex); free(ey); free(bz); free(jx); free(jy);
    return 0;
}
