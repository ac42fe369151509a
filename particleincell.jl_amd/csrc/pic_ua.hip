// pic_ua.hip -- the uniformly accurate (UA) PIC step of src/ua_steps.jl on
// the device: preparation!, update_particles_e! (the tau-fold M6 gather of
// interpolation_m6.jl:3-123), compute_f!, both ua_step! forms,
// update_particles_x! (the reconstruction and M6 deposit of
// compute_rho_m6.jl:29-179) and compute_v!, and the loop of test/bupdate.jl.
//
// A particle is worked by ntau consecutive lanes of one wave, lane m holding
// the tau-sample m of its (ntau, 2) arrays; the reference's (ntau, 2, n)
// layout is then read and written coalesced.  The tau-FFTs are radix-2
// transforms across those lanes with wave shuffles (FFTW's signs; the inverse
// scales by 1/ntau as Julia's ifft does).  The fused loop keeps between the
// prediction and the correction only the positions at the step's start and
// the tau-spectra of fx, fy; everything else is recomputed from them with the
// same operations, so it is bit-identical to what the reference holds.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_internal.h"

namespace pic {

namespace {

constexpr int UB = 256;  // threads per block

struct cd {
    double re, im;
};
__device__ __forceinline__ cd mk(double r, double i = 0.0) { return {r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd operator*(double s, cd a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ cd operator/(cd a, double s) { return {a.re / s, a.im / s}; }
__device__ __forceinline__ cd operator+(double s, cd a) { return {s + a.re, a.im}; }

__device__ __forceinline__ cd shfl_xor(cd z, int mask) {
    return {__shfl_xor(z.re, mask), __shfl_xor(z.im, mask)};
}
__device__ __forceinline__ cd shfl(cd z, int src) { return {__shfl(z.re, src), __shfl(z.im, src)}; }

template <int NT>
__device__ __forceinline__ int bitrev(int m) {
    int r = 0;
#pragma unroll
    for (int b = 1; b < NT; b <<= 1) r = (r << 1) | ((m & b) ? 1 : 0);
    return r;
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// This lane's twiddle of every FFT stage: stage h = NT/2, NT/4, ..., 1 uses
// cos(pi j / h) and sin(pi j / h), j = m mod h (computed once per kernel).
template <int NT>
struct Twiddles {
    double c[ilog2(NT)], s[ilog2(NT)];
    __device__ explicit Twiddles(int m) {
        int q = 0;
#pragma unroll
        for (int h = NT / 2; h >= 1; h >>= 1, ++q)
            sincospi((double)(m & (h - 1)) / (double)h, &s[q], &c[q]);
    }
};

// Radix-2 decimation-in-frequency FFT over the NT lanes of a group (lane m
// holds element m in and out): forward exp(-2 pi i jk / NT), or the inverse
// with exp(+...) scaled by 1/NT.
template <int NT, bool INV>
__device__ __forceinline__ cd fft_lanes(cd z, int m, int base, const Twiddles<NT> &w) {
    int q = 0;
#pragma unroll
    for (int h = NT / 2; h >= 1; h >>= 1, ++q) {
        // lower lane: z + p; upper lane: (p - z) w, without a branch
        const cd p = shfl_xor(z, h);
        const bool up = (m & h) != 0;
        const cd d = up ? p - z : z + p;
        z = up ? d * mk(w.c[q], INV ? w.s[q] : -w.s[q]) : d;
    }
    z = shfl(z, base + bitrev<NT>(m));
    if (INV) z = mk(z.re * (1.0 / NT), z.im * (1.0 / NT));
    return z;
}

// Sum over the NT lanes of a group (every lane receives it).
template <int NT>
__device__ __forceinline__ cd lane_sum(cd z) {
#pragma unroll
    for (int h = 1; h < NT; h <<= 1) z = z + shfl_xor(z, h);
    return z;
}

// Per-lane constants of UA(ntau, eps): tau = m dtau, ltau (ua_type.jl:18-24),
// and the lane's FFT twiddles.
template <int NT>
struct UaLane {
    int m;
    double l, c, s;  // ltau[m], cos(tau[m]), sin(tau[m])
    Twiddles<NT> w;
    __device__ explicit UaLane(int m_) : m(m_), w(m_) {
        l = m < NT / 2 ? (double)m : (double)(m - NT);
        const double tau = (double)m * (2 * M_PI / NT);
        sincos(tau, &s, &c);
    }
};

// exp(sg 1im l t / eps) as Julia evaluates it: cis(((sg l) t) / eps).
__device__ __forceinline__ cd ua_cis(int sg, double l, double t, double eps) {
    const double th = (((double)sg * l) * t) / eps;
    double s, c;
    sincos(th, &s, &c);
    return mk(c, s);
}

// pl[m], ql[m] of preparation! (ua_steps.jl:26-33); elt = exp(-1im l t / eps).
template <int NT>
__device__ __forceinline__ void ua_plql(const UaLane<NT> &L, double t, double eps, cd elt, cd &pl,
                                        cd &ql) {
    if (L.m == 0) {
        pl = mk(t);
        ql = mk(t * t / 2);
        return;
    }
    pl = (mk(0.0, eps) * (elt - mk(1.0))) / L.l;                         // eps*1im*(elt-1)/l
    ql = (eps * (eps * (mk(1.0) - elt) - mk(0.0 * t, L.l * t))) / (L.l * L.l);  // eps*(eps*(1-elt)-1im*l*t)/l^2
}

// preparation! for one particle (ua_steps.jl:3-83): b, t, xt (real) and yt
// at this lane's tau.
template <int NT>
__device__ __forceinline__ void ua_prep(const UaLane<NT> &L, int base, double x1, double x2,
                                        double vx, double vy, double ex, double ey, double dt,
                                        double eps, double &b, double &t, double xt[2], cd yt[2],
                                        double &interv_out) {
    b = 1 + 0.5 * sin(x1) * sin(x2);
    t = dt * b;
    const double vxb = vx / b, vyb = vy / b;
    const double c = L.c, s = L.s;
    const double h1 = eps * (s * vxb - c * vyb);
    const double h2 = eps * (s * vyb + c * vxb);
    xt[0] = x1 + h1 + eps * vyb;
    xt[1] = x2 + h2 - eps * vxb;
    const double interv = (1 + 0.5 * sin(xt[0]) * sin(xt[1]) - b) / eps;
    interv_out = interv;
    const double exb = ((c * vy - s * vx) * interv + ex) / b;
    const double eyb = ((-c * vx - s * vy) * interv + ey) / b;
    cd r[2] = {mk(c * exb - s * eyb), mk(s * exb + c * eyb)};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        cd R = fft_lanes<NT, false>(r[q], L.m, base, L.w);
        if (L.m != 0) R = mk(R.im / L.l, -R.re / L.l);  // -1im * R / l
        r[q] = fft_lanes<NT, true>(R, L.m, base, L.w);
    }
    const cd r10 = shfl(r[0], base), r20 = shfl(r[1], base);
    yt[0] = vx + eps * (r[0] - r10);
    yt[1] = vy + eps * (r[1] - r20);
}

// (1 + sin(x1) sin(x2)/2 - b)/eps at the lane's position (ua_steps.jl:55, :130).
__device__ __forceinline__ double ua_interv(const double xt[2], double b, double eps) {
    return (1 + 0.5 * sin(xt[0]) * sin(xt[1]) - b) / eps;
}

// compute_f! at this lane's tau (ua_steps.jl:117-145), before its FFTs;
// interv from ua_interv at xt.
template <int NT>
__device__ __forceinline__ void ua_f(const UaLane<NT> &L, double b, double interv,
                                     const cd yt[2], const double et[2], cd fx[2], cd fy[2]) {
    const double c = L.c, s = L.s;
    fx[0] = (c * yt[0] + s * yt[1]) / b;
    fx[1] = (-s * yt[0] + c * yt[1]) / b;
    const cd tmp1 = et[0] + interv * (c * yt[1] - s * yt[0]);
    const cd tmp2 = et[1] + interv * (-c * yt[0] - s * yt[1]);
    fy[0] = (c * tmp1 - s * tmp2) / b;
    fy[1] = (s * tmp1 + c * tmp2) / b;
}

// The gather of update_particles_e! at real(xt) (interpolation_m6.jl:22-120).
__device__ __forceinline__ void ua_gather(const double xt[2], int nx, int ny, double xmin,
                                          double ymin, double dimx, double dimy, double dx,
                                          double dy, const double *ex, const double *ey,
                                          double et[2]) {
    const double xn = picd::jl_mod_pos(xt[0] - xmin, dimx);
    const double yn = picd::jl_mod_pos(xt[1] - ymin, dimy);
    const double px = xn / dx, py = yn / dy;
    const int i = (int)px, j = (int)py;
    picd::mf_gather_m6(i, j, px - (double)i, py - (double)j, nx, ny, ex, ey, et[0], et[1]);
}

struct UaGrid {
    int nx, ny;
    double xmin, ymin, dimx, dimy, dx, dy;
    const double *ex, *ey;  // node fields, nx x ny
    double *rho_acc;        // raw charge, nx x ny
    double w;               // particle weight
};

// update_particles_x! for one particle (compute_rho_m6.jl:37-164): x(t) from
// its tau-samples, into the domain, stored; the M6 deposit (nodes i-2..i+3,
// wrapped, as in :70-81) spread over the group's lanes.
// eplus = exp(+1im l t / eps), the conjugate of elt (cos is even and sin odd
// in the device library, and the two angles differ only in sign).
template <int NT>
__device__ __forceinline__ void ua_reconstruct_deposit(const UaLane<NT> &L, int base, cd eplus,
                                                       const cd xt[2], const UaGrid &g, bool live,
                                                       double *xo, double *yo) {
    double pos[2];
    const cd e = eplus / (double)NT;
#pragma unroll
    for (int q = 0; q < 2; ++q)
        pos[q] = lane_sum<NT>(fft_lanes<NT, false>(xt[q], L.m, base, L.w) * e).re;
    const double xt1 = picd::jl_mod_pos(pos[0] - g.xmin, g.dimx);
    const double xt2 = picd::jl_mod_pos(pos[1] - g.ymin, g.dimy);
    if (!live) return;
    if (L.m == 0) {
        *xo = xt1 + g.xmin;
        *yo = xt2 + g.ymin;
    }
    const double px = xt1 / g.dx, py = xt2 / g.dy;
    const int i = (int)px, j = (int)py;
    double wx[6], wy[6];
    picd::m6_weights_d(px - (double)i, wx);
    picd::m6_weights_d(py - (double)j, wy);
    for (int q = L.m; q < 36; q += NT) {
        const int a = q / 6, bb = q % 6;
        const int ii = picd::wrap2(i - 2 + a, g.nx), jj = picd::wrap2(j - 2 + bb, g.ny);
        unsafeAtomicAdd(&g.rho_acc[ii + (size_t)g.nx * jj], wx[a] * wy[bb] * g.w);
    }
}

// compute_v! for one particle (ua_steps.jl:206-224) from the tau-spectrum yt.
template <int NT>
__device__ __forceinline__ void ua_v(const UaLane<NT> &L, double t, double eps, cd eplus,
                                     const cd yt[2], double &v1, double &v2) {
    const cd e = eplus;
    const cd px = lane_sum<NT>((yt[0] / (double)NT) * e);
    const cd py = lane_sum<NT>((yt[1] / (double)NT) * e);
    const double ct = cos(t / eps), st = sin(t / eps);
    v1 = ct * px.re + st * py.re;
    v2 = ct * py.re - st * px.re;
}

// Lane, group and particle of this thread.
template <int NT>
struct Slot {
    int m, base;
    int64_t k;
    bool live;
    __device__ Slot(int64_t n) {
        const int lane = threadIdx.x & 63;
        m = lane & (NT - 1);
        base = lane & ~(NT - 1);
        const int64_t k0 = (int64_t)blockIdx.x * (UB / NT) + threadIdx.x / NT;
        live = k0 < n;
        k = live ? k0 : n - 1;  // idle groups shadow the last particle
    }
};

// (m, c, k) of a (ntau, 2, n) array.
__device__ __forceinline__ size_t ux(int m, int c, int64_t k, int nt) {
    return (size_t)m + (size_t)nt * ((size_t)c + 2 * (size_t)k);
}

// (m, q, k) of the (ntau, 4, n) spectra of fx (q = 0, 1) and fy (q = 2, 3).
__device__ __forceinline__ size_t uf(int m, int q, int64_t k, int nt) {
    return (size_t)m + (size_t)nt * ((size_t)q + 4 * (size_t)k);
}

struct UaArgs {
    int64_t n;
    double eps, dt;
    double *x, *y, *vx, *vy;  // Particles.x, .v (SoA)
    const double *epx, *epy;  // Particles.e
    double *b, *t;            // Particles.b, .t
    double *x0, *y0;          // positions at the step's start
    cd *fhat;                 // tau-spectra of fx, fy: (ntau, 4, n)
    UaGrid g;
};

// ---------------------------------------------------------------- the loop

// Prediction of test/bupdate.jl:64-84: preparation!, update_particles_e!,
// compute_f!, x~t = fft(xt), ua_step!, the same for yt, ifft!, then
// update_particles_x!.  Keeps x at the step's start and the spectra of fx, fy.
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_predict(UaArgs a) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    const double x1 = a.x[k], x2 = a.y[k];
    double b, t, xt[2], et[2], interv;
    cd yt[2], fx[2], fy[2];
    ua_prep<NT>(L, S.base, x1, x2, a.vx[k], a.vy[k], a.epx[k], a.epy[k], a.dt, a.eps, b, t, xt,
                yt, interv);
    ua_gather(xt, a.g.nx, a.g.ny, a.g.xmin, a.g.ymin, a.g.dimx, a.g.dimy, a.g.dx, a.g.dy, a.g.ex,
              a.g.ey, et);
    ua_f<NT>(L, b, interv, yt, et, fx, fy);
    const cd elt = ua_cis(-1, L.l, t, a.eps);
    cd pl, ql;
    ua_plql<NT>(L, t, a.eps, elt, pl, ql);
    cd xn[2], yn[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        fx[q] = fft_lanes<NT, false>(fx[q], L.m, S.base, L.w);
        fy[q] = fft_lanes<NT, false>(fy[q], L.m, S.base, L.w);
        const cd xs = fft_lanes<NT, false>(mk(xt[q]), L.m, S.base, L.w);
        const cd ys = fft_lanes<NT, false>(yt[q], L.m, S.base, L.w);
        xn[q] = fft_lanes<NT, true>(elt * xs + pl * fx[q], L.m, S.base, L.w);
        yn[q] = fft_lanes<NT, true>(elt * ys + pl * fy[q], L.m, S.base, L.w);
    }
    if (S.live) {
        if (S.m == 0) {
            a.x0[k] = x1;
            a.y0[k] = x2;
            a.b[k] = b;
            a.t[k] = t;
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            a.fhat[uf(S.m, q, k, NT)] = fx[q];
            a.fhat[uf(S.m, 2 + q, k, NT)] = fy[q];
        }
    }
    ua_reconstruct_deposit<NT>(L, S.base, mk(elt.re, -elt.im), xn, a.g, S.live, &a.x[k],
                               &a.y[k]);
}

// Correction of test/bupdate.jl:90-114: the predicted xt, yt rebuilt from the
// step's start, update_particles_e! on the new field, compute_f!, the
// corrected ua_step!s, ifft!(xt), update_particles_x! and compute_v!.
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_correct(UaArgs a) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    double b, t, xt[2], et[2], interv;
    cd yt[2];
    ua_prep<NT>(L, S.base, a.x0[k], a.y0[k], a.vx[k], a.vy[k], a.epx[k], a.epy[k], a.dt, a.eps,
                b, t, xt, yt, interv);
    const cd elt = ua_cis(-1, L.l, t, a.eps);
    cd pl, ql;
    ua_plql<NT>(L, t, a.eps, elt, pl, ql);
    cd fx[2], fy[2], xs[2], ys[2], xp[2], yp[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        fx[q] = a.fhat[uf(S.m, q, k, NT)];
        fy[q] = a.fhat[uf(S.m, 2 + q, k, NT)];
        xs[q] = fft_lanes<NT, false>(mk(xt[q]), L.m, S.base, L.w);
        ys[q] = fft_lanes<NT, false>(yt[q], L.m, S.base, L.w);
        xp[q] = fft_lanes<NT, true>(elt * xs[q] + pl * fx[q], L.m, S.base, L.w);
        yp[q] = fft_lanes<NT, true>(elt * ys[q] + pl * fy[q], L.m, S.base, L.w);
    }
    const double xr[2] = {xp[0].re, xp[1].re};
    ua_gather(xr, a.g.nx, a.g.ny, a.g.xmin, a.g.ymin, a.g.dimx, a.g.dimy, a.g.dx, a.g.dy, a.g.ex,
              a.g.ey, et);
    cd gx[2], gy[2];
    ua_f<NT>(L, b, ua_interv(xr, b, a.eps), yp, et, gx, gy);
    cd xn[2], yn[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        gx[q] = fft_lanes<NT, false>(gx[q], L.m, S.base, L.w);
        gy[q] = fft_lanes<NT, false>(gy[q], L.m, S.base, L.w);
        xn[q] = elt * xs[q] + pl * fx[q];
        xn[q] = xn[q] + (ql * (gx[q] - fx[q])) / t;
        xn[q] = fft_lanes<NT, true>(xn[q], L.m, S.base, L.w);
        yn[q] = elt * ys[q] + pl * fy[q];
        yn[q] = yn[q] + (ql * (gy[q] - fy[q])) / t;
    }
    const cd eplus = mk(elt.re, -elt.im);
    ua_reconstruct_deposit<NT>(L, S.base, eplus, xn, a.g, S.live, &a.x[k], &a.y[k]);
    double v1, v2;
    ua_v<NT>(L, t, a.eps, eplus, yn, v1, v2);
    if (S.live && S.m == 0) {
        a.vx[k] = v1;
        a.vy[k] = v2;
    }
}

// ------------------------------------------------------- per-call pieces

// preparation!: b, t, pl, ql, xt, yt (any output may be null).
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_prep_call(UaArgs a, cd *pl, cd *ql, cd *xt, cd *yt) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    double b, t, x[2], interv;
    cd y[2];
    ua_prep<NT>(L, S.base, a.x[k], a.y[k], a.vx[k], a.vy[k], a.epx[k], a.epy[k], a.dt, a.eps, b,
                t, x, y, interv);
    cd p, q;
    ua_plql<NT>(L, t, a.eps, ua_cis(-1, L.l, t, a.eps), p, q);
    if (!S.live) return;
    if (S.m == 0) {
        a.b[k] = b;
        a.t[k] = t;
    }
    if (pl) pl[S.m + (size_t)NT * k] = p;
    if (ql) ql[S.m + (size_t)NT * k] = q;
    for (int c = 0; c < 2; ++c) {
        if (xt) xt[ux(S.m, c, k, NT)] = mk(x[c]);
        if (yt) yt[ux(S.m, c, k, NT)] = y[c];
    }
}

// update_particles_e!: et from real(xt).
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_gather_call(UaArgs a, const cd *xt, double *et) {
    const Slot<NT> S(a.n);
    const int64_t k = S.k;
    const double x[2] = {xt[ux(S.m, 0, k, NT)].re, xt[ux(S.m, 1, k, NT)].re};
    double e[2];
    ua_gather(x, a.g.nx, a.g.ny, a.g.xmin, a.g.ymin, a.g.dimx, a.g.dimy, a.g.dx, a.g.dy, a.g.ex,
              a.g.ey, e);
    if (!S.live) return;
    et[ux(S.m, 0, k, NT)] = e[0];
    et[ux(S.m, 1, k, NT)] = e[1];
}

// compute_f!, with its closing FFTs.
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_f_call(UaArgs a, const cd *xt, const cd *yt,
                                                  const double *et, cd *fx, cd *fy) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    const double x[2] = {xt[ux(S.m, 0, k, NT)].re, xt[ux(S.m, 1, k, NT)].re};
    const cd y[2] = {yt[ux(S.m, 0, k, NT)], yt[ux(S.m, 1, k, NT)]};
    const double e[2] = {et[ux(S.m, 0, k, NT)], et[ux(S.m, 1, k, NT)]};
    const double b = a.b[k];
    cd f[2], h[2];
    ua_f<NT>(L, b, ua_interv(x, b, a.eps), y, e, f, h);
    for (int c = 0; c < 2; ++c) {
        f[c] = fft_lanes<NT, false>(f[c], L.m, S.base, L.w);
        h[c] = fft_lanes<NT, false>(h[c], L.m, S.base, L.w);
    }
    if (!S.live) return;
    for (int c = 0; c < 2; ++c) {
        fx[ux(S.m, c, k, NT)] = f[c];
        fy[ux(S.m, c, k, NT)] = h[c];
    }
}

// ua_step! (gx null) or its corrected form.
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_step_call(UaArgs a, const cd *xtil, const cd *fx,
                                                     const cd *gx, cd *xt) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    const double t = a.t[k];
    const cd elt = ua_cis(-1, L.l, t, a.eps);
    cd pl, ql;
    ua_plql<NT>(L, t, a.eps, elt, pl, ql);
    if (!S.live) return;
    for (int c = 0; c < 2; ++c) {
        const size_t u = ux(S.m, c, k, NT);
        const cd f = fx[u];
        cd z = elt * xtil[u] + pl * f;
        if (gx) z = z + (ql * (gx[u] - f)) / t;
        xt[u] = z;
    }
}

// update_particles_x!: reconstruction into Particles.x and the deposit.
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_rho_call(UaArgs a, const cd *xt) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    const cd x[2] = {xt[ux(S.m, 0, k, NT)], xt[ux(S.m, 1, k, NT)]};
    ua_reconstruct_deposit<NT>(L, S.base, ua_cis(+1, L.l, a.t[k], a.eps), x, a.g, S.live,
                               &a.x[k], &a.y[k]);
}

// compute_v!.
template <int NT>
__global__ __launch_bounds__(UB) void k_ua_v_call(UaArgs a, const cd *yt) {
    const Slot<NT> S(a.n);
    const UaLane<NT> L(S.m);
    const int64_t k = S.k;
    const cd y[2] = {yt[ux(S.m, 0, k, NT)], yt[ux(S.m, 1, k, NT)]};
    const double t = a.t[k];
    double v1, v2;
    ua_v<NT>(L, t, a.eps, ua_cis(+1, L.l, t, a.eps), y, v1, v2);
    if (S.live && S.m == 0) {
        a.vx[k] = v1;
        a.vy[k] = v2;
    }
}

// ------------------------------------------------------------------ host

#define UA_NT_SWITCH(nt, ...)                          \
    switch (nt) {                                      \
        case 2: { constexpr int NT = 2; __VA_ARGS__; } break;   \
        case 4: { constexpr int NT = 4; __VA_ARGS__; } break;   \
        case 8: { constexpr int NT = 8; __VA_ARGS__; } break;   \
        case 16: { constexpr int NT = 16; __VA_ARGS__; } break; \
        case 32: { constexpr int NT = 32; __VA_ARGS__; } break; \
        case 64: { constexpr int NT = 64; __VA_ARGS__; } break; \
        default: return fail(PIC_ERR_ARG, "ntau %d", nt);       \
    }

int ua_alloc(double **p, size_t count) {
    if (*p) return PIC_OK;
    const hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 2) * 8);
    if (e != hipSuccess) return fail(PIC_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * 8, hipGetErrorString(e));
    return PIC_OK;
}

void ua_free(Ctx *c) {
    for (double **p : {&c->ua_b, &c->ua_t, &c->ua_x0, &c->ua_y0, &c->ua_fhat}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    c->ua_cap = 0;
}

// The per-particle UA buffers for the context's particle count.
int ua_ensure(Ctx *c) {
    PIC_CHECK(c->ua_ntau > 0, PIC_ERR_STATE, "pic_ua_init has not been called");
    PIC_CHECK(c->state == PIC_STATE_F64, PIC_ERR_STATE, "the UA step runs in the FP64 state");
    PIC_CHECK(c->uniform_w, PIC_ERR_STATE, "the UA step needs one particle weight (Particles.w)");
    if (c->ua_cap < c->n) ua_free(c);
    if (!c->ua_b) {
        const size_t n = (size_t)std::max<int64_t>(c->n, 1);
        PIC_TRY(ua_alloc(&c->ua_b, n));
        PIC_TRY(ua_alloc(&c->ua_t, n));
        PIC_TRY(ua_alloc(&c->ua_x0, n));
        PIC_TRY(ua_alloc(&c->ua_y0, n));
        PIC_TRY(ua_alloc(&c->ua_fhat, n * (size_t)c->ua_ntau * 8));
        c->ua_cap = (int64_t)n;
    }
    PIC_TRY(ensure_pfields(c));
    return PIC_OK;
}

UaArgs ua_args(Ctx *c, double dt) {
    UaArgs a;
    a.n = c->n;
    a.eps = c->ua_eps;
    a.dt = dt;
    a.x = c->x; a.y = c->y; a.vx = c->vx; a.vy = c->vy;
    a.epx = c->epx; a.epy = c->epy;
    a.b = c->ua_b; a.t = c->ua_t;
    a.x0 = c->ua_x0; a.y0 = c->ua_y0;
    a.fhat = reinterpret_cast<cd *>(c->ua_fhat);
    a.g.nx = c->nx; a.g.ny = c->ny;
    a.g.xmin = c->xmin; a.g.ymin = c->ymin;
    a.g.dimx = c->dimx; a.g.dimy = c->dimy;
    a.g.dx = c->dx; a.g.dy = c->dy;
    a.g.ex = c->ex; a.g.ey = c->ey;
    a.g.rho_acc = c->rho_acc;
    a.g.w = c->w0;
    return a;
}

inline dim3 ua_grid(int64_t n, int nt) {
    return dim3((unsigned)((n + UB / nt - 1) / (UB / nt)));
}

// A (ntau, 2, n) complex host array (interleaved doubles) on the device.
struct DevC {
    cd *p = nullptr;
    ~DevC() {
        if (p) (void)hipFree(p);
    }
};

int to_dev(Ctx *c, DevC &d, const double *host, size_t count_c) {
    if (!host) return PIC_OK;
    PIC_HIP(hipMalloc((void **)&d.p, std::max<size_t>(count_c, 1) * 16));
    PIC_HIP(hipMemcpyAsync(d.p, host, count_c * 16, hipMemcpyHostToDevice, c->stream));
    return PIC_OK;
}

int alloc_dev(DevC &d, size_t count_c) {
    PIC_HIP(hipMalloc((void **)&d.p, std::max<size_t>(count_c, 1) * 16));
    return PIC_OK;
}

int from_dev(Ctx *c, double *host, const DevC &d, size_t count_c) {
    if (!host) return PIC_OK;
    PIC_HIP(hipMemcpyAsync(host, d.p, count_c * 16, hipMemcpyDeviceToHost, c->stream));
    return PIC_OK;
}

// Particles.b or .t from the host into the context (NULL: keep the ones of
// the last preparation).
int load_bt(Ctx *c, double *dst, const double *src) {
    if (src && c->n)
        PIC_HIP(hipMemcpyAsync(dst, src, (size_t)c->n * 8, hipMemcpyHostToDevice, c->stream));
    return PIC_OK;
}

int ua_begin(pic_ctx *ctx, Ctx *&c) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    return ua_ensure(c);
}

}  // namespace

}  // namespace pic

using namespace pic;

extern "C" {

int pic_ua_init(pic_ctx *ctx, int ntau, double eps) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(ntau >= 2 && ntau <= 64 && (ntau & (ntau - 1)) == 0, PIC_ERR_ARG,
              "ntau must be a power of two from 2 to 64 (got %d)", ntau);
    PIC_CHECK(eps > 0.0, PIC_ERR_ARG, "eps must be positive");
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    if (c->ua_ntau != ntau) ua_free(c);
    c->ua_ntau = ntau;
    c->ua_eps = eps;
    return PIC_OK;
}

int pic_ua_preparation(pic_ctx *ctx, double dt, double *b, double *t, double *pl, double *ql,
                       double *xt, double *yt) {
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    const int nt = c->ua_ntau;
    const size_t n = (size_t)c->n, nc = (size_t)nt * 2 * n;
    DevC dpl, dql, dxt, dyt;
    if (pl) PIC_TRY(alloc_dev(dpl, (size_t)nt * n));
    if (ql) PIC_TRY(alloc_dev(dql, (size_t)nt * n));
    if (xt) PIC_TRY(alloc_dev(dxt, nc));
    if (yt) PIC_TRY(alloc_dev(dyt, nc));
    if (n) {
        const UaArgs a = ua_args(c, dt);
        UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_prep_call<NT>, ua_grid(c->n, NT), dim3(UB), 0,
                                            c->stream, a, dpl.p, dql.p, dxt.p, dyt.p));
        PIC_HIP(hipGetLastError());
    }
    if (b) PIC_HIP(hipMemcpyAsync(b, c->ua_b, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (t) PIC_HIP(hipMemcpyAsync(t, c->ua_t, n * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_TRY(from_dev(c, pl, dpl, (size_t)nt * n));
    PIC_TRY(from_dev(c, ql, dql, (size_t)nt * n));
    PIC_TRY(from_dev(c, xt, dxt, nc));
    PIC_TRY(from_dev(c, yt, dyt, nc));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_ua_interpol(pic_ctx *ctx, const double *xt, double *et) {
    PIC_CHECK(xt && et, PIC_ERR_ARG, "null pointer");
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    const int nt = c->ua_ntau;
    const size_t n = (size_t)c->n, nc = (size_t)nt * 2 * n;
    DevC dxt, det;
    PIC_TRY(to_dev(c, dxt, xt, nc));
    PIC_TRY(alloc_dev(det, (nc + 1) / 2));
    if (n) {
        const UaArgs a = ua_args(c, 0.0);
        UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_gather_call<NT>, ua_grid(c->n, NT), dim3(UB), 0,
                                            c->stream, a, dxt.p, (double *)det.p));
        PIC_HIP(hipGetLastError());
        PIC_HIP(hipMemcpyAsync(et, det.p, nc * 8, hipMemcpyDeviceToHost, c->stream));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_ua_compute_f(pic_ctx *ctx, const double *b, const double *xt, const double *yt,
                     const double *et, double *fx, double *fy) {
    PIC_CHECK(xt && yt && et && fx && fy, PIC_ERR_ARG, "null pointer");
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    PIC_TRY(load_bt(c, c->ua_b, b));
    const int nt = c->ua_ntau;
    const size_t n = (size_t)c->n, nc = (size_t)nt * 2 * n;
    DevC dxt, dyt, det, dfx, dfy;
    PIC_TRY(to_dev(c, dxt, xt, nc));
    PIC_TRY(to_dev(c, dyt, yt, nc));
    PIC_TRY(alloc_dev(det, (nc + 1) / 2));
    if (n) PIC_HIP(hipMemcpyAsync(det.p, et, nc * 8, hipMemcpyHostToDevice, c->stream));
    PIC_TRY(alloc_dev(dfx, nc));
    PIC_TRY(alloc_dev(dfy, nc));
    if (n) {
        const UaArgs a = ua_args(c, 0.0);
        UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_f_call<NT>, ua_grid(c->n, NT), dim3(UB), 0,
                                            c->stream, a, dxt.p, dyt.p, (const double *)det.p,
                                            dfx.p, dfy.p));
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(from_dev(c, fx, dfx, nc));
    PIC_TRY(from_dev(c, fy, dfy, nc));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_ua_step(pic_ctx *ctx, const double *t, const double *xtilde, const double *fx,
                const double *gx, double *xt) {
    PIC_CHECK(xtilde && fx && xt, PIC_ERR_ARG, "null pointer");
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    PIC_TRY(load_bt(c, c->ua_t, t));
    const int nt = c->ua_ntau;
    const size_t n = (size_t)c->n, nc = (size_t)nt * 2 * n;
    DevC dxs, dfx, dgx, dxt;
    PIC_TRY(to_dev(c, dxs, xtilde, nc));
    PIC_TRY(to_dev(c, dfx, fx, nc));
    PIC_TRY(to_dev(c, dgx, gx, nc));
    PIC_TRY(alloc_dev(dxt, nc));
    if (n) {
        const UaArgs a = ua_args(c, 0.0);
        UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_step_call<NT>, ua_grid(c->n, NT), dim3(UB), 0,
                                            c->stream, a, dxs.p, dfx.p, dgx.p, dxt.p));
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(from_dev(c, xt, dxt, nc));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_ua_compute_rho(pic_ctx *ctx, const double *t, const double *xt) {
    PIC_CHECK(xt, PIC_ERR_ARG, "null pointer");
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    PIC_TRY(load_bt(c, c->ua_t, t));
    const int nt = c->ua_ntau;
    const size_t n = (size_t)c->n, nc = (size_t)nt * 2 * n;
    DevC dxt;
    PIC_TRY(to_dev(c, dxt, xt, nc));
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, (int64_t)c->nx * c->ny));
    if (n) {
        const UaArgs a = ua_args(c, 0.0);
        UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_rho_call<NT>, ua_grid(c->n, NT), dim3(UB), 0,
                                            c->stream, a, dxt.p));
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(finalize_rho(c, PIC_M6));
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->seg_valid = false;
    c->sort_fresh = false;
    return PIC_OK;
}

int pic_ua_compute_v(pic_ctx *ctx, const double *t, const double *yt) {
    PIC_CHECK(yt, PIC_ERR_ARG, "null pointer");
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    PIC_TRY(load_bt(c, c->ua_t, t));
    const int nt = c->ua_ntau;
    const size_t n = (size_t)c->n, nc = (size_t)nt * 2 * n;
    DevC dyt;
    PIC_TRY(to_dev(c, dyt, yt, nc));
    if (n) {
        const UaArgs a = ua_args(c, 0.0);
        UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_v_call<NT>, ua_grid(c->n, NT), dim3(UB), 0,
                                            c->stream, a, dyt.p));
        PIC_HIP(hipGetLastError());
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_ua_run(pic_ctx *ctx, double dt, int nsteps, double *energy) {
    PIC_CHECK(nsteps >= 0, PIC_ERR_ARG, "nsteps %d", nsteps);
    Ctx *c = nullptr;
    PIC_TRY(ua_begin(ctx, c));
    const int nt = c->ua_ntau;
    const int64_t m = (int64_t)c->nx * c->ny;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (c->timing) {
        PIC_HIP(hipEventCreate(&ev[0]));
        PIC_HIP(hipEventCreate(&ev[1]));
    }
    struct EvFree {
        hipEvent_t *e;
        ~EvFree() {
            for (int k = 0; k < 2; ++k)
                if (e[k]) (void)hipEventDestroy(e[k]);
        }
    } evfree{ev};
    // energies of every half step queued on the device (pairs of interior
    // sums, then the ghost-node terms) and read once after the loop
    const int64_t nh = 2 * (int64_t)nsteps;
    if (energy) PIC_TRY(ensure_energy(c, nh + (nh + 1) / 2));
    double *ghost = c->energy_hist + 2 * nh;
    for (int s = 0; s < nsteps; ++s) {
        for (int half = 0; half < 2; ++half) {
            PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
            if (c->n) {
                const UaArgs a = ua_args(c, dt);
                if (c->timing) PIC_HIP(hipEventRecord(ev[0], c->stream));
                if (half == 0) {
                    UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_predict<NT>, ua_grid(c->n, NT),
                                                        dim3(UB), 0, c->stream, a));
                } else {
                    UA_NT_SWITCH(nt, hipLaunchKernelGGL(k_ua_correct<NT>, ua_grid(c->n, NT),
                                                        dim3(UB), 0, c->stream, a));
                }
                PIC_HIP(hipGetLastError());
                if (c->timing) {
                    // timing waits for each kernel (the untimed loop does not wait)
                    PIC_HIP(hipEventRecord(ev[1], c->stream));
                    PIC_HIP(hipEventSynchronize(ev[1]));
                    float ms = 0.0f;
                    PIC_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
                    c->t_ua_ms[half] += ms;
                    c->t_ua_n[half] += 1;
                }
            }
            PIC_TRY(finalize_rho(c, PIC_M6));
            // pic_solve_fields without its wait: E, then the energy sums
            PIC_TRY(poisson_solve_e(c, c->rho, 1.0 / ((double)c->nx * (double)c->ny), 1, -1));
            if (energy) PIC_TRY(field_energy_async(c, 2 * s + half, ghost + 2 * s + half));
        }
    }
    if (energy && nh > 0) {
        std::vector<double> h(3 * (size_t)nh);
        PIC_HIP(hipMemcpyAsync(h.data(), c->energy_hist, h.size() * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        for (int64_t k = 0; k < nh; ++k)
            energy[k] = h[2 * k] + h[2 * k + 1] + h[2 * nh + k] * c->dx * c->dy;
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->seg_valid = false;
    c->sort_fresh = false;
    return PIC_OK;
}

int pic_get_ua_timing(pic_ctx *ctx, double *ms, int64_t *launches) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    for (int k = 0; k < 2; ++k) {
        if (ms) ms[k] = c->t_ua_ms[k];
        if (launches) launches[k] = c->t_ua_n[k];
    }
    return PIC_OK;
}

}  // extern "C"
