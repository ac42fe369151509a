// pic_run.hip -- the hot-path driver's time loop on the device.
//
// examples/landau_damping_poisson_cic_m6.jl:70-83 runs, per iteration,
//   push_x!(dt/2); rho = compute_rho; solve!(ex, ey); energy; push_v!(dt);
//   push_x!(dt/2)
// Consecutive half drifts belong to consecutive iterations, so the device
// loop is
//   prologue:   drift(dt/2) + deposit                       (one pass)
//   step s:     solve(rho) -> E, energy[s]
//               kick(dt) + drift(dt/2) + drift(dt/2) + deposit (s < last)
//               kick(dt) + drift(dt/2)                          (s = last)
// which is the same sequence of operations on every particle.  In the FP32
// state the two half drifts are one drift of dt.
#include <vector>

#include "pic_internal.h"

namespace pic {

int launch_step64(Ctx *c, int kernel, int mode, double dt, double dtd);
int launch_step32(Ctx *c, int kernel, int mode, double dt, double dtd);
int launch_seg64(Ctx *c, int kernel, int mode, double dt, double dtd);
int launch_seg32(Ctx *c, int kernel, int mode, double dt, double dtd);

// A pass of pic_run: segment-persistent kernels over the lane-sorted slots.
static int seg_pass(Ctx *c, int kernel, int mode, double dt) {
    if (mode == 3)
        return c->state == PIC_STATE_F64 ? launch_seg64(c, kernel, 3, 0.0, 0.0)
                                         : launch_seg32(c, kernel, 3, 0.0, 0.0);
    if (c->state == PIC_STATE_F64) return launch_seg64(c, kernel, mode, dt, 0.5 * dt);
    return launch_seg32(c, kernel, mode, dt, mode == 1 ? dt : 0.5 * dt);
}

// A pass of pic_step_pass: chunk kernels, any particle order.
static int step_pass(Ctx *c, int kernel, int mode, double dt) {
    if (mode == 3) {
        return c->state == PIC_STATE_F64 ? launch_step64(c, kernel, 3, 0.0, 0.0)
                                         : launch_step32(c, kernel, 3, 0.0, 0.0);
    }
    if (c->state == PIC_STATE_F64) {
        // modes 0 and 2 drift once by dt/2, mode 1 twice by dt/2
        return launch_step64(c, kernel, mode, dt, 0.5 * dt);
    }
    return launch_step32(c, kernel, mode, dt, mode == 1 ? dt : 0.5 * dt);
}

// pic_run's loop as separate kernels: the E solve, then a segment pass over
// the lane-sorted particles, a lane sort every sort_every steps.
static int run_passes(Ctx *c, int kernel, double dt, int nsteps, int sort_every,
                      std::vector<hipEvent_t> &ev) {
    PIC_HIP(hipMemsetAsync(c->step_counter, 0, sizeof(int), c->stream));
    const int64_t m = (int64_t)c->nx * c->ny;
    const double scale = 1.0 / ((double)c->nx * (double)c->ny * c->dx * c->dy);
    const int f32 = c->state == PIC_STATE_F32;
    // the passes work segment by segment on lane-sorted slots (a lane sort
    // whose particles have not moved since needs no repeat)
    if (!(c->seg_valid && c->sort_fresh)) PIC_TRY(sort_by_cell(c, nullptr, true, false));
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
    PIC_TRY(seg_pass(c, kernel, 0, dt));
    auto mark = [&]() -> int {
        hipEvent_t e;
        PIC_HIP(hipEventCreate(&e));
        ev.push_back(e);
        PIC_HIP(hipEventRecord(e, c->stream));
        return PIC_OK;
    };
    int rc = PIC_OK;
    for (int s = 0; s < nsteps && rc == PIC_OK; ++s) {
        const bool last = s == nsteps - 1;
        if (last) rc = finalize_rho(c, kernel);  // leave rho of the last step
        // post pass: FP32 field copy (FP32 state), accumulator cleared for the
        // next deposit (all but the last step), energy appended
        const int post = (f32 ? 1 : 0) | (last ? 0 : 2);
        const bool timed = c->timing && s % c->timing_every == 0;
        if (rc == PIC_OK && timed) rc = mark();
        if (rc == PIC_OK) rc = poisson_solve_e(c, c->rho_acc, scale, post, -2);
        if (rc == PIC_OK && timed) rc = mark();
        if (rc == PIC_OK) rc = seg_pass(c, kernel, last ? 2 : 1, dt);
        if (rc == PIC_OK && timed) rc = mark();
        if (rc == PIC_OK && !last && sort_every > 0 && (s + 1) % sort_every == 0)
            rc = sort_by_cell(c, nullptr, true, false);
    }
    // a sort whose permutation no pass consumed (an error stopped the loop)
    if (c->pending_perm) {
        const int rs = settle_state(c);
        if (rc == PIC_OK) rc = rs;
    }
    return rc;
}

}  // namespace pic

using namespace pic;

extern "C" int pic_run(pic_ctx *ctx, int kernel, double dt, int nsteps, int sort_every,
                       double *energy) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(valid_shape(kernel), PIC_ERR_ARG, "kernel %d", kernel);
    PIC_CHECK(nsteps >= 0, PIC_ERR_ARG, "nsteps %d", nsteps);
    PIC_CHECK(sort_every >= 0, PIC_ERR_ARG, "sort_every %d", sort_every);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    if (nsteps == 0) return PIC_OK;
    PIC_CHECK(c->state == PIC_STATE_F64 || c->uniform_w, PIC_ERR_STATE, "FP32 state needs uniform w");
    PIC_TRY(ensure_energy(c, nsteps));
    // timing: events around every solve and every particle pass
    std::vector<hipEvent_t> ev;
    int rc = PIC_OK;
    if (es_step_ok(c)) {
        // small grids: one launch per iteration (pic_small.hip)
        rc = es_run(c, kernel, dt, nsteps);
    } else {
        rc = run_passes(c, kernel, dt, nsteps, sort_every, ev);
    }
    if (rc == PIC_OK && energy)
        rc = hipMemcpyAsync(energy, c->energy_hist, (size_t)nsteps * 2 * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream) == hipSuccess
                 ? PIC_OK
                 : fail(PIC_ERR_HIP, "energy history copy failed");
    hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == PIC_OK && se != hipSuccess)
        rc = fail(PIC_ERR_HIP, "pic_run: %s", hipGetErrorString(se));
    if (rc == PIC_OK && c->timing) {
        for (size_t k = 0; k + 2 < ev.size(); k += 3) {  // (solve, pass) per step
            float a = 0, b = 0;
            (void)hipEventElapsedTime(&a, ev[k], ev[k + 1]);
            (void)hipEventElapsedTime(&b, ev[k + 1], ev[k + 2]);
            c->t_solve_ms += a;
            c->t_pass_ms += b;
            c->t_passes += 1;
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int pic_set_timing(pic_ctx *ctx, int enable) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    c->timing = enable != 0;
    c->timing_every = enable > 1 ? enable : 1;
    c->t_pass_ms = c->t_solve_ms = 0.0;
    c->t_passes = 0;
    for (int k = 0; k < 4; ++k) {
        c->t_adj_ms[k] = 0.0;
        c->t_adj_n[k] = 0;
    }
    for (int k = 0; k < 2; ++k) {
        c->t_ua_ms[k] = 0.0;
        c->t_ua_n[k] = 0;
    }
    return PIC_OK;
}

extern "C" int pic_get_timing(pic_ctx *ctx, double *pass_ms, int64_t *passes, double *solve_ms) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    Ctx *c = ctx_of(ctx);
    if (pass_ms) *pass_ms = c->t_pass_ms;
    if (passes) *passes = c->t_passes;
    if (solve_ms) *solve_ms = c->t_solve_ms;
    return PIC_OK;
}

extern "C" int pic_step_pass(pic_ctx *ctx, int kernel, int mode, double dt, double *rho_raw) {
    PIC_CHECK(ctx, PIC_ERR_ARG, "null context");
    PIC_CHECK(valid_shape(kernel), PIC_ERR_ARG, "kernel %d", kernel);
    PIC_CHECK(mode >= 0 && mode <= 3, PIC_ERR_ARG, "pass mode %d", mode);
    Ctx *c = ctx_of(ctx);
    PIC_HIP(hipSetDevice(c->device));
    PIC_CHECK(c->state == PIC_STATE_F64 || c->uniform_w, PIC_ERR_STATE, "FP32 state needs uniform w");
    const int64_t m = (int64_t)c->nx * c->ny;
    PIC_TRY(launch_fill(c, c->rho_acc, 0.0, m));
    // particles in the lane order of a pic_sort_lanes: the segment kernels of
    // pic_run; any other order: the chunk kernels.  With timing on, the pass
    // is timed like pic_run's (HIP events on the context stream).
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (c->timing) {
        PIC_HIP(hipEventCreate(&ev[0]));
        PIC_HIP(hipEventCreate(&ev[1]));
        PIC_HIP(hipEventRecord(ev[0], c->stream));
    }
    const int rc = c->seg_valid ? seg_pass(c, kernel, mode, dt) : step_pass(c, kernel, mode, dt);
    if (c->timing) {
        if (rc == PIC_OK && hipEventRecord(ev[1], c->stream) == hipSuccess &&
            hipEventSynchronize(ev[1]) == hipSuccess) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
            c->t_pass_ms += ms;
            c->t_passes += 1;
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    PIC_TRY(rc);
    if (rho_raw)
        PIC_HIP(hipMemcpyAsync(rho_raw, c->rho_acc, (size_t)m * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}
