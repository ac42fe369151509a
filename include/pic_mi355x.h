/*
 * pic_mi355x.h -- C ABI of the MI355X particle-mesh library (libpicmi355x.so).
 *
 * Drop-in for the 2D particle-mesh step of ParticleInCell.jl: charge
 * deposition, field gather, the CIC/M6 pushers and the periodic FFT Poisson
 * solve.  Every function returns 0 (PIC_OK) or a negative PIC_ERR_* code and
 * leaves a message for pic_last_error().  Plain pointers and sizes only.
 *
 * Layouts are the reference's Julia layouts (column-major, x fastest):
 *   particle group  p.array, 5 x n doubles: x, y, vx, vy, w per particle
 *                   (GEMPIC ParticleGroup{2,2}, used at pushers_cic.jl:14-17)
 *   grid            nx x ny doubles, element (i,j) at i + ld*j (0-based),
 *                   ld = nx for the Julia methods, ld = nx+1 for the ghosted
 *                   arrays of the Fortran twin and of MeshFields2D
 *   MeshFields2D.e  2 x (nx+1) x (ny+1) doubles (component fastest),
 *                   meshfields.jl:3-21
 *   Particles.x/.e  2 x n doubles, particles.jl:12-35
 *
 * Particles and grids live on the GPU inside a context; calls on one context
 * are ordered on the context's HIP stream and are not thread-safe.  Entry
 * points taking host pointers copy through pageable memory and synchronise.
 */
#ifndef PIC_MI355X_H
#define PIC_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pic_ctx pic_ctx;

/* status codes */
#define PIC_OK 0
#define PIC_ERR_ARG -1        /* invalid argument (size, pointer, enum) */
#define PIC_ERR_STATE -2      /* call not valid in the context's current state */
#define PIC_ERR_HIP -3        /* HIP runtime error */
#define PIC_ERR_FFT -4        /* rocFFT error */
#define PIC_ERR_NOMEM -5      /* device allocation failed */

/* particle shapes: the reference's dispatch tags CloudInCell
 * (src/cloud_in_cell.jl:6), M6 (src/pushers_m6.jl:3), NearestGridPoint
 * (src/pushers_ngp.jl:3), TriangularShapeCloud (src/pushers_tsc.jl:14) and
 * M4 (src/pushers_m4.jl:6).  As in the reference, TriangularShapeCloud
 * deposits with the TSC weights f_tsc but its push_v! gathers with the M4
 * weights f_m4 (pushers_tsc.jl:46-53).  pic_adjoint takes CIC and M6. */
#define PIC_CIC 0
#define PIC_M6 1
#define PIC_NGP 2
#define PIC_TSC 3
#define PIC_M4 4

/* particle state held on the device */
#define PIC_STATE_F64 0       /* x, y, vx, vy (+w) FP64 structure of arrays */
#define PIC_STATE_F32 1       /* int32 cell + FP32 in-cell offsets + FP32 velocity */

/* grids of a context */
#define PIC_GRID_RHO 0
#define PIC_GRID_EX 1
#define PIC_GRID_EY 2
#define PIC_GRID_BZ 3
#define PIC_GRID_PHI 4

/* magnetic field seen by push_v */
#define PIC_BZ_NONE 0         /* bz == 0 everywhere */
#define PIC_BZ_UNIFORM 1      /* bz == constant */
#define PIC_BZ_GRID 2         /* bz grid of the context */

/* ------------------------------------------------------------ library */

/* Message of the last failed call on this thread ("" if none). */
const char *pic_last_error(void);

/* Library version, e.g. 100 for 0.1.0. */
int pic_version(void);

/* Number of visible GPUs. */
int pic_device_count(int *count);

/* ------------------------------------------------------------ context */

/* Creates a context for the periodic mesh TwoDGrid(xmin, xmax, nx, ymin,
 * ymax, ny) (GEMPIC's TwoDGrid; dx = (xmax-xmin)/nx) on GPU `device`, with
 * room for `capacity` particles.  Replaces the host allocations the reference
 * makes per call (compute_rho's zeros(nx, ny), pushers_cic.jl:80) and the
 * Poisson tables TwoDPoissonPeriodic(grid) (poisson_2d.jl:28-61) and
 * Poisson2D(mesh) (meshfields.jl:27-73).  nx, ny >= 4 and nx * ny <=
 * 2^24; any sizes for the FP64 calls (the Fortran twin's 100 x 100 too), but
 * the FP32 state (pic_set_state) and the E solve's LDS-FFT form need powers
 * of two, as every reference configuration has. */
int pic_ctx_create(pic_ctx **out, int device, int nx, int ny, double xmin, double xmax,
                   double ymin, double ymax, int64_t capacity);
int pic_ctx_destroy(pic_ctx *ctx);

/* Waits for all work queued on the context's stream. */
int pic_synchronize(pic_ctx *ctx);

/* The context's HIP stream (a hipStream_t), for callers that order their own
 * work against it. */
int pic_stream(pic_ctx *ctx, void **stream);

/* --------------------------------------------------------- particles */

/* Copies a host particle group p.array (5 x n) into the context's FP64 state.
 * Sets the uniform-weight fast path when all w are equal. */
int pic_set_particles(pic_ctx *ctx, const double *p, int64_t n);
/* Copies the context's particles back into p.array (5 x n).  In FP32 state
 * the positions are rebuilt as (cell + offset) * dx. */
int pic_get_particles(pic_ctx *ctx, double *p);
/* The same for a 5 x n particle array already in device memory. */
int pic_set_particles_device(pic_ctx *ctx, const double *d_p, int64_t n);
int pic_get_particles_device(pic_ctx *ctx, double *d_p);
int pic_num_particles(pic_ctx *ctx, int64_t *n);

/* Particles.x / Particles.v (2 x n each, particles.jl:12-35) and the scalar
 * Particles.w: set or read the positions and velocities of the MeshFields
 * API.  v may be NULL (velocities then 0 on set, skipped on get). */
int pic_set_particles_xv(pic_ctx *ctx, const double *x, const double *v, double w, int64_t n);
int pic_get_particles_xv(pic_ctx *ctx, double *x, double *v);
/* Particles.e (2 x n): fields gathered by pic_interpol_eb_m6. */
int pic_get_particle_fields(pic_ctx *ctx, double *e);

/* Switches the device state between FP64 and FP32 cell-relative storage
 * (conversion on the GPU).  The FP32 state needs uniform particle weights
 * and nx, ny powers of two (PIC_ERR_STATE otherwise). */
int pic_set_state(pic_ctx *ctx, int state);
/* Raw FP32 cell-relative state: cell (int32), ox, oy, vx, vy (float), n
 * each, host memory. Any pointer may be NULL to skip it. */
int pic_get_state_f32(pic_ctx *ctx, int32_t *cell, float *ox, float *oy, float *vx, float *vy);
int pic_set_state_f32(pic_ctx *ctx, const int32_t *cell, const float *ox, const float *oy,
                      const float *vx, const float *vy, double w, int64_t n);

/* -------------------------------------------------------------- grids */

/* Copies grid `which` (PIC_GRID_*) from/to host memory, nx x ny with leading
 * dimension ld >= nx (element (i,j) at host[stride*(i + ld*j)], stride >= 1;
 * stride 2 addresses one component of MeshFields2D.e).  On get with
 * ld >= nx+1 and at least ny+1 columns (ghost = 1) the periodic ghost row and
 * column are filled too, as meshfields.jl:100-103 and
 * compute_rho_m6.jl:303-305 do. */
int pic_set_grid(pic_ctx *ctx, int which, const double *host, int ld, int stride);
int pic_get_grid(pic_ctx *ctx, int which, double *host, int ld, int stride, int ghost);
/* Device pointer of a grid (nx*ny doubles, x fastest) for zero-copy use. */
int pic_grid_device_ptr(pic_ctx *ctx, int which, double **d_ptr);
/* Selects the magnetic field push_v sees: PIC_BZ_NONE, PIC_BZ_UNIFORM with
 * value bz0, or PIC_BZ_GRID (the PIC_GRID_BZ grid). */
int pic_set_bz(pic_ctx *ctx, int mode, double bz0);

/* ---------------------------------------------- the reference's calls */

/* compute_rho(p, kernel, mesh) -> rho (PIC_GRID_RHO).
 * CIC: pushers_cic.jl:75-114 (normalised by dx*dy, mean removed).
 * M6 : pushers_m6.jl:226-348 (normalised, rho_total/dimx/dimy removed).
 * FP64 state. */
int pic_compute_rho(pic_ctx *ctx, int kernel);

/* compute_rho(p, kernel, mesh) as pic_compute_rho computes it, on
 * caller-owned device memory and allocating nothing: no counterpart in the
 * reference, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_p is the reference's p.array (5 x n doubles: x, y, vx, vy, w) and d_rho
 * the nx x ny result, both in device memory of the current device, on the
 * mesh TwoDGrid(xmin, xmax, nx, ymin, ymax, ny); any particle weights.  The
 * deposit accumulates in a static device array of the library (one per
 * GPU, held until the process ends) that limits the grid to nx * ny <=
 * 512 * 512 nodes (PIC_ERR_ARG beyond).  The kernels are enqueued on
 * `stream` (a hipStream_t, NULL for the null stream) and the call returns
 * without waiting; calls that overlap on one GPU share the array and race. */
int pic_compute_rho_noalloc(int kernel, const double *d_p, int64_t n, int nx, int ny, double xmin,
                            double xmax, double ymin, double ymax, double *d_rho, void *stream);

/* push_v!(p, kernel, mesh, ex, ey, bz, dt): pushers_cic.jl:3-53 (CIC) and
 * pushers_m6.jl:5-224 (M6), with the grids EX, EY and the field chosen by
 * pic_set_bz.  FP64 state; CIC results are bit-identical to the reference's
 * Fortran twin push_v (pic.f90:20-77). */
int pic_push_v(pic_ctx *ctx, int kernel, double dt);

/* push_x!(p, mesh, dt): pushers_cic.jl:55-72 (single conditional periodic
 * wrap).  FP64 state. */
int pic_push_x(pic_ctx *ctx, double dt);

/* operator_t!(split, dt): poisson_2d.jl:265-277, the drift of the splitting
 * operator, x += dt v with no periodic wrap.  FP64 state. */
int pic_operator_t(pic_ctx *ctx, double dt);

/* interpolate_cic!(epx, epy, mesh, p, ex, ey): interpolate_cic.jl:3-33.
 * Writes n values each into host epx, epy (either may be NULL to keep the
 * result on the device for pic_get_particle_fields). */
int pic_interpolate_cic(pic_ctx *ctx, double *epx, double *epy);

/* solve!(ex, ey, poisson, rho): poisson_2d.jl:97-113, RHO -> EX, EY. */
int pic_solve_e(pic_ctx *ctx);
/* solve!(phi, poisson, rho): poisson_2d.jl:78-84, RHO -> PHI. */
int pic_solve_phi(pic_ctx *ctx);

/* The hot-path driver's energy: out[0] = sum(ex.^2)*dx*dy, out[1] =
 * sum(ey.^2)*dx*dy, FP64 (examples/landau_damping_poisson_cic_m6.jl:65 takes
 * sqrt(out[0])). */
int pic_field_energy(pic_ctx *ctx, double out[2]);

/* compute_rho_m6!(fields, particles): compute_rho_m6.jl:181-316, positions
 * from pic_set_particles_xv, result in RHO (read it with ghost = 1).
 * compat = 1 reproduces the reference exactly: the swapped xmin/xmax
 * (:190-191), so the rewritten positions land in [xmax, xmax+dimx), and its
 * neighbour indices mod1(i+k, nx) on the 0-based cell (:216-227), which put
 * every off-centre weight one node low.  compat = 0 uses the unswapped
 * bounds and the neighbours of compute_rho(p, M6(), mesh). */
int pic_compute_rho_m6_fields(pic_ctx *ctx, int compat);

/* interpol_eb_m6!(particles, fields): interpolation_m6.jl:125-247, EX/EY ->
 * Particles.e (pic_get_particle_fields); positions rewritten into
 * [xmin, xmax). */
int pic_interpol_eb_m6(pic_ctx *ctx);

/* solve!(fields, poisson::Poisson2D): meshfields.jl:87-108, RHO -> EX, EY,
 * returns sum(ex^2 + ey^2)*dx*dy over the (nx+1)(ny+1) ghosted nodes. */
int pic_solve_fields(pic_ctx *ctx, double *energy);

/* ------------------------------------------------ fused device loop */

/* Stable counting sort of the particles by cell (0-based i + nx*j of the
 * current position), reordering the device state.  perm (host, n entries,
 * may be NULL) receives the source index of each sorted slot: the
 * permutation std::stable_sort by cell gives.  No reference counterpart
 * (build addition). */
int pic_sort(pic_ctx *ctx, int32_t *perm);
/* pic_sort, then within every segment of 64 consecutive cells (s = cell/64,
 * l = cell%64) the particles are ordered by (rank k within their cell, l):
 * round k holds the k-th particle of each cell of the segment that has more
 * than k.  A wave of 64 consecutive particles then covers 64 different
 * adjacent cells, the order pic_run keeps between its re-sorts.  perm
 * receives the source index of each slot.  No reference counterpart. */
int pic_sort_lanes(pic_ctx *ctx, int32_t *perm);
/* After pic_sort: exclusive offsets of each cell's run (nx*ny+1 entries). */
int pic_cell_offsets(pic_ctx *ctx, int64_t *offsets);

/* Runs `nsteps` iterations of the hot-path driver's time loop
 * (examples/landau_damping_poisson_cic_m6.jl:70-83):
 *   drift dt/2; rho = compute_rho(p, kernel); solve!(ex, ey, poisson, rho);
 *   energy; push_v!(p, kernel, ex, ey, bz, dt); drift dt/2
 * fused on the GPU into one particle kernel per step plus the FFT solve.
 * Works in either state.  energy (host, may be NULL) receives per step
 * out[2*s] = sum(ex^2)*dx*dy and out[2*s+1] = sum(ey^2)*dx*dy.
 * sort_every > 0 re-sorts the particles by cell every that many steps
 * (0 = never).  Leaves RHO, EX, EY of the last step in the context.
 * On power-of-two grids of at most 64 x 64 nodes (bz none or uniform, FP64
 * state or FP32 with fixed-point charge) the loop runs one launch per
 * iteration instead (pic_set_fused_step): every block solves the field of
 * the previous deposit in LDS, then moves its particles; no sort (sort_every
 * is ignored, the particles keep their order). */
int pic_run(pic_ctx *ctx, int kernel, double dt, int nsteps, int sort_every, double *energy);

/* pic_run on power-of-two grids of at most 64 x 64 nodes: fused = 1 (the
 * default) one launch per iteration (the E solve redundantly in every
 * block's LDS, then gather, push and deposit of the block's particles),
 * fused = 0 the separate solve, particle-pass and sort kernels.  Measured
 * on MI355X, Landau CIC 64 x 64 with 204,800 particles, 100 steps: one
 * launch 21.5 us (FP32 state) and 24.0 us (FP64) per step, 24.0 and 26.5 us
 * per step of wall time; the separate kernels 37.7 and 38.0 us.  The FP32 state's
 * one-launch charge is int32 fixed point with one quantum for every block,
 * summed as int64 (exact and run-to-run identical).  No reference
 * counterpart. */
int pic_set_fused_step(pic_ctx *ctx, int fused);

/* Charge accumulation of pic_run's FP32 passes in their LDS windows:
 * 0 compensated FP32 (hi, lo) pairs;
 * 1 FP64 accumulation of the FP32 contributions;
 * 2 (default) fixed point: every contribution rounded to a multiple of a
 *   power-of-two quantum u chosen per block (2^-30 of the block's largest
 *   possible node sum, error <= u/2 per contribution) and summed exactly as
 *   int32.
 * The grid totals are FP64 in every mode.  No reference counterpart. */
int pic_set_f32_accumulation(pic_ctx *ctx, int mode);

/* Copies of the fixed-point charge tile in the LDS window of pic_run's
 * FP32 passes that drift and then deposit (accumulation mode 2):
 * 1, 2 or 4 (default 4).  Lane l adds into copy l mod tiles, so that neighbouring
 * lanes whose particles drifted into one cell do not add to one LDS word in
 * one instruction; the copies are summed exactly at the flush, so the
 * result does not depend on the count.  More copies leave a smaller drift
 * margin in the window (pic_set_f32_window), which also holds the field
 * twice: for M6 in 160 KiB 25 cells with 2 copies, 20 with 4; in 80 KiB 14
 * and 10.  Measured on MI355X with 160 KiB windows: Landau M6, 1e7
 * particles on 128^2 (~0.5 cell of drift per step), 4 copies and a lane
 * sort every 12 steps, 4.48e10 particle-steps/s (2 copies 4.31e10); 5e8
 * particles on 512^2 (~2 cells per step), 2 copies and a sort every 4 steps,
 * 3.93e10 (4 copies, every 3 steps, 3.82e10).  No reference counterpart. */
int pic_set_f32_tiles(pic_ctx *ctx, int tiles);

/* LDS window per block, in KiB (40 to 160, default 160), of the same passes
 * as pic_set_f32_tiles: 160 runs one 1024-thread block per CU, 80 two.  A
 * larger window keeps a wider drift margin (M6 with two tile copies: 25 cells
 * at 160 KiB, 14 at 80), so fewer particles leave it between lane sorts and
 * take the global-memory path, and the sorts can be further apart.
 * Measured on MI355X (M6, two copies, 5e8 particles on 512^2): 80 KiB with a
 * sort every 3 steps 2.76e10 particle-steps/s, 160 KiB every 4 steps 3.93e10.
 * No reference counterpart. */
int pic_set_f32_window(pic_ctx *ctx, int kib);

/* One particle pass of pic_run with the fields currently in the context
 * (EX, EY as last set or solved, and the pic_set_bz field), for callers that
 * compose their own loop:
 *   mode 0  drift dt/2, deposit                 (the loop's prologue)
 *   mode 1  kick dt, drift dt (two half drifts in FP64 state), deposit
 *   mode 2  kick dt, drift dt/2                 (the loop's last pass)
 *   mode 3  deposit only
 * The deposit is the raw charge sum of w times the shape weights (before the
 * 1/(dx dy) normalisation and the mean removal of compute_rho); rho_raw
 * (host, nx*ny, may be NULL) receives it.  After pic_sort_lanes (or
 * pic_run) the pass runs pic_run's segment kernels, otherwise kernels that
 * accept any particle order. */
int pic_step_pass(pic_ctx *ctx, int kernel, int mode, double dt, double *rho_raw);

/* Kernel timing of pic_run: with timing enabled, pic_run records HIP events
 * on the context's stream around every particle pass (kick, drift,
 * deposit) and around every Poisson solve, and pic_step_pass around its
 * pass.  enable = k > 1 records them only at every k-th step (the
 * one-launch loop of small grids, pic_run_maxwell and the adjoint's passes
 * likewise): each event costs a few microseconds of stream time, which
 * would otherwise be part of a timed loop of short steps (measured on
 * MI355X, Landau CIC 64^2: 22.7 us per step without events, 26.4 with an
 * event pair at every step).  pic_get_timing returns the summed durations
 * in milliseconds and the number of timed particle passes since timing was
 * enabled (any pointer may be NULL). */
int pic_set_timing(pic_ctx *ctx, int enable);

/* The E solve of pic_solve_e and pic_run on power-of-two grids up to
 * 512 x 512: fused = 1 (the default) runs it as three kernels with the FFTs
 * across the lanes of a wave (64 to 512 nodes a side; in LDS below 64),
 * fused = 0 as rocFFT transforms around the spectral kernel.  Both
 * compute the same transforms and spectral values (poisson_2d.jl:97-113);
 * results agree to rounding.  Other grids always use rocFFT.  No reference
 * counterpart. */
int pic_set_fused_solve(pic_ctx *ctx, int fused);
int pic_get_timing(pic_ctx *ctx, double *pass_ms, int64_t *passes, double *solve_ms);

/* Gradient of the field energy after `nsteps` iterations of pic_run's loop,
 * J = sum(ex^2 + ey^2) dx dy of the last step's field (the energy the loop
 * records at t = nsteps*dt), with respect to the initial positions and
 * velocities of every particle: the exact discrete adjoint of the loop,
 * bz = 0, uniform weights.  In the FP64 state every operation is FP64; in
 * the FP32 state the loop and the particle adjoints are FP32 (cell-relative
 * positions, as pic_run) while the charge sums, the Poisson solve and its
 * adjoint stay FP64.  The forward pass keeps the solved field of every step
 * and the particle state every `ckpt` steps; the reverse sweep recomputes
 * each segment from its checkpoint.  The particles are also lane-sorted
 * every `sort_every` steps (0: at checkpoints only), as pic_run's
 * sort_every, which keeps the segment passes of both sweeps (the deposit,
 * and in reverse the gather and deposit adjoints) on their LDS windows.  gx, gv
 * (host, 2 x n, may be NULL) receive dJ/d(x0, y0) and dJ/d(vx0, vy0) in the
 * context's particle order; the particles are left as they were.  The
 * sweep's forward pass is the separate-kernel form of pic_run's loop
 * (pic_set_fused_step(ctx, 0)); in the FP32 state the one-launch form of
 * small grids rounds its fixed-point charge with another quantum, so its
 * energies differ from J at the 1e-7 level.  No reference counterpart (the
 * Julia rrule of the drop-in wraps this call). */
int pic_adjoint(pic_ctx *ctx, int kernel, double dt, int nsteps, int ckpt, int sort_every,
                double *J, double *gx, double *gv);

/* Device memory of pic_adjoint's reverse sweep in the FP32 state.  Inside a
 * checkpoint segment the sweep keeps the positions of `block` consecutive
 * steps at once: it recomputes the segment once to store restart points
 * (state, sort offsets and permutation) every `block` steps, then each block
 * from its restart point.  block = 0 (the default) takes the largest block,
 * up to `ckpt`, whose buffers fit the free device memory (PIC_ERR_NOMEM when
 * even one step does not); the gradients do not depend on it beyond the
 * order of floating-point sums.  pic_get_adjoint_plan returns the block and
 * the scratch bytes the last pic_adjoint used.  No reference counterpart. */
int pic_set_adjoint_block(pic_ctx *ctx, int block);
int pic_get_adjoint_plan(pic_ctx *ctx, int *block, int64_t *bytes);

/* With timing enabled (pic_set_timing, which also clears these), the FP32
 * adjoint records HIP events around its particle passes: ms[k] (host, 4)
 * receives the summed durations and passes[k] (host, 4) the counts of
 * k = 0 forward passes, 1 gather-adjoint passes, 2 deposit-adjoint passes,
 * 3 recomputation passes (each with its deferred-particle kernel).  No
 * reference counterpart. */
int pic_get_adjoint_timing(pic_ctx *ctx, double *ms, int64_t *passes);

/* Number of particle chunks that took the global-memory slow path in the
 * fused kernels since the context was created (diagnostic). */
int pic_slow_chunks(pic_ctx *ctx, uint64_t *count);

/* ------------------------------------------ Vlasov-Maxwell (FDTD) */

/* The FDTD solver of src/fdtd.jl on the Yee grid and the CIC current
 * deposit of src/cloud_in_cell.jl, for the electromagnetic loop of
 * examples/landau_damping_maxwell_fdtd.jl:58-83.  The context holds the
 * FDTD fields (fdtd.jl:3-20: ex nx x (ny+1), ey (nx+1) x ny, bz nx x ny,
 * column-major, zero until set) and a periodic nx x ny current.  The node
 * fields that update_fields! produces live in the context's ex, ey, bz
 * grids (nx x ny; the reference's ghost row and column are their periodic
 * images, which they equal whenever fdtd.ex's row ny+1 equals its row 1 and
 * fdtd.ey's column nx+1 its column 1, as ampere_maxwell! keeps them).
 * Particle work needs the FP64 state (pic_set_state). */

/* Copy the FDTD fields from / to the host (fdtd.ex, fdtd.ey, fdtd.bz; any
 * pointer may be NULL to skip that field). */
int pic_fdtd_set(pic_ctx *ctx, const double *ex, const double *ey, const double *bz);
int pic_fdtd_get(pic_ctx *ctx, double *ex, double *ey, double *bz);

/* faraday!(fdtd, m, dt): fdtd.jl:24-36. */
int pic_faraday(pic_ctx *ctx, double dt);

/* ampere_maxwell!(fdtd, m, jx, jy, dt): fdtd.jl:39-54, with the context's
 * current (pic_compute_current or pic_set_current). */
int pic_ampere_maxwell(pic_ctx *ctx, double dt);

/* update_fields!(ex, ey, bz, m, fdtd): fdtd.jl:58-77, into the context's
 * node grids (read them with pic_get_grid; bz mode becomes PIC_BZ_GRID). */
int pic_update_fields(pic_ctx *ctx);

/* compute_current!(jx, jy, m, ::CloudInCell, p): cloud_in_cell.jl:10-81,
 * into the context's current; when jx, jy are not NULL also out to them as
 * the reference's (nx+1) x (ny+1) arrays with folded ghosts.  kernel must be
 * PIC_CIC. */
int pic_compute_current(pic_ctx *ctx, int kernel, double *jx, double *jy);

/* Load the context's current from (nx+1) x (ny+1) host arrays (their
 * first nx x ny nodes; the rest is the periodic image). */
int pic_set_current(pic_ctx *ctx, const double *jx, const double *jy);

/* compute_energy(fdtd, m) = 0.5 log(sum(fdtd.ex .^ 2) dx dy): fdtd.jl:79-80. */
int pic_fdtd_energy(pic_ctx *ctx, double *energy);

/* nsteps iterations of the loop of examples/landau_damping_maxwell_fdtd.jl
 * :58-83 on the resident particles and FDTD fields: faraday!(dt/2) from the
 * second iteration, update_fields!, push_v!(dt), push_x!(dt/2),
 * compute_current!, push_x!(dt/2), faraday!(dt/2), ampere_maxwell!(dt);
 * energy[s] = compute_energy after iteration s+1 (examples/vm2d.jl:55; may
 * be NULL).  One launch per iteration when 8 (nx(ny+1) + (nx+1)ny + 6 nx ny)
 * bytes fit in 150 KiB of LDS (132 KiB at 128 x 16), else two up to 3072
 * cells, else one kernel per reference function.  kernel must be PIC_CIC;
 * FP64 state. */
int pic_run_maxwell(pic_ctx *ctx, int kernel, double dt, int nsteps, double *energy);

/* How pic_run_maxwell launches an iteration: 0 (the default) the most
 * fused form the grid allows in LDS (one launch per iteration up to 2048
 * cells, two up to 3072), 2 at most two launches, 3 one kernel per
 * reference function.  All compute the same arithmetic; for tests and
 * measurements.  No reference counterpart. */
int pic_set_maxwell_path(pic_ctx *ctx, int path);

/* ------------------------------------------ uniformly accurate (UA) PIC */

/* The UA step of src/ua_steps.jl for the 2D Vlasov-Poisson problem with the
 * strong inhomogeneous magnetic field b(x) = 1 + sin(x1) sin(x2)/2, run as in
 * test/bupdate.jl.  The context's particles are the reference's `Particles`
 * (x, v through pic_set_particles_xv with the common weight w; Particles.e
 * from pic_interpol_eb_m6 or pic_set_particle_fields); the fields are
 * MeshFields2D's (rho from these calls, E from pic_solve_fields).  FP64
 * state.
 *
 * Complex arrays cross the boundary as interleaved (re, im) doubles in the
 * reference's column-major (ntau, 2, n) layout: element (m, c, k), 0-based,
 * at complex index m + ntau (c + 2 k); pl and ql are (ntau, n); et is a real
 * (ntau, 2, n) array. */

/* UA(ntau, eps, nbpart): ua_type.jl:16-41.  ntau is a power of two from 2 to
 * 64 (the tau-FFTs run across ntau lanes of one wave). */
int pic_ua_init(pic_ctx *ctx, int ntau, double eps);

/* preparation!(ua, dt, particles, xt, yt): ua_steps.jl:3-83.  Keeps b and t
 * (Particles.b, .t) in the context for the calls below; any output pointer
 * may be NULL (b, t: n doubles).  The calls below take Particles.b or .t as
 * an n-vector, or NULL for the context's from the last preparation. */
int pic_ua_preparation(pic_ctx *ctx, double dt, double *b, double *t, double *pl, double *ql,
                       double *xt, double *yt);

/* update_particles_e!(particles, et, fields, ua, xt) = interpol_eb_m6!(et,
 * fields, xt, nbpart, ntau): ua_steps.jl:87-93, interpolation_m6.jl:3-123,
 * with the context's E (pic_set_grid or pic_solve_fields). */
int pic_ua_interpol(pic_ctx *ctx, const double *xt, double *et);

/* compute_f!(fx, fy, ua, particles, xt, yt, et): ua_steps.jl:107-148,
 * including its closing fft!(fx, 1), fft!(fy, 1). */
int pic_ua_compute_f(pic_ctx *ctx, const double *b, const double *xt, const double *yt,
                     const double *et, double *fx, double *fy);

/* ua_step!(xt, xtilde, ua, particles, fx) (ua_steps.jl:152-172), or with gx
 * not NULL ua_step!(xt, xtilde, ua, particles, fx, gx) (:174-202). */
int pic_ua_step(pic_ctx *ctx, const double *t, const double *xtilde, const double *fx,
                const double *gx, double *xt);

/* update_particles_x!(particles, fields, ua, xt) = compute_rho_m6!(fields,
 * particles, xt, ua): compute_rho_m6.jl:29-179.  Writes Particles.x and the
 * context's rho (read it with pic_get_grid). */
int pic_ua_compute_rho(pic_ctx *ctx, const double *t, const double *xt);

/* compute_v!(yt, particles, ua): ua_steps.jl:206-224, into Particles.v. */
int pic_ua_compute_v(pic_ctx *ctx, const double *t, const double *yt);

/* nsteps iterations of the loop of test/bupdate.jl:62-116 (prediction,
 * solve!, correction, solve!, compute_v!) on the resident particles, from
 * the state the prologue (:56-60: compute_rho_m6!, solve!, interpol_eb_m6!)
 * left.  energy (may be NULL) receives the two solve! energies of every
 * iteration, 2 x nsteps.  Per particle the loop keeps only the positions at
 * the step's start and the tau-spectra of fx, fy (64 ntau + 32 bytes). */
int pic_ua_run(pic_ctx *ctx, double dt, int nsteps, double *energy);

/* Summed durations (ms, HIP events) and launch counts of pic_ua_run's two
 * particle kernels, [0] prediction and [1] correction, since timing was
 * enabled (pic_set_timing); either pointer may be NULL.  No reference
 * counterpart. */
int pic_get_ua_timing(pic_ctx *ctx, double *ms, int64_t *launches);

/* Particles.e (2 x n) from the host; the counterpart of
 * pic_get_particle_fields. */
int pic_set_particle_fields(pic_ctx *ctx, const double *e);

/* ------------------------------------- initial conditions and files */

/* sample!(pg, mesh, LandauDamping(alpha, kx)): landau_damping.jl:77-97, the
 * quiet start, made on the device: n particles with x = newton(r1) (:19-28),
 * y = ymin + r2 dimy, |v| = sqrt(-2 log((i - 1/2)/n)), theta = 2 pi r3,
 * w = dimx dimy / n, where (r1, r2, r3) is the i-th point (1-based) of the
 * unscrambled 3-dimensional Sobol sequence (Joe-Kuo direction numbers, 30
 * bits; the all-zero point 0 is not used).  The mesh must have dimx =
 * 2 pi / kx (:81).  The particles replace the context's, in `state`
 * (PIC_STATE_F64 or PIC_STATE_F32). */
int pic_sample_landau(pic_ctx *ctx, int64_t n, double alpha, double kx, int state);

/* n particles of density 1 + alpha cos(kx (x - xmin)) cos(ky (y - ymin))
 * and Maxwellian velocities of width vth, w = dimx dimy / n, seeded and
 * reproducible: random numbers from the counter-based Philox4x32-10
 * generator (key = seed), x uniform, y by inverting its conditional
 * distribution (Newton), velocities by the Box-Muller transform.  The
 * mesh's dimy must hold whole periods 2 pi / ky; |alpha| < 1.  The build's
 * start for the strong-field workload (1 + 0.05 cos(x/2) cos(y/2)), in
 * place of the rejection loop of plasma(mesh, n) (src/plasma.jl:17-31),
 * which samples a density of the same kind.  The particles replace the
 * context's, in `state`. */
int pic_sample_maxwellian(pic_ctx *ctx, int64_t n, double alpha, double kx, double ky,
                          double vth, uint64_t seed, int state);

/* read_particles(filename, mesh): read_particles.jl:3-35.  One particle per
 * line, "ix iy dpx dpy vx vy" (cell indices, offsets in the cell,
 * velocity); blank lines are skipped.  In PIC_STATE_F64 the positions are
 * (dpx + ix) dx, (dpy + iy) dy as in the reference; PIC_STATE_F32 stores the
 * cell ix + nx iy and the offsets rounded to FP32 (cells must lie in the
 * mesh and offsets in [0, 1); an offset that rounds up to 1 moves to the
 * next cell with offset 0).  w = dimx dimy / n.  *n_out (may be NULL)
 * receives n. */
int pic_read_particles(pic_ctx *ctx, const char *path, int state, int64_t *n_out);

/* Writes the context's particles in the same format: from the FP32 state
 * the cells and offsets as stored (9 significant digits: read back into the
 * FP32 state they are bit-identical); from the FP64 state ix = floor(x/dx),
 * dpx = x/dx - ix with 17 digits.  No reference counterpart (the reference
 * only reads). */
int pic_write_particles(pic_ctx *ctx, const char *path);

/* ------------------------------------------ Fortran-twin drop-ins */

/* The symbols of src/pic.f90, same arguments (scalars by reference, arrays
 * column-major on the host), for F90.push_v!/push_x!/compute_current!
 * (src/pic.jl:18-114): a ccall to (:push_v, piclib) becomes a ccall to
 * (:pic_f90_push_v, libpicmi355x).  Fields are (nx+1) x (ny+1).  They use a
 * private per-process context; not thread-safe.  Return PIC_OK or an error
 * code (the Fortran originals return nothing). */
int pic_f90_push_v(const int32_t *nbpart, double *p, const double *dt, const int32_t *nx,
                   const int32_t *ny, const double *dx, const double *dy, const double *ex,
                   const double *ey, const double *bz);
int pic_f90_push_x(const int32_t *nbpart, const double *dimx, const double *dimy, double *p,
                   const double *dt);
int pic_f90_deposition(const int32_t *nbpart, const int32_t *nx, const int32_t *ny,
                       const double *dx, const double *dy, const double *p, double *jx,
                       double *jy);

#ifdef __cplusplus
}
#endif

#endif /* PIC_MI355X_H */
