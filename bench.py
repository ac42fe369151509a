"""Benchmark of the particle-mesh step on one MI355X.

Default workload (--workload landau_m6, the headline): 2D2V Landau damping
with M6 (quintic spline) deposition and gather, grid 128 x 128 on
[0, 4pi]^2, 1e7 particles, alpha = 0.01, k = 0.5, dt = 0.05, 400 steps; the
FP32 cell-relative state and the FP64 state run side by side from the same
quiet start, made on the GPU by the library (pic_sample_landau).  A step is
one iteration of the reference's time loop
(examples/landau_damping_poisson_cic_m6.jl:70-83): Poisson solve, kick,
drift, deposit, on device-resident particles.

Warmup: the W warmup steps run on the workload itself and continue, in
blocks of W steps, until the device has been busy --settle-ms (default 50)
milliseconds; the initial state is then made again (on the device, for the
workloads sampled there), sorted, and the K timed steps run from it at
once.  From idle, MI355X's shader
clock takes ~15 ms of load to come up to the clock it holds under load, so
timed steps that follow a host-side reload or a warmup of a millisecond run
their first ~80 steps ~7 % slower (measured: 207 us per step after the
bench's former host reload, 209-215 after 100 ms idle, 192-194 after 40 ms
of load).

The UA workloads ua_bupdate and ua_large time the uniformly accurate loop
of test/bupdate.jl (FP64, 128 x 64 cells, ntau 16, eps 0.1, dt pi/16, 8
steps; 204,800 or 4,096,000 particles).

The Vlasov-Maxwell workloads vm2d and vm2d_npm500 time the FDTD + CIC
current loop of examples/vm2d.jl (FP64, 128 x 16 cells, 204,800 or
1,024,000 particles, dt = 0.01, 1000 steps) the same way.

Other workloads, measured the same way: landau_cic (CIC, 64 x 64, 204,800
particles, alpha 0.05, dt 0.1, 100 steps; on a grid this small pic_run runs
one launch per step, and the line also gives the loop as separate solve,
pass and sort kernels), strong_b (CIC, 256 x 256, 1e8
particles, density 1 + 0.05 cos(x/2) cos(y/2), Maxwellian v, uniform bz =
1/eps = 20, dt = eps/4, 1000 steps) and large_m6 (M6, 512 x 512, 5e8
particles, alpha 0.05, 200 steps); the last two are sampled on the GPU by
the library (pic_sample_maxwellian, pic_sample_landau), like the 1e7
Landau workloads (landau_m6, landau_ngp, landau_tsc, landau_m4).

Prints one JSON line.  value = particle-steps/s of the FP32 path over the
timed steps; roofline = the fused FP32 particle pass (the dominant kernel),
algorithmic bytes (read 20 B + write 20 B per particle) over its average
duration from HIP events on the context stream; cpu_baseline = the oracle's
C restatement of the reference loop on a bounded sample of the same
particles.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Least device time of the warmup phase, milliseconds (see above).
SETTLE_MS = 50.0
# The timed runs record HIP events at every 5th step: coprime to the
# workloads' sort intervals (3, 4, 8, 12, 16, 32), so the passes that apply a
# sort's permutation are sampled in proportion
TIMING_EVERY = 5
BYTES_F32 = 40.0   # per particle per fused pass: cell, ox, oy, vx, vy read and written
BYTES_F64 = 64.0   # x, y, vx, vy read and written


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count(), model or platform.processor()


def pmc_traffic(kernel, profile="pmc_traffic_landau_m6_128x128_1e7.txt"):
    """HBM bytes per launch of `kernel` from the committed byte-counter profile
    of this workload (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each scaled
    by the same counter on a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        for line in open(path):
            f = line.split()
            if len(f) >= 3 and f[0] == kernel:
                return float(f[1]) + float(f[2]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None


# MI355X issue rates behind the pass's non-HBM roofs: 1024 SIMDs issuing one
# wave64 VALU instruction per 4 clocks each (FP32 without packed math), LDS
# 128 bytes per clock per CU on 256 CUs (the ds_read_b128 rate measured on
# the card), at the 2.4 GHz peak engine clock.
CLOCK_HZ = 2.4e9
VALU_WAVE_INSTS_PER_S = 1024 * CLOCK_HZ / 4
LDS_BYTES_PER_S = 256 * 128 * CLOCK_HZ


def issue_counts(kernel, profile="sq_counters_dominant_kernels.txt"):
    """(VALU wave instructions, LDS bytes) per particle of `kernel` and the
    shader clock (Hz) measured during it, from the committed SQ-counter
    profile's 'issue' lines, or None."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        for line in open(path):
            f = line.split()
            if len(f) >= 5 and f[0] == "issue" and f[1] == kernel:
                return float(f[2]), float(f[3]), float(f[4]) * 1e9
    except OSError:
        pass
    return None


def stream_ceiling():
    """GB/s of the in-place state stream in profiles/hbm_stream_ceiling.txt."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_stream_ceiling.txt")) as f:
            for line in f:
                line = line.strip()
                if line.startswith('{"state_stream_particles"'):
                    return float(json.loads(line)["GBs"])
    except (OSError, ValueError, KeyError):
        pass
    return None


def copy_probe(mib, reps=5, particles=100_000_000):
    """Plain device copies of `mib` MiB (torch's elementwise copy kernel), the
    calibration workload for the HBM byte counters: a copy of B bytes reads B
    and writes B.  Then the streaming ceiling of the FP32 passes' access
    pattern: in-place updates of an int32 array and a float4 array of
    `particles` entries (the cell and (ox, oy, vx, vy) state: 20 B read and
    20 B written per particle), no other work."""
    import torch

    def best_of(fn):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dt = best_of(lambda: dst.copy_(src))
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4,
                      "GBs": 2 * n * 4 / dt / 1e9}), flush=True)
    del src, dst
    cell = torch.zeros(particles, dtype=torch.int32, device="cuda")
    ov = torch.zeros((particles, 4), dtype=torch.float32, device="cuda")

    def state_update():
        cell.add_(1)
        ov.add_(1.0)

    dt = best_of(state_update)
    print(json.dumps({"state_stream_particles": particles, "bytes_read": 20 * particles,
                      "bytes_written": 20 * particles, "us": dt * 1e6,
                      "GBs": 40 * particles / dt / 1e9}), flush=True)


class DeviceLandau:
    """The Landau quiet start of sample!(pg, mesh, LandauDamping(alpha, kx))
    made on the device (pic_sample_landau) whenever a run needs it, so that
    the particle array never exists in host memory."""

    def __init__(self, n, alpha, kx):
        self.n, self.alpha, self.kx = n, alpha, kx

    def simulation(self, pic, mesh, kernel, state, bz=None):
        sim = pic.Simulation.landau(mesh, self.n, self.alpha, self.kx, kernel, state=state)
        sim._set_bz(bz)
        return sim

    def reload(self, sim, state):
        sim.sample_landau(self.n, self.alpha, self.kx, state)


class DeviceMaxwellian:
    """Density 1 + alpha cos(kx x) cos(ky y), Maxwellian velocities of width
    1, made on the device from a seed (pic_sample_maxwellian)."""

    def __init__(self, n, alpha, kx, ky, seed):
        self.n, self.alpha, self.kx, self.ky, self.seed = n, alpha, kx, ky, seed

    def simulation(self, pic, mesh, kernel, state, bz=None):
        return pic.Simulation.maxwellian(mesh, self.n, self.alpha, self.kx, self.ky, kernel,
                                         seed=self.seed, state=state, bz=bz)

    def reload(self, sim, state):
        sim.sample_maxwellian(self.n, self.alpha, self.kx, self.ky, self.seed, 1.0, state)


DEVICE_SAMPLES = (DeviceLandau, DeviceMaxwellian)


def _new_sim(pic, mesh, p0, kernel, state, bz=None):
    if isinstance(p0, DEVICE_SAMPLES):
        return p0.simulation(pic, mesh, kernel, state, bz)
    return pic.Simulation(mesh, p0, kernel, state=state, bz=bz)


def _reload(sim, p0, state):
    if isinstance(p0, DEVICE_SAMPLES):
        p0.reload(sim, state)
    else:
        sim.set_particles(p0, state=state)


def run_path(pic, mesh, p0, kernel, state, dt, steps, warmup, sort_every, f32_acc=2, bz=None,
             tiles=4, acc_steps=0, window=160, fused_step=True, settle_ms=SETTLE_MS):
    """Warmup, then the timed run from the initial state, then (acc_steps >
    0) an untimed accuracy run of acc_steps from the same initial state,
    whose energy history the damping rate is fitted on. Returns dict."""
    sim = _new_sim(pic, mesh, p0, kernel, state, bz)
    sim.set_f32_accumulation(f32_acc)
    sim.set_f32_tiles(tiles)
    sim.set_f32_window(window)
    sim.set_fused_step(fused_step)
    sim.sort(lanes=True)
    # warmup steps in blocks of `warmup` until the device has been busy
    # settle_ms, so that the timed steps start at the clock it holds under
    # load (each pic_run call ends with a stream synchronisation)
    warm_steps = 0
    if warmup:
        t_w = time.perf_counter()
        while True:
            sim.run(dt, warmup, sort_every=sort_every)
            warm_steps += warmup
            if (time.perf_counter() - t_w) * 1e3 >= settle_ms:
                break
    # restart from the initial state for the timed (and fitted) run; the
    # initial ordering (sort by cell, lane spread) is part of the set-up
    _reload(sim, p0, state)
    sim.sort(lanes=True)
    # HIP events around every TIMING_EVERY-th step's solve and pass: each
    # event costs the stream a few microseconds, which would otherwise be
    # part of the timed steps
    sim.set_timing(True, every=TIMING_EVERY)
    slow0 = sim.slow_chunks()
    sim.ctx.sync()
    t0 = time.perf_counter()
    e = sim.run(dt, steps, sort_every=sort_every)
    sim.ctx.sync()
    t1 = time.perf_counter()
    pass_ms, passes, solve_ms = sim.timing()
    out = {
        "seconds": t1 - t0,
        "warmup_steps": warm_steps,
        "energy": e,
        "pass_ms_avg": pass_ms / max(passes, 1),
        "solve_ms_avg": solve_ms / max(passes, 1),
        "passes": passes,
        "slow_chunks": sim.slow_chunks() - slow0,
        # the last step's grids, for the FP32-vs-FP64 comparison
        "grids": {g: sim.grid(i) for g, i in (("rho", 0), ("ex", 1), ("ey", 2))},
    }
    # the fused pass split into its halves, each on freshly lane-sorted
    # particles: deposit only, and gather + kick + drift
    split = {}
    for mode, name in ((3, "deposit"), (2, "gather_push"), (1, "fused")):
        ms_sum = 0.0
        for _ in range(3):
            sim.sort(lanes=True)
            sim.set_timing(True)
            sim.step_pass(mode, dt)
            ms, cnt, _ = sim.timing()
            ms_sum += ms / max(cnt, 1)
        split[name] = ms_sum / 3
    sim.set_timing(False)
    out["split_ms"] = split
    if state == 1:
        # the fused pass under each FP32 charge accumulation: 0 compensated
        # FP32 (hi, lo) pair tiles, 1 FP64 tiles, 2 int32 fixed-point tiles
        accs = {}
        for acc in (0, 1, 2):
            sim.set_f32_accumulation(acc)
            ms_sum = 0.0
            for _ in range(3):
                sim.sort(lanes=True)
                sim.set_timing(True)
                sim.step_pass(1, dt)
                ms, cnt, _ = sim.timing()
                ms_sum += ms / max(cnt, 1)
            accs[acc] = ms_sum / 3
        sim.set_timing(False)
        sim.set_f32_accumulation(f32_acc)
        out["acc_ms"] = accs
    if acc_steps:
        if acc_steps == steps:
            out["energy_acc"] = e
        else:
            _reload(sim, p0, state)
            sim.sort(lanes=True)
            out["energy_acc"] = sim.run(dt, acc_steps, sort_every=sort_every)
    sim.close()
    return out


# The configurations the benchmark knows.  landau_m6 is the headline
# (particle-steps/s of the FP32 state beside the FP64 state); the others are
# the remaining hot-path configurations, measured the same way.
WORKLOADS = {
    "landau_m6": dict(name="landau2d2v-m6-128x128-1e7-dt0.05-400steps", grid=128, n=10_000_000,
                      kernel="m6", alpha=0.01, k=0.5, dt=0.05, steps=400, warmup=20,
                      sampler="landau_device", sort=(12, 12)),
    "landau_cic": dict(name="landau2d2v-cic-64x64-204800-dt0.1-100steps", grid=64, n=204_800,
                       kernel="cic", alpha=0.05, k=0.5, dt=0.1, steps=100, warmup=10,
                       sampler="quiet", sort=(8, 8)),
    "strong_b": dict(name="strongB-cic-256x256-1e8-eps0.05-dt0.0125-1000steps", grid=256,
                     n=100_000_000, kernel="cic", alpha=0.05, k=0.5, dt=0.0125, steps=1000,
                     warmup=10, sampler="maxwellian_device", bz=20.0, sort=(32, 32)),
    "large_m6": dict(name="landau2d2v-m6-512x512-5e8-dt0.05-200steps", grid=512,
                     n=500_000_000, kernel="m6", alpha=0.05, k=0.5, dt=0.05, steps=200,
                     warmup=5, sampler="landau_device", sort=(4, 4), tiles=2),
    # the other shapes of the reference's Landau example
    # (examples/landau_damping_poisson_cic_m6.jl:88-103) on the M6 workload
    "landau_ngp": dict(name="landau2d2v-ngp-128x128-1e7-dt0.05-400steps", grid=128,
                       n=10_000_000, kernel="ngp", alpha=0.01, k=0.5, dt=0.05, steps=400,
                       warmup=20, sampler="landau_device", sort=(8, 12)),
    "landau_tsc": dict(name="landau2d2v-tsc-128x128-1e7-dt0.05-400steps", grid=128,
                       n=10_000_000, kernel="tsc", alpha=0.01, k=0.5, dt=0.05, steps=400,
                       warmup=20, sampler="landau_device", sort=(12, 12)),
    "landau_m4": dict(name="landau2d2v-m4-128x128-1e7-dt0.05-400steps", grid=128,
                      n=10_000_000, kernel="m4", alpha=0.01, k=0.5, dt=0.05, steps=400,
                      warmup=20, sampler="landau_device", sort=(12, 12)),
    # the Vlasov-Maxwell loop of examples/vm2d.jl (FDTD fields + CIC current
    # deposit, FP64): 100 particles per cell (the README's 204,800
    # particles) and 500 per cell (vm2d.jl's timed run), 1000 steps
    "vm2d": dict(name="vm2d-fdtd-cic-128x16-204800-dt0.01-1000steps", grid=(128, 16),
                 dimy=1.0, n=204_800, kernel="cic", alpha=0.1, k=0.5, dt=0.01, steps=1000,
                 warmup=20, sampler="quiet", sort=(0, 0), maxwell=True),
    "vm2d_npm500": dict(name="vm2d-fdtd-cic-128x16-1024000-dt0.01-1000steps", grid=(128, 16),
                        dimy=1.0, n=1_024_000, kernel="cic", alpha=0.1, k=0.5, dt=0.01,
                        steps=1000, warmup=20, sampler="quiet", sort=(0, 0), maxwell=True),
    # the uniformly accurate (UA) loop of test/bupdate.jl: 128 x 64 cells on
    # [0, 4pi] x [0, 2pi], ntau 16, eps 0.1, dt = pi/16 (8 steps to pi/2),
    # particles from plasma(mesh, n)'s densities
    "ua_bupdate": dict(name="ua-bupdate-m6-128x64-204800-ntau16-eps0.1-8steps", grid=(128, 64),
                       n=204_800, ntau=16, eps=0.1, dt=math.pi / 16, steps=8, warmup=2,
                       ua=True, sort=(0, 0)),
    "ua_large": dict(name="ua-bupdate-m6-128x64-4096000-ntau16-eps0.1-8steps", grid=(128, 64),
                     n=4_096_000, ntau=16, eps=0.1, dt=math.pi / 16, steps=8, warmup=2,
                     ua=True, sort=(0, 0)),
    # adjoint sweeps: forward loop with checkpoints + reverse sweep, timed as
    # one call; value = particle-steps/s of the whole sweep
    "adjoint_cic": dict(name="adjoint-landau2d2v-cic-128x128-2e6-dt0.1-200steps-ckpt20",
                        grid=128, n=2_000_000, kernel="cic", alpha=0.05, k=0.5, dt=0.1,
                        steps=200, warmup=20, sampler="quiet", sort=(4, 0),
                        adjoint=dict(ckpt=20, fd=True, f64=True)),
    "adjoint_large_m6": dict(name="adjoint-landau2d2v-m6-512x512-5e8-dt0.05-200steps-ckpt25",
                             grid=512, n=500_000_000, kernel="m6", alpha=0.05, k=0.5, dt=0.05,
                             steps=200, warmup=0, sampler="landau_device", sort=(3, 0), tiles=2,
                             adjoint=dict(ckpt=25, fd=False, f64=False)),
}

SHAPES = {"cic": lambda pic: pic.CloudInCell(), "m6": lambda pic: pic.M6(),
          "ngp": lambda pic: pic.NearestGridPoint(),
          "tsc": lambda pic: pic.TriangularShapeCloud(), "m4": lambda pic: pic.M4()}
SHAPE_NAMES = {"cic": "CIC", "m6": "M6", "ngp": "NGP", "tsc": "TSC", "m4": "M4"}

# Algorithmic bytes per particle of the adjoint sweep's passes (FP32 state):
# forward pass and recompute pass read and write cell + 4 floats; the
# adjoint passes read the stored position (cell + 2 floats) and the
# particle adjoint (4 floats) and write the adjoint back.
ADJ_PASS_BYTES = {"forward": 40.0, "gather_adjoint": 44.0, "deposit_adjoint": 44.0,
                  "recompute": 40.0}


def bench_adjoint(args, W, pic, entry, mesh, held, n, steps, warmup):
    """The adjoint workloads: FP32 sweep (and the FP64 one), gradient checks
    against central differences of FP64 runs, per-pass roofline.  `held`
    holds the only reference to the sampled particles: a GPU-sampled set is
    dropped once the context holds its copy."""
    A = W["adjoint"]
    p0 = held.pop()
    ckpt = A["ckpt"]
    dt, alpha, kx = W["dt"], W["alpha"], W["k"]
    nx, ny = mesh.nx, mesh.ny
    kernel = pic.M6() if W["kernel"] == "m6" else pic.CloudInCell()
    kname = "M6" if W["kernel"] == "m6" else "CIC"
    se = args.sort_every_f32
    on_device = isinstance(p0, DeviceLandau)
    dxa = None
    if not on_device:  # dx0/dalpha of the sampler at fixed uniform numbers
        dxa = pic.sampling.dnewton_dalpha(p0[0], alpha, kx)
    res = {}
    for name, state in (("f32", 1), ("f64", 0)):
        if name == "f64" and not A["f64"]:
            continue
        if args.only not in ("both", name):
            continue
        sim = _new_sim(pic, mesh, p0, kernel, state)
        sim.set_f32_tiles(W.get("tiles", 4))
        sim.set_f32_window(W.get("window", 160))
        if dxa is None:  # the start made on the device: x0 read back once
            dxa = pic.sampling.dnewton_dalpha(sim.particles()[0], alpha, kx)
        # warmup sweeps of `warmup` steps until the device has been busy
        # settle_ms, as in run_path (pic_adjoint leaves the particles as
        # they were)
        warm_steps = 0
        if warmup:
            log(f"bench: {name} adjoint warmup, {warmup} steps a sweep, {args.settle_ms} ms")
            t_w = time.perf_counter()
            while True:
                sim.adjoint(dt, warmup, min(ckpt, warmup), se)
                warm_steps += warmup
                if (time.perf_counter() - t_w) * 1e3 >= args.settle_ms:
                    break
        sim.set_timing(True, every=TIMING_EVERY)
        log(f"bench: {name} adjoint, {steps} steps, checkpoints every {ckpt}, sort every {se}")
        sim.ctx.sync()
        t0 = time.perf_counter()
        J, gx, gv = sim.adjoint(dt, steps, ckpt, se)
        t1 = time.perf_counter()
        r = {"seconds": t1 - t0, "J": J, "dJ_dalpha": float(np.sum(gx[0] * dxa)),
             "warmup_steps": warm_steps}
        u = np.random.default_rng(7).choice([-1.0, 1.0], size=n)
        r["dJ_dvx_dir"] = float(np.sum(gv[0] * u))
        del gx, gv, u
        if name == "f32":
            r["plan"] = sim.adjoint_plan()
            r["passes"] = sim.adjoint_timing()
        sim.close()
        res[name] = r
        log(f"bench: {name} adjoint {r['seconds']:.3f} s, J {J:.10e}, dJ/dalpha "
            f"{r['dJ_dalpha']:.10e}")
    fd = {}
    if A["fd"]:
        h = 1e-6

        def fwd(q):
            s_ = pic.Simulation(mesh, q, kernel)
            e = s_.run(dt, steps, sort_every=max(se, 1))
            s_.close()
            return e[-1, 0] + e[-1, 1]

        log("bench: central differences of FP64 runs, h = 1e-6")
        sob = pic.sampling.sobol3(n)
        fd["dJ_dalpha"] = (fwd(pic.sampling.landau_damping(n, mesh, alpha + h, kx, sobol=sob))
                           - fwd(pic.sampling.landau_damping(n, mesh, alpha - h, kx,
                                                             sobol=sob))) / (2 * h)
        u = np.random.default_rng(7).choice([-1.0, 1.0], size=n)
        qp, qm = p0.copy(order="F"), p0.copy(order="F")
        qp[2] += h * u
        qm[2] -= h * u
        fd["dJ_dvx_dir"] = (fwd(qp) - fwd(qm)) / (2 * h)

    head = "f32" if "f32" in res else "f64"
    r = res[head]
    value = n * steps / r["seconds"]
    out = {
        "metric": "particle-steps/s FP32 vs FP64 at <=1% damping-rate error; % HBM roofline",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup,
        "warmup_settle": {"settle_ms": args.settle_ms,
                          "steps_run": {k: v["warmup_steps"] for k, v in res.items()}},
        "ms_per_step": r["seconds"] * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": head,
        "data": "synthetic (Landau quiet start, Sobol)" if W["sampler"] == "quiet" else
                "synthetic (Landau quiet start, Sobol, made on the GPU)",
        "config": {"workload": W["name"], "grid": [nx, ny], "particles": n, "kernel": kname,
                   "alpha": alpha, "k": kx, "dt": dt, "checkpoint_every": ckpt,
                   "sort_every": se, "sweep": "forward with checkpoints + reverse adjoint",
                   "state": "int32 cell + fp32 offset + fp32 velocity" if head == "f32" else
                   "fp64 soa", "parallelism": "single-gpu"},
    }
    if "passes" in r:
        passes = {}
        for k, (ms, cnt) in r["passes"].items():
            if cnt:
                avg = ms / cnt
                gbs = ADJ_PASS_BYTES[k] * n / (avg * 1e-3) / 1e9
                passes[k] = {"avg_us": avg * 1e3, "count": cnt, "bytes_per_particle":
                             ADJ_PASS_BYTES[k], "GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
        dom = max(passes, key=lambda k: passes[k]["avg_us"] * passes[k]["count"])
        d = passes[dom]
        out["roofline"] = {"bound": "hbm", "achieved": d["GBs"], "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": d["frac"], "traffic": None,
                           "kernel": f"{dom} pass <{kname}> (FP32 state)",
                           "bytes_per_particle": d["bytes_per_particle"], "avg_us": d["avg_us"]}
        out["adjoint_passes"] = passes
        out["adjoint_plan"] = {"positions_kept_steps": r["plan"][0],
                               "scratch_bytes": r["plan"][1]}
    grad = {}
    for name, rr in res.items():
        grad[name] = {"J": rr["J"], "dJ_dalpha": rr["dJ_dalpha"], "dJ_dvx_dir": rr["dJ_dvx_dir"]}
        if fd:
            grad[name]["rel_err_vs_fd_dalpha"] = abs(rr["dJ_dalpha"] - fd["dJ_dalpha"]) / abs(
                fd["dJ_dalpha"])
            grad[name]["rel_err_vs_fd_dvx_dir"] = abs(rr["dJ_dvx_dir"] - fd["dJ_dvx_dir"]) / abs(
                fd["dJ_dvx_dir"])
    if fd:
        grad["fd_fp64_h1e-6"] = fd
    if "f32" in res and "f64" in res:
        grad["f32_vs_f64_dalpha_rel"] = abs(res["f32"]["dJ_dalpha"] - res["f64"]["dJ_dalpha"]) / \
            abs(res["f64"]["dJ_dalpha"])
        v64 = n * steps / res["f64"]["seconds"]
        out["fp64"] = {"value": v64, "ms_per_step": res["f64"]["seconds"] * 1e3 / steps,
                       "sort_every": se}
        out["speedup_fp32_vs_fp64"] = value / v64
    out["gradient"] = grad

    if not args.no_cpu and args.cpu_steps > 0:
        O = entry.load_oracle()
        m = min(n, 200_000)
        steps_cpu = max(args.cpu_steps, 4)
        q = (np.asfortranarray(p0[:, :m]) if not on_device else
             np.asfortranarray(pic.sampling.landau_damping(m, mesh, alpha, kx)))
        q[4] = mesh.dimx * mesh.dimy / m
        log(f"bench: cpu baseline, oracle adjoint ({kname}), {m} particles, {steps_cpu} steps")
        c0 = time.perf_counter()
        O.adjoint(q, W["kernel"], nx, ny, mesh.dimx, mesh.dimy, dt, steps_cpu,
                  min(ckpt, steps_cpu))
        c1 = time.perf_counter()
        cores, model = cpu_info()
        out["cpu_baseline"] = {
            "value": m * steps_cpu / (c1 - c0), "unit": "particle-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"first {m} particles of the same workload, {steps_cpu} steps of the "
                      f"FP64 forward loop + adjoint sweep (oracle C restatement, serial), "
                      f"{model}, nproc={cores}"}
    print(json.dumps(out), flush=True)


VM_BYTES = 64.0  # per particle per pass: x, y, vx, vy read and written (uniform weights)


def bench_maxwell(args, W, pic, entry, steps, warmup):
    """The Vlasov-Maxwell workloads: pic_run_maxwell over `steps` iterations
    from the initial state of examples/vm2d.jl (Landau alpha 0.1, kx 0.5,
    fdtd.ex = alpha/kx sin(kx x) at the cell midpoints), timed as one call;
    the particle pass timed with HIP events in a second run; the energy
    history of the first 100 iterations against the oracle's loop."""
    nx, ny = W["grid"]
    n = W["n"] if args.particles is None else args.particles
    alpha, kx, dt = W["alpha"], W["k"], W["dt"]
    mesh = pic.TwoDGrid(2 * math.pi / kx, nx, W["dimy"], ny)
    log(f"bench: {W['name']}: sampling {n} particles (quiet)")
    p0 = np.asfortranarray(pic.sampling.landau_damping(n, mesh, alpha, kx))
    f0 = pic.FDTD(mesh)
    xn = np.linspace(0.0, mesh.dimx, nx + 1)
    f0.ex[...] = (alpha / kx * np.sin(kx * (xn[:-1] + xn[1:]) / 2))[:, None]
    sim = pic.Simulation(mesh, p0, pic.CloudInCell())

    def reset():
        sim.set_particles(p0)
        sim.set_fdtd(f0)
        sim.ctx.sync()

    # warmup in blocks of `warmup` steps until the device has been busy
    # settle_ms (as in run_path)
    warm_steps = 0
    if warmup:
        reset()
        t_w = time.perf_counter()
        while True:
            sim.run_maxwell(dt, warmup)
            warm_steps += warmup
            if (time.perf_counter() - t_w) * 1e3 >= args.settle_ms:
                break
    reset()
    log(f"bench: maxwell loop, {steps} steps")
    t0 = time.perf_counter()
    energy = sim.run_maxwell(dt, steps)
    t1 = time.perf_counter()
    reset()
    sim.set_timing(True)
    nt = min(steps, 200)
    sim.run_maxwell(dt, nt)
    pass_ms, npass, _ = sim.timing()
    sim.set_timing(False)
    pass_ms /= max(npass, 1)
    sim.close()
    seconds = t1 - t0
    value = n * steps / seconds
    achieved = VM_BYTES * n / (pass_ms * 1e-3) / 1e9
    log(f"bench: maxwell {seconds * 1e3:.2f} ms for {steps} steps, pass {pass_ms * 1e3:.2f} us")
    out = {
        "metric": "particle-steps/s of the Vlasov-Maxwell FDTD + CIC loop",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "warmup_settle": {"settle_ms": args.settle_ms, "steps_run": warm_steps},
        "ms_per_step": seconds * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Landau quiet start, Sobol)",
        "config": {"workload": W["name"], "grid": [nx, ny], "particles": n, "kernel": "CIC",
                   "alpha": alpha, "k": kx, "dt": dt, "dimy": W["dimy"],
                   "state": "fp64 soa", "parallelism": "single-gpu"},
        "seconds_total": seconds,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_vm_step<P> (one launch per iteration: field work, update_fields, push_v, push_x, current)",
                     "bytes_per_particle": VM_BYTES, "avg_us": pass_ms * 1e3},
        "energy_first_last": [float(energy[0]), float(energy[-1])],
    }
    if args.workload == "vm2d" and n == W["n"]:
        traffic = pmc_traffic("k_vm_step", "pmc_traffic_vm2d_128x16_204800.txt")
        if traffic is not None:
            out["roofline"]["traffic"] = traffic[0]
            out["roofline"]["traffic_source"] = traffic[1]
    if not args.no_cpu and args.cpu_steps > 0:
        O = entry.load_oracle()
        steps_cpu = min(steps, 100)
        q = p0.copy(order="F")
        fo = O.FDTD(nx, ny)
        fo.ex[...] = f0.ex
        log(f"bench: cpu baseline, oracle Maxwell loop, {n} particles, {steps_cpu} steps, "
            f"{O.num_threads()} threads")
        c0 = time.perf_counter()
        eo = O.maxwell_loop(q, fo, nx, ny, mesh.dimx, mesh.dimy, dt, steps_cpu)
        c1 = time.perf_counter()
        out["energy_rel_diff_vs_oracle"] = float(np.max(np.abs(energy[:steps_cpu] - eo))
                                                 / np.max(np.abs(eo)))
        cores, model = cpu_info()
        out["cpu_baseline"] = {
            "value": n * steps_cpu / (c1 - c0), "unit": "particle-steps/s",
            "cores": O.num_threads(), "kind": "port",
            "sample": f"all {n} particles of the same workload, first {steps_cpu} steps of the "
                      f"loop (oracle C restatement: OpenMP push_v, serial current deposit and "
                      f"FDTD), {model}, nproc={cores}"}
    print(json.dumps(out), flush=True)


def plasma_sample(n, dimx, dimy, seed):
    """Particles with plasma(mesh, n)'s densities (src/plasma.jl:3-58):
    x, y by rejection from 1 + sin(y) + 0.05 cos(x/2), v by rejection from
    the two Maxwellians at (+-2, 0) on [-5, 5]^2; numpy's generator."""
    rng = np.random.default_rng(seed)
    kx, alpha = 0.5, 0.05
    out = np.zeros((2, 0))
    while out.shape[1] < n:
        m = 2 * n
        xi, yi = rng.uniform(0, dimx, m), rng.uniform(0, dimy, m)
        zi = (2.0 + alpha) * rng.uniform(size=m)
        keep = 1.0 + np.sin(yi) + alpha * np.cos(kx * xi) >= zi
        out = np.concatenate([out, np.stack([xi[keep], yi[keep]])], axis=1)
    x = out[:, :n]
    vel = np.zeros((2, 0))
    while vel.shape[1] < n:
        m = 4 * n
        xi, yi = (rng.uniform(size=m) - 0.5) * 10, (rng.uniform(size=m) - 0.5) * 10
        zi = rng.uniform(size=m)
        tem = (np.exp(-((xi - 2) ** 2 + yi ** 2) / 2) + np.exp(-((xi + 2) ** 2 + yi ** 2) / 2)) / 2
        keep = tem >= zi
        vel = np.concatenate([vel, np.stack([xi[keep], yi[keep]])], axis=1)
    return np.asfortranarray(x), np.asfortranarray(vel[:, :n])


def bench_ua(args, W, pic, entry, steps, warmup):
    """The UA workloads: UASimulation (prologue of test/bupdate.jl) then
    pic_ua_run over `steps` iterations, timed as one call; the prediction and
    correction kernels timed with HIP events in a second run."""
    nx, ny = W["grid"]
    n = W["n"] if args.particles is None else args.particles
    ntau, eps, dt = W["ntau"], W["eps"], W["dt"]
    dimx, dimy = 4 * math.pi, 2 * math.pi
    mesh = pic.TwoDGrid(0.0, dimx, nx, 0.0, dimy, ny)
    log(f"bench: {W['name']}: sampling {n} particles (plasma densities)")
    x, v = plasma_sample(n, dimx, dimy, 1234)
    p = pic.Particles(n, dimx * dimy / n)
    p.x[...], p.v[...] = x, v

    def fresh():
        return pic.UASimulation(mesh, p, ntau, eps)

    # the timed simulation is set up first, so that the warmup (on a second
    # one, in blocks of `warmup` steps until the device has been busy
    # settle_ms, as in run_path) runs right before the timed steps
    sim = fresh()
    warm_steps = 0
    if warmup:
        s0 = fresh()
        t_w = time.perf_counter()
        while True:
            s0.run(dt, warmup)
            warm_steps += warmup
            if (time.perf_counter() - t_w) * 1e3 >= args.settle_ms:
                break
        s0.close()
    sim.ctx.sync()
    log(f"bench: UA loop, {steps} steps")
    t0 = time.perf_counter()
    en = sim.run(dt, steps)
    t1 = time.perf_counter()
    sim.close()
    sim = fresh()
    sim.set_timing(True)
    sim.run(dt, steps)
    tm = sim.timing()
    sim.close()
    seconds = t1 - t0
    # algorithmic bytes per particle: x, y, vx, vy and Particles.e read by both
    # kernels; the prediction writes x0, y0, b, t, x, y and the tau-spectra of
    # fx, fy (4 ntau complex), the correction reads those spectra and writes
    # x, y, vx, vy
    kb = {"predict": 48.0 + 48.0 + 64.0 * ntau, "correct": 48.0 + 32.0 + 64.0 * ntau}
    kern = {}
    for k, (ms, cnt) in tm.items():
        if cnt:
            avg = ms / cnt
            gbs = kb[k] * n / (avg * 1e-3) / 1e9
            kern[k] = {"avg_us": avg * 1e3, "launches": cnt, "bytes_per_particle": kb[k],
                       "GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    dom = max(kern, key=lambda k: kern[k]["avg_us"])
    d = kern[dom]
    out = {
        "metric": "particle-steps/s of the UA PIC loop (test/bupdate.jl)",
        "value": n * steps / seconds, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "warmup_settle": {"settle_ms": args.settle_ms, "steps_run": warm_steps},
        "ms_per_step": seconds * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (plasma(mesh, n) densities by rejection, numpy seed 1234)",
        "config": {"workload": W["name"], "grid": [nx, ny], "particles": n, "ntau": ntau,
                   "eps": eps, "dt": dt, "kernel": "M6", "state": "fp64 soa",
                   "parallelism": "single-gpu"},
        "seconds_total": seconds,
        "roofline": {"bound": "hbm", "achieved": d["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": d["frac"], "traffic": None,
                     "kernel": f"k_ua_{dom}<{ntau}> (tau-FFTs across lanes, M6 gather, deposit)",
                     "bytes_per_particle": d["bytes_per_particle"], "avg_us": d["avg_us"]},
        "ua_kernels": kern,
        "energy_last": [float(en[0, -1]), float(en[1, -1])],
    }
    if not args.no_cpu and args.cpu_steps > 0:
        O = entry.load_oracle()
        m = min(n, 20_000)
        steps_cpu = 2
        xs = np.asfortranarray(x[:, :m].copy())
        vs = np.asfortranarray(v[:, :m].copy())
        w = dimx * dimy / m
        rho = O.compute_rho_m6_fields(xs, w, nx, ny, 0.0, dimx, 0.0, dimy, compat=True)
        e_f, _ = O.solve_fields(rho, 0.0, dimx, 0.0, dimy)
        eo = O.interpol_eb_m6(xs, e_f, nx, ny, 0.0, dimx, 0.0, dimy)
        log(f"bench: cpu baseline, oracle UA loop, {m} particles, {steps_cpu} steps")
        c0 = time.perf_counter()
        O.ua_run(ntau, eps, dt, steps_cpu, xs, vs, eo, w, e_f, 0.0, dimx, 0.0, dimy)
        c1 = time.perf_counter()
        cores, model = cpu_info()
        out["cpu_baseline"] = {
            "value": m * steps_cpu / (c1 - c0), "unit": "particle-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"first {m} particles of the same workload (weight dimx dimy / {m}), "
                      f"{steps_cpu} steps of the UA loop (oracle C restatement, serial), "
                      f"{model}, nproc={cores}"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="landau_m6")
    ap.add_argument("--steps", type=int, default=None, help="default: the workload's")
    ap.add_argument("--warmup", type=int, default=None, help="default: the workload's")
    ap.add_argument("--particles", type=int, default=None, help="default: the workload's")
    ap.add_argument("--sort-every-f32", type=int, default=None,
                    help="steps between lane sorts on the FP32 path (default: the workload's)")
    ap.add_argument("--sort-every-f64", type=int, default=None,
                    help="steps between lane sorts on the FP64 path (default: the workload's)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["f32", "f64", "both"], default="both")
    ap.add_argument("--f32-tiles", type=int, choices=[1, 2, 4], default=None,
                    help="copies of the fixed-point charge tile of the FP32 "
                         "drift-and-deposit passes (default: the workload's, else 4)")
    ap.add_argument("--f32-window", type=int, default=None, metavar="KIB",
                    help="LDS window per block of the FP32 drift-and-deposit "
                         "passes, 40 to 160 KiB (default: the workload's, else 160)")
    ap.add_argument("--f32-acc", type=int, choices=[0, 1, 2], default=2,
                    help="FP32 charge tiles: 0 compensated FP32 pairs, 1 FP64 sums, "
                         "2 int32 fixed point")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run plain device copies of MIB MiB (byte-counter calibration)")
    ap.add_argument("--dump", default="", help="write the energy histories to this .npz")
    ap.add_argument("--no-fused-step", action="store_true",
                    help="grids up to 64 x 64: time the separate solve, pass and sort kernels "
                         "instead of the one-launch iteration")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the untimed full-length run the damping rate is fitted on")
    ap.add_argument("--settle-ms", type=float, default=SETTLE_MS, metavar="MS",
                    help="least device time of the warmup phase: the warmup steps repeat in "
                         "blocks of --warmup steps until it has passed (0: --warmup steps only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   f"(got {args.gpus})"}))
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import __graft_entry__ as entry
    pic = entry.load_package()
    from picmi355x import diagnostics

    W = dict(WORKLOADS[args.workload])
    if args.f32_tiles is not None:
        W["tiles"] = args.f32_tiles
    if args.f32_window is not None:
        W["window"] = args.f32_window
    if W.get("ua"):
        bench_ua(args, W, pic, entry, W["steps"] if args.steps is None else args.steps,
                 W["warmup"] if args.warmup is None else args.warmup)
        return
    if W.get("maxwell"):
        bench_maxwell(args, W, pic, entry, W["steps"] if args.steps is None else args.steps,
                      W["warmup"] if args.warmup is None else args.warmup)
        return
    # steps between lane sorts: each path at the interval that suits it (the
    # FP32 windows keep a wider drift margin than the FP64 ones; particles
    # cross cells faster on finer grids and barely drift in a strong field)
    if args.sort_every_f32 is None:
        args.sort_every_f32 = W["sort"][0]
    if args.sort_every_f64 is None:
        args.sort_every_f64 = W["sort"][1]
    steps = W["steps"] if args.steps is None else args.steps
    warmup = W["warmup"] if args.warmup is None else args.warmup
    n = W["n"] if args.particles is None else args.particles
    nx = ny = W["grid"]
    alpha, kx, dt = W["alpha"], W["k"], W["dt"]
    bz = W.get("bz")
    kernel = SHAPES[W["kernel"]](pic)
    mesh = pic.TwoDGrid(0.0, 2 * math.pi / kx, nx, 0.0, 2 * math.pi / kx, ny)
    log(f"bench: {args.workload}: sampling {n} particles ({W['sampler']})")
    if W["sampler"] == "quiet":
        p0 = pic.sampling.landau_damping(n, mesh, alpha, kx)
    elif W["sampler"] == "landau_device":
        p0 = DeviceLandau(n, alpha, kx)
    else:
        p0 = DeviceMaxwellian(n, alpha, kx, kx, seed=1234)

    if "adjoint" in W:
        held = [p0]
        del p0
        bench_adjoint(args, W, pic, entry, mesh, held, n, steps, warmup)
        return

    # the damping rate is fitted on an untimed run of the workload's full
    # length (the reference's loop, examples/landau_damping_poisson_cic_m6.jl
    # :69-91), whatever --steps is: a short timed run sees no maxima
    landau = W["sampler"] in ("quiet", "landau_device")
    acc_steps = W["steps"] if (landau and not args.no_accuracy) else 0
    res = {}
    for name, state in (("f32", 1), ("f64", 0)):
        if args.only not in ("both", name):
            continue
        log(f"bench: {name} path, warmup {warmup}, timed {steps} steps"
            + (f", accuracy run {acc_steps} steps" if acc_steps else ""))
        sort_every = args.sort_every_f32 if name == "f32" else args.sort_every_f64
        res[name] = run_path(pic, mesh, p0, kernel, state, dt, steps, warmup, sort_every,
                             args.f32_acc, bz, W.get("tiles", 4), acc_steps,
                             W.get("window", 160), fused_step=not args.no_fused_step,
                             settle_ms=args.settle_ms)
        r = res[name]
        log(f"bench: {name} {r['seconds']*1e3:.2f} ms total, pass {r['pass_ms_avg']*1e3:.1f} us,"
            f" solve {r['solve_ms_avg']*1e3:.1f} us, deferred {r['slow_chunks']}")

    if args.dump:
        np.savez(args.dump, t=np.arange(1, steps + 1) * dt,
                 **{k: v["energy"] for k, v in res.items()},
                 **{k + "_acc": v["energy_acc"] for k, v in res.items() if "energy_acc" in v})
    gam = {}
    gam_skipped = None
    if acc_steps:
        t_acc = np.arange(1, acc_steps + 1) * dt
        for name, r in res.items():
            ea = r["energy_acc"]
            g, w, nm, sig = diagnostics.landau_damping_rate(t_acc, ea[:, 0], ea[:, 1])
            log(f"bench: {name} gamma {g:.5f} omega {w:.4f} from {nm} maxima "
                f"({acc_steps}-step accuracy run, noise {sig:.3e})")
            if math.isfinite(g):
                gam[name] = (g, w, nm, sig)
            else:
                gam_skipped = (f"fewer than {diagnostics.MIN_MAXIMA} maxima of the mode above "
                               f"2x the particle noise ({nm} found): no damping-rate fit")

    head = "f32" if "f32" in res else "f64"
    r = res[head]
    # pic_run's one-launch iteration (power-of-two grids up to 64 x 64, bz none
    # or uniform, FP64 or fixed-point FP32 charge): one kernel per step, whose
    # launch time is the pass time (the E solve inside it)
    pow2 = lambda v: v & (v - 1) == 0
    one_launch = (not args.no_fused_step and nx <= 64 and ny <= 64 and pow2(nx) and pow2(ny)
                  and (bz is None or np.ndim(bz) == 0) and (head == "f64" or args.f32_acc == 2))
    value = n * steps / r["seconds"]
    bytes_pp = BYTES_F32 if head == "f32" else BYTES_F64
    achieved = bytes_pp * n / (r["pass_ms_avg"] * 1e-3) / 1e9
    kname = SHAPE_NAMES[W["kernel"]]
    tiles = ({0: " (FP32 pair tiles)", 1: " (FP64 tiles)", 2: " (int32 fixed-point tiles)"}
             [args.f32_acc] if head == "f32" else " (FP64 tiles)")
    out = {
        "metric": "particle-steps/s FP32 vs FP64 at <=1% damping-rate error; % HBM roofline",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        # the warmup steps repeat in blocks of `warmup` until the device has
        # been busy settle_ms; these are the untimed steps that ran
        "warmup_settle": {"settle_ms": args.settle_ms,
                          "steps_run": {k: v["warmup_steps"] for k, v in res.items()}},
        "ms_per_step": r["seconds"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "f32" else "f64",
        "data": {"quiet": "synthetic (Landau quiet start, Sobol)",
                 "landau_device": "synthetic (Landau quiet start, Sobol, made on the GPU)"}.get(
                     W["sampler"], "synthetic (density 1 + 0.05 cos(x/2) cos(y/2), Maxwellian v, "
                                   "Philox seed 1234, sampled on the GPU)"),
        "config": {"workload": W["name"], "grid": [nx, ny], "particles": n, "kernel": kname,
                   "alpha": alpha, "k": kx, "dt": dt, "bz": bz,
                   "state": "int32 cell + fp32 offset + fp32 velocity" if head == "f32" else
                   "fp64 soa", "sort_every": args.sort_every_f32 if head == "f32" else
                   args.sort_every_f64, "f32_charge_tiles": W.get("tiles", 4),
                   "f32_window_kib": W.get("window", 160),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": (f"k_es_step<{kname}> one launch per step: E solve in LDS + "
                                "gather+kick+drift+deposit" if one_launch else
                                f"k_seg_{head}<{kname}> gather+kick+drift+deposit pass") + tiles,
                     "bytes_per_particle": bytes_pp, "avg_us": r["pass_ms_avg"] * 1e3},
    }
    if "f32" in res and "f64" in res:
        v64 = n * steps / res["f64"]["seconds"]
        out["fp64"] = {"value": v64, "ms_per_step": res["f64"]["seconds"] * 1e3 / steps,
                       "pass_us": res["f64"]["pass_ms_avg"] * 1e3,
                       "achieved_GBs": BYTES_F64 * n / (res["f64"]["pass_ms_avg"] * 1e-3) / 1e9,
                       "sort_every": args.sort_every_f64,
                       "deferred_particles": res["f64"]["slow_chunks"]}
        out["speedup_fp32_vs_fp64"] = value / v64
        if gam_skipped:
            out["gamma"] = {"fit": None, "reason": gam_skipped}
        if "f32" in gam and "f64" in gam:
            g32, g64 = gam["f32"][0], gam["f64"][0]
            rel = abs(g32 - g64) / abs(g64)
            out["gamma"] = {"fp32": g32, "fp64": g64, "rel_err": rel, "analytic": -0.153359,
                            "fp64_vs_analytic": abs(g64 + 0.153359) / 0.153359,
                            "omega": {"fp32": gam["f32"][1], "fp64": gam["f64"][1]},
                            "maxima": {"fp32": gam["f32"][2], "fp64": gam["f64"][2]},
                            "noise": gam["f64"][3], "accuracy_steps": acc_steps,
                            "within_1pct": bool(rel <= 0.01),
                            "fit": "ln of the local maxima of sqrt(sum ex^2 dx dy), noise "
                                   "(median of sum ey^2 dx dy) removed in quadrature, up to the "
                                   "first below 2x noise, at least 4"}
        e32, e64 = res["f32"]["energy"][:, 0], res["f64"]["energy"][:, 0]
        out["energy_rel_diff_fp32_fp64"] = float(np.max(np.abs(e32 - e64)) / np.max(e64))
        # normwise relative difference of the last step's rho and E
        g32, g64 = res["f32"]["grids"], res["f64"]["grids"]
        out["grid_rel_diff_fp32_fp64"] = {
            "rho": float(np.linalg.norm(g32["rho"] - g64["rho"]) / np.linalg.norm(g64["rho"])),
            "e": float(np.sqrt(np.sum((g32["ex"] - g64["ex"]) ** 2 + (g32["ey"] - g64["ey"]) ** 2)
                               / np.sum(g64["ex"] ** 2 + g64["ey"] ** 2)))}
    out["solve_us"] = r["solve_ms_avg"] * 1e3
    out["deferred_particles"] = r["slow_chunks"]
    if one_launch:
        # the same loop as separate kernels (solve, segment pass, lane sorts)
        sep = {}
        for name, state in (("f32", 1), ("f64", 0)):
            if name not in res:
                continue
            sort_every = args.sort_every_f32 if name == "f32" else args.sort_every_f64
            log(f"bench: {name} path as separate kernels, timed {steps} steps")
            rr = run_path(pic, mesh, p0, kernel, state, dt, steps, warmup, sort_every,
                          args.f32_acc, bz, W.get("tiles", 4), 0, W.get("window", 160),
                          fused_step=False, settle_ms=args.settle_ms)
            sep[name] = {"value": n * steps / rr["seconds"],
                         "ms_per_step": rr["seconds"] * 1e3 / steps,
                         "pass_us": rr["pass_ms_avg"] * 1e3, "solve_us": rr["solve_ms_avg"] * 1e3,
                         "sort_every": sort_every}
        out["separate_kernels"] = sep
    # the halves of the fused pass timed on their own (HIP events), against
    # their own algorithmic bytes: the deposit reads the state (FP32 20 B,
    # FP64 x, y 16 B); gather + kick + drift reads and writes it
    split_bytes = {"f32": {"deposit": 20.0, "gather_push": 40.0, "fused": 40.0},
                   "f64": {"deposit": 16.0, "gather_push": 64.0, "fused": 64.0}}
    if "acc_ms" in res.get("f32", {}):
        names = {0: "fp32_pairs", 1: "fp64_tiles", 2: "int32_fixed_point"}
        out["f32_accumulation_pass_us"] = {
            names[k]: {"avg_us": ms * 1e3,
                       "frac": BYTES_F32 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            for k, ms in res["f32"]["acc_ms"].items()}
    out["passes_split"] = {}
    for name, rr in res.items():
        out["passes_split"][name] = {
            k: {"avg_us": ms * 1e3, "bytes_per_particle": split_bytes[name][k],
                "GBs": split_bytes[name][k] * n / (ms * 1e-3) / 1e9,
                "frac": split_bytes[name][k] * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            for k, ms in rr["split_ms"].items()}
    # the pass against its VALU-issue and LDS roofs (instruction counts of the
    # committed SQ profile over the live pass time)
    ic = None if one_launch else issue_counts(f"k_seg_{head}<{kname}>")
    if ic is not None:
        valu_us = n * ic[0] / VALU_WAVE_INSTS_PER_S * 1e6
        lds_us = n * ic[1] / LDS_BYTES_PER_S * 1e6
        pass_us = r["pass_ms_avg"] * 1e3
        # the same floor at the shader clock measured during the pass
        valu_clk_us = valu_us * CLOCK_HZ / ic[2]
        out["roofline"]["issue"] = {
            "valu": {"wave_insts_per_particle": ic[0], "floor_us": valu_us,
                     "frac": valu_us / pass_us, "clock_ghz": ic[2] / 1e9,
                     "floor_us_at_clock": valu_clk_us, "frac_at_clock": valu_clk_us / pass_us},
            "lds": {"bytes_per_particle": ic[1], "floor_us": lds_us, "frac": lds_us / pass_us},
            "source": "profiles/sq_counters_dominant_kernels.txt"}
    # the HBM-bound passes against the measured streaming ceiling of their
    # access pattern (in-place update of the state arrays)
    ceiling = stream_ceiling()
    if ceiling is not None and not one_launch:
        out["roofline"]["stream_ceiling"] = {
            "GBs": ceiling, "frac": out["roofline"]["achieved"] / ceiling,
            "source": "profiles/hbm_stream_ceiling.txt"}
    traffic_file = {"landau_m6": "pmc_traffic_landau_m6_128x128_1e7.txt",
                    "strong_b": "pmc_traffic_strong_b_256x256_1e8.txt"}.get(args.workload)
    if traffic_file and n == W["n"]:
        traffic = pmc_traffic(f"k_seg_{head}", traffic_file)
        if traffic is not None:
            out["roofline"]["traffic"] = traffic[0]
            out["roofline"]["traffic_source"] = traffic[1]

    if not args.no_cpu and args.cpu_steps > 0:
        O = entry.load_oracle()
        steps_cpu = args.cpu_steps
        m = min(n, 10_000_000)
        if isinstance(p0, DeviceLandau):  # the same start made on the host at m particles
            q = np.asfortranarray(pic.sampling.landau_damping(m, mesh, alpha, kx))
            sample = f"a {m}-particle quiet start of the same configuration"
        elif isinstance(p0, DeviceMaxwellian):
            q = np.asfortranarray(pic.sampling.perturbed_maxwellian(m, mesh, alpha, kx, kx,
                                                                    seed=1234))
            sample = f"a {m}-particle sample of the same density (numpy, seed 1234)"
        else:
            q = np.asfortranarray(p0[:, :m])
            sample = f"first {m} particles of the same workload"
        q[4] = mesh.dimx * mesh.dimy / m
        log(f"bench: cpu baseline, oracle loop ({kname}), {m} particles, {steps_cpu} steps, "
            f"{O.num_threads()} threads")
        c0 = time.perf_counter()
        O.landau_loop(q, W["kernel"], nx, ny, mesh.dimx, mesh.dimy, dt, steps_cpu)
        c1 = time.perf_counter()
        cores, model = cpu_info()
        out["cpu_baseline"] = {
            "value": m * steps_cpu / (c1 - c0), "unit": "particle-steps/s",
            "cores": O.num_threads(), "kind": "port",
            "sample": f"{sample}, first {steps_cpu} steps of "
                      f"the reference loop without magnetic field (oracle C restatement: "
                      f"serial deposit, OpenMP push), {model}, nproc={cores}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
