"""The hot path's five workloads on the GPU, each at its full size, one test
per workload that stops at its first failed check:

1. Landau damping, CIC, 64 x 64, 204,800 particles, alpha 0.05, dt 0.1,
   100 steps, FP64 (examples/landau_damping_poisson_cic_m6.jl:35-87):
   against the oracle's restatement of the reference loop.
2. Landau damping, M6, 128 x 128, 1e7 particles, alpha 0.01, dt 0.05, 400
   steps, FP64 and FP32 side by side from one quiet start: the damping
   rate of the FP32 state within 1% of the FP64 state's, both near the
   Landau root -0.153359 (k = 0.5, v_th = 1).
3. Strong uniform Bz = 1/eps = 20, CIC, 256 x 256, 1e8 particles, density
   1 + 0.05 cos(x/2) cos(y/2), dt = eps/4, 1000 steps, FP32 and FP64:
   total charge every 100 steps, cells and offsets in range, FP32 tracking
   FP64 in energy and rho.
4. Adjoint, CIC, 128 x 128, 2e6 particles, dt 0.1, 200 steps, checkpoints
   every 20: FP32 against FP64 gradients, J against the forward loop, and
   central differences where the problem's conditioning allows 1e-4.
5. M6, 512 x 512, 5e8 particles in the FP32 state, alpha 0.05, dt 0.05,
   200 steps, then the adjoint sweep with checkpoints every 25: total
   charge, J against the forward loop, gradients finite.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA_LANDAU = -0.153359  # Landau root for k = 0.5, v_th = 1


def _mesh(pic, nx, kx=0.5):
    L = 2 * math.pi / kx
    return pic.TwoDGrid(0.0, L, nx, 0.0, L, nx)


def _raw_charge(pic, sim):
    """The raw deposit (sum of w times the shape weights) of the resident
    particles, from one deposit-only pass."""
    nx, ny = sim.mesh.nx, sim.mesh.ny
    raw = np.zeros((nx, ny), order="F")
    pic._lib.call("pic_step_pass", sim.ctx.h, sim.kernel, 3, 0.0, raw.ctypes.data)
    return raw


def test_config1_landau_cic_full(gpu, O):
    """Config 1 at its full size against the oracle loop, FP64: the energy
    of all 100 steps within 1e-9 relative (atomics reorder the deposit sums
    and the FFTs round differently, and 100 steps let those roundings
    grow); the FP32 state's energies within 1e-4 of the FP64 ones.  Both
    forms of the loop on this grid: one launch per iteration (the default
    up to 64 x 64) and the separate solve, pass and sort kernels."""
    pic = gpu
    n, nx, dt, steps = 204_800, 64, 0.1, 100
    mesh = _mesh(pic, nx)
    p = pic.sampling.landau_damping(n, mesh, 0.05, 0.5)
    eo = O.landau_loop(p.copy(order="F"), "cic", nx, nx, mesh.dimx, mesh.dimy, dt, steps)
    for fused in (True, False):
        print(f"config 1: CIC {nx}x{nx}, {n} particles, {steps} steps, "
              f"{'one launch per step' if fused else 'separate kernels'}", flush=True)
        sim = pic.Simulation(mesh, p, pic.CloudInCell())
        sim.set_fused_step(fused)
        e64 = sim.run(dt, steps, sort_every=4)
        sim.close()
        rel = np.max(np.abs(np.sqrt(e64[:, 0]) - eo) / eo)
        print(f"config 1: FP64 energy vs oracle {rel:.3e}", flush=True)
        assert rel < 1e-9
        sim = pic.Simulation(mesh, p, pic.CloudInCell(), state=pic._lib.PIC_STATE_F32)
        sim.set_fused_step(fused)
        e32 = sim.run(dt, steps, sort_every=8)
        sim.close()
        rel32 = np.max(np.abs(np.sqrt(e32[:, 0]) - eo) / eo)
        print(f"config 1: FP32 energy vs oracle {rel32:.3e}", flush=True)
        assert rel32 < 1e-4


def test_config2_landau_m6_damping_rate(gpu):
    """Config 2 at its full size: both states from the same quiet start
    (pic_sample_landau), 400 steps (t = 20).  The damping rate is fitted
    with diagnostics.landau_damping_rate on at least 4 maxima.  The bar is
    the metric's: |gamma32 - gamma64| <= 1% of |gamma64|.  Both rates lie
    within 3% of the Landau root: the fit sees the mode only until it sinks
    into the particle noise (~6% of the initial field), and its slope
    through those maxima sits 0.8% from the root in FP64."""
    pic = gpu
    n, nx, alpha, dt, steps = 10_000_000, 128, 0.01, 0.05, 400
    mesh = _mesh(pic, nx)
    t = np.arange(1, steps + 1) * dt
    fit = {}
    for name, state, sort_every in (("f64", pic._lib.PIC_STATE_F64, 4),
                                    ("f32", pic._lib.PIC_STATE_F32, 8)):
        print(f"config 2: M6 {nx}x{nx}, {n} particles, {steps} steps, {name}", flush=True)
        sim = pic.Simulation.landau(mesh, n, alpha, 0.5, pic.M6(), state=state)
        sim.sort(lanes=True)
        e = sim.run(dt, steps, sort_every=sort_every)
        sim.close()
        fit[name] = pic.diagnostics.landau_damping_rate(t, e[:, 0], e[:, 1])
        print(f"config 2: {name} gamma {fit[name][0]:.6f} omega {fit[name][1]:.4f} "
              f"from {fit[name][2]} maxima, noise {fit[name][3]:.3e}", flush=True)
    g32, g64 = fit["f32"][0], fit["f64"][0]
    assert fit["f64"][2] >= pic.diagnostics.MIN_MAXIMA
    assert fit["f32"][2] >= pic.diagnostics.MIN_MAXIMA
    assert abs(g32 - g64) <= 0.01 * abs(g64)
    assert abs(g64 - GAMMA_LANDAU) <= 0.03 * abs(GAMMA_LANDAU)
    assert abs(g32 - GAMMA_LANDAU) <= 0.03 * abs(GAMMA_LANDAU)
    assert abs(fit["f32"][1] - 1.415662) < 0.05


def test_config3_strong_b(gpu):
    """Config 3 at its full size, both states from one sample, 1000 steps in
    10 calls of 100: after each call the raw charge sums to N w (FP32: the
    int32 fixed-point tiles round each contribution to 2^-30 of a block's
    largest node sum; FP64 to rounding); at the end every FP32 cell lies in
    the grid and every offset in [0, 1), and FP32 tracks FP64 in the field
    energy and in rho."""
    pic = gpu
    n, nx, eps = 100_000_000, 256, 0.05
    dt, bz = eps / 4, 1 / eps
    mesh = _mesh(pic, nx)
    w = mesh.dimx * mesh.dimy / n
    res = {}
    for name, state, sort_every in (("f32", pic._lib.PIC_STATE_F32, 32),
                                    ("f64", pic._lib.PIC_STATE_F64, 16)):
        print(f"config 3: CIC {nx}x{nx}, {n} particles, bz {bz}, 1000 steps, {name}",
              flush=True)
        sim = pic.Simulation.maxwellian(mesh, n, 0.05, 0.5, 0.5, pic.CloudInCell(), seed=1234,
                                        state=state, bz=bz)
        sim.sort(lanes=True)
        energies = []
        for chunk in range(10):
            energies.append(sim.run(dt, 100, sort_every=sort_every))
            q = _raw_charge(pic, sim).sum()
            assert q == pytest.approx(n * w, rel=1e-6 if name == "f32" else 1e-12), chunk
        res[name] = {"energy": np.concatenate(energies),
                     "rho": sim.grid(pic._lib.PIC_GRID_RHO)}
        if name == "f32":
            cell, ox, oy, vx, vy = sim.state_f32()
            assert cell.min() >= 0 and cell.max() < nx * nx
            assert ox.min() >= 0 and ox.max() < 1 and oy.min() >= 0 and oy.max() < 1
            assert np.all(np.isfinite(vx)) and np.all(np.isfinite(vy))
            del cell, ox, oy, vx, vy
        sim.close()
    e32, e64 = res["f32"]["energy"], res["f64"]["energy"]
    de = np.max(np.abs(e32 - e64)) / np.max(e64)
    r32, r64 = res["f32"]["rho"], res["f64"]["rho"]
    drho = np.linalg.norm(r32 - r64) / np.linalg.norm(r64)
    print(f"config 3: FP32 vs FP64 energy {de:.3e}, rho {drho:.3e}", flush=True)
    assert de < 1e-2
    assert drho < 1e-2


def test_config4_adjoint_cic(gpu):
    """Config 4 at its full size.  The FP32 and FP64 sweeps agree on
    dJ/dalpha to 1e-2: at T = 20 with CIC's kinked weights a relative
    change of 2^-24 in the initial velocities already moves the FP64
    gradient by ~1.6e-3, so the two states' trajectories carry gradients
    that differ at that level.  J of each sweep equals its own forward
    loop's.  Where that conditioning allows 1e-4 the sweeps meet central
    differences (h = 1e-6) of FP64 forward runs to 1e-4: CIC over 50 steps
    (T = 5), and M6 over the full 200 steps (T = 20)."""
    pic = gpu
    n, nx, alpha, dt, ck = 2_000_000, 128, 0.05, 0.1, 20
    mesh = _mesh(pic, nx)
    kx = 0.5

    def sweep(kernel, state, steps):
        sim = pic.Simulation.landau(mesh, n, alpha, kx, kernel, state=state)
        x0 = sim.particles()[0]
        J, gx, gv = sim.adjoint(dt, steps, ck, 4 if state == pic._lib.PIC_STATE_F32 else 0)
        sim.close()
        dja = float(np.sum(gx[0] * pic.sampling.dnewton_dalpha(x0, alpha, kx)))
        return J, dja

    def forward(kernel, state, steps, a):
        sim = pic.Simulation.landau(mesh, n, a, kx, kernel, state=state)
        e = sim.run(dt, steps, sort_every=4 if state == pic._lib.PIC_STATE_F32 else 0)
        sim.close()
        return e[-1, 0] + e[-1, 1]

    F64, F32 = pic._lib.PIC_STATE_F64, pic._lib.PIC_STATE_F32
    cic, m6 = pic.CloudInCell(), pic.M6()
    print(f"config 4: CIC {nx}x{nx}, {n} particles, 200 steps, checkpoints every {ck}",
          flush=True)
    J64, g64 = sweep(cic, F64, 200)
    J32, g32 = sweep(cic, F32, 200)
    print(f"config 4: dJ/dalpha FP64 {g64:.10e} FP32 {g32:.10e}", flush=True)
    assert J64 == pytest.approx(forward(cic, F64, 200, alpha), rel=1e-12)
    assert J32 == pytest.approx(forward(cic, F32, 200, alpha), rel=1e-5)
    assert abs(g32 - g64) <= 1e-2 * abs(g64)
    h = 1e-6
    for kernel, steps, name in ((cic, 50, "CIC T=5"), (m6, 200, "M6 T=20")):
        fd = (forward(kernel, F64, steps, alpha + h) - forward(kernel, F64, steps, alpha - h)) \
            / (2 * h)
        _, a64 = sweep(kernel, F64, steps)
        _, a32 = sweep(kernel, F32, steps)
        print(f"config 4: {name} FD {fd:.10e} FP64 {a64:.10e} FP32 {a32:.10e}", flush=True)
        assert a64 == pytest.approx(fd, rel=1e-4)
        assert a32 == pytest.approx(fd, rel=1e-4)


@pytest.mark.timeout(900)
def test_config5_large_m6_forward_and_adjoint(gpu):
    """Config 5 at its full size in the FP32 state (10 GB of particles):
    200 forward steps, then the adjoint sweep with checkpoints every 25 and
    a lane sort every 3 steps.  The raw charge after the forward run sums to
    N w; the sweep's J equals the forward loop's; the gradient is finite and
    dJ/dalpha has the sign and order of the FP64 central difference of
    the same start over the same 200 steps (within 1e-2: at 5e8 particles
    the FP32 trajectory's rounding sets the level)."""
    pic = gpu
    n, nx, alpha, dt, steps, ck, se = 500_000_000, 512, 0.05, 0.05, 200, 25, 3
    mesh = _mesh(pic, nx)
    kx = 0.5
    F32 = pic._lib.PIC_STATE_F32
    print(f"config 5: M6 {nx}x{nx}, {n} particles, FP32 forward {steps} steps", flush=True)
    sim = pic.Simulation.landau(mesh, n, alpha, kx, pic.M6(), state=F32)
    sim.set_f32_tiles(2)
    sim.sort(lanes=True)
    e = sim.run(dt, steps, sort_every=se)
    J_fwd = e[-1, 0] + e[-1, 1]
    q = _raw_charge(pic, sim).sum()
    w = mesh.dimx * mesh.dimy / n
    print(f"config 5: charge {q:.12e} vs {n * w:.12e}", flush=True)
    assert q == pytest.approx(n * w, rel=1e-6)
    sim.close()
    print("config 5: adjoint sweep", flush=True)
    sim = pic.Simulation.landau(mesh, n, alpha, kx, pic.M6(), state=F32)
    sim.set_f32_tiles(2)
    x0 = sim.particles()[0]
    dxa = pic.sampling.dnewton_dalpha(x0, alpha, kx)
    del x0
    J, gx, gv = sim.adjoint(dt, steps, ck, se)
    sim.close()
    print(f"config 5: J sweep {J:.12e} forward {J_fwd:.12e}", flush=True)
    assert J == pytest.approx(J_fwd, rel=1e-5)
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gv))
    dja = float(np.sum(gx[0] * dxa))
    del gx, gv, dxa
    h = 1e-6

    def forward64(a):
        s = pic.Simulation.landau(mesh, n, a, kx, pic.M6())
        ee = s.run(dt, steps, sort_every=1)
        s.close()
        return ee[-1, 0] + ee[-1, 1]

    fd = (forward64(alpha + h) - forward64(alpha - h)) / (2 * h)
    print(f"config 5: dJ/dalpha FP32 sweep {dja:.10e}, FP64 FD {fd:.10e}", flush=True)
    assert dja == pytest.approx(fd, rel=1e-2)
