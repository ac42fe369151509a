"""Diagnostics of the Landau damping runs.

The reference fits the electric-energy history with the external
DispersionRelations.fit_complex_frequency (examples/landau_damping_poisson_cic_m6.jl:91),
which is not available here.  The damping rate is instead the slope of a
least-squares line through ln of the local maxima of
sqrt(sum(ex^2) dx dy) over the linear phase, and the frequency is pi over
the mean spacing of those maxima.  The analytic reference for k = 0.5,
v_th = 1 is omega = 1.415662 - 0.153359 i (root of the plasma dispersion
relation).
"""
import numpy as np

LANDAU_K05 = complex(1.415662, -0.153359)


def local_maxima(t, e, min_sep=0.0):
    """Local maxima of e(t), at least min_sep apart in t (the larger of two
    close maxima is kept)."""
    t = np.asarray(t, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    idx = [i for i in range(1, len(e) - 1) if e[i] > e[i - 1] and e[i] >= e[i + 1]]
    kept = []
    for i in sorted(idx, key=lambda k: -e[k]):
        if all(abs(t[i] - t[j]) >= min_sep for j in kept):
            kept.append(i)
    kept.sort()
    return t[kept], e[kept]


def damping_rate(t, e, tmin=0.0, tmax=None, min_sep=1.5, noise=None, snr=3.0):
    """(gamma, omega, n_maxima): least-squares fit of ln(local maxima) against
    t over [tmin, tmax], maxima at least min_sep apart.  With `noise` (the
    level of a pure-noise series such as sqrt(sum(ey^2) dx dy)), the fit stops
    at the first maximum below snr * noise."""
    tm, em = local_maxima(t, e, min_sep)
    sel = tm >= tmin
    if tmax is not None:
        sel &= tm <= tmax
    tm, em = tm[sel], em[sel]
    if noise is not None:
        keep = em >= snr * noise
        if not keep.all():
            stop = int(np.argmin(keep))
            tm, em = tm[:stop], em[:stop]
    if len(tm) < 2:
        return float("nan"), float("nan"), len(tm)
    gamma = np.polyfit(tm, np.log(em), 1)[0]
    omega = np.pi / np.mean(np.diff(tm))
    return float(gamma), float(omega), len(tm)
