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


def local_maxima(t, e):
    t = np.asarray(t)
    e = np.asarray(e)
    idx = [i for i in range(1, len(e) - 1) if e[i] > e[i - 1] and e[i] >= e[i + 1]]
    return t[idx], e[idx]


def damping_rate(t, e, tmin=0.0, tmax=None):
    """(gamma, omega, n_maxima): fit of ln(local maxima) against t in
    [tmin, tmax]."""
    tm, em = local_maxima(t, e)
    sel = tm >= tmin
    if tmax is not None:
        sel &= tm <= tmax
    tm, em = tm[sel], em[sel]
    if len(tm) < 2:
        return float("nan"), float("nan"), len(tm)
    gamma = np.polyfit(tm, np.log(em), 1)[0]
    omega = np.pi / np.mean(np.diff(tm))
    return float(gamma), float(omega), len(tm)
