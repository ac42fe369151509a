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


# Fewest maxima landau_damping_rate fits a rate through: four maxima of |E|
# span three half periods (3 pi / omega, about 6.7 time units for k = 0.5),
# over which the Landau mode decays by about one e-fold, so the slope rests
# on a change of ~1 in ln|E| rather than on two neighbouring points.
MIN_MAXIMA = 4


def landau_damping_rate(t, ex2, ey2, tmin=0.5, snr=2.0, min_maxima=MIN_MAXIMA):
    """(gamma, omega, n_maxima, noise) of a Landau run in x from its energy
    history ex2 = sum(ex^2) dx dy and ey2 = sum(ey^2) dx dy per step.

    The perturbation is along x only, so ey is particle noise alone; its
    level sigma^2 = median(ey2) stands for the noise that ex carries too.
    Each local maximum e of sqrt(ex2) (from tmin on) is corrected for it in
    quadrature, sqrt(e^2 - sigma^2), and the fit keeps the maxima up to the
    first whose corrected value falls below snr * sigma: the linear phase
    before the mode sinks into the noise.  gamma is the least-squares slope
    of ln(corrected maxima) against t, omega pi over their mean spacing.
    NaN when fewer than min_maxima maxima qualify."""
    t = np.asarray(t, dtype=np.float64)
    ex2 = np.asarray(ex2, dtype=np.float64)
    sigma = float(np.sqrt(np.median(np.asarray(ey2, dtype=np.float64))))
    tm, em = local_maxima(t, np.sqrt(ex2), 1.5)
    sel = tm >= tmin
    tm, em = tm[sel], em[sel]
    ec = np.sqrt(np.maximum(em * em - sigma * sigma, 0.0))
    keep = ec >= snr * sigma
    stop = len(keep) if keep.all() else int(np.argmin(keep))
    tm, ec = tm[:stop], ec[:stop]
    if len(tm) < max(min_maxima, 2):
        return float("nan"), float("nan"), len(tm), sigma
    gamma = np.polyfit(tm, np.log(ec), 1)[0]
    omega = np.pi / np.mean(np.diff(tm))
    return float(gamma), float(omega), len(tm), sigma
