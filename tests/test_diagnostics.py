"""The damping-rate fit of the accuracy leg (diagnostics.landau_damping_rate)
on synthetic histories whose rate is known: a damped oscillation
|A e^(gamma t) cos(omega t)| with particle noise of level sigma added in
quadrature, as the Landau runs' ex carries it, and a pure-noise ey."""
import numpy as np
import pytest


def _history(gamma, omega, sigma, dt=0.05, steps=400, amp=0.178, seed=3):
    rng = np.random.default_rng(seed)
    t = np.arange(1, steps + 1) * dt
    sig = amp * np.exp(gamma * t) * np.abs(np.cos(omega * t))
    # noise of rms sigma in each field component, uncorrelated with the mode
    nx = sigma * (1 + 0.02 * rng.standard_normal(steps))
    ny = sigma * (1 + 0.02 * rng.standard_normal(steps))
    return t, sig ** 2 + nx ** 2, ny ** 2


@pytest.mark.parametrize("gamma", [-0.153359, -0.1, -0.2])
def test_fit_recovers_rate_above_noise(pic, gamma):
    t, ex2, ey2 = _history(gamma, 1.415662, 0.0117)
    g, w, nm, s = pic.diagnostics.landau_damping_rate(t, ex2, ey2)
    assert nm >= pic.diagnostics.MIN_MAXIMA
    assert g == pytest.approx(gamma, rel=0.01)
    assert w == pytest.approx(1.415662, rel=0.02)
    assert s == pytest.approx(0.0117, rel=0.05)


def test_fit_without_noise_correction_is_biased(pic):
    """Without the quadrature correction the maxima near the noise floor
    flatten the slope: the corrected fit is what brings it to the rate."""
    t, ex2, ey2 = _history(-0.153359, 1.415662, 0.0117)
    noise = float(np.median(np.sqrt(ey2)))
    g_raw, _, nm_raw = pic.diagnostics.damping_rate(t, np.sqrt(ex2), tmin=0.5, noise=noise,
                                                     snr=2.0)
    g, _, nm, _ = pic.diagnostics.landau_damping_rate(t, ex2, ey2)
    assert abs(g - (-0.153359)) < abs(g_raw - (-0.153359))


def test_fit_needs_min_maxima(pic):
    """A history that reaches the noise after two maxima gives no rate."""
    t, ex2, ey2 = _history(-0.6, 1.415662, 0.0117)
    g, w, nm, _ = pic.diagnostics.landau_damping_rate(t, ex2, ey2)
    assert nm < pic.diagnostics.MIN_MAXIMA
    assert np.isnan(g) and np.isnan(w)
