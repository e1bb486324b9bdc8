"""tau_int estimator (restating lib/m_variance.f90:716-750 auto_cor) against closed forms."""
import numpy as np

from sunhubbardmy_amd.autocorr import auto_cor, tau_int


def test_auto_cor_matches_reference_formula_on_small_series():
    x = np.array([1.0, 3.0, 2.0, 5.0, 4.0, 4.5, 1.5])
    r = auto_cor(x, 3)
    for t in (1, 2, 3):
        a, b = x[:-t], x[t:]
        ref = (np.mean(a * b) - np.mean(a) ** 2) / (np.mean(a * a) - np.mean(a) ** 2)
        assert abs(r[t - 1] - ref) < 1e-15


def test_tau_int_ar1_closed_form():
    # AR(1) x_t = phi x_{t-1} + noise: rho(t) = phi^t, tau_int = (1 + phi) / (2 (1 - phi))
    rng = np.random.default_rng(0)
    for phi in (0.0, 0.5, 0.8):
        chains = np.zeros((64, 4000))
        for c in range(64):
            e = rng.standard_normal(4000)
            x = np.zeros(4000)
            for t in range(1, 4000):
                x[t] = phi * x[t - 1] + e[t]
            chains[c] = x
        tau, w, rho = tau_int(chains, max_lag=200)
        exact = (1 + phi) / (2 * (1 - phi))
        assert abs(tau - exact) < 0.08 * exact + 0.03, (phi, tau, exact)
