"""tau_int estimator (restating lib/m_variance.f90:716-750 auto_cor) against closed forms."""
import numpy as np

from sunhubbardmy_amd.autocorr import auto_cor, tau_estimate, tau_int, tau_int_batch


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


def test_batch_estimator_removes_a_common_relaxation():
    # 128 AR(1) chains of length 100 (phi = 0.4, tau_int = 7/6) on top of a slow relaxation
    # common to all chains, as a batch started from random fields shows
    rng = np.random.default_rng(1)
    phi, n, steps = 0.4, 128, 100
    x = np.zeros((n, steps))
    e = rng.standard_normal((n, steps))
    x[:, 0] = e[:, 0] / np.sqrt(1 - phi ** 2)
    for t in range(1, steps):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    drift = 3.0 * np.exp(-np.arange(steps) / 60.0)
    exact = (1 + phi) / (2 * (1 - phi))
    tau_raw, _, _ = tau_int(x + drift)
    tau_b, _, _ = tau_int_batch(x + drift)
    tau_clean, _, _ = tau_int_batch(x)
    assert tau_raw > tau_b + 0.5                      # the per-chain estimator reads the drift
    # 100-step series: auto_cor's per-chain means bias rho low by O((1 + 2 tau) / steps)
    assert abs(tau_b - exact) < 0.2 * exact, (tau_b, exact)
    assert abs(tau_b - tau_clean) < 1e-9              # a common drift drops out exactly


def _ar1(rng, phi, n, steps):
    x = np.zeros((n, steps))
    e = rng.standard_normal((n, steps))
    x[:, 0] = e[:, 0] / np.sqrt(1 - phi ** 2)
    for t in range(1, steps):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    return x


def test_pooled_batch_estimator_has_no_short_series_bias():
    # per-chain means bias rho low by ~(1 + 2 tau) / steps; the pooled deviations do not
    rng = np.random.default_rng(4)
    phi = 0.6
    exact = (1 + phi) / (2 * (1 - phi))          # 2.0
    x = _ar1(rng, phi, 256, 60)
    r = tau_estimate(x, batch=True)
    t_chain = tau_estimate(x, batch=False)["tau"]
    assert r["resolved"]
    assert abs(r["tau"] - exact) < 3 * r["sigma"] + 0.05, (r["tau"], r["sigma"])
    assert t_chain < r["tau"] - 0.2                # the per-chain estimate is biased low


def test_sokal_sigma_matches_the_spread_over_replicas():
    rng = np.random.default_rng(5)
    phi = 0.5
    est = [tau_estimate(_ar1(rng, phi, 64, 200), batch=True) for _ in range(40)]
    taus = np.array([e["tau"] for e in est])
    sig = np.mean([e["sigma"] for e in est])
    assert 0.5 * sig < taus.std() < 2.0 * sig, (taus.std(), sig)


def test_unresolved_when_the_window_reaches_max_lag():
    # tau = 19.5 on a 40-step series: the window cannot close within max_lag = 10
    rng = np.random.default_rng(6)
    r = tau_estimate(_ar1(rng, 0.95, 64, 40), batch=True)
    assert not r["resolved"] and r["window"] == r["max_lag"] == 10
    # no clip: an uncorrelated series may estimate slightly below 1/2 and says so
    r0 = tau_estimate(rng.standard_normal((64, 400)), batch=True)
    assert r0["resolved"] and abs(r0["tau"] - 0.5) < 4 * r0["sigma"] + 0.02
