"""Integrated autocorrelation time of per-sweep observable series.

rho(t) is the reference's estimator ``auto_cor`` (lib/m_variance.f90:716-750): for lag t the
means over the first nb - t points of x_i x_{i+t}, x_i^2 and x_i give
rho(t) = (<x x_t> - <x>^2) / (<x^2> - <x>^2).  tau_int = 1/2 + sum_{t=1}^{W} rho(t) with Sokal's
automatic window, the smallest W with W >= c tau_int(W) (c = 5).  Sokal's variance estimate of
tau_int over N samples is 2 (2 W + 1) / N tau_int^2.

An estimate is *resolved* only when the window closed before the largest lag the series allows
(max_lag, a quarter of its length by default).  A window that runs into max_lag means the series
is too short for its own correlations, and the number is then a lower bound, not a measurement;
no value is clipped.

``tau_int`` averages auto_cor's rho over the chains of a batch (per-chain means).  On a series of
nb sweeps those means bias rho(t) low by about (1 + 2 tau) / nb at every lag.

``tau_int_batch`` first removes the batch mean of every sweep: the chains of a batch are
independent and identically distributed, so the mean over chains at sweep t estimates the
(possibly still relaxing) ensemble mean at t, and a slow relaxation common to every chain drops
out.  auto_cor's three means are then taken over the lag pairs of all chains together (the pooled
estimator): the deviations have mean zero by construction, so no per-chain mean is subtracted and
the 1/nb bias of the per-chain estimator disappears (removing the batch mean scales every lag's
covariance by the same factor 1 - 1/nchains, which cancels in rho).
"""
import numpy as np


def auto_cor(data, nt):
    """rho(1..nt) of one series, as auto_cor computes it."""
    data = np.asarray(data, dtype=np.float64)
    nb = data.size
    if nb < nt:
        raise ValueError("series shorter than the number of lags")
    res = np.zeros(nt)
    for ntau in range(1, nt + 1):
        a, b = data[: nb - ntau], data[ntau:]
        x1 = np.mean(a * b)
        x2 = np.mean(a * a)
        x3 = np.mean(a)
        den = x2 - x3 * x3
        res[ntau - 1] = (x1 - x3 * x3) / den if den > 0 else 0.0
    return res


def auto_cor_pooled(x, nt):
    """auto_cor's rho(1..nt) with its three means taken over the lag pairs of every row of the
    array [nchains, nsteps] together."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    nb = x.shape[1]
    if nb <= nt:
        raise ValueError("series shorter than the number of lags")
    res = np.zeros(nt)
    for ntau in range(1, nt + 1):
        a, b = x[:, : nb - ntau], x[:, ntau:]
        x1 = np.mean(a * b)
        x2 = np.mean(a * a)
        x3 = np.mean(a)
        den = x2 - x3 * x3
        res[ntau - 1] = (x1 - x3 * x3) / den if den > 0 else 0.0
    return res


def _window(rho, c, max_lag):
    tau = 0.5
    for w in range(1, max_lag + 1):
        tau = 0.5 + float(np.sum(rho[:w]))
        if w >= c * tau:
            return tau, w, True
    return tau, max_lag, False


def tau_estimate(series, batch=True, c=5.0, max_lag=None):
    """tau_int of an array [nchains, nsteps] (or one series) in units of the series step.

    batch = True: deviations from the batch mean of every step, pooled rho (needs >= 2 chains);
    batch = False: auto_cor per chain, rho averaged over chains.  Returns a dict with tau, its
    Sokal standard error sigma, the window W, max_lag, resolved (the window closed before
    max_lag) and rho."""
    x = np.atleast_2d(np.asarray(series, dtype=np.float64))
    nch, nsteps = x.shape
    if max_lag is None:
        max_lag = max(1, nsteps // 4)
    if batch:
        if nch < 2:
            raise ValueError("the batch estimator needs at least two chains")
        rho = auto_cor_pooled(x - x.mean(axis=0, keepdims=True), max_lag)
    else:
        rho = np.mean([auto_cor(row, max_lag) for row in x], axis=0)
    tau, w, closed = _window(rho, c, max_lag)
    sigma = abs(tau) * np.sqrt(2.0 * (2 * w + 1) / float(nch * nsteps))
    return {"tau": tau, "sigma": float(sigma), "window": int(w), "max_lag": int(max_lag), "resolved": bool(closed),
            "rho": rho}


def tau_int(series, c=5.0, max_lag=None):
    """Per-chain estimator; returns (tau, window, rho)."""
    r = tau_estimate(series, batch=False, c=c, max_lag=max_lag)
    return r["tau"], r["window"], r["rho"]


def tau_int_batch(series, c=5.0, max_lag=None):
    """Batch (pooled, batch-mean-removed) estimator; returns (tau, window, rho)."""
    r = tau_estimate(series, batch=True, c=c, max_lag=max_lag)
    return r["tau"], r["window"], r["rho"]
