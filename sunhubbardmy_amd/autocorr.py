"""Integrated autocorrelation time of per-sweep observable series.

rho(t) is the reference's estimator ``auto_cor`` (lib/m_variance.f90:716-750): for lag t the
means over the first nb - t points of x_i x_{i+t}, x_i^2 and x_i give
rho(t) = (<x x_t> - <x>^2) / (<x^2> - <x>^2).  For a batch of independent chains of equal
length rho(t) is averaged over the chains.  tau_int = 1/2 + sum_{t=1}^{W} rho(t) with Sokal's
automatic window, the smallest W with W >= c tau_int(W) (c = 5).

``tau_int_batch`` first removes the batch mean of every sweep: the chains of a batch are
independent and identically distributed, so the mean over chains at sweep t estimates the
(possibly still relaxing) ensemble mean at t.  A slow relaxation common to every chain, which
a per-chain estimator reads as correlation at every lag, then drops out, and rho(t) measures
the equilibrium fluctuations only (removing the mean scales every lag's covariance by the
same factor 1 - 1/nchains, which cancels in rho).
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


def tau_int(series, c=5.0, max_lag=None):
    """tau_int in units of the series step for an array [nchains, nsteps] (or one series).

    Returns (tau, window, rho)."""
    x = np.atleast_2d(np.asarray(series, dtype=np.float64))
    nsteps = x.shape[1]
    if max_lag is None:
        max_lag = max(1, nsteps // 4)
    rho = np.mean([auto_cor(row, max_lag) for row in x], axis=0)
    tau = 0.5
    for w in range(1, max_lag + 1):
        tau = 0.5 + float(np.sum(rho[:w]))
        if w >= c * tau:
            return max(tau, 0.5), w, rho
    return max(tau, 0.5), max_lag, rho


def tau_int_batch(series, c=5.0, max_lag=None):
    """tau_int of a batch [nchains, nsteps] from the deviations from the batch mean of every
    step (see the module docstring); returns (tau, window, rho)."""
    x = np.atleast_2d(np.asarray(series, dtype=np.float64))
    if x.shape[0] < 2:
        raise ValueError("the batch estimator needs at least two chains")
    return tau_int(x - x.mean(axis=0, keepdims=True), c=c, max_lag=max_lag)
