"""Convergence and efficiency diagnostics of MCMC output (host, numpy).

* `ess`: multi-chain effective sample size: autocorrelations of the chains
  pooled through the between/within variance estimate (Gelman et al., BDA3
  section 11.5), summed as Geyer's initial monotone sequence (Geyer 1992).
* `rhat`: split-chain potential scale reduction (BDA3 section 11.4).

Arrays are [iterations, chains] for one parameter.
"""
from __future__ import annotations

import numpy as np


def _autocov(x: np.ndarray) -> np.ndarray:
    """Autocovariance at every lag of each column of x [n, m] (biased, FFT)."""
    n = x.shape[0]
    xc = x - x.mean(axis=0)
    size = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(xc, n=size, axis=0)
    ac = np.fft.irfft(f * np.conj(f), n=size, axis=0)[:n]
    return ac / n


def ess(samples: np.ndarray) -> float:
    x = np.asarray(samples, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n, m = x.shape
    if n < 4:
        return float(n * m)
    acov = _autocov(x)                      # [n, m]
    chain_var = acov[0] * n / (n - 1.0)
    W = chain_var.mean()
    B_over_n = x.mean(axis=0).var(ddof=1) if m > 1 else 0.0
    var_plus = W * (n - 1.0) / n + B_over_n
    if not var_plus > 0.0:
        return float(n * m)
    rho = 1.0 - (W - acov.mean(axis=1)) / var_plus
    rho[0] = 1.0
    # Geyer: sums of adjacent pairs, truncated at the first non-positive pair,
    # forced monotone non-increasing
    pairs = rho[0:n - 1:2] + rho[1:n:2]
    k = 0
    while k < pairs.size and pairs[k] > 0.0:
        k += 1
    p = pairs[:k].copy()
    for i in range(1, p.size):
        p[i] = min(p[i], p[i - 1])
    tau = -1.0 + 2.0 * p.sum() if p.size else 1.0
    tau = max(tau, 1.0 / np.log10(n * m + 10.0))
    return float(n * m / tau)


def rhat(samples: np.ndarray) -> float:
    x = np.asarray(samples, dtype=np.float64)
    n = x.shape[0] // 2
    if n < 2:
        return float("nan")
    s = np.concatenate([x[:n], x[n:2 * n]], axis=1)
    W = s.var(axis=0, ddof=1).mean()
    B = n * s.mean(axis=0).var(ddof=1)
    var_plus = (n - 1.0) / n * W + B / n
    return float(np.sqrt(var_plus / W)) if W > 0 else float("nan")


def circular_unwrap(samples: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Samples of a circular parameter re-centred on their circular mean, so
    that linear diagnostics see no jump at the wrap point."""
    w = hi - lo
    ang = (np.asarray(samples) - lo) / w * 2 * np.pi
    mean = np.arctan2(np.sin(ang).mean(), np.cos(ang).mean())
    d = np.angle(np.exp(1j * (ang - mean)))
    return d / (2 * np.pi) * w
