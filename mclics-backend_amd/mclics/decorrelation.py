"""Integrated autocorrelation time and decorrelated-sweep throughput.

tau_int follows Sokal's automatic windowing ("Monte Carlo methods in
statistical mechanics", Cargese lectures 1996, sec. 3): the normalised
autocorrelation rho(t) is summed up to the first window M with
M >= c * tau_int(M), c = 5.  Decorrelated sweeps per second are
sweeps/s / (2 tau_int), the rate of effectively independent samples.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np


def autocorr(x: np.ndarray) -> np.ndarray:
    """Normalised autocorrelation rho(t), t = 0..n-1 (FFT, zero padded)."""
    x = np.asarray(x, np.float64)
    n = len(x)
    d = x - x.mean()
    f = np.fft.rfft(d, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    if acf[0] <= 0:
        return np.zeros(n)
    return acf / acf[0]


def tau_int(x: np.ndarray, c: float = 5.0) -> Tuple[float, int]:
    """(tau_int, window M) by automatic windowing."""
    rho = autocorr(x)
    tau = 0.5
    for M in range(1, len(rho)):
        tau += rho[M]
        if M >= c * tau:
            return max(tau, 0.5), M
    return max(tau, 0.5), len(rho) - 1


def tau_error(tau: float, M: int, n: int) -> float:
    """Madras-Sokal estimate of the standard error of tau_int."""
    return tau * math.sqrt(2.0 * (2 * M + 1) / n)
