"""tau_int estimator against AR(1) processes, whose integrated
autocorrelation time is (1 + phi) / (2 (1 - phi)) in closed form."""
import math

import numpy as np
import pytest

from mclics.decorrelation import autocorr, tau_error, tau_int


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.9, 0.98])
def test_tau_int_of_ar1(phi):
    rng = np.random.default_rng(7)
    n = 400000
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0]
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    want = (1 + phi) / (2 * (1 - phi))
    tau, M = tau_int(x)
    err = tau_error(tau, M, n)
    assert abs(tau - want) < 4 * err + 0.02 * want, (tau, want, err)


def test_autocorr_of_alternating_series():
    x = np.array([1.0, -1.0] * 500)
    rho = autocorr(x)
    assert rho[0] == pytest.approx(1.0)
    assert rho[1] == pytest.approx(-1.0, abs=2e-3)
    assert rho[2] == pytest.approx(1.0, abs=3e-3)
