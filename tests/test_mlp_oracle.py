"""Pins the learned-proposal network oracle: finite-difference gradients,
normalisation of the proposal density, feature symmetry, bf16 rounding."""
import math

import numpy as np
import pytest

import _oracle as O


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def rand_params(rng, scale=0.3):
    return (rng.normal(size=O.MLP_NPARAM) * scale).astype(np.float32)


def batch(rng, B, T=1.1):
    X, base, n = [], [], []
    for _ in range(B):
        nb = unit(rng, 6).reshape(-1)
        x, b = O.mlp_features(nb, rng.normal(size=3) * 0.1, T)
        X.append(x)
        base.append(b)
        n.append(unit(rng, 1)[0])
    return np.array(X), np.array(base), np.array(n)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    P = rand_params(rng)
    X, base, n = batch(rng, 16)
    loss, g = O.mlp_nll_grad(P, X, base, n, 0)
    assert np.isfinite(loss)
    idx = rng.choice(O.MLP_NPARAM, 120, replace=False)
    # include biases and the output layer explicitly
    idx = np.concatenate([idx, [64 * 22 + 3, 64 * 22 + 64 + 64 * 64 + 7, O.MLP_NPARAM - 1, O.MLP_NPARAM - 4]])
    Pd = P.astype(np.float64)
    for p in idx:
        h = 1e-3 * max(1.0, abs(Pd[p]))
        a = P.copy()
        b = P.copy()
        a[p] = np.float32(Pd[p] + h)
        b[p] = np.float32(Pd[p] - h)
        ha = float(a[p]) - Pd[p]
        hb = Pd[p] - float(b[p])
        la, _ = O.mlp_nll_grad(a, X, base, n, 0)
        lb, _ = O.mlp_nll_grad(b, X, base, n, 0)
        fd = (la - lb) / (ha + hb)
        assert g[p] == pytest.approx(fd, rel=2e-3, abs=2e-6), p


@pytest.mark.parametrize("kappa", [0.01, 0.7, 5.0, 80.0, 2000.0])
def test_proposal_density_is_normalised(kappa):
    # integral over the sphere = 2*pi * int_{-1}^{1} q(c) dc
    mu = [0.0, 0.0, 1.0]
    c = np.linspace(-1, 1, 400001)
    lq = np.array([O.vmf_logq(mu, kappa, [math.sqrt(max(0, 1 - x * x)), 0.0, x]) for x in c[::100]])
    # fine grid through the closed form of the same expression
    cc = c
    z = kappa * cc
    ls = kappa + np.log1p(-np.exp(-2 * kappa)) - math.log(2)
    lc = np.abs(z) + np.log1p(np.exp(-2 * np.abs(z))) - math.log(2)
    lq_fine = math.log(kappa) - math.log(4 * math.pi) - ls + lc
    np.testing.assert_allclose(lq, lq_fine[::100], rtol=1e-12, atol=1e-12)
    from scipy.integrate import simpson
    total = 2 * math.pi * simpson(np.exp(lq_fine), x=cc)
    assert total == pytest.approx(1.0, rel=1e-6)


def test_density_is_axial_and_peaks_on_mu():
    rng = np.random.default_rng(2)
    mu = unit(rng, 1)[0].astype(np.float64)
    for k in (0.5, 10.0):
        assert O.vmf_logq(mu, k, mu) == pytest.approx(O.vmf_logq(mu, k, -mu), abs=1e-12)
        other = unit(rng, 1)[0].astype(np.float64)
        assert O.vmf_logq(mu, k, mu) > O.vmf_logq(mu, k, other)


def test_features_are_sign_invariant():
    rng = np.random.default_rng(3)
    for _ in range(50):
        nb = unit(rng, 6)
        f = rng.normal(size=3).astype(np.float32)
        x0, b0 = O.mlp_features(nb.reshape(-1), f, 1.13)
        flips = rng.choice([-1.0, 1.0], size=(6, 1)).astype(np.float32)
        x1, b1 = O.mlp_features((nb * flips).reshape(-1), f, 1.13)
        np.testing.assert_array_equal(x0, x1)
        np.testing.assert_array_equal(b0, b1)
        assert x0[21] == np.float32(1.0) / np.float32(1.13)
        assert x0[2] >= 0
        for b in range(1, 6):
            assert np.dot(x0[:3], x0[3 * b:3 * b + 3]) >= -1e-7


def test_head_normalises_and_clamps():
    mu, lk = O.mlp_head([0.3, -0.2, 0.1, 50.0], np.array([1.0, 0.0, 0.0], np.float32))
    assert np.linalg.norm(mu) == pytest.approx(1.0)
    assert lk == 8.0
    mu, lk = O.mlp_head([0.0, 0.0, 0.0, -50.0], np.array([0.0, 2.0, 0.0], np.float32))
    assert list(mu) == pytest.approx([0.0, 1.0, 0.0]) and lk == -6.0


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-9, -2.5, 1e-40, 3.0e38], np.float32)
    got = O.bf16(vals)
    want = np.array([1.0, 1.0, 1 + 2**-7, -2.5, 0.0, 2.9980945e38], np.float32)
    np.testing.assert_array_equal(got[:4], want[:4])
    assert abs(got[4]) < 1.2e-38
    assert got[5] == pytest.approx(3.0e38, rel=4e-3)


def test_bf16_forward_close_to_fp64():
    rng = np.random.default_rng(4)
    P = rand_params(rng, 0.2)
    Pb = O.bf16(P)
    X, base, n = batch(rng, 8)
    for x in X:
        a = O.mlp_forward(Pb, x, 0)
        b = O.mlp_forward(Pb, x, 1)
        assert np.allclose(a, b, rtol=5e-2, atol=5e-2)
