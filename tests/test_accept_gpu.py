"""GPU accept decisions at the inputs where glibc's expf is not correctly
rounded.

The reference accepts a move when u < exp(-dE/T) with exp of a float, i.e.
glibc's expf (src/evolve.cpp:160).  For each trial the test picks a site of
an injected pass, tunes T one ulp at a time until x = -(E_new - E_old)/T is
an input where glibc's expf differs from exp rounded once to FP32, and
gives that site the smaller of the two values as its uniform: glibc and the
correctly rounded exp then decide the move differently.  The engine (whose
hardware exp2 cannot separate the two, so the exact path decides) must agree
with the oracle, which calls glibc's expf, on that site and on every other
site of the pass."""
import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


def in_colour(nx, ny, nz, c):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i + j + k) % 2 == c).reshape(-1)


def tune_T(dE, T0, want_glibc_larger, span=400_000):
    """First T >= T0 (stepping one FP32 ulp) at which x = -dE/T falls on a
    glibc/correct-rounding mismatch with the wanted ordering and exp(x) in
    (0.02, 0.98); returns (T, x, glibc value, correctly rounded value)."""
    Ts = np.float32(T0) + np.arange(span, dtype=np.float32) * np.spacing(np.float32(T0))
    Ts = np.unique(Ts.astype(np.float32))
    xs = (-np.float32(dE)) / Ts  # FP32 division, as on both sides
    g = O.expf(xs)
    c = np.exp(xs.astype(np.float64)).astype(np.float32)
    ok = (g != c) & (g > 0.02) & (g < 0.98) & ((g > c) == want_glibc_larger)
    idx = np.flatnonzero(ok)
    if idx.size == 0:
        return None
    t = idx[0]
    return float(Ts[t]), xs[t], g[t], c[t]


@pytest.mark.parametrize("pot", [O.LL, O.GHRL])
def test_accept_matches_glibc_expf_at_its_rounding_errors(pot):
    rng = np.random.default_rng(11 + pot)
    nx = ny = nz = 4
    nt = nx * ny * nz
    kw = {"k": (1.0, 0.7, 1.3, 16.0)} if pot == O.GHRL else {"A": 1.0}
    po = O.params(nx, ny, nz, potential=pot, **kw)
    cfg = mclics.Config(nx, ny, nz, potential=pot, A=po.A, ghrl=O.ghrl_tuple(po), T=1.0, seed=3)
    pt = np.ones(nt, np.int32)
    n0 = unit(rng, nt)
    sel = in_colour(nx, ny, nz, 0)
    sites = np.flatnonzero(sel)
    checked = 0
    with mclics.Lattice(cfg, n0, pt) as lat:
        for trial in range(16):
            prop = unit(rng, nt)
            # energies of the proposals (independent of T and u)
            _, _, eo, en = O.pass_inject(po, n0.copy(), pt, None, 8, np.float32(1.0), prop,
                                         np.ones(nt, np.float32))
            dE = (en - eo).astype(np.float32)
            cand = [s for s in sites if 0.05 < abs(float(dE[s])) < 8.0 and dE[s] > 0]
            s = cand[trial % len(cand)]
            hit = tune_T(dE[s], abs(float(dE[s])) / (0.3 + 0.2 * trial), want_glibc_larger=bool(trial % 2))
            assert hit is not None, "no glibc/correct-rounding mismatch found near the start T"
            T, x, g, c = hit
            u = rng.random(nt).astype(np.float32)
            u[s] = min(g, c)
            # the case separates the two exps
            assert (u[s] < g) != (u[s] < c)
            lat.set_directors(n0)
            lat.set_T(T)
            acc, geo, gen = lat.pass_inject(8, prop, u)
            _, racc, reo, ren = O.pass_inject(po, n0.copy(), pt, None, 8, np.float32(T), prop, u)
            assert np.array_equal(geo[sel], reo[sel]) and np.array_equal(gen[sel], ren[sel])
            assert acc[s] == racc[s] == (1 if u[s] < g else 0), (trial, T, float(x), float(g), float(c))
            assert np.array_equal(acc[sel], racc[sel]), trial
            checked += 1
    assert checked == 16
