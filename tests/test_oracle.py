"""Pins the CPU oracle against closed forms, invariants, known answers and
brute force (the reference ships no tests or golden outputs; its CPU path
cannot be built here because GSL is absent).

Each check is chosen so a plausible transcription slip (dropped term, wrong
sign, swapped operand, wrong constant) fails it.
"""
import json
import math
import os

import numpy as np
import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


# --------------------------------------------------------------------------
# Philox4x32-10 known answers.

# Published known-answer vectors of the Random123 distribution
# (kat_vectors, philox4x32 with 10 rounds).
RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_random123_kat(ctr, key, out):
    assert O.philox(ctr, key) == list(out)


def test_philox_matches_rocrand_engine():
    vecs = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    assert len(vecs) >= 20
    for v in vecs:
        assert O.philox(v["ctr"], v["key"]) == v["out"], v


def test_python_ic_philox_matches_oracle():
    from mclics.ic import philox4x32_10
    rng = np.random.default_rng(5)
    c = rng.integers(0, 2**32, size=(4, 64), dtype=np.uint64).astype(np.uint32)
    k = (123456789, 987654321)
    got = philox4x32_10(c[0], c[1], c[2], c[3], *k)
    for t in range(64):
        ref = O.philox([int(c[w][t]) for w in range(4)], k)
        assert [int(g[t]) for g in got] == ref


# --------------------------------------------------------------------------
# Bond energies: closed forms (Appendix B1, B4, B5) and symmetries.

def test_ll_parallel_closed_form():
    # B1: all directors parallel -> bond -1.5A, site -9A (6 bonds), mean -9A.
    for A in (1.0, 0.37, 2.5):
        p = O.params(4, 4, 4, A=A)
        n = np.tile(np.array([0.6, 0.0, 0.8], np.float32), 64)
        pt = np.ones(64, np.int32)
        assert O.bond(p, [0.6, 0, 0.8], [0.6, 0, 0.8], [1, 0, 0]) == pytest.approx(-1.5 * A, rel=1e-6)
        e = O.site_energies(p, n, pt)
        np.testing.assert_allclose(e, -9 * A, rtol=1e-6)
        assert O.energy_mean(p, n, pt, gpu=0) == pytest.approx(-9 * A, rel=1e-6)
        # perpendicular neighbours contribute nothing
        assert O.bond(p, [1, 0, 0], [0, 1, 0], [1, 0, 0]) == 0.0


def test_ll_isotropic_mean_is_minus_three_A():
    # B2: <nij^2> = 1/3 for independent uniform directors -> site E -> -3A.
    rng = np.random.default_rng(1)
    p = O.params(24, 24, 24, A=1.0)
    n = unit(rng, 24**3).reshape(-1)
    pt = np.ones(24**3, np.int32)
    assert O.energy_mean(p, n, pt) == pytest.approx(-3.0, abs=0.03)


def test_ghrl_coefficients_closed_forms():
    # B5: the slab configuration (k11,k22,k33) = (1.0, 0.7, 1.3).
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.0, 0.7, 1.3, 0.0))
    rho, lam, mu, nu, sig = O.ghrl_tuple(p)
    assert (rho, lam, mu, nu, sig) == pytest.approx((-0.125, 0.5, -1.125, -1.0, 0.0), rel=1e-6)
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.0, 0.7, 1.3, 16.0))
    assert p.ghrl_sigma == pytest.approx(-3.75 * 0.7 * 2 * math.pi / 16, rel=1e-6)
    assert p.ghrl_sigma == pytest.approx(-1.030835, rel=1e-6)
    # test1.txt E7 constants (1.48, 1, 2.28): S = 9/|1.48-3-2.28| = 2.368421
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.48, 1.0, 2.28, 0.0))
    rho, lam, mu, nu, sig = O.ghrl_tuple(p)
    assert (rho, lam, mu, nu) == pytest.approx((-0.210526, 0.589474, -1.136842, -1.0), rel=2e-6)
    # equal elastic constants: only nu survives and it is -1 (B4)
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.0, 1.0, 1.0, 0.0))
    assert O.ghrl_tuple(p) == pytest.approx((0.0, 0.0, 0.0, -1.0, 0.0), abs=1e-7)


def test_ghrl_equal_constants_is_minus_P2():
    # B4: k11=k22=k33 -> bond = -P2(nij), so site E_GHRL = site E_LL(A=1) + 3.
    rng = np.random.default_rng(2)
    pg = O.params(6, 6, 6, potential=O.GHRL, k=(1.0, 1.0, 1.0, 0.0))
    pl = O.params(6, 6, 6, potential=O.LL, A=1.0)
    for _ in range(50):
        a, b = unit(rng, 2)
        nij = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
        assert O.bond(pg, a, b, [0, 1, 0]) == pytest.approx(-(1.5 * nij * nij - 0.5), abs=2e-7)
    n = unit(rng, 216).reshape(-1)
    pt = np.ones(216, np.int32)
    np.testing.assert_allclose(O.site_energies(pg, n, pt), O.site_energies(pl, n, pt) + 3.0, atol=3e-6)


@pytest.mark.parametrize("k", [(1.0, 0.7, 1.3, 0.0), (1.48, 1.0, 2.28, 0.0), (1.0, 0.7, 1.3, 16.0)])
def test_ghrl_parallel_pair_closed_forms(k):
    # Parallel directors (nij=1, pij=1, cross=0) at angle a=n.r to the bond:
    # E = (3a^2-1)(rho+lambda) + mu(a^2 - 1/9) + nu.
    p = O.params(2, 2, 2, potential=O.GHRL, k=k)
    rho, lam, mu, nu, _ = (float(x) for x in O.ghrl_tuple(p))
    for n, r, a in (([1, 0, 0], [1, 0, 0], 1.0), ([0, 1, 0], [1, 0, 0], 0.0), ([0, 0, 1], [0, 0, -1], -1.0)):
        want = (3 * a * a - 1) * (rho + lam) + mu * (a * a - 1.0 / 9) + nu
        assert O.bond(p, n, n, r) == pytest.approx(want, rel=1e-6, abs=1e-6)


def test_ghrl_chiral_twist_sign():
    # A twist between neighbours along z: (ni x nj).z = sin(phi).  With p0 > 0,
    # sigma < 0 and the cross term sigma*sign(nij)*(-(ni x nj).r) favours the
    # twist of one handedness: E(+phi) - E(-phi) = -2*sigma*sin(phi) for nij>0.
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.0, 0.7, 1.3, 16.0))
    phi = 0.3
    ni = [1.0, 0.0, 0.0]
    jp = [math.cos(phi), math.sin(phi), 0.0]
    jm = [math.cos(phi), -math.sin(phi), 0.0]
    d = O.bond(p, ni, jp, [0, 0, 1]) - O.bond(p, ni, jm, [0, 0, 1])
    assert d == pytest.approx(-2 * p.ghrl_sigma * math.sin(phi), rel=1e-5)


def test_pear_closed_forms():
    # nj = ni: polar splay vanishes -> -A - B1.  nj = -ni: (1+2nij+nij^2)=0 -> -A + B1.
    p = O.params(2, 2, 2, potential=O.PEAR, A=1.0, B1=0.3, Cc=0.3)
    n = [0.0, 0.6, 0.8]
    m = [-x for x in n]
    assert O.bond(p, n, n, [1, 0, 0]) == pytest.approx(-1.0 - 0.3, rel=1e-6)
    assert O.bond(p, n, m, [0, 1, 0]) == pytest.approx(-1.0 + 0.3, rel=1e-6)
    # splay term: ni=(1,0,0), nj=(0,1,0), r=(0,1,0): nij=0, (nj-ni).r = 1 -> -C/4
    assert O.bond(p, [1, 0, 0], [0, 1, 0], [0, 1, 0]) == pytest.approx(-0.3 * 0.25, rel=1e-6)


@pytest.mark.parametrize("pot,kw", [(O.LL, {}), (O.GHRL, {"k": (1.0, 0.7, 1.3, 16.0)}),
                                    (O.PEAR, {"A": 1.0, "B1": 0.3, "Cc": 0.3})])
def test_bond_pair_symmetry(pot, kw):
    # bond(i,j,r) = bond(j,i,-r): the site-energy Metropolis needs it.
    rng = np.random.default_rng(3)
    p = O.params(2, 2, 2, potential=pot, **kw)
    for _ in range(200):
        a, b = unit(rng, 2)
        r = [0, 0, 0]
        ax = rng.integers(3)
        r[ax] = 1
        mr = [-x for x in r]
        assert O.bond(p, a, b, r) == pytest.approx(O.bond(p, b, a, mr), rel=2e-6, abs=2e-6)


@pytest.mark.parametrize("pot,kw", [(O.LL, {}), (O.GHRL, {"k": (1.0, 0.7, 1.3, 0.0)})])
def test_bond_head_tail_invariance(pot, kw):
    # n -> -n for LL and achiral GHRL
    rng = np.random.default_rng(4)
    p = O.params(2, 2, 2, potential=pot, **kw)
    for _ in range(200):
        a, b = unit(rng, 2)
        assert O.bond(p, a, b, [0, 0, 1]) == pytest.approx(O.bond(p, -a, b, [0, 0, 1]), rel=1e-6, abs=1e-6)


def test_ll_rotation_invariance():
    rng = np.random.default_rng(5)
    p = O.params(2, 2, 2)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    for _ in range(100):
        a, b = unit(rng, 2)
        ra, rb = (q @ a).astype(np.float32), (q @ b).astype(np.float32)
        assert O.bond(p, a, b, [1, 0, 0]) == pytest.approx(O.bond(p, ra, rb, [1, 0, 0]), abs=2e-6)


def test_field_closed_form_and_unnormalised_direction():
    # -elecA*(elecE*(n.e))^2 ; e is not normalised by the reference (A18).
    p = O.params(1, 1, 1, elec=(0.0, 0.0, 2.0), elecA=0.7, elecE=0.5)
    assert O.field(p, [0, 0, 1]) == pytest.approx(-0.7 * (0.5 * 2.0) ** 2, rel=1e-6)
    assert O.field(p, [1, 0, 0]) == 0.0


def test_surface_kinds_closed_forms():
    s = [0.0, 0.0, 1.0]
    d = [1.0, 0.0, 0.0]
    kinds = [(k, 2.0, d) for k in ("homeotropic", "rp", "fg", "strong", "homeotropic_ghrl", "rp_ghrl", "fg_ghrl",
                                  "strong_ghrl")]
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.0, 1.0, 1.0, 0.0), surfaces=kinds)
    ez, ex = [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]
    assert O.surface(p, 0, ez, s) == pytest.approx(-2.0)       # homeotropic, n || s
    assert O.surface(p, 0, ex, s) == 0.0                       # homeotropic, n _|_ s
    assert O.surface(p, 1, ex, s) == pytest.approx(-2.0)       # rp uses its own direction d
    assert O.surface(p, 1, ez, s) == 0.0
    assert O.surface(p, 2, ez, s) == pytest.approx(+2.0)       # fg: +W (n.s)^2
    assert O.surface(p, 3, ex, s) == pytest.approx(-2.0)       # strong: -W (n.d)^2
    assert O.surface(p, 7, ex, s) == pytest.approx(-2.0)       # strong_ghrl == strong on the CPU
    # GHRL forms with equal constants reduce to W*(-P2(n.partner)):
    assert O.surface(p, 4, ez, s) == pytest.approx(-2.0)       # partner s, nij = 1
    assert O.surface(p, 4, ex, s) == pytest.approx(+1.0)       # nij = 0 -> -P2 = 0.5
    assert O.surface(p, 5, ex, s) == pytest.approx(-2.0)       # partner d, nij = 1
    # fg_ghrl with n || s: the projected partner vanishes, nij = 0 -> W*0.5
    assert O.surface(p, 6, ez, s) == pytest.approx(+1.0)


@pytest.mark.parametrize("p0", [0.0, 16.0])
def test_ghrl_anchoring_closed_forms_unequal_constants(p0):
    # The three GHRL surface forms at k = (1.0, 0.7, 1.3), where every GHRL
    # coefficient is nonzero (S = 3.75: rho = -0.125, lambda = 0.5,
    # mu = -1.125, nu = -1; sigma = 0 or -3.75*0.7*2pi/16), with the site
    # normal s = z and the rp direction d = x (theta_s = 90, phi_s = 0).
    # Values worked out by hand from src/anchoring_homeotropic.cpp:51-68,
    # src/anchoring_rp.cpp:77-98 and src/anchoring_fg.cpp:51-74
    # (E = W [E1 (rho P2 + lambda) + mu a_i a_j n_ij - (1/9 in integers = 0)
    #  + nu P2 + sigma sgn(n_ij) cross], cross = (n_i x n_j).(-s)).
    W = 2.5
    rho, lam, mu, nu = -0.125, 0.5, -1.125, -1.0
    sig = -3.75 * 0.7 * 2 * math.pi / 16 if p0 else 0.0
    kinds = [("homeotropic_ghrl", W, (1, 0, 0)), ("rp_ghrl", W, (1, 0, 0)), ("fg_ghrl", W, (1, 0, 0))]
    p = O.params(2, 2, 2, potential=O.GHRL, k=(1.0, 0.7, 1.3, p0), surfaces=kinds)
    assert O.ghrl_tuple(p) == pytest.approx((rho, lam, mu, nu, sig), rel=1e-6)
    s = [0.0, 0.0, 1.0]
    c = math.sqrt(0.5)
    E = lambda kind, n: O.surface(p, kind, n, s)
    # homeotropic_ghrl: partner and bond direction are s, so a_j = 1, n_ij = a_i = n.s
    # and the chiral term vanishes ((n x s).s = 0) for either p0.
    assert E(0, [0, 0, 1]) == pytest.approx(W * (2 * (rho + lam) + mu + nu), rel=1e-6)   # -1.375 W
    assert E(0, [0, 0, 1]) == pytest.approx(-1.375 * W, rel=1e-6)
    assert E(0, [1, 0, 0]) == pytest.approx(0.78125 * W, rel=1e-6)                      # a = 0
    assert E(0, [c, 0, c]) == pytest.approx(-0.2265625 * W, rel=1e-6)                   # a = 1/sqrt2
    for a in (0.1, 0.35, 0.8, -0.6):
        n = [math.sqrt(1 - a * a), 0.0, a]
        P2 = 1.5 * a * a - 0.5
        want = W * ((1.5 * a * a + 0.5) * (rho * P2 + lam) + mu * a * a + nu * P2)
        assert E(0, n) == pytest.approx(want, rel=2e-6, abs=2e-6)
    # rp_ghrl: partner d = x, bond direction s; a_j = d.s = 0
    assert E(1, [1, 0, 0]) == pytest.approx(-1.375 * W, rel=1e-6)                       # n = d
    assert E(1, [0, 0, 1]) == pytest.approx(0.78125 * W, rel=1e-6)                      # n = s
    # n = y: n_ij = 0 (sgn -1), cross = 1: W (-(lambda - rho/2) - nu/2 - sigma)
    assert E(1, [0, 1, 0]) == pytest.approx(W * (-0.0625 - sig), rel=1e-6, abs=1e-6)
    # n = (c, +-c, 0): n_ij = c, cross = +-c
    assert E(1, [c, c, 0]) == pytest.approx(W * (-0.71875 + sig * c), rel=1e-6, abs=2e-6)
    assert E(1, [c, -c, 0]) == pytest.approx(W * (-0.71875 - sig * c), rel=1e-6, abs=2e-6)
    assert E(1, [-c, c, 0]) == pytest.approx(W * (-0.71875 - sig * c), rel=1e-6, abs=2e-6)       # n_ij < 0 flips sgn
    # fg_ghrl: partner = n - s sqrt|n.s| normalised (in the n, s plane: no chiral term)
    assert E(2, [0, 0, 1]) == pytest.approx(0.78125 * W, rel=1e-6)                      # partner vanishes
    assert E(2, [1, 0, 0]) == pytest.approx(-1.375 * W, rel=1e-6)                       # partner = n
    m = math.sqrt(c)
    nj = [c, 0.0, c - m]
    nn = math.sqrt(nj[0] ** 2 + nj[2] ** 2)
    nj = [nj[0] / nn, 0.0, nj[2] / nn]
    ai, aj, nij = c, nj[2], c * nj[0] + c * nj[2]
    P2 = 1.5 * nij * nij - 0.5
    want = W * ((1.5 * ai * ai + 1.5 * aj * aj - 1) * (rho * P2 + lam) + mu * ai * aj * nij + nu * P2)
    assert E(2, [c, 0, c]) == pytest.approx(want, rel=2e-6, abs=3e-6)  # terms of ~0.5 cancel to 0.02 in FP32


# --------------------------------------------------------------------------
# Accept rule, ordering and adaptation.

def test_accept_rule_invariants():
    rng = np.random.default_rng(6)
    L = 6
    nt = L**3
    p = O.params(L, L, L)
    pt = np.ones(nt, np.int32)
    n0 = unit(rng, nt).reshape(-1)
    # proposals equal to the current state: dE = 0 -> exp(0) = 1 > u always
    u = np.full(nt, 1 - 2**-24, np.float32)
    n = n0.copy()
    cnt, acc, eo, en = O.pass_inject(p, n, pt, None, 8, 1.0, n0.copy(), u)
    assert cnt == nt // 2 and np.all(acc[eo != 0] == 1)
    # T -> infinity accepts every proposal
    prop = unit(rng, nt).reshape(-1)
    n = n0.copy()
    cnt, acc, eo, en = O.pass_inject(p, n, pt, None, 9, 1e30, prop, u)
    assert cnt == nt // 2
    # T -> 0 accepts exactly the moves with dE < 0 (u = 0 accepts iff exp > 0)
    n = n0.copy()
    cnt, acc, eo, en = O.pass_inject(p, n, pt, None, 9, 1e-30, prop, np.zeros(nt, np.float32))
    black = np.array([(i + j + k) % 2 == 1 for k in range(L) for j in range(L) for i in range(L)])
    assert np.array_equal(acc[black].astype(bool), (en < eo)[black])


@pytest.mark.parametrize("set_id", list(range(10)))
def test_pass_equals_sequential_any_order(set_id):
    # Sites of one pass do not interact: updating them one at a time in a
    # shuffled order (with the oracle's own single-site energy) gives the
    # same lattice.  Brute force on 6x4x8, periodic.
    rng = np.random.default_rng(10 + set_id)
    nx, ny, nz = 6, 4, 8
    nt = nx * ny * nz
    p = O.params(nx, ny, nz, potential=O.GHRL, k=(1.0, 0.7, 1.3, 16.0))
    pt = np.ones(nt, np.int32)
    n0 = unit(rng, nt).reshape(-1)
    prop = unit(rng, nt).reshape(-1)
    u = rng.random(nt).astype(np.float32)
    T = 0.7
    n1 = n0.copy()
    cnt, acc, _, _ = O.pass_inject(p, n1, pt, None, set_id, T, prop, u)
    sites = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if set_id < 8:
                    ok = (i % 2, j % 2, k % 2) == (set_id % 2, (set_id // 2) % 2, set_id // 4)
                else:
                    ok = (i + j + k) % 2 == set_id - 8
                if ok:
                    sites.append((i, j, k))
    rng.shuffle(sites)
    n2 = n0.copy()
    import ctypes as C
    for (i, j, k) in sites:
        s = i + nx * (j + ny * k)
        eo = O.lib().orc_site_energy(C.byref(p), O.f32(n2), O.i32(pt), None, i, j, k, None)
        q = prop[3 * s:3 * s + 3].copy()
        en = O.lib().orc_site_energy(C.byref(p), O.f32(n2), O.i32(pt), None, i, j, k, O.f32(q))
        x = np.float32(-(np.float32(en) - np.float32(eo))) / np.float32(T)
        if float(u[s]) < math.exp(float(x)):
            n2[3 * s:3 * s + 3] = q
    np.testing.assert_array_equal(n1, n2)
    idx = [i + nx * (j + ny * k) for i, j, k in sites]
    changed = np.any(n1.reshape(-1, 3)[idx] != n0.reshape(-1, 3)[idx], axis=1).sum()
    assert cnt == changed


def test_adapt_ang_var_rules():
    assert O.lib().orc_adapt_ang_var(0.5, 40, 100) == pytest.approx(0.5 * 0.99, rel=1e-7)
    assert O.lib().orc_adapt_ang_var(0.5, 60, 100) == pytest.approx(0.5 / 0.99, rel=1e-7)
    assert O.lib().orc_adapt_ang_var(0.5, 50, 100) == 0.5
    assert O.lib().orc_adapt_ang_var(0.0100001, 0, 100) == pytest.approx(0.01, rel=1e-6)
    assert O.lib().orc_adapt_ang_var(0.995, 100, 100) == pytest.approx(0.995 / 0.99 - 0.5, rel=1e-6)


def test_energy_mean_semantics_with_inactive_sites():
    rng = np.random.default_rng(7)
    p = O.params(4, 4, 4, bound=(O.FREE, O.FREE, O.FREE))
    nt = 64
    pt = np.ones(nt, np.int32)
    pt[::5] = 0
    n = unit(rng, nt).reshape(-1)
    e = O.site_energies(p, n, pt)
    cpu = O.energy_mean(p, n, pt, gpu=0)
    gpu = O.energy_mean(p, n, pt, gpu=1)
    assert cpu == pytest.approx(e.astype(np.float64).sum() / nt, rel=1e-12)
    assert gpu == pytest.approx(e[pt != 0].astype(np.float64).sum() / (pt != 0).sum(), rel=1e-12)


# --------------------------------------------------------------------------
# Stationary distributions (B6, B7): the whole proposal + accept machinery.

def _quad_moment(a, npts=200001):
    x = np.linspace(0.0, 1.0, npts)
    w = np.exp(a * x * x)
    return np.trapezoid(x * x * w, x) / np.trapezoid(w, x)


def _batch_mean(x, nb=50):
    x = np.asarray(x)
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / math.sqrt(nb)


@pytest.mark.parametrize("bound,bonds", [(O.FREE, 1), (O.PERIODIC, 2)])
def test_two_site_stationary_distribution(bound, bonds):
    # Two sites along x: x = n1.n2 has density exp(bonds*1.5*A*x^2/T) on [-1,1].
    A, T = 1.0, 0.8
    p = O.params(2, 1, 1, bound=(bound, O.FREE, O.FREE), A=A)
    pt = np.ones(2, np.int32)
    n = np.array([0, 0, 1, 1, 0, 0], np.float32)
    av = 0.5
    xs = []
    for w in range(60000):
        _, av = O.sweep_rng(p, n, pt, None, T, av, seed=11, sweep=w, nthreads=1)
        if w >= 1000:
            a, b = n[:3].astype(np.float64), n[3:].astype(np.float64)
            xs.append(float(np.dot(a, b)) ** 2)
    m, se = _batch_mean(xs)
    want = _quad_moment(bonds * 1.5 * A / T)
    assert abs(m - want) < 4 * se + 1e-3, (m, want, se)


@pytest.mark.parametrize("kind", ["field", "homeotropic"])
def test_one_site_stationary_distribution(kind):
    # One free site with only a field (kappa = elecA*elecE^2) or only
    # homeotropic anchoring (kappa = W): c = n.e has density exp(kappa c^2/T).
    T = 0.6
    if kind == "field":
        p = O.params(1, 1, 1, bound=(O.FREE,) * 3, elec=(0.0, 0.0, 1.0), elecA=0.9, elecE=1.2)
        pt = np.ones(1, np.int32)
        ns = None
        kappa = 0.9 * 1.2**2
    else:
        p = O.params(1, 1, 1, bound=(O.FREE,) * 3, surfaces=[("homeotropic", 1.3, (0, 0, 1))])
        pt = np.full(1, 2, np.int32)
        ns = np.array([0, 0, 1], np.float32)
        kappa = 1.3
    n = np.array([1, 0, 0], np.float32)
    av = 0.5
    cs = []
    for w in range(60000):
        _, av = O.sweep_rng(p, n, pt, ns, T, av, seed=12, sweep=w, nthreads=1)
        if w >= 1000:
            cs.append(float(n[2]) ** 2)
    m, se = _batch_mean(cs)
    want = _quad_moment(kappa / T)
    assert abs(m - want) < 4 * se + 1e-3, (m, want, se)


# --------------------------------------------------------------------------
# Order parameter.

def test_eigvals_match_lapack():
    rng = np.random.default_rng(8)
    for _ in range(500):
        m = rng.normal(size=(3, 3))
        m = m + m.T
        if rng.random() < 0.2:
            m = np.diag(rng.normal(size=3))
        q6 = [m[0, 0], m[0, 1], m[0, 2], m[1, 1], m[1, 2], m[2, 2]]
        np.testing.assert_allclose(O.eigvals(q6), np.sort(np.linalg.eigvalsh(m))[::-1], atol=1e-12, rtol=1e-12)


def test_order_parameter_closed_forms():
    # B3: parallel -> 1; half x / half y -> 1/4; equal thirds on x,y,z -> 0.
    nt = 600
    pt = np.ones(nt, np.int32)
    n = np.tile(np.array([0, 0, 1], np.float32), nt)
    assert O.eigvals(O.qtensor(n, pt))[0] == pytest.approx(1.0, abs=1e-12)
    n = np.concatenate([np.tile(np.array([1, 0, 0], np.float32), nt // 2),
                        np.tile(np.array([0, 1, 0], np.float32), nt // 2)])
    assert O.eigvals(O.qtensor(n, pt))[0] == pytest.approx(0.25, abs=1e-12)
    n = np.concatenate([np.tile(np.array(v, np.float32), nt // 3) for v in ([1, 0, 0], [0, 1, 0], [0, 0, 1])])
    assert O.eigvals(O.qtensor(n, pt))[0] == pytest.approx(0.0, abs=1e-12)
    # inactive sites are excluded
    pt2 = pt.copy()
    pt2[nt // 3:] = 0
    assert O.eigvals(O.qtensor(n, pt2))[0] == pytest.approx(1.0, abs=1e-12)


def test_local_order_parameter():
    # Uniform field -> 1; a site and its 6 neighbours from an isotropic sample
    # -> numpy eigvalsh of the 7-director Q.
    L = 5
    p = O.params(L, L, L, bound=(O.FREE,) * 3)
    nt = L**3
    pt = np.ones(nt, np.int32)
    n = np.tile(np.array([0, 1, 0], np.float32), nt)
    assert O.local_S(p, n, pt, 2, 2, 2) == pytest.approx(1.0, abs=1e-12)
    rng = np.random.default_rng(9)
    n = unit(rng, nt).reshape(-1)
    for (i, j, k) in ((2, 2, 2), (0, 0, 0), (4, 2, 0)):
        vs = []
        for d in ((0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            a, b, c = i + d[0], j + d[1], k + d[2]
            if 0 <= a < L and 0 <= b < L and 0 <= c < L:
                s = a + L * (b + L * c)
                vs.append(n[3 * s:3 * s + 3].astype(np.float64))
        V = np.array(vs)
        Q = (3 * V.T @ V / len(vs) - np.eye(3)) / 2
        assert O.local_S(p, n, pt, i, j, k) == pytest.approx(np.linalg.eigvalsh(Q)[-1], abs=1e-12)


# --------------------------------------------------------------------------
# Mode loops (A14, B10).

def test_thermal_float_loop_rows():
    # C1: T 1.2 -> 0.0 by -0.02 in float gives 61 rows, the last ~6.26e-7.
    rows, xs = O.float_loop(1.2, 0.0, -0.02)
    assert rows == 61
    assert xs[0] == np.float32(1.2)
    assert 0 < xs[-1] < 1e-6
    assert xs[-1] == pytest.approx(6.26e-7, rel=0.01)


def test_electric_float_loop_rows():
    # C3: E 0 -> 1.0 by 0.05 gives 20 rows; the 21st value is 1.0000001 > 1.
    rows, xs = O.float_loop(0.0, 1.0, 0.05)
    assert rows == 20
    assert xs[-1] == pytest.approx(0.95, abs=1e-6)
