"""The learned proposal on the GPU: bf16-MFMA inference and training
gradients against the FP64 oracle, training on plain-sampler neighbourhoods,
the exact Metropolis-Hastings decisions of a learned sweep, and agreement of
E and S between learned and plain sweeps (3 sigma over replicas)."""
import math

import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def rand_params(rng, scale=0.25):
    p = (rng.normal(size=O.MLP_NPARAM) * scale).astype(np.float32)
    return p


def feature_batch(rng, n, T=1.1):
    X, B, D = [], [], []
    for _ in range(n):
        x, b = O.mlp_features(unit(rng, 6).reshape(-1), (rng.normal(size=3) * 0.3).astype(np.float32), T)
        X.append(x)
        B.append(b)
        D.append(unit(rng, 1)[0])
    return np.array(X, np.float32), np.array(B, np.float32), np.array(D, np.float32)


def test_inference_matches_oracle():
    rng = np.random.default_rng(1)
    P = rand_params(rng)
    X, B, _ = feature_batch(rng, 300)
    with mclics.Network(P) as net:
        mu, lk = net.infer(X, B)
    Pb = P.copy()
    nw = 64 * 22 + 64 + 64 * 64 + 64 + 4 * 64 + 4
    # the device uses bf16 weights, FP32 biases
    idx_w = np.r_[0:64 * 22, 64 * 22 + 64:64 * 22 + 64 + 64 * 64, 64 * 22 + 128 + 64 * 64:nw - 4]
    Pb[idx_w] = O.bf16(P[idx_w])
    err_mu, err_lk = [], []
    for s in range(len(X)):
        out = O.mlp_forward(Pb, X[s], 1)
        m, l = O.mlp_head(out, B[s])
        err_mu.append(np.abs(mu[s] - m).max())
        err_lk.append(abs(lk[s] - l))
    assert np.median(err_mu) < 2e-3 and max(err_mu) < 5e-2
    assert np.median(err_lk) < 2e-3 and max(err_lk) < 5e-2


def test_gradient_matches_oracle():
    rng = np.random.default_rng(2)
    P = rand_params(rng, 0.2)
    X, B, D = feature_batch(rng, 256)
    with mclics.Network(P) as net:
        loss, g = net.grad(X, B, D)
    nw = O.MLP_NPARAM
    Pb = P.copy()
    idx_w = np.r_[0:64 * 22, 64 * 22 + 64:64 * 22 + 64 + 64 * 64, 64 * 22 + 128 + 64 * 64:nw - 4]
    Pb[idx_w] = O.bf16(P[idx_w])
    lo, go = O.mlp_nll_grad(Pb, X, B, D, 1)
    assert loss == pytest.approx(lo, rel=2e-3)
    cos = float(np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go)))
    assert cos > 0.995, cos
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 0.1
    # each layer's block agrees on its own
    for a, b in ((0, 64 * 22), (64 * 22, 64 * 22 + 64), (64 * 22 + 64, 64 * 22 + 64 + 4096),
                 (64 * 22 + 128 + 4096, nw)):
        ga, gb = g[a:b], go[a:b]
        assert np.dot(ga, gb) / (np.linalg.norm(ga) * np.linalg.norm(gb) + 1e-30) > 0.99


def equilibrated(L, T, sweeps, seed):
    lat = mclics.Lattice(mclics.Config(L, L, L, T=T, seed=seed), mclics.random_ic(L, L, L, seed=seed))
    lat.sweeps(sweeps)
    lat.sync()
    return lat


def test_training_lowers_the_loss_and_raises_acceptance():
    L = 16
    with mclics.Network(seed=3) as net:
        for T, seed in ((1.05, 11), (1.1, 12), (1.2, 13)):
            lat = equilibrated(L, T, 300, seed)
            for k in range(4):
                lat.sweeps(20)
                net.add_samples(lat, 8192, seed=100 * seed + k)
            lat.close()
        assert net.dataset_size() == 3 * 4 * 8192
        lat = equilibrated(L, 1.1, 300, 21)
        lat.set_learned(net)
        acc0 = lat.record_learned_sweep()["accept"].mean()
        losses = net.train(epochs=6, batch=1024, lr=2e-3)
        assert losses[-1] < losses[0] - 0.05, losses
        acc1 = lat.record_learned_sweep()["accept"].mean()
        lat.close()
    assert acc1 > acc0 + 0.05, (acc0, acc1)


U32 = 2.0 ** -24  # FP32 unit roundoff


def mh_window(dE_T, kap, lk, co, cn, x):
    """Bound on |x_device - x| for one learned move, where the device forms
    x = -dE*RN(1/T) + lcosh2(kappa*co) - lcosh2(kappa*cn) in FP32 with
    kappa = __expf(log kappa) and compares u with __expf(x), and x is the
    FP64 value from the recorded dE, log kappa, mu and proposal (u = 2^-24):
      -dE/T in FP32: <= 2u |dE/T|;
      kappa: relative <= (|log kappa| + 3)u (x*log2e rounded, then v_exp_f32
        within one ulp); co, cn: three-term FP32 dots, <= 3u each; so each
        z = kappa*c is off by <= kappa((|log kappa| + 4)|c| + 3u) + u|z|;
      lcosh2(z) = |z| + log(1 + e^-2|z|): e^-2|z| off by <= 3u absolute,
        the rounding of 1 + y by 2u, the hardware log by 4u, the final
        addition by u(|z| + 1); d lcosh2/dz = tanh z, so a z error passes
        through with factor <= 1;
      the difference and the sum x: u|lr| + u|x|; __expf(x) relative
        (|x| + 3)u, which is the same in log space.
    A factor 2 covers the second-order terms."""
    zo, zn = kap * abs(co), kap * abs(cn)
    zerr = kap * ((abs(lk) + 4) * (abs(co) + abs(cn)) + 6) * U32 + U32 * (zo + zn)
    lc = (20 + zo + zn) * U32
    return 2.0 * (2 * abs(dE_T) * U32 + zerr + lc + U32 * (zo + zn) + (2 * abs(x) + 3) * U32)


GHRL_K = (1.0, 0.7, 1.3, 16.0)
PHYS = {"ll": {},
        "ghrl_field": {"potential": O.GHRL, "k": GHRL_K, "elec": (0.3, -0.2, 1.1), "elecA": 0.8, "elecE": 0.7},
        "pear": {"potential": O.PEAR, "A": 1.0, "B1": 0.3, "Cc": 0.3}}


@pytest.mark.parametrize("dims,bound,reflect,phys", [((12, 12, 12), (1, 1, 1), True, "ll"),
                                                     ((7, 6, 6), (0, 1, 1), True, "ll"),
                                                     ((12, 12, 12), (1, 1, 1), False, "ll"),
                                                     ((128, 6, 4), (1, 1, 1), True, "ll"),
                                                     ((8, 6, 4), (1, 1, 1), True, "ghrl_field"),
                                                     ((7, 6, 6), (0, 1, 1), True, "pear")])
def test_learned_sweep_decisions_are_exact_mh(dims, bound, reflect, phys):
    # Replays one recorded learned sweep: the oracle recomputes the energy
    # change from the reference site energies (the recorded FP32 dE must be
    # within the FMA form's proven distance of it) and, in FP64 from the
    # recorded dE, mu, log kappa and proposal, the log acceptance ratio; every
    # move outside its own FP32 error window (mh_window, a few 1e-6 at
    # kappa ~ 1) must be decided the same way.
    # The second move (reflect) is replayed the same way: the reflected
    # director must be 2 (mu.c) mu - c of the director c the first move left,
    # its dE the reference's, and its decision u2 < exp(-dE2/T) outside the
    # FP32 window of -dE2 RN(1/T) and the hardware exp.
    # The 7 x 6 x 6 lattice with a free x axis has colour-set slots past the
    # last column.  On the 128 x 6 x 4 lattice a 64-site block is one row's
    # colour sites, which the pass decodes and gathers with scalar row offsets.
    # GHRL (with a field) and pear take the FP64 bond sums of the site energy
    # for both moves.
    nx, ny, nz = dims
    T = 1.1
    rng = np.random.default_rng(4)
    P = rand_params(rng, 0.1)
    po = O.params(nx, ny, nz, bound=bound, **PHYS[phys])
    pt = np.ones(nx * ny * nz, np.int32)
    cfg = mclics.Config(nx, ny, nz, bound=bound, potential=po.potential, A=po.A, B1=po.B1, B2=po.B2, C=po.C,
                        ghrl=O.ghrl_tuple(po), elec=(po.elecX, po.elecY, po.elecZ), elecA=po.elecA,
                        elecE=po.elecE, T=T, seed=5)
    with mclics.Network(P) as net, mclics.Lattice(cfg, mclics.random_ic(nx, ny, nz, seed=5)) as lat:
        lat.sweeps(50)
        start = lat.directors()
        lat.set_learned(net, reflect=reflect)
        rec = lat.record_learned_sweep()
        end = lat.directors()
    if not reflect:
        for key in ("refl", "u2", "accept2", "dE2"):
            assert not rec[key].any(), key
    state = start.copy().reshape(-1, 3)
    T32 = float(np.float32(T))
    near, checked, widest = 0, 0, 0.0
    # distance of the FMA form of dE from the reference's site energies:
    # 300u * 1.5|A| (the bound the FP32 screen of the plain sweep uses)
    dE_tol = 300 * U32 * 1.5 * abs(po.A)
    import ctypes as C
    for colour in (0, 1):
        sites = [i + nx * (j + ny * k) for k in range(nz) for j in range(ny) for i in range(nx)
                 if (i + j + k) % 2 == colour]
        flat = state.reshape(-1).copy()
        for s in sites:
            i, j, k = s % nx, (s // nx) % ny, s // (nx * ny)
            q = rec["prop"][3 * s:3 * s + 3]
            assert abs(np.linalg.norm(q.astype(np.float64)) - 1) < 1e-5
            eo = O.lib().orc_site_energy(C.byref(po), O.f32(flat), O.i32(pt), None, i, j, k, None)
            en = O.lib().orc_site_energy(C.byref(po), O.f32(flat), O.i32(pt), None, i, j, k, O.f32(q.copy()))
            dE_ref = float(en) - float(eo)
            dE = float(rec["dE"][s])
            assert abs(dE - dE_ref) <= dE_tol, (s, dE, dE_ref)
            mu = rec["mu"][3 * s:3 * s + 3].astype(np.float64)
            lk = float(rec["logk"][s])
            kap = math.exp(lk)
            o = flat[3 * s:3 * s + 3].astype(np.float64)
            qd = q.astype(np.float64)
            x = -dE / T32 + (O.vmf_logq(mu, kap, o) - O.vmf_logq(mu, kap, qd))
            win = mh_window(dE / T32, kap, lk, float(mu @ o), float(mu @ qd), x)
            widest = max(widest, win)
            u = float(rec["u"][s])
            if (u > 0 and abs(math.log(u) - x) <= win) or (u == 0 and x < -80):
                near += 1
                acc = bool(rec["accept"][s])
            else:
                acc = u < math.exp(min(x, 700.0))
                assert acc == bool(rec["accept"][s]), (colour, s, x, u, win)
            checked += 1
            c = qd if acc else o
            if acc:
                state[s] = q
            if not reflect:
                continue
            r = rec["refl"][3 * s:3 * s + 3]
            r_ref = 2.0 * float(mu @ c) * mu - c
            assert np.abs(r.astype(np.float64) - r_ref).max() < 1e-5, (s, r, r_ref)
            cf = state[s].copy()
            ec = O.lib().orc_site_energy(C.byref(po), O.f32(flat), O.i32(pt), None, i, j, k, O.f32(cf))
            er = O.lib().orc_site_energy(C.byref(po), O.f32(flat), O.i32(pt), None, i, j, k, O.f32(r.copy()))
            dE2 = float(rec["dE2"][s])
            assert abs(dE2 - (float(er) - float(ec))) <= dE_tol, (s, dE2, float(er) - float(ec))
            x2 = -dE2 / T32
            win2 = 2.0 * (2 * abs(x2) + abs(x2) + 3) * U32
            u2 = float(rec["u2"][s])
            if (u2 > 0 and abs(math.log(u2) - x2) <= win2) or (u2 == 0 and x2 < -80):
                near += 1
                acc2 = bool(rec["accept2"][s])
            else:
                acc2 = u2 < math.exp(min(x2, 700.0))
                assert acc2 == bool(rec["accept2"][s]), (colour, s, x2, u2, win2)
            checked += 1
            if acc2:
                state[s] = r
        # next colour sees this colour's accepted moves
    print(f"learned MH replay: {checked} moves, {near} inside their window, widest window {widest:.2e}")
    assert widest < 1e-4
    assert near <= max(1, checked // 2000)
    np.testing.assert_array_equal(state.reshape(-1), end)


def test_learned_proposals_follow_their_density():
    # The sampler must draw from the density the acceptance ratio uses:
    # c = mu.q has density kappa cosh(kappa c) / (2 sinh kappa) on [-1, 1], so
    # U = (sinh(kappa c) + sinh kappa) / (2 sinh kappa) is uniform (probability
    # integral transform, each site with its own kappa).
    from scipy import stats
    L = 32
    with mclics.Network(seed=9) as net:
        lat = equilibrated(L, 1.1, 100, 41)
        for k in range(2):
            lat.sweeps(10)
            net.add_samples(lat, 8192, seed=900 + k)
        net.train(epochs=3, batch=1024, lr=2e-3)
        lat.set_learned(net)
        rec = lat.record_learned_sweep()
        lat.close()
    mu = rec["mu"].reshape(-1, 3).astype(np.float64)
    q = rec["prop"].reshape(-1, 3).astype(np.float64)
    kap = np.exp(rec["logk"].astype(np.float64))
    c = np.clip(np.sum(mu * q, axis=1), -1.0, 1.0)
    # (sinh(k c) + sinh k) / (2 sinh k), written stably for large kappa
    U = 0.5 * (np.exp(kap * (c - 1)) - np.exp(-kap * (c + 1))) / (1 - np.exp(-2 * kap)) + 0.5
    assert kap.min() > 0 and np.isfinite(U).all()
    res = stats.kstest(U, "uniform")
    assert res.pvalue > 1e-3, (res, np.percentile(kap, [5, 50, 95]))
    # azimuth and sign: the projection on a fixed axis orthogonal to mu has mean 0
    assert abs(np.mean(np.sign(np.sum(mu * q, axis=1)))) < 4 / math.sqrt(len(c))


@pytest.mark.slow
def test_learned_and_plain_agree_within_3_sigma():
    # Same lattice, temperatures on both sides of the transition; R replicas
    # per method: |mean_A - mean_B| <= 3 sqrt(se_A^2 + se_B^2).
    L, R, nequil, nmeas = 8, 16, 400, 1600
    with mclics.Network(seed=6) as net:
        lat = equilibrated(16, 1.1, 300, 31)
        for k in range(3):
            lat.sweeps(20)
            net.add_samples(lat, 8192, seed=700 + k)
        lat.close()
        net.train(epochs=4, batch=1024, lr=2e-3)
        for T in (0.9, 1.3):
            res = {}
            for method in ("plain", "learned"):
                Em, Sm = [], []
                for r in range(R):
                    seed = 1000 + r + (0 if method == "plain" else 500)
                    with mclics.Lattice(mclics.Config(L, L, L, T=T, seed=seed), mclics.random_ic(L, L, L, seed=seed)) as lt:
                        if method == "learned":
                            lt.set_learned(net)
                        lt.sweeps(nequil)
                        E, S = lt.measure_series(nmeas)
                    Em.append(E.mean())
                    Sm.append(S.mean())
                res[method] = (np.mean(Em), np.std(Em, ddof=1) / math.sqrt(R), np.mean(Sm), np.std(Sm, ddof=1) / math.sqrt(R))
            (Ea, ea, Sa, sa), (Eb, eb, Sb, sb) = res["plain"], res["learned"]
            assert abs(Ea - Eb) <= 3 * math.hypot(ea, eb), (T, res)
            assert abs(Sa - Sb) <= 3 * math.hypot(sa, sb), (T, res)


@pytest.mark.parametrize("kernel", ["auto", "slab"])
def test_learned_sweeps_in_graphs_equal_eager_sweeps(kernel):
    # Captured learned sweeps (a 64-sweep chunk replayed, then a remainder)
    # give the eager chain bit for bit, before and after the reflection move
    # is switched off (which drops the captured graphs).  The slab kernel's
    # plane swaps must not be replayed for learned sweeps, which keep the
    # plane set; plain graph sweeps after the network is detached must also
    # match.
    L = 16
    with mclics.Network(seed=12) as net:
        src = equilibrated(L, 1.1, 100, 51)
        src.sweeps(10)
        net.add_samples(src, 8192, seed=1200)
        n0 = src.directors()
        src.close()
        net.train(epochs=2, batch=1024, lr=2e-3)
        lats = []
        for graph in (False, True):
            lt = mclics.Lattice(mclics.Config(L, L, L, T=1.1, seed=52), n0)
            if kernel == "slab":
                lt.set_kernel(mclics.KERNEL_SLAB)
            lt.sweeps(3, graph=graph)  # plain sweeps first (the slab kernel flips the plane set)
            lt.set_learned(net)
            lt.sweeps(64 + 5, graph=graph)
            lt.set_learned(net, reflect=False)
            lt.sweeps(7, graph=graph)
            lt.set_learned(net)
            lt.sweeps(3, graph=graph)
            lt.set_learned(None)
            lt.sweeps(5, graph=graph)
            lt.sync()
            lats.append(lt)
        a, b = lats
        assert np.array_equal(a.directors(), b.directors())
        assert a.acceptance() == b.acceptance()
        for lt in lats:
            lt.close()
