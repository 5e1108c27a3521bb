"""GPU parity: the HIP engine against the CPU oracle through the C ABI.

Injected proposals and uniforms (mc_pass) and recorded production sweeps
(mc_record_sweep, replayed pass by pass in the oracle) must give bit-identical
accept decisions, site energies and lattice state.  Energies are compared
bit-exactly, which is stricter than the 1e-5 relative bound required.
"""
import math
import zlib

import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


def in_set(nx, ny, nz, set_id):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).reshape(-1) for a in (i, j, k))
    if set_id < 8:
        return (i % 2 == set_id % 2) & (j % 2 == (set_id // 2) % 2) & (k % 2 == set_id // 4)
    return (i + j + k) % 2 == set_id - 8


GHRL_K = (1.0, 0.7, 1.3, 16.0)


def make_case(name, rng):
    """(oracle params, engine config, pt, ns) for one geometry/physics case."""
    surf = []
    if name == "ll_bulk":
        dims, bound, pot, kw = (6, 4, 8), (1, 1, 1), O.LL, {"A": 1.3}
    elif name == "ghrl_bulk_field":
        dims, bound, pot, kw = (8, 6, 4), (1, 1, 1), O.GHRL, {"k": GHRL_K, "elec": (0.3, -0.2, 1.1), "elecA": 0.8,
                                                              "elecE": 0.7}
    elif name == "pear_free":
        dims, bound, pot, kw = (5, 7, 4), (0, 0, 0), O.PEAR, {"A": 1.0, "B1": 0.3, "Cc": 0.3}
    elif name == "slab_chiral":
        # the cholesteric slab workload at a small size: GHRL with p0 = 16
        # (sigma from src/io.cpp:284), homeotropic GHRL faces, field along z
        dims, bound, pot = (8, 6, 6), (1, 1, 0), O.GHRL
        surf = [("homeotropic_ghrl", 5.0, (0.0, 0.0, 1.0))] * 2
        kw = {"k": GHRL_K, "elec": (0.0, 0.0, 1.0), "elecA": 1.0, "elecE": 0.5}
    elif name.startswith("slab_"):
        kind = name[5:]
        dims, bound, pot = (8, 6, 6), (1, 1, 0), O.GHRL
        d = (0.6, 0.0, 0.8)
        surf = [(kind, 4.0, d), ("homeotropic_ghrl" if kind != "homeotropic_ghrl" else "fg", 2.5, d)]
        kw = {"k": (1.48, 1.0, 2.28, 0.0), "elec": (0.0, 0.0, 1.0), "elecA": 1.0, "elecE": 0.45}
    elif name == "ll_holes":
        dims, bound, pot, kw = (6, 6, 6), (1, 1, 1), O.LL, {}
    elif name == "ll_slab":
        # LL between homeotropic (k = 0) and planar (k = Nz-1) faces, with a field
        dims, bound, pot = (16, 8, 10), (1, 1, 0), O.LL
        surf = [("homeotropic", 3.0, (0.0, 0.0, 1.0)), ("rp", 2.0, (1.0, 0.0, 0.0))]
        kw = {"A": 1.0, "elec": (0.0, 0.0, 1.0), "elecA": 0.6, "elecE": 0.5}
    elif name == "ll_freez":
        dims, bound, pot, kw = (16, 8, 9), (1, 1, 0), O.LL, {}
    else:
        raise KeyError(name)
    nx, ny, nz = dims
    nt = nx * ny * nz
    po = O.params(nx, ny, nz, bound=bound, potential=pot, surfaces=surf, **kw)
    pt = np.ones(nt, np.int32)
    ns = None
    if name.startswith("slab_") or name == "ll_slab":  # slab_chiral included
        ns = np.zeros(3 * nt, np.float32)
        for j in range(ny):
            for i in range(nx):
                s0 = i + nx * j
                s1 = i + nx * (j + ny * (nz - 1))
                pt[s0], pt[s1] = 2, 3
                ns[3 * s0 + 2], ns[3 * s1 + 2] = -1.0, 1.0
    if name == "ll_holes":
        pt[rng.random(nt) < 0.2] = 0
    cfg = mclics.Config(nx, ny, nz, bound=bound, potential=pot, A=po.A, B1=po.B1, B2=po.B2, C=po.C,
                        ghrl=O.ghrl_tuple(po), elec=(po.elecX, po.elecY, po.elecZ), elecA=po.elecA, elecE=po.elecE,
                        surfaces=[mclics.Surface(k, w, d) for k, w, d in surf], T=0.9, seed=7)
    return po, cfg, pt, ns


CASES = ["ll_bulk", "ghrl_bulk_field", "pear_free", "ll_holes"] + [
    "slab_" + k for k in ("homeotropic", "rp", "fg", "strong", "homeotropic_ghrl", "rp_ghrl", "fg_ghrl",
                          "strong_ghrl")]


@pytest.mark.parametrize("case", CASES)
def test_injected_pass_parity(case):
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    po, cfg, pt, ns = make_case(case, rng)
    nx, ny, nz = po.nx, po.ny, po.nz
    nt = nx * ny * nz
    n = unit(rng, nt)
    n[np.repeat(pt == 0, 3)] = np.tile([0.0, 0.0, 1.0], int((pt == 0).sum())).astype(np.float32)
    ref = n.copy()
    with mclics.Lattice(cfg, n, pt, ns) as lat:
        for rep in range(2):
            for set_id in range(10):
                # a mix of small rotations and fresh directions, and a spread of T
                prop = unit(rng, nt)
                small = rng.random(nt) < 0.5
                jitter = (ref.reshape(-1, 3) + 0.2 * rng.normal(size=(nt, 3))).astype(np.float64)
                jitter /= np.linalg.norm(jitter, axis=1, keepdims=True)
                prop.reshape(-1, 3)[small] = jitter[small].astype(np.float32)
                u = rng.random(nt).astype(np.float32)
                T = float(np.float32([0.05, 0.9, 3.0][(set_id + rep) % 3]))
                lat.set_T(T)
                acc, eo, en = lat.pass_inject(set_id, prop, u)
                _, racc, reo, ren = O.pass_inject(po, ref, pt, ns, set_id, np.float32(T), prop, u)
                sel = in_set(nx, ny, nz, set_id) & (pt != 0)
                assert np.array_equal(acc[sel], racc[sel]), (case, set_id)
                assert np.array_equal(eo[sel], reo[sel]), (case, set_id, np.max(np.abs(eo[sel] - reo[sel])))
                assert np.array_equal(en[sel], ren[sel]), (case, set_id)
                assert np.array_equal(lat.directors(), ref), (case, set_id)
        np.testing.assert_array_equal(lat.site_energies()[pt != 0], O.site_energies(po, ref, pt, ns)[pt != 0])


@pytest.mark.parametrize("case,order", [("ll_bulk", mclics.ORDER_CHECKER), ("ll_bulk", mclics.ORDER_REF8),
                                        ("ghrl_bulk_field", mclics.ORDER_CHECKER),
                                        ("slab_homeotropic_ghrl", mclics.ORDER_REF8),
                                        ("pear_free", mclics.ORDER_CHECKER), ("ll_holes", mclics.ORDER_REF8)])
def test_recorded_sweep_replay(case, order):
    rng = np.random.default_rng(11)
    po, cfg, pt, ns = make_case(case, rng)
    cfg.order = order
    nx, ny, nz = po.nx, po.ny, po.nz
    nt = nx * ny * nz
    n = unit(rng, nt)
    n[np.repeat(pt == 0, 3)] = np.tile([0.0, 0.0, 1.0], int((pt == 0).sum())).astype(np.float32)
    ref = n.copy()
    sets = list(range(8)) if order == mclics.ORDER_REF8 else [8, 9]
    with mclics.Lattice(cfg, n, pt, ns) as lat:
        lat.set_kernel(mclics.KERNEL_GENERIC)
        for w in range(4):
            av = lat.ang_var()
            rec = lat.record_sweep()
            total = 0
            for set_id in sets:
                cnt, racc, reo, ren = O.pass_inject(po, ref, pt, ns, set_id, np.float32(cfg.T), rec["prop"],
                                                    rec["u2"])
                sel = in_set(nx, ny, nz, set_id) & (pt != 0)
                assert np.array_equal(rec["accept"][sel], racc[sel]), (case, w, set_id)
                assert np.array_equal(rec["e_old"][sel], reo[sel]), (case, w, set_id)
                assert np.array_equal(rec["e_new"][sel], ren[sel]), (case, w, set_id)
                total += cnt
            assert np.array_equal(lat.directors(), ref), (case, w)
            acc, sweeps = lat.acceptance()
            assert acc == total and sweeps == w + 1
            assert lat.ang_var() == O.lib().orc_adapt_ang_var(av, total, int((pt != 0).sum()))


def test_proposals_are_normalised_rotations():
    # Each proposal is the current director rotated by pi*va (va = (2u0-1)*ang_var)
    # about x, y or z chosen by u1 against 0.333/0.666, then renormalised
    # (src/evolve.cu:143-163).  Compared with an FP64 rotation.
    rng = np.random.default_rng(12)
    nx = ny = nz = 12
    nt = nx * ny * nz
    cfg = mclics.Config(nx, ny, nz, T=1.1, seed=99)
    n = unit(rng, nt)
    with mclics.Lattice(cfg, n) as lat:
        lat.set_kernel(mclics.KERNEL_GENERIC)
        lat.set_ang_var(0.8)
        old = lat.directors().reshape(-1, 3).astype(np.float64)
        rec = lat.record_sweep()
    u0, u1 = rec["u0"].astype(np.float64), rec["u1"].astype(np.float64)
    prop = rec["prop"].reshape(-1, 3).astype(np.float64)
    black = in_set(nx, ny, nz, 9)
    # black sites see red neighbours already updated; their own director is unchanged before the pass
    va = (2 * u0 - 1) * np.float32(0.8)
    a = np.pi * va
    c, s = np.cos(a), np.sin(a)
    x, y, z = old[:, 0], old[:, 1], old[:, 2]
    want = np.where((u1 < 0.333)[:, None], np.stack([x, y * c + z * s, z * c - y * s], 1),
                    np.where((u1 < 0.666)[:, None], np.stack([x * c - z * s, y, z * c + x * s], 1),
                             np.stack([x * c + y * s, y * c - x * s, z], 1)))
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    err = np.abs(prop - want).max(1)
    assert err.max() < 5e-6, err.max()
    np.testing.assert_allclose(np.linalg.norm(prop, axis=1), 1.0, atol=3e-7)
    # axis frequencies follow the 0.333/0.333/0.334 thresholds
    frac = [(u1 < 0.333).mean(), ((u1 >= 0.333) & (u1 < 0.666)).mean()]
    assert abs(frac[0] - 0.333) < 0.03 and abs(frac[1] - 0.333) < 0.03
    assert black.sum() > 0


def test_measure_matches_oracle():
    rng = np.random.default_rng(13)
    for case in ("ll_bulk", "ghrl_bulk_field", "slab_rp_ghrl", "ll_holes"):
        po, cfg, pt, ns = make_case(case, rng)
        nt = po.nx * po.ny * po.nz
        n = unit(rng, nt)
        n[np.repeat(pt == 0, 3)] = np.tile([0.0, 0.0, 1.0], int((pt == 0).sum())).astype(np.float32)
        with mclics.Lattice(cfg, n, pt, ns) as lat:
            lat.sweeps(3)
            d = lat.directors()
            E, S, q = lat.measure()
        assert E == pytest.approx(O.energy_mean(po, d, pt, ns, gpu=1), rel=1e-12, abs=1e-12)
        qo = O.qtensor(d, pt)
        np.testing.assert_allclose(q, qo, rtol=1e-11, atol=1e-13)
        Q = np.array([[qo[0], qo[1], qo[2]], [qo[1], qo[3], qo[4]], [qo[2], qo[4], qo[5]]])
        assert S == pytest.approx(np.linalg.eigvalsh(Q)[-1], abs=1e-12)


def _measure_case(dims, pot, zfree, faces, rng, **kw):
    nx, ny, nz = dims
    nt = nx * ny * nz
    surf = [("homeotropic_ghrl", 3.0, (0.0, 0.0, 1.0)), ("homeotropic_ghrl", 2.0, (0.0, 0.0, 1.0))] if faces else []
    bound = (1, 1, 0 if zfree else 1)
    po = O.params(nx, ny, nz, bound=bound, potential=pot, surfaces=surf, **kw)
    pt = np.ones(nt, np.int32)
    ns = None
    if faces:
        ns = np.zeros(3 * nt, np.float32)
        pt[: nx * ny] = 2
        pt[nx * ny * (nz - 1):] = 3
        ns[2: 3 * nx * ny: 3] = -1.0
        ns[3 * nx * ny * (nz - 1) + 2::3] = 1.0
    cfg = mclics.Config(nx, ny, nz, bound=bound, potential=pot, A=po.A, B1=po.B1, B2=po.B2, C=po.C,
                        ghrl=O.ghrl_tuple(po), elec=(po.elecX, po.elecY, po.elecZ), elecA=po.elecA, elecE=po.elecE,
                        surfaces=[mclics.Surface(k, w, d) for k, w, d in surf], T=0.9, seed=11)
    return po, cfg, pt if faces else None, ns, unit(rng, nt)


@pytest.mark.parametrize("dims,pot,zfree,faces,kw", [
    ((4, 2, 2), O.LL, False, False, {"A": 1.0}),                       # one tile, one segment per row
    ((12, 6, 10), O.LL, False, False, {"A": 1.3}),                     # 6-row tiles, 3 segments per row
    ((128, 12, 4), O.LL, False, False, {"A": 1.0, "elec": (0.2, 0.1, 1.0), "elecA": 0.5, "elecE": 0.8}),
    ((64, 16, 8), O.GHRL, False, False, {"k": GHRL_K, "elec": (0.3, -0.2, 1.1), "elecA": 0.8, "elecE": 0.7}),
    ((16, 10, 7), O.PEAR, True, False, {"A": 1.0, "B1": 0.3, "Cc": 0.3}),  # free z, odd plane count
    ((32, 8, 6), O.GHRL, True, True, {"k": (1.48, 1.0, 2.28, 0.0), "elec": (0.0, 0.0, 1.0), "elecA": 1.0,
                                      "elecE": 0.45}),                 # anchoring faces
])
def test_measure_shapes_match_oracle(dims, pot, zfree, faces, kw):
    """E, Q and S over tile shapes of the plane-streaming measurement (row
    tiles of 1-8, plane chunks, partial waves, free z, surface faces)."""
    rng = np.random.default_rng(sum(dims) + pot)
    po, cfg, pt, ns, n = _measure_case(dims, pot, zfree, faces, rng, **kw)
    ptn = pt if pt is not None else np.ones(len(n) // 3, np.int32)
    with mclics.Lattice(cfg, n, pt, ns) as lat:
        for _ in range(2):
            lat.sweeps(2)
            d = lat.directors()
            E, S, q = lat.measure()
            assert E == pytest.approx(O.energy_mean(po, d, ptn, ns, gpu=1), rel=1e-12, abs=1e-12)
            qo = O.qtensor(d, ptn)
            np.testing.assert_allclose(q, qo, rtol=1e-11, atol=1e-13)
            Q = np.array([[qo[0], qo[1], qo[2]], [qo[1], qo[3], qo[4]], [qo[2], qo[4], qo[5]]])
            assert S == pytest.approx(np.linalg.eigvalsh(Q)[-1], abs=1e-12)


def test_graph_replay_matches_eager():
    rng = np.random.default_rng(14)
    nx, ny, nz = 16, 12, 10
    nt = nx * ny * nz
    n = unit(rng, nt)
    cfg = mclics.Config(nx, ny, nz, T=0.95, seed=21)
    with mclics.Lattice(cfg, n) as a, mclics.Lattice(cfg, n) as b:
        for rep in range(3):
            a.sweeps(5)
            b.sweeps(5, graph=True)
            a.sync()
            b.sync()
            # eager work between replays
            assert np.array_equal(a.directors(), b.directors()), rep
            assert a.measure()[0] == b.measure()[0]
            assert a.ang_var() == b.ang_var()
            assert a.acceptance() == b.acceptance()


@pytest.mark.parametrize("case", ["slab", "x4", "generic", "generic_ref8"])
def test_chunked_graphs_match_eager(case):
    # mc_sweeps_graph splits n sweeps into graphs of at most 64 sweeps (the
    # full chunk replayed, then the remainder).  Odd counts above the chunk,
    # an exact chunk and a single sweep, in a row, must leave the chain the
    # eager launches give: directors, counters, ang_var and energy.
    rng = np.random.default_rng(15)
    dims = (8, 6, 6) if case == "generic_ref8" else (16, 12, 10)
    nt = dims[0] * dims[1] * dims[2]
    n = unit(rng, nt)
    kind = {"slab": mclics.KERNEL_SLAB, "x4": mclics.KERNEL_X4, "generic": mclics.KERNEL_GENERIC,
            "generic_ref8": mclics.KERNEL_GENERIC}[case]
    order = mclics.ORDER_REF8 if case == "generic_ref8" else mclics.ORDER_CHECKER
    bound = (0, 1, 0) if case == "generic_ref8" else (1, 1, 1)
    cfg = mclics.Config(*dims, bound=bound, T=1.05, seed=23, order=order)
    with mclics.Lattice(cfg, n) as a, mclics.Lattice(cfg, n) as b:
        for lat in (a, b):
            lat.set_kernel(kind)
        for count in (165, 64, 1, 129):
            a.sweeps(count)
            b.sweeps(count, graph=True)
            a.sync()
            b.sync()
            assert np.array_equal(a.directors(), b.directors()), (case, count)
            assert a.acceptance() == b.acceptance()
            assert a.ang_var() == b.ang_var()
            assert a.measure()[0] == b.measure()[0]
        ms, launches = b.time_sweeps(131, graph=True)
        a.sweeps(131)
        assert ms > 0 and np.array_equal(a.directors(), b.directors())


def test_batch_chunked_graphs_keep_the_chain():
    # A batch's sweeps run as graphs of at most 64 sweeps; 133 sweeps in one
    # call equal 7 calls of 19 and, for each replica, the lattice run alone.
    sizes = [(16, 16, 16), (8, 8, 8), (12, 6, 10)]
    cfgs = [mclics.Config(*d, T=1.0 + 0.1 * r, seed=500 + r, stream=r) for r, d in enumerate(sizes)]
    ics = [mclics.random_ic(*d, seed=500 + r) for r, d in enumerate(sizes)]
    with mclics.Batch(cfgs, ics) as b1, mclics.Batch(cfgs, ics) as b2:
        b1.sweeps(133)
        for _ in range(7):
            b2.sweeps(19)
        for r in range(len(sizes)):
            assert np.array_equal(b1.directors(r), b2.directors(r)), r
        a1, s1, v1 = b1.state()
        a2, s2, v2 = b2.state()
        assert np.array_equal(a1, a2) and np.array_equal(s1, s2) and np.array_equal(v1, v2)
        assert np.all(s1 == 133)
        with mclics.Lattice(cfgs[2], ics[2]) as lat:
            lat.set_kernel(mclics.KERNEL_X4)
            lat.sweeps(133)
            assert np.array_equal(lat.directors(), b1.directors(2))


def test_full_size_256_properties():
    # The headline lattice: 256^3 LL, random start (seed 2).  Size-independent
    # properties: unit norms kept, counters, E inside [-9A, 0], S in [0, 1],
    # ordering at T=0.6 lowers E, measurement agrees with a host recount of Q.
    L = 256
    n = mclics.random_ic(L, L, L, seed=2)
    cfg = mclics.Config(L, L, L, T=0.6, seed=2)
    with mclics.Lattice(cfg, n) as lat:
        E0, S0, _ = lat.measure()
        lat.sweeps(20, graph=True)
        E1, S1, q = lat.measure()
        acc, sweeps = lat.acceptance()
        d = lat.directors().reshape(-1, 3)
    assert sweeps == 20 and 0 < acc < L**3
    norms = np.linalg.norm(d.astype(np.float64), axis=1)
    assert np.abs(norms - 1).max() < 2e-6
    assert -9.0 <= E1 < E0 < -2.9
    assert 0.0 <= S0 < 0.01 and S0 < S1 <= 1.0
    # Q from the host in FP64 (chunked to bound memory)
    acc6 = np.zeros(6)
    for a in range(0, len(d), 1 << 22):
        x = d[a:a + (1 << 22)].astype(np.float64)
        acc6 += [np.sum(3 * x[:, 0] ** 2 - 1), np.sum(3 * x[:, 0] * x[:, 1]), np.sum(3 * x[:, 0] * x[:, 2]),
                 np.sum(3 * x[:, 1] ** 2 - 1), np.sum(3 * x[:, 1] * x[:, 2]), np.sum(3 * x[:, 2] ** 2 - 1)]
    np.testing.assert_allclose(q, acc6 / (2 * len(d)), rtol=1e-9, atol=1e-12)


def test_create_rejects_bad_input():
    n = unit(np.random.default_rng(0), 64)
    with pytest.raises(mclics.EngineError):
        mclics.Lattice(mclics.Config(4, 4, 4), n * 2)  # non-unit directors
    with pytest.raises(mclics.EngineError):
        mclics.Lattice(mclics.Config(5, 4, 4), unit(np.random.default_rng(0), 80))  # odd periodic axis
    with mclics.Lattice(mclics.Config(4, 4, 4), n) as lat:
        with pytest.raises(mclics.EngineError):
            lat.set_T(-1.0)
        with pytest.raises(mclics.EngineError):
            lat.pass_inject(11, n, np.zeros(64, np.float32))


@pytest.mark.parametrize("kernel", ["x4", "slab"])
@pytest.mark.parametrize("pot,T,field", [(O.LL, 1.13, False), (O.LL, 0.6, True), (O.LL, 1e-4, False),
                                         (O.LL, 40.0, False), (O.GHRL, 0.5, True), (O.PEAR, 0.8, False)])
def test_segment_kernel_replay(pot, T, field, kernel):
    # The x4 segment kernel and the slab kernel (fast FP32 screen + exact
    # fallback for LL) must make exactly the CPU path's decisions: replay
    # recorded sweeps in the oracle, colour 0 then colour 1.  T = 1e-4 sends
    # nearly every move through the exact fallback, T = 40 makes most of them
    # fast accepts.
    rng = np.random.default_rng(15)
    nx, ny, nz = 16, 8, 12
    nt = nx * ny * nz
    kw = {}
    if pot == O.GHRL:
        kw["k"] = GHRL_K
    if pot == O.PEAR:
        kw.update(A=1.0, B1=0.3, Cc=0.3)
    if field:
        kw.update(elec=(0.2, 0.1, 1.0), elecA=0.9, elecE=0.8)
    po = O.params(nx, ny, nz, potential=pot, **kw)
    cfg = mclics.Config(nx, ny, nz, potential=pot, A=po.A, B1=po.B1, C=po.C, ghrl=O.ghrl_tuple(po),
                        elec=(po.elecX, po.elecY, po.elecZ), elecA=po.elecA, elecE=po.elecE,
                        T=float(np.float32(T)), seed=31)
    n = unit(rng, nt)
    ref = n.copy()
    pt = np.ones(nt, np.int32)
    kind = mclics.KERNEL_X4 if kernel == "x4" else mclics.KERNEL_SLAB
    with mclics.Lattice(cfg, n) as lat:
        lat.set_kernel(kind)
        assert lat.kernel() == kind
        if kernel == "slab":
            lat.set_slab_tile(4, 3)  # 2 row tiles x 4 plane tiles: halos cross tiles on y and z
        for w in range(3):
            av = lat.ang_var()
            pre = ref.copy()
            rec = lat.record_sweep()
            total = 0
            for colour in (0, 1):
                cnt, racc, reo, ren = O.pass_inject(po, ref, pt, None, 8 + colour, np.float32(T), rec["prop"],
                                                    rec["u2"])
                sel = in_set(nx, ny, nz, 8 + colour)
                assert np.array_equal(rec["accept"][sel], racc[sel]), (w, colour)
                assert np.array_equal(rec["e_old"][sel], reo[sel]), (w, colour)
                assert np.array_equal(rec["e_new"][sel], ren[sel]), (w, colour)
                total += cnt
            assert np.array_equal(lat.directors(), ref), w
            assert lat.acceptance() == (total, w + 1)
            assert lat.ang_var() == O.lib().orc_adapt_ang_var(av, total, nt)
    # the proposals are normalised rotations by pi*va about the chosen axis
    u0, u1 = rec["u0"].astype(np.float64), rec["u1"].astype(np.float64)
    assert ((u1 >= 0) & (u1 < 1)).all() and ((u0 >= 0) & (u0 < 1)).all()
    old = pre.reshape(-1, 3).astype(np.float64)
    a = np.pi * (2 * u0 - 1) * np.float64(np.float32(av))
    c, s_ = np.cos(a), np.sin(a)
    x, y, z = old[:, 0], old[:, 1], old[:, 2]
    want = np.where((u1 < 0.333)[:, None], np.stack([x, y * c + z * s_, z * c - y * s_], 1),
                    np.where((u1 < 0.666)[:, None], np.stack([x * c - z * s_, y, z * c + x * s_], 1),
                             np.stack([x * c + y * s_, y * c - x * s_, z], 1)))
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.abs(rec["prop"].reshape(-1, 3) - want).max() < 5e-6
    np.testing.assert_allclose(np.linalg.norm(rec["prop"].reshape(-1, 3), axis=1), 1.0, atol=3e-7)


@pytest.mark.parametrize("pot,T,geo", [(O.LL, 1.13, False), (O.LL, 1e-4, False), (O.LL, 40.0, False),
                                       (O.GHRL, 0.5, True), (O.PEAR, 0.8, False), (O.LL, 0.9, True)])
def test_slab_fixed_shape_replay(pot, T, geo):
    """Rows of 256 sites in tiles of 8 rows take the slab kernel's
    compile-time instantiation (a row is one wave).  Its recorded sweeps must
    replay exactly in the oracle, on a periodic lattice and on the slab
    geometry (free z, homeotropic GHRL faces, field), with the tiles meeting
    inside the lattice on y and z."""
    rng = np.random.default_rng(int(T * 1000) + 7 * pot + geo)
    nx, ny, nz = 256, 16, 10
    nt = nx * ny * nz
    kw = {"k": GHRL_K} if pot == O.GHRL else ({"A": 1.0, "B1": 0.3, "Cc": 0.3} if pot == O.PEAR else {})
    surf, pt, ns, bound = [], np.ones(nt, np.int32), None, (1, 1, 1)
    if geo:
        bound = (1, 1, 0)
        kw.update(elec=(0.0, 0.0, 1.0), elecA=1.0, elecE=0.5)
        surf = [("homeotropic_ghrl" if pot == O.GHRL else "homeotropic", 5.0, (0, 0, 1)), ("rp", 2.0, (1, 0, 0))]
        ns = np.zeros(3 * nt, np.float32)
        for j in range(ny):
            for i in range(nx):
                s0, s1 = i + nx * j, i + nx * (j + ny * (nz - 1))
                pt[s0], pt[s1] = 2, 3
                ns[3 * s0 + 2], ns[3 * s1 + 2] = -1.0, 1.0
    po = O.params(nx, ny, nz, bound=bound, potential=pot, surfaces=surf, **kw)
    cfg = mclics.Config(nx, ny, nz, bound=bound, potential=pot, A=po.A, B1=po.B1, C=po.C, ghrl=O.ghrl_tuple(po),
                        elec=(po.elecX, po.elecY, po.elecZ), elecA=po.elecA, elecE=po.elecE,
                        surfaces=[mclics.Surface(k, w, d) for k, w, d in surf], T=float(np.float32(T)), seed=77)
    n = unit(rng, nt)
    ref = n.copy()
    with mclics.Lattice(cfg, n, pt, ns) as lat:
        lat.set_kernel(mclics.KERNEL_SLAB)
        lat.set_slab_tile(8, 5)  # 2 row tiles x 2 plane tiles
        for w in range(3):
            av = lat.ang_var()
            rec = lat.record_sweep()
            total = 0
            for colour in (0, 1):
                cnt, racc, reo, ren = O.pass_inject(po, ref, pt, ns, 8 + colour, np.float32(T), rec["prop"],
                                                    rec["u2"])
                sel = in_set(nx, ny, nz, 8 + colour)
                assert np.array_equal(rec["accept"][sel], racc[sel]), (w, colour)
                assert np.array_equal(rec["e_old"][sel], reo[sel]), (w, colour)
                assert np.array_equal(rec["e_new"][sel], ren[sel]), (w, colour)
                total += cnt
            assert np.array_equal(lat.directors(), ref), w
            assert lat.acceptance() == (total, w + 1)
            assert lat.ang_var() == O.lib().orc_adapt_ang_var(av, total, nt)
        # the production instantiation of the same shape continues the same chain as the segment kernel
        with mclics.Lattice(cfg, ref, pt, ns) as seg:
            seg.set_kernel(mclics.KERNEL_X4)
            seg.set_sweep_index(3)
            seg.set_ang_var(lat.ang_var())
            lat.sweeps(2)
            seg.sweeps(2)
            assert np.array_equal(lat.directors(), seg.directors())
            assert lat.acceptance() == seg.acceptance()


@pytest.mark.parametrize("dims,pot,zfree,faces,kw", [
    ((32, 16, 16), O.LL, False, False, {"A": 1.0}),
    ((32, 16, 12), O.GHRL, True, True, {"k": GHRL_K, "elec": (0.0, 0.0, 1.0), "elecA": 1.0, "elecE": 0.5}),
    ((32, 16, 16), O.PEAR, False, False, {"A": 1.0, "B1": 0.3, "Cc": 0.3}),
])
def test_segment_kernel_pipelined_grid_keeps_the_chain(dims, pot, zfree, faces, kw):
    """With fewer workgroups than tiles the segment kernel runs its
    pipelined loop (the next tile's loads in flight during an update); the
    chain must equal the one-tile-per-workgroup kernel's, which the replay
    tests pin to the oracle."""
    rng = np.random.default_rng(sum(dims) + 3 * pot)
    po, cfg, pt, ns, n = _measure_case(dims, pot, zfree, faces, rng, **kw)
    with mclics.Lattice(cfg, n, pt, ns) as a, mclics.Lattice(cfg, n, pt, ns) as b:
        a.set_kernel(mclics.KERNEL_X4)
        b.set_kernel(mclics.KERNEL_X4)
        b.set_tuning(3, 0)  # 8 tiles over 3 workgroups: 3, 3 and 2 tiles each
        for T in (0.9, 0.3, 2.0):
            a.set_T(T)
            b.set_T(T)
            a.sweeps(3)
            b.sweeps(3)
            assert np.array_equal(a.directors(), b.directors()), T
            assert a.acceptance() == b.acceptance() and a.ang_var() == b.ang_var(), T


SLAB_GEOMETRY = ["slab_" + k for k in ("homeotropic", "rp", "fg", "strong", "homeotropic_ghrl", "rp_ghrl",
                                        "fg_ghrl", "strong_ghrl", "chiral")] + ["ll_slab", "ll_freez"]


@pytest.mark.parametrize("kernel", ["x4", "slab"])
@pytest.mark.parametrize("case", SLAB_GEOMETRY)
def test_fast_kernels_slab_geometry_replay(case, kernel):
    # x, y periodic, z free, surfaces on the two faces (or none): the segment
    # and slab kernels drop the missing z bond on each face and add the
    # surface term there; recorded sweeps must replay exactly in the oracle.
    rng = np.random.default_rng(zlib.crc32(case.encode()) + 7)
    po, cfg, pt, ns = make_case(case, rng)
    nx, ny, nz = po.nx, po.ny, po.nz
    nt = nx * ny * nz
    n = unit(rng, nt)
    ref = n.copy()
    tiles = {(8, 6, 6): (3, 2), (16, 8, 10): (4, 5), (16, 8, 9): (4, 3)}
    kind = mclics.KERNEL_X4 if kernel == "x4" else mclics.KERNEL_SLAB
    with mclics.Lattice(cfg, n, pt, ns) as lat:
        # AUTO takes a fast kernel for this geometry: the slab for LL, the segment kernel otherwise
        assert lat.kernel() == (mclics.KERNEL_SLAB if po.potential == O.LL else mclics.KERNEL_X4)
        lat.set_kernel(kind)
        if kernel == "slab":
            lat.set_slab_tile(*tiles[(nx, ny, nz)])  # tiles meet inside the lattice on y and z
        for w, T in enumerate((0.9, 0.05, 2.5)):
            lat.set_T(float(np.float32(T)))
            av = lat.ang_var()
            rec = lat.record_sweep()
            total = 0
            for colour in (0, 1):
                cnt, racc, reo, ren = O.pass_inject(po, ref, pt, ns, 8 + colour, np.float32(T), rec["prop"],
                                                    rec["u2"])
                sel = in_set(nx, ny, nz, 8 + colour)
                assert np.array_equal(rec["accept"][sel], racc[sel]), (case, w, colour)
                assert np.array_equal(rec["e_old"][sel], reo[sel]), (case, w, colour)
                assert np.array_equal(rec["e_new"][sel], ren[sel]), (case, w, colour)
                total += cnt
            assert np.array_equal(lat.directors(), ref), (case, w)
            assert lat.acceptance() == (total, w + 1)
            assert lat.ang_var() == O.lib().orc_adapt_ang_var(av, total, nt)
        E, S, q = lat.measure()
        assert E == pytest.approx(O.energy_mean(po, ref, pt, ns, gpu=1), rel=1e-12, abs=1e-12)
        np.testing.assert_allclose(q, O.qtensor(ref, pt), rtol=1e-11, atol=1e-13)


def test_segment_kernel_selection_rules():
    rng = np.random.default_rng(16)
    n = unit(rng, 6 * 4 * 4)
    with mclics.Lattice(mclics.Config(6, 4, 4), n) as lat:  # Nx % 4 != 0
        assert lat.kernel() == mclics.KERNEL_GENERIC
        with pytest.raises(mclics.EngineError):
            lat.set_kernel(mclics.KERNEL_X4)
        with pytest.raises(mclics.EngineError):
            lat.set_kernel(mclics.KERNEL_SLAB)
    n = unit(rng, 8 * 4 * 4)
    with mclics.Lattice(mclics.Config(8, 4, 4), n) as lat:
        assert lat.kernel() == mclics.KERNEL_SLAB
        lat.set_kernel(mclics.KERNEL_X4)
        assert lat.kernel() == mclics.KERNEL_X4
        lat.set_kernel(mclics.KERNEL_GENERIC)
        assert lat.kernel() == mclics.KERNEL_GENERIC
        lat.set_kernel(mclics.KERNEL_AUTO)
        assert lat.kernel() == mclics.KERNEL_SLAB
        for bad in ((3, 4), (4, 3), (0, 4), (64, 4)):
            with pytest.raises(mclics.EngineError):
                lat.set_slab_tile(*bad)
    with mclics.Lattice(mclics.Config(8, 4, 4, order=mclics.ORDER_REF8), n) as lat:
        assert lat.kernel() == mclics.KERNEL_GENERIC


def twin_run(cfg, n, sweeps, shape=None, graph=False):
    """(directors, acceptance, ang_var, E) after `sweeps` sweeps with the
    segment kernel and with the slab kernel (tile shape `shape`)."""
    out = []
    for kind in (mclics.KERNEL_X4, mclics.KERNEL_SLAB):
        with mclics.Lattice(cfg, n) as lat:
            lat.set_kernel(kind)
            if kind == mclics.KERNEL_SLAB and shape is not None:
                lat.set_slab_tile(*shape)
            lat.sweeps(sweeps, graph=graph)
            lat.sync()
            out.append((lat.directors(), lat.acceptance(), lat.ang_var(), lat.measure()[0]))
    return out


@pytest.mark.parametrize("dims,shape", [((16, 8, 12), (8, 12, 0)), ((16, 8, 12), (2, 2, 1)),
                                        ((16, 8, 12), (1, 4, 0)), ((8, 2, 2), (2, 2, 1)),
                                        ((4, 6, 10), (6, 5, 0)), ((32, 16, 16), (4, 4, 1)),
                                        ((12, 10, 8), (10, 8, 1)), ((64, 16, 32), None)])
def test_slab_matches_segment_kernel(dims, shape):
    # The slab kernel recomputes the red updates of its halo rows and planes;
    # with any tile shape (one tile, tiles of one row or plane pair, tiles
    # that wrap onto themselves) the chain must equal the segment kernel's
    # bit for bit: state, acceptance counts, ang_var and energy.
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 10000 + ny * 100 + nz)
    n = unit(rng, nx * ny * nz)
    for T in (0.7, 1.15):
        cfg = mclics.Config(nx, ny, nz, T=T, seed=40 + nx)
        a, b = twin_run(cfg, n, 7, shape)
        assert np.array_equal(a[0], b[0]), (dims, shape, T)
        assert a[1:] == b[1:]


def test_slab_matches_segment_kernel_ghrl_field():
    nx, ny, nz = 16, 12, 8
    n = unit(np.random.default_rng(17), nx * ny * nz)
    po = O.params(nx, ny, nz, potential=O.GHRL, k=GHRL_K, elec=(0.2, 0.1, 1.0), elecA=0.9, elecE=0.8)
    cfg = mclics.Config(nx, ny, nz, potential=O.GHRL, ghrl=O.ghrl_tuple(po), elec=(0.2, 0.1, 1.0), elecA=0.9,
                        elecE=0.8, T=0.5, seed=3)
    a, b = twin_run(cfg, n, 5, (4, 2, 1))
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


def test_slab_full_size_256_matches_segment_kernel():
    # The headline lattice with the default (device-chosen) tile, an odd
    # number of sweeps through a graph (the plane sets end swapped).
    L = 256
    n = mclics.random_ic(L, L, L, seed=2)
    a, b = twin_run(mclics.Config(L, L, L, T=1.13, seed=2), n, 3, graph=True)
    assert a[1:] == b[1:]
    assert np.array_equal(a[0], b[0])


def test_slab_autotune_leaves_chain_and_keeps_parity():
    nx, ny, nz = 32, 16, 16
    n = unit(np.random.default_rng(18), nx * ny * nz)
    cfg = mclics.Config(nx, ny, nz, T=0.9, seed=8)
    with mclics.Lattice(cfg, n) as a, mclics.Lattice(cfg, n) as b:
        assert a.kernel() == mclics.KERNEL_SLAB
        a.sweeps(3)
        b.sweeps(3)
        a.set_kernel(mclics.KERNEL_SLAB)  # an explicit kernel: only its shapes are tuned
        grid, xcd, us = a.autotune()
        ty, cz, x = a.slab_tile()
        assert grid == (ny // ty) * (nz // cz) and x == xcd and us > 0
        a.sweeps(4, graph=True)
        b.sweeps(4)
        a.sync()
        b.sync()
        assert np.array_equal(a.directors(), b.directors())
        assert a.acceptance() == b.acceptance() and a.ang_var() == b.ang_var()
        # AUTO: both fast kernels are timed and the faster one kept; the chain goes on unchanged
        a.set_kernel(mclics.KERNEL_AUTO)
        grid, xcd, us = a.autotune()
        assert a.kernel() in (mclics.KERNEL_SLAB, mclics.KERNEL_X4) and grid > 0 and us > 0
        a.sweeps(5, graph=True)
        b.sweeps(5)
        a.sync()
        b.sync()
        assert np.array_equal(a.directors(), b.directors())
        assert a.acceptance() == b.acceptance() and a.ang_var() == b.ang_var()


def test_slab_measure_series_reads_the_current_planes():
    # mc_measure_series interleaves sweeps and measurements; with the slab
    # kernel the lattice alternates between two plane sets.
    nx, ny, nz = 16, 8, 8
    n = unit(np.random.default_rng(19), nx * ny * nz)
    cfg = mclics.Config(nx, ny, nz, T=0.8, seed=9)
    with mclics.Lattice(cfg, n) as a, mclics.Lattice(cfg, n) as b:
        E, S = a.measure_series(5)
        for w in range(5):
            b.sweeps(1)
            e, s_, _ = b.measure()
            assert E[w] == e and S[w] == s_, w


@pytest.mark.parametrize("kernel", ["slab", "x4", "generic"])
def test_measure_series_every_equals_sweeps_then_measure(kernel):
    # mc_measure_series_every(n, k): k sweeps before each of n measurements,
    # odd k so the slab kernel's plane set alternates between samples.
    nx, ny, nz = 16, 8, 8
    n = unit(np.random.default_rng(23), nx * ny * nz)
    cfg = mclics.Config(nx, ny, nz, T=1.05, seed=13)
    kind = {"slab": mclics.KERNEL_SLAB, "x4": mclics.KERNEL_X4, "generic": mclics.KERNEL_GENERIC}[kernel]
    with mclics.Lattice(cfg, n) as a, mclics.Lattice(cfg, n) as b:
        a.set_kernel(kind)
        b.set_kernel(kind)
        E, S = a.measure_series(6, every=3)
        for w in range(6):
            b.sweeps(3)
            e, s_, _ = b.measure()
            assert E[w] == e and S[w] == s_, (kernel, w)
        assert a.acceptance() == b.acceptance() and a.ang_var() == b.ang_var()
        assert np.array_equal(a.directors(), b.directors())


@pytest.mark.parametrize("kernel", ["slab", "generic", "ref8"])
def test_checkpoint_resumes_the_chain_exactly(kernel):
    # 12 sweeps in one go equal 5 sweeps, a checkpoint (directors, ang_var,
    # sweep index), a fresh handle restored from it and 7 more sweeps.
    nx, ny, nz = 16, 8, 12
    n0 = unit(np.random.default_rng(23), nx * ny * nz)
    order = mclics.ORDER_REF8 if kernel == "ref8" else mclics.ORDER_CHECKER
    cfg = mclics.Config(nx, ny, nz, T=1.05, seed=77, order=order)

    def make(n):
        lat = mclics.Lattice(cfg, n)
        if kernel == "generic":
            lat.set_kernel(mclics.KERNEL_GENERIC)
        return lat

    with make(n0) as a:
        a.sweeps(12, graph=True)
        want = a.checkpoint()
    with make(n0) as b:
        b.sweeps(5)
        d, av, sw = b.checkpoint()
    assert sw == 5
    with make(d) as c:
        c.set_ang_var(av)
        c.set_sweep_index(sw)
        c.sweeps(7, graph=True)
        got = c.checkpoint()
    assert np.array_equal(got[0], want[0])
    assert got[1:] == want[1:]
