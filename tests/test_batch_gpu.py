"""Batched replicas on the GPU: every replica of a batch must follow exactly
the chain the same lattice follows alone (segment kernel, same seed and
stream), and the per-replica measurements must equal the single-lattice ones.
The single-lattice path is itself pinned to the CPU oracle elsewhere."""
import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu


def test_batch_replicas_match_single_lattices():
    sizes = [(8, 8, 8), (12, 6, 10), (16, 16, 16), (8, 4, 6), (20, 8, 8)]
    rng = np.random.default_rng(3)
    cfgs, ics = [], []
    for r, (x, y, z) in enumerate(sizes):
        T = float(np.float32(rng.uniform(0.9, 1.3)))
        cfgs.append(mclics.Config(x, y, z, T=T, seed=100 + r, stream=r))
        ics.append(mclics.random_ic(x, y, z, seed=100 + r))
    with mclics.Batch(cfgs, ics) as b:
        b.sweeps(7)
        b.sweeps(5)
        b.sync()
        E, S = b.measure()
        acc, sw, av = b.state()
        for r, cfg in enumerate(cfgs):
            with mclics.Lattice(cfg, ics[r]) as lat:
                # AUTO picks the slab kernel here; it draws the same numbers and
                # evolves the lattice exactly as the segment kernel does
                assert lat.kernel() in (mclics.KERNEL_SLAB, mclics.KERNEL_X4)
                lat.sweeps(12)
                d = lat.directors()
                E1, S1, _ = lat.measure()
                a1, s1 = lat.acceptance()
                av1 = lat.ang_var()
            assert np.array_equal(b.directors(r), d), r
            assert E[r] == pytest.approx(E1, rel=1e-12) and S[r] == pytest.approx(S1, abs=1e-12)
            assert (acc[r], sw[r], av[r]) == (a1, s1, av1), r


def test_batch_temperature_control_and_series():
    cfgs = [mclics.Config(8, 8, 8, T=0.3, seed=5 + r, stream=r) for r in range(3)]
    ics = [mclics.random_ic(8, 8, 8, seed=5 + r) for r in range(3)]
    with mclics.Batch(cfgs, ics) as b:
        b.set_T(1, 50.0)  # replica 1 hot: stays disordered
        E, S = b.measure_series(40)
        assert E.shape == (40, 3)
        assert S[-1, 1] < S[-1, 0] and S[-1, 1] < S[-1, 2]
        assert E[-1, 0] < E[0, 0]
        g, x, us = b.autotune()
        assert g > 0 and us > 0
        # tuning changes the launch shape, not the chain: rerun from the same start
    with mclics.Batch(cfgs, ics) as b1, mclics.Batch(cfgs, ics) as b2:
        b2.set_tuning(mclics.device_count() and 256 * 4, 1)
        b1.sweeps(6)
        b2.sweeps(6)
        for r in range(3):
            assert np.array_equal(b1.directors(r), b2.directors(r))


def test_batch_slab_and_segment_kernels_agree():
    """The batched slab kernel (one launch per sweep, planes swapped after
    each) and the two-pass segment kernel give every replica the same chain,
    across graph replays of odd and even lengths and measured sweeps."""
    sizes = [(32, 32, 32), (8, 8, 8), (12, 6, 10), (16, 16, 16), (20, 8, 8), (24, 12, 16)]
    rng = np.random.default_rng(9)
    cfgs, ics = [], []
    for r, (x, y, z) in enumerate(sizes):
        T = float(np.float32(rng.uniform(0.9, 1.3)))
        cfgs.append(mclics.Config(x, y, z, T=T, seed=300 + r, stream=r))
        ics.append(mclics.random_ic(x, y, z, seed=300 + r))
    with mclics.Batch(cfgs, ics) as b1, mclics.Batch(cfgs, ics) as b2:
        assert b1.kernel() == mclics.KERNEL_SLAB
        b2.set_kernel(mclics.KERNEL_X4)
        assert b2.kernel() == mclics.KERNEL_X4
        for b in (b1, b2):
            b.sweeps(3)
            b.sweeps(4)
            b.sweeps(3)
        E1, S1 = b1.measure_series(5)
        E2, S2 = b2.measure_series(5)
        # the slab batch measures on its tiles, the segment batch per segment
        # tile: the same sums in another order
        np.testing.assert_allclose(E1, E2, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(S1, S2, rtol=0, atol=1e-12)
        for r in range(len(sizes)):
            assert np.array_equal(b1.directors(r), b2.directors(r)), r
        a1, s1, v1 = b1.state()
        a2, s2, v2 = b2.state()
        assert np.array_equal(a1, a2) and np.array_equal(s1, s2) and np.array_equal(v1, v2)
        assert np.all(s1 == 15)


def test_batch_autotuned_shape_keeps_the_chain():
    """Whatever shape the batch autotuner keeps (segment grid, slab budget,
    planes per tile, parity-aligned tile heights), the replicas follow the
    same chains as with the default slab shape."""
    sizes = [(96, 24, 8), (64, 16, 8), (48, 12, 8), (32, 32, 4), (8, 8, 8)]
    rng = np.random.default_rng(17)
    cfgs, ics = [], []
    for r, (x, y, z) in enumerate(sizes):
        cfgs.append(mclics.Config(x, y, z, T=float(np.float32(rng.uniform(0.9, 1.3))), seed=500 + r, stream=r))
        ics.append(mclics.random_ic(x, y, z, seed=500 + r))
    with mclics.Batch(cfgs, ics) as b1, mclics.Batch(cfgs, ics) as b2:
        g, x, us = b2.autotune()
        assert g > 0 and us > 0
        for b in (b1, b2):
            b.sweeps(5)
            b.sweeps(4)
        for r in range(len(sizes)):
            assert np.array_equal(b1.directors(r), b2.directors(r)), r
        assert all(np.array_equal(u, v) for u, v in zip(b1.state(), b2.state()))
        E1, S1 = b1.measure()
        E2, S2 = b2.measure()
        assert np.array_equal(E1, E2) and np.array_equal(S1, S2)


@pytest.mark.parametrize("pot", [O.GHRL, O.PEAR])
def test_batch_other_potentials_match_single_lattices(pot):
    """GHRL (chiral, with a field) and pear replicas in a batch, on the slab
    kernel and on the segment kernel, against the same lattices alone."""
    sizes = [(8, 8, 8), (12, 6, 10), (16, 8, 12)]
    kw = {"k": (1.0, 0.7, 1.3, 16.0)} if pot == O.GHRL else {"A": 1.1, "B1": 0.05, "Cc": 0.3}
    po = O.params(8, 8, 8, potential=pot, elec=(0.2, -0.1, 1.0), elecA=0.7, elecE=0.6, **kw)
    cfgs, ics = [], []
    for r, (x, y, z) in enumerate(sizes):
        cfgs.append(mclics.Config(x, y, z, potential=pot, A=po.A, B1=po.B1, C=po.C, ghrl=O.ghrl_tuple(po),
                                  elec=(0.2, -0.1, 1.0), elecA=0.7, elecE=0.6, T=0.8 + 0.1 * r, seed=40 + r,
                                  stream=r))
        ics.append(mclics.random_ic(x, y, z, seed=40 + r))
    singles = []
    for r, cfg in enumerate(cfgs):
        with mclics.Lattice(cfg, ics[r]) as lat:
            lat.sweeps(9)
            singles.append((lat.directors(), lat.acceptance(), lat.ang_var()))
    for kind in (mclics.KERNEL_SLAB, mclics.KERNEL_X4):
        with mclics.Batch(cfgs, ics) as b:
            b.set_kernel(kind)
            b.sweeps(4)
            b.sweeps(5)
            acc, sw, av = b.state()
            E, S = b.measure()
            for r in range(len(sizes)):
                d, (a1, s1), av1 = singles[r]
                assert np.array_equal(b.directors(r), d), (kind, r)
                assert (acc[r], sw[r], av[r]) == (a1, s1, av1), (kind, r)
                # per-replica E and S against the oracle on the replica's lattice
                x, y, z = sizes[r]
                pr = O.params(x, y, z, potential=pot, elec=(0.2, -0.1, 1.0), elecA=0.7, elecE=0.6, **kw)
                pt = np.ones(x * y * z, np.int32)
                assert E[r] == pytest.approx(O.energy_mean(pr, d, pt, None, gpu=1), rel=1e-12, abs=1e-12), (kind, r)
                qo = O.qtensor(d, pt)
                Q = np.array([[qo[0], qo[1], qo[2]], [qo[1], qo[3], qo[4]], [qo[2], qo[4], qo[5]]])
                assert S[r] == pytest.approx(np.linalg.eigvalsh(Q)[-1], abs=1e-12), (kind, r)


def test_batch_rejects_unsupported_replicas():
    with pytest.raises(mclics.EngineError):
        mclics.Batch([mclics.Config(6, 4, 4)], [mclics.random_ic(6, 4, 4, seed=1)])  # Nx % 4
    with pytest.raises(mclics.EngineError):
        mclics.Batch([mclics.Config(8, 4, 4, bound=(1, 1, 0))], [mclics.random_ic(8, 4, 4, seed=1)])


def test_batch_persistent_launch_keeps_the_chain():
    """MC_KERNEL_SLAB_PERSISTENT (up to 64 sweeps per launch, tiles taken
    from a per-sweep counter, a grid barrier between sweeps) gives every
    replica the chain of the one-launch-per-sweep slab kernel: runs of odd
    and even length, one across the 64-sweep chunk, and measured sweeps."""
    sizes = [(32, 32, 32), (8, 8, 8), (12, 6, 10), (16, 16, 16), (20, 8, 8), (24, 12, 16)]
    rng = np.random.default_rng(29)
    cfgs, ics = [], []
    for r, (x, y, z) in enumerate(sizes):
        T = float(np.float32(rng.uniform(0.9, 1.3)))
        cfgs.append(mclics.Config(x, y, z, T=T, seed=600 + r, stream=r))
        ics.append(mclics.random_ic(x, y, z, seed=600 + r))
    with mclics.Batch(cfgs, ics) as b1, mclics.Batch(cfgs, ics) as b2:
        b1.set_kernel(mclics.KERNEL_SLAB)
        b2.set_kernel(mclics.KERNEL_SLAB_PERSISTENT)
        assert b2.kernel() == mclics.KERNEL_SLAB_PERSISTENT
        for n in (3, 4, 1, 67):
            print(f"persistent batch: {n} sweeps", flush=True)
            b1.sweeps(n)
            b2.sweeps(n)
        E1, S1 = b1.measure_series(5)
        E2, S2 = b2.measure_series(5)
        assert np.array_equal(E1, E2) and np.array_equal(S1, S2)
        for r in range(len(sizes)):
            assert np.array_equal(b1.directors(r), b2.directors(r)), r
        a1, s1, v1 = b1.state()
        a2, s2, v2 = b2.state()
        assert np.array_equal(a1, a2) and np.array_equal(s1, s2) and np.array_equal(v1, v2)
        assert np.all(s2 == 80)
        ms = b2.time_sweeps(20)
        assert ms > 0


def test_batch_persistent_barrier_timeout_is_reported():
    """A grid barrier that gives up (forced from the host with a one-poll
    limit) ends the launch and is reported as an error by the next
    synchronising call; the kernel's waits are bounded, so nothing hangs."""
    sizes = [(16, 16, 16), (32, 32, 32)]
    cfgs = [mclics.Config(*d, T=1.1, seed=700 + r, stream=r) for r, d in enumerate(sizes)]
    ics = [mclics.random_ic(*d, seed=700 + r) for r, d in enumerate(sizes)]
    with mclics.Batch(cfgs, ics) as b:
        b.set_kernel(mclics.KERNEL_SLAB_PERSISTENT)
        b.set_barrier_limit(1)
        b.sweeps(8)
        with pytest.raises(mclics.EngineError) as ei:
            b.sync()
        assert ei.value.code == -2 and "grid barrier" in str(ei.value)
        b.sync()  # the fault is reported once
