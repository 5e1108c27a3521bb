"""The device generator and the production sweep instantiations.

* Known answers of the device Philox4x32-10: the uniforms a recorded sweep
  reports must be the oracle's Philox (itself pinned by Random123/rocRAND
  known-answer vectors) on the documented counters: (site, sweep, stream,
  0x4D43) for the generic kernel, (first site of the pair, sweep, stream,
  0x4D44) for the segment and slab kernels, keyed by the 64-bit seed.  This
  replaces cuRAND's per-thread XORWOW (src/evolve.cu:25-31,143,145,193).
* A production sweep (the kernels without recording) leaves the lattice, the
  accepted count and ang_var exactly as a recorded sweep from the same state,
  whose every decision the parity tests replay in the oracle.
"""
import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu

SEED = 0x1234_5678_9ABC_DEF1
STREAM = 5


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


def u01(w):
    return np.float32((w >> 8) * (1.0 / 16777216.0))


def colour_of(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i + j + k) % 2).reshape(-1)


@pytest.mark.parametrize("kernel", [mclics.KERNEL_GENERIC, mclics.KERNEL_X4, mclics.KERNEL_SLAB])
def test_device_philox_known_answers(kernel):
    nx, ny, nz = 16, 8, 6
    nt = nx * ny * nz
    cfg = mclics.Config(nx, ny, nz, T=1.1, seed=SEED, stream=STREAM)
    key = [SEED & 0xFFFFFFFF, SEED >> 32]
    rng = np.random.default_rng(3)
    with mclics.Lattice(cfg, unit(rng, nt)) as lat:
        lat.set_kernel(kernel)
        for sweep in (0, 7, 0xFFFFFFF0):
            lat.set_sweep_index(sweep)
            rec = lat.record_sweep()
            for s in range(nt):
                if kernel == mclics.KERNEL_GENERIC:
                    w = O.philox([s, sweep, STREAM, 0x4D43], key)
                    want = (u01(w[0]), u01(w[1]), u01(w[2]))
                else:
                    e = s % 4
                    t = 1 if e >= 2 else 0
                    w = O.philox([s - 2 * t, sweep, STREAM, 0x4D44], key)
                    ang, acc = (w[2], w[3]) if t else (w[0], w[1])
                    ax = (((w[2] if t else w[0]) & 0xFF) << 8) | ((w[3] if t else w[1]) & 0xFF)
                    want = (u01(ang), np.float32(ax / 65536.0), u01(acc))
                got = (rec["u0"][s], rec["u1"][s], rec["u2"][s])
                assert got == want, (kernel, sweep, s, got, want)


def slab_case(nx, ny, nz):
    nt = nx * ny * nz
    pt = np.ones(nt, np.int32)
    ns = np.zeros(3 * nt, np.float32)
    for j in range(ny):
        for i in range(nx):
            s0, s1 = i + nx * j, i + nx * (j + ny * (nz - 1))
            pt[s0], pt[s1] = 2, 3
            ns[3 * s0 + 2], ns[3 * s1 + 2] = -1.0, 1.0
    po = O.params(nx, ny, nz, bound=(1, 1, 0), potential=O.GHRL, k=(1.0, 0.7, 1.3, 16.0), elec=(0.0, 0.0, 1.0),
                  elecA=1.0, elecE=0.5, surfaces=[("homeotropic_ghrl", 5.0, (0, 0, 1)), ("homeotropic_ghrl", 5.0, (0, 0, 1))])
    cfg = mclics.Config(nx, ny, nz, bound=(1, 1, 0), potential=mclics.GHRL, ghrl=O.ghrl_tuple(po),
                        elec=(0.0, 0.0, 1.0), elecA=1.0, elecE=0.5, T=0.5,
                        surfaces=[mclics.Surface("homeotropic_ghrl", 5.0, (0, 0, 1))] * 2, seed=21)
    return cfg, pt, ns


@pytest.mark.parametrize("what", ["ll", "ghrl_slab", "pear"])
@pytest.mark.parametrize("kernel", [mclics.KERNEL_GENERIC, mclics.KERNEL_X4, mclics.KERNEL_SLAB])
def test_production_sweep_equals_recorded_sweep(what, kernel):
    rng = np.random.default_rng(17)
    if what == "ll":
        nx, ny, nz = 16, 12, 10
        cfg, pt, ns = mclics.Config(nx, ny, nz, T=1.05, seed=4), None, None
    elif what == "pear":
        nx, ny, nz = 12, 8, 8
        cfg, pt, ns = mclics.Config(nx, ny, nz, potential=mclics.PEAR, A=1.0, B1=0.3, C=0.3, T=0.8, seed=5), None, None
    else:
        nx, ny, nz = 12, 8, 10
        cfg, pt, ns = slab_case(nx, ny, nz)
    nt = nx * ny * nz
    n0 = unit(rng, nt)
    with mclics.Lattice(cfg, n0, pt, ns) as a, mclics.Lattice(cfg, n0, pt, ns) as b:
        a.set_kernel(kernel)
        b.set_kernel(kernel)
        for w in range(6):
            rec = a.record_sweep()
            if w % 2:
                b.sweeps(1)
            else:
                b.sweeps(1, graph=True)
            assert np.array_equal(a.directors(), b.directors()), (what, kernel, w)
            assert a.acceptance() == b.acceptance(), (what, kernel, w)
            assert a.ang_var() == b.ang_var(), (what, kernel, w)
            assert int(rec["accept"].sum()) == a.acceptance()[0]
