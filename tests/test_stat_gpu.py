"""S(T) and E(T) of the GPU production chains against the reference CPU
chain, within 3 sigma of replica scatter.

The GPU chains (checkerboard order, pair-keyed or site-keyed Philox,
hardware sin/cos proposals renormalised) and the oracle's restatement of the
reference CPU sweep (src/evolve.cpp:64-179: 2x2x2 sublattices in the CPU
order, the CPU proposal without renormalisation, glibc expf) are different
Markov chains with the same stationary distribution.  For each case R = 8
independent replicas of each chain (distinct seeds) are equilibrated and
then averaged; the means must satisfy
    |mean_A - mean_B| <= 3 sqrt(sigma_A^2 + sigma_B^2),  sigma = std(X_r) / sqrt(R)
for the mean site energy E and the order parameter S (the rule of the
measurement plan).  Every chain is deterministic for its seed, so the test
gives the same verdict on every run.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu

R = 8
N_EQ = 2000
N_MEAS = 4000
ORC_EVERY = 4  # the oracle measures every 4th sweep (same interval, fewer samples)


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


def start(kind, nt, seed):
    if kind == "random":
        return unit(np.random.default_rng(seed), nt)
    axis = {"x": 0, "z": 2}[kind]
    n = np.zeros((nt, 3), np.float32)
    n[:, axis] = 1.0
    return n.reshape(-1)


def slab_layout(nx, ny, nz):
    nt = nx * ny * nz
    pt = np.ones(nt, np.int32)
    ns = np.zeros(3 * nt, np.float32)
    for j in range(ny):
        for i in range(nx):
            s0, s1 = i + nx * j, i + nx * (j + ny * (nz - 1))
            pt[s0], pt[s1] = 2, 3
            ns[3 * s0 + 2], ns[3 * s1 + 2] = -1.0, 1.0
    return pt, ns


def case(name):
    """(oracle params, engine config, pt, ns, T, initial state)."""
    if name.startswith("ll_"):
        T = float(name[3:])
        nx = ny = nz = 12
        po = O.params(nx, ny, nz, potential=O.LL, A=1.0)
        cfg = mclics.Config(nx, ny, nz, potential=mclics.LL, A=1.0, T=T)
        ic = "x" if T < 1.0 else "random"
        return po, cfg, np.ones(nx * ny * nz, np.int32), None, T, ic
    # GHRL slab: periodic x/y, free z, homeotropic GHRL anchoring on both faces, field along z
    nx, ny, nz = 12, 12, 8
    surf = [("homeotropic_ghrl", 5.0, (0.0, 0.0, 1.0)), ("homeotropic_ghrl", 5.0, (0.0, 0.0, 1.0))]
    po = O.params(nx, ny, nz, bound=(1, 1, 0), potential=O.GHRL, k=(1.0, 0.7, 1.3, 16.0 if name.endswith("16") else 0.0),
                  elec=(0.0, 0.0, 1.0), elecA=1.0, elecE=0.5, surfaces=surf)
    cfg = mclics.Config(nx, ny, nz, bound=(1, 1, 0), potential=mclics.GHRL, ghrl=O.ghrl_tuple(po),
                        elec=(0.0, 0.0, 1.0), elecA=1.0, elecE=0.5,
                        surfaces=[mclics.Surface(k, w, d) for k, w, d in surf], T=0.5)
    pt, ns = slab_layout(nx, ny, nz)
    return po, cfg, pt, ns, 0.5, "z"


def oracle_replica(po, pt, ns, T, ic, seed):
    nt = po.nx * po.ny * po.nz
    n = start(ic, nt, seed)
    av = 0.5
    for s in range(N_EQ):
        _, av = O.sweep_rng(po, n, pt, ns, np.float32(T), av, seed, s, 0, 0, 1)
    Es, Ss = [], []
    for s in range(N_EQ, N_EQ + N_MEAS):
        _, av = O.sweep_rng(po, n, pt, ns, np.float32(T), av, seed, s, 0, 0, 1)
        if (s - N_EQ) % ORC_EVERY == ORC_EVERY - 1:
            Es.append(O.energy_mean(po, n, pt, ns, gpu=1, nthreads=1))
            Ss.append(O.eigvals(O.qtensor(n, pt))[0])
    return float(np.mean(Es)), float(np.mean(Ss))


def gpu_replica(cfg, pt, ns, ic, seed, kernel):
    nt = cfg.nx * cfg.ny * cfg.nz
    cfg.seed = seed
    with mclics.Lattice(cfg, start(ic, nt, seed), pt, ns) as lat:
        lat.set_kernel(kernel)
        lat.sweeps(N_EQ)
        E, S = lat.measure_series(N_MEAS)
    return float(np.mean(E)), float(np.mean(S))


def agree(a, b):
    a, b = np.asarray(a), np.asarray(b)
    sa, sb = a.std(ddof=1) / np.sqrt(len(a)), b.std(ddof=1) / np.sqrt(len(b))
    d = abs(a.mean() - b.mean())
    return d <= 3.0 * np.hypot(sa, sb), d, np.hypot(sa, sb)


_ORACLE = {}


def oracle_ensemble(name):
    if name not in _ORACLE:
        po, cfg, pt, ns, T, ic = case(name)
        seeds = [0x5EED0000 + 97 * r for r in range(R)]
        with ThreadPoolExecutor(max_workers=R) as ex:  # ctypes releases the GIL
            _ORACLE[name] = list(ex.map(lambda s: oracle_replica(po, pt, ns, T, ic, s), seeds))
    return _ORACLE[name]


KERNELS = {"slab": mclics.KERNEL_SLAB, "x4": mclics.KERNEL_X4, "generic": mclics.KERNEL_GENERIC}


@pytest.mark.parametrize("name", ["ll_0.9", "ll_1.1", "ll_1.3", "ghrl_slab", "ghrl_slab_p16"])
@pytest.mark.parametrize("kernel", ["slab", "x4", "generic"])
def test_gpu_chain_matches_cpu_chain_within_3_sigma(name, kernel):
    po, cfg, pt, ns, T, ic = case(name)
    kseed = {"slab": 1, "x4": 2, "generic": 3}[kernel]
    gpu = [gpu_replica(case(name)[1], pt, ns, ic, 1000 * kseed + 13 * r + 1, KERNELS[kernel]) for r in range(R)]
    orc = oracle_ensemble(name)
    for c, what in enumerate(("E", "S")):
        ok, d, sig = agree([g[c] for g in gpu], [o[c] for o in orc])
        print(f"{name} {kernel} {what}: gpu {np.mean([g[c] for g in gpu]):.5f} "
              f"cpu {np.mean([o[c] for o in orc]):.5f} |d| {d:.2e} sigma {sig:.2e}", flush=True)
        assert ok, (name, kernel, what, d, sig)
