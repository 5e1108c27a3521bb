"""The allocation-free measurement (mc_measure_noalloc) against mc_measure.

The caller owns the device buffers (allocated here through the HIP runtime
the engine itself links), hands the entry point three SoA planes and an
output, and must get mc_measure's E, S and Q bit for bit for the same
lattice: on two calls in a row, the second with a smaller lattice, and once
more on the first (the shared scratch keeps the last call's contents)."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    mclics.load_library()  # loads libamdhip64 for the engine; the same copy serves the test
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h


def entry():
    lib = mclics.load_library()
    f = lib.mc_measure_noalloc
    f.argtypes = [C.POINTER(mclics.McParams), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


class DevBuf:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def download(self, a):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.p)


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


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


def lattice(name, rng):
    if name == "ll_big":
        cfg = mclics.Config(64, 48, 40, potential=mclics.LL, A=1.0, T=1.1)
        pt = ns = None
    elif name == "ghrl_slab_small":
        po = O.params(16, 12, 10, bound=(1, 1, 0), potential=O.GHRL, k=(1.0, 0.7, 1.3, 16.0))
        cfg = mclics.Config(16, 12, 10, bound=(1, 1, 0), potential=mclics.GHRL, ghrl=O.ghrl_tuple(po),
                            elec=(0.0, 0.0, 1.0), elecA=1.0, elecE=0.5, T=0.5,
                            surfaces=[mclics.Surface("homeotropic_ghrl", 5.0, (0, 0, 1)),
                                      mclics.Surface("rp_ghrl", 2.0, (0.6, 0.0, 0.8))])
        pt, ns = slab_layout(16, 12, 10)
    else:  # pear bulk, small
        cfg = mclics.Config(8, 8, 6, potential=mclics.PEAR, A=1.0, B1=0.3, C=0.3, elec=(0.3, 0.1, 1.0),
                            elecA=0.7, elecE=0.9, T=0.8)
        pt = ns = None
    n = unit(rng, cfg.nx * cfg.ny * cfg.nz)
    return cfg, n, pt, ns


def test_noalloc_measure_equals_mc_measure():
    h = hip()
    f = entry()
    rng = np.random.default_rng(29)
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    big = lattice("ll_big", rng)
    cases = [big, lattice("ghrl_slab_small", rng), lattice("pear_small", rng), big]
    for ci, (cfg, n, pt, ns) in enumerate(cases):
        nt = cfg.nx * cfg.ny * cfg.nz
        with mclics.Lattice(cfg, n, pt, ns) as lat:
            lat.sweeps(3)
            d = lat.directors()
            E, S, q6 = lat.measure()
        planes = [DevBuf(h, 4 * nt) for _ in range(3)]
        out = DevBuf(h, 9 * 8)
        try:
            soa = d.reshape(-1, 3)
            for c in range(3):
                planes[c].upload(np.ascontiguousarray(soa[:, c]))
            out.upload(np.full(9, np.nan))
            params = cfg.to_struct()
            rc = f(C.byref(params), planes[0].p, planes[1].p, planes[2].p, out.p, stream if ci % 2 else None)
            assert rc == 0, mclics.load_library().mc_last_error()
            assert h.hipDeviceSynchronize() == 0
            r = out.download(np.zeros(9))
        finally:
            for b in planes + [out]:
                b.free()
        assert r[0] == E and r[1] == S, (ci, r[:2], E, S)
        assert np.array_equal(r[2:8], q6), ci
        assert r[8] == nt
    h.hipStreamDestroy(stream)


def test_noalloc_measure_rejects_unsupported_lattices():
    f = entry()
    cfg = mclics.Config(6, 8, 8)  # Nx not a multiple of 4
    params = cfg.to_struct()
    dummy = C.c_void_p(16)
    assert f(C.byref(params), dummy, dummy, dummy, dummy, None) == -1  # MC_EINVAL
    cfg = mclics.Config(8, 8, 8, bound=(1, 1, 0), surfaces=[mclics.Surface("homeotropic", 1.0)])  # one surface
    params = cfg.to_struct()
    assert f(C.byref(params), dummy, dummy, dummy, dummy, None) == -1
    assert f(None, dummy, dummy, dummy, dummy, None) == -1
