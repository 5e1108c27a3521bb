"""Initial director fields (host side), following src/ic.cpp of the reference.

``random_ic`` replaces GSL's taus + gsl_ran_dir_3d (src/ic.cpp:27-59) with the
Archimedes sampler (z = 2u - 1, phi = 2*pi*v) driven by Philox4x32-10 keyed by
the seed: the same distribution (uniform on the sphere), a different stream.
"""
from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)
IC_TAG = 0x4943  # counter word 3 of initial-condition draws


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox4x32-10 (Salmon et al., SC'11) over uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(x, np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = np.uint32(k0 + _W0)
                k1 = np.uint32(k1 + _W1)
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
            n1 = (p1 & _MASK).astype(np.uint32)
            n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
            n3 = (p0 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = n0, n1, n2, n3
    return c0, c1, c2, c3


def u01_24(x: np.ndarray) -> np.ndarray:
    return (x >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


def random_ic(nx: int, ny: int, nz: int, seed: int, pt=None, chunk: int = 1 << 22) -> np.ndarray:
    """AoS float32[3*Nt] unit directors uniform on S^2; inactive sites get (0,0,1)."""
    nt = nx * ny * nz
    out = np.empty((nt, 3), np.float32)
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    for a in range(0, nt, chunk):
        b = min(nt, a + chunk)
        s = np.arange(a, b, dtype=np.uint32)
        x0, x1, _, _ = philox4x32_10(s, np.zeros_like(s), np.zeros_like(s), np.full_like(s, IC_TAG), k0, k1)
        z = 2.0 * u01_24(x0) - 1.0
        phi = 2.0 * np.pi * u01_24(x1)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        out[a:b, 0] = r * np.cos(phi)
        out[a:b, 1] = r * np.sin(phi)
        out[a:b, 2] = z
    if pt is not None:
        off = np.asarray(pt).reshape(nt) == 0
        out[off] = (0.0, 0.0, 1.0)
    return out.reshape(3 * nt)


def homogeneous_ic(nx: int, ny: int, nz: int, theta0: float, phi0: float, pt=None) -> np.ndarray:
    """src/ic.cpp:61-86: every active site along (theta0, phi0) in degrees."""
    nt = nx * ny * nz
    th = np.float32(theta0) * np.pi / 180
    ph = np.float32(phi0) * np.pi / 180
    v = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], np.float32)
    out = np.tile(v, (nt, 1))
    if pt is not None:
        out[np.asarray(pt).reshape(nt) == 0] = (0.0, 0.0, 1.0)
    return out.reshape(3 * nt)
