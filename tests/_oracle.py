"""ctypes binding of the CPU oracle (oracle/build/liboracle.so) for the tests.

The oracle is test infrastructure: a C restatement of the reference CPU path.
Only tests/, the smoke check and bench.py's cpu_baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")

LL, GHRL, PEAR = 0, 1, 2
FREE, PERIODIC = 0, 1
ANCHORING = {"homeotropic": 0, "rp": 1, "fg": 2, "strong": 3, "homeotropic_ghrl": 4, "rp_ghrl": 5, "fg_ghrl": 6,
             "strong_ghrl": 7}


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("bound", C.c_int32 * 3),
        ("potential", C.c_int32),
        ("A", C.c_float), ("B1", C.c_float), ("B2", C.c_float), ("C", C.c_float),
        ("ghrl_rho", C.c_float), ("ghrl_lambda", C.c_float), ("ghrl_mu", C.c_float), ("ghrl_nu", C.c_float),
        ("ghrl_sigma", C.c_float),
        ("elecX", C.c_float), ("elecY", C.c_float), ("elecZ", C.c_float), ("elecA", C.c_float),
        ("elecE", C.c_float),
        ("n_surf", C.c_int32),
        ("surf_kind", C.c_int32 * 8),
        ("surf_W", C.c_float * 8),
        ("surf_dir", (C.c_float * 3) * 8),
    ]


_lib = None


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        L = C.CDLL(ORACLE_LIB)
        f32p, i32p, u8p, dp = (C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint8),
                               C.POINTER(C.c_double))
        P = C.POINTER(OrcParams)
        L.orc_bond.argtypes = [P, f32p, f32p, f32p]
        L.orc_bond.restype = C.c_float
        L.orc_surface.argtypes = [P, C.c_int, f32p, f32p]
        L.orc_surface.restype = C.c_float
        L.orc_field.argtypes = [P, f32p]
        L.orc_field.restype = C.c_float
        L.orc_set_ghrl.argtypes = [P, C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_site_energy.argtypes = [P, f32p, i32p, f32p, C.c_int, C.c_int, C.c_int, f32p]
        L.orc_site_energy.restype = C.c_float
        L.orc_site_energies.argtypes = [P, f32p, i32p, f32p, f32p]
        L.orc_pass_inject.argtypes = [P, f32p, i32p, f32p, C.c_int, C.c_float, f32p, f32p, u8p, f32p, f32p]
        L.orc_pass_inject.restype = C.c_int64
        L.orc_adapt_ang_var.argtypes = [C.c_float, C.c_int64, C.c_int64]
        L.orc_adapt_ang_var.restype = C.c_float
        L.orc_sweep_rng.argtypes = [P, f32p, i32p, f32p, C.c_float, f32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_int, C.c_int]
        L.orc_sweep_rng.restype = C.c_int64
        L.orc_energy_mean.argtypes = [P, f32p, i32p, f32p, C.c_int, C.c_int]
        L.orc_energy_mean.restype = C.c_double
        L.orc_qtensor.argtypes = [f32p, i32p, C.c_int64, dp]
        L.orc_sym3_eigvals.argtypes = [dp, dp]
        L.orc_local_S.argtypes = [P, f32p, i32p, C.c_int, C.c_int, C.c_int]
        L.orc_local_S.restype = C.c_double
        L.orc_float_loop.argtypes = [C.c_float, C.c_float, C.c_float, f32p, C.c_int]
        L.orc_float_loop.restype = C.c_int
        L.orc_expf_n.argtypes = [f32p, f32p, C.c_int64]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def f32(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def u8(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint8))


def params(nx, ny, nz, bound=(PERIODIC, PERIODIC, PERIODIC), potential=LL, A=1.0, B1=0.04, B2=0.4, Cc=0.3,
           ghrl=None, k=None, elec=(0.0, 0.0, 0.0), elecA=0.0, elecE=1.0, surfaces=()):
    p = OrcParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    for a in range(3):
        p.bound[a] = bound[a]
    p.potential = potential
    p.A, p.B1, p.B2, p.C = A, B1, B2, Cc
    p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma = 0.0, 0.0, 0.0, 1.0, 0.0
    if ghrl is not None:
        p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma = ghrl
    if k is not None:
        lib().orc_set_ghrl(C.byref(p), *k)
    p.elecX, p.elecY, p.elecZ = elec
    p.elecA, p.elecE = elecA, elecE
    p.n_surf = len(surfaces)
    for s, (kind, W, d) in enumerate(surfaces):
        p.surf_kind[s] = ANCHORING[kind]
        p.surf_W[s] = W
        for c in range(3):
            p.surf_dir[s][c] = d[c]
    return p


def ghrl_tuple(p):
    return (p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma)


def bond(p, ni, nj, r):
    a, b, c = (np.ascontiguousarray(x, np.float32) for x in (ni, nj, r))
    return lib().orc_bond(C.byref(p), f32(a), f32(b), f32(c))


def surface(p, s, ni, sv):
    a, b = (np.ascontiguousarray(x, np.float32) for x in (ni, sv))
    return lib().orc_surface(C.byref(p), s, f32(a), f32(b))


def field(p, ni):
    a = np.ascontiguousarray(ni, np.float32)
    return lib().orc_field(C.byref(p), f32(a))


def site_energies(p, n, pt, ns=None):
    out = np.empty(p.nx * p.ny * p.nz, np.float32)
    lib().orc_site_energies(C.byref(p), f32(n), i32(pt), f32(ns), f32(out))
    return out


def pass_inject(p, n, pt, ns, set_id, T, prop, u):
    nt = p.nx * p.ny * p.nz
    acc = np.zeros(nt, np.uint8)
    eo = np.zeros(nt, np.float32)
    en = np.zeros(nt, np.float32)
    cnt = lib().orc_pass_inject(C.byref(p), f32(n), i32(pt), f32(ns), set_id, T, f32(prop), f32(u), u8(acc),
                                f32(eo), f32(en))
    return cnt, acc, eo, en


def sweep_rng(p, n, pt, ns, T, ang_var, seed, sweep, stream=0, order=0, nthreads=0):
    av = np.array([ang_var], np.float32)
    acc = lib().orc_sweep_rng(C.byref(p), f32(n), i32(pt), f32(ns), T, f32(av), seed & 0xFFFFFFFF,
                              (seed >> 32) & 0xFFFFFFFF, sweep, stream, order, nthreads)
    return int(acc), float(av[0])


def energy_mean(p, n, pt, ns=None, gpu=1, nthreads=0):
    return lib().orc_energy_mean(C.byref(p), f32(n), i32(pt), f32(ns), gpu, nthreads)


def qtensor(n, pt):
    q = (C.c_double * 6)()
    lib().orc_qtensor(f32(n), i32(pt), len(pt), q)
    return np.array(q[:])


def eigvals(q6):
    q = (C.c_double * 6)(*q6)
    ev = (C.c_double * 3)()
    lib().orc_sym3_eigvals(q, ev)
    return np.array(ev[:])


def local_S(p, n, pt, i, j, k):
    return lib().orc_local_S(C.byref(p), f32(n), i32(pt), i, j, k)


def float_loop(xi, xf, dx, max_rows=100000):
    xs = np.zeros(max_rows, np.float32)
    rows = lib().orc_float_loop(xi, xf, dx, f32(xs), max_rows)
    return rows, xs[:min(rows, max_rows)]


def expf(x):
    """glibc's expf of every element (the reference's exp(float))."""
    a = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(a)
    lib().orc_expf_n(f32(a), f32(y), a.size)
    return y


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


# ---------------------------------------------------------------------------
# Learned-proposal network oracle (oracle/mlp_oracle.c).

MLP_NPARAM = 64 * 22 + 64 + 64 * 64 + 64 + 4 * 64 + 4


def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_ready", False):
        f32p, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.orc_mlp_features.argtypes = [f32p, f32p, C.c_float, f32p, f32p]
        L.orc_bf16.argtypes = [C.c_float]
        L.orc_bf16.restype = C.c_float
        L.orc_mlp_forward.argtypes = [f32p, f32p, C.c_int, dp]
        L.orc_mlp_head.argtypes = [dp, f32p, dp, dp]
        L.orc_mlp_nll_grad.argtypes = [f32p, C.c_int, f32p, f32p, f32p, C.c_int, dp]
        L.orc_mlp_nll_grad.restype = C.c_double
        L.orc_vmf_logq.argtypes = [dp, C.c_double, dp]
        L.orc_vmf_logq.restype = C.c_double
        L._mlp_ready = True
    return L


def mlp_features(nbr18, field, T):
    L = _mlp_lib()
    nb = np.ascontiguousarray(nbr18, np.float32)
    fd = np.ascontiguousarray(field, np.float32)
    x = np.zeros(22, np.float32)
    base = np.zeros(3, np.float32)
    L.orc_mlp_features(f32(nb), f32(fd), T, f32(x), f32(base))
    return x, base


def bf16(v):
    L = _mlp_lib()
    return np.array([L.orc_bf16(float(a)) for a in np.ravel(v)], np.float32).reshape(np.shape(v))


def mlp_forward(params, x, use_bf16=0):
    L = _mlp_lib()
    out = (C.c_double * 4)()
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    L.orc_mlp_forward(f32(p), f32(xx), use_bf16, out)
    return np.array(out[:])


def mlp_head(out, base):
    L = _mlp_lib()
    o = (C.c_double * 4)(*out)
    mu = (C.c_double * 3)()
    lk = C.c_double()
    b = np.ascontiguousarray(base, np.float32)
    L.orc_mlp_head(o, f32(b), mu, C.byref(lk))
    return np.array(mu[:]), lk.value


def mlp_nll_grad(params, X, base, n, use_bf16=0):
    L = _mlp_lib()
    g = np.zeros(MLP_NPARAM, np.float64)
    p = np.ascontiguousarray(params, np.float32)
    XX = np.ascontiguousarray(X, np.float32)
    bb = np.ascontiguousarray(base, np.float32)
    nn = np.ascontiguousarray(n, np.float32)
    loss = L.orc_mlp_nll_grad(f32(p), len(XX), f32(XX), f32(bb), f32(nn), use_bf16,
                              g.ctypes.data_as(C.POINTER(C.c_double)))
    return loss, g


def vmf_logq(mu, kappa, n):
    L = _mlp_lib()
    m = (C.c_double * 3)(*mu)
    v = (C.c_double * 3)(*n)
    return L.orc_vmf_logq(m, kappa, v)
