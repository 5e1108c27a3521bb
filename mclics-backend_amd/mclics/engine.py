"""ctypes binding of the HIP engine's C ABI (include/mclics_hip.h).

The engine is the shared library built from ``csrc/`` into
``lib/libmclics_hip.so``.  Loading fails loudly when it has not been built:
there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libmclics_hip.so")

MC_MAX_SURF = 8
FREE, PERIODIC = 0, 1
LL, GHRL, PEAR = 0, 1, 2
ORDER_CHECKER, ORDER_REF8 = 0, 1
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_X4, KERNEL_SLAB, KERNEL_SLAB_PERSISTENT = 0, 1, 2, 3, 4
ANCHORING = {
    "homeotropic": 0,
    "rp": 1,
    "fg": 2,
    "strong": 3,
    "homeotropic_ghrl": 4,
    "rp_ghrl": 5,
    "fg_ghrl": 6,
    "strong_ghrl": 7,
}

# Every symbol the public header declares.
EXPORTS = (
    "mc_last_error",
    "mc_device_count",
    "mc_create",
    "mc_destroy",
    "mc_set_T",
    "mc_set_field",
    "mc_set_ang_var",
    "mc_get_ang_var",
    "mc_set_sweep_kernel",
    "mc_get_sweep_kernel",
    "mc_sweeps",
    "mc_sweeps_graph",
    "mc_sync",
    "mc_time_sweeps",
    "mc_time_passes",
    "mc_measure",
    "mc_measure_series",
    "mc_measure_series_every",
    "mc_measure_noalloc",
    "mc_site_energies",
    "mc_local_order",
    "mc_get_directors_aos",
    "mc_set_directors_aos",
    "mc_get_acceptance",
    "mc_set_sweep_index",
    "mc_pass",
    "mc_record_sweep",
    "mc_set_tuning",
    "mc_set_slab_tile",
    "mc_get_slab_tile",
    "mc_autotune",
    "mc_batch_create",
    "mc_batch_destroy",
    "mc_batch_set_T",
    "mc_batch_sweeps",
    "mc_batch_sync",
    "mc_batch_time_sweeps",
    "mc_batch_measure",
    "mc_batch_measure_series",
    "mc_batch_set_barrier_limit",
    "mc_batch_get_directors_aos",
    "mc_batch_get_state",
    "mc_batch_set_tuning",
    "mc_batch_set_kernel",
    "mc_batch_get_kernel",
    "mc_batch_autotune",
    "mc_mlp_create",
    "mc_mlp_destroy",
    "mc_mlp_get_params",
    "mc_mlp_set_params",
    "mc_mlp_infer",
    "mc_mlp_grad",
    "mc_mlp_dataset_add",
    "mc_mlp_dataset_size",
    "mc_mlp_dataset_clear",
    "mc_mlp_train",
    "mc_set_learned_proposal",
    "mc_set_learned_reflection",
    "mc_record_learned_sweep",
)


class McParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("bound", C.c_int32 * 3),
        ("potential", C.c_int32),
        ("A", C.c_float),
        ("B1", C.c_float),
        ("B2", C.c_float),
        ("C", C.c_float),
        ("ghrl_rho", C.c_float),
        ("ghrl_lambda", C.c_float),
        ("ghrl_mu", C.c_float),
        ("ghrl_nu", C.c_float),
        ("ghrl_sigma", C.c_float),
        ("elecX", C.c_float),
        ("elecY", C.c_float),
        ("elecZ", C.c_float),
        ("elecA", C.c_float),
        ("elecE", C.c_float),
        ("T", C.c_float),
        ("n_surf", C.c_int32),
        ("surf_kind", C.c_int32 * MC_MAX_SURF),
        ("surf_W", C.c_float * MC_MAX_SURF),
        ("surf_dir", (C.c_float * 3) * MC_MAX_SURF),
        ("seed", C.c_uint64),
        ("stream", C.c_uint32),
        ("order", C.c_int32),
        ("ang_var", C.c_float),
    ]


class EngineError(RuntimeError):
    """A non-zero status from the C ABI, with mc_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib: Optional[C.CDLL] = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the engine library; raises FileNotFoundError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"HIP engine not built: {path} is missing (run __graft_entry__.build())")
    lib = C.CDLL(path)
    f32p = C.POINTER(C.c_float)
    u8p = C.POINTER(C.c_uint8)
    vp = C.c_void_p
    lib.mc_last_error.restype = C.c_char_p
    lib.mc_device_count.restype = C.c_int
    lib.mc_create.argtypes = [C.POINTER(McParams), f32p, C.POINTER(C.c_int32), f32p, C.POINTER(vp)]
    lib.mc_destroy.argtypes = [vp]
    lib.mc_set_T.argtypes = [vp, C.c_float]
    lib.mc_set_field.argtypes = [vp, C.c_float]
    lib.mc_set_ang_var.argtypes = [vp, C.c_float]
    lib.mc_get_ang_var.argtypes = [vp, f32p]
    lib.mc_set_sweep_kernel.argtypes = [vp, C.c_int]
    lib.mc_get_sweep_kernel.argtypes = [vp, C.POINTER(C.c_int)]
    lib.mc_sweeps.argtypes = [vp, C.c_int]
    lib.mc_sweeps_graph.argtypes = [vp, C.c_int]
    lib.mc_sync.argtypes = [vp]
    lib.mc_time_sweeps.argtypes = [vp, C.c_int, C.c_int, f32p, C.POINTER(C.c_int)]
    lib.mc_time_passes.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]
    lib.mc_measure.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.mc_measure_series.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.mc_measure_series_every.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.mc_measure_noalloc.argtypes = [C.POINTER(McParams), vp, vp, vp, vp, vp]
    lib.mc_site_energies.argtypes = [vp, f32p]
    lib.mc_local_order.argtypes = [vp, f32p]
    lib.mc_get_directors_aos.argtypes = [vp, f32p]
    lib.mc_set_directors_aos.argtypes = [vp, f32p]
    lib.mc_get_acceptance.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.mc_set_sweep_index.argtypes = [vp, C.c_uint32]
    lib.mc_pass.argtypes = [vp, C.c_int, f32p, f32p, u8p, f32p, f32p]
    lib.mc_record_sweep.argtypes = [vp, f32p, f32p, f32p, f32p, u8p, f32p, f32p]
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    u32p = C.POINTER(C.c_uint32)
    lib.mc_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    lib.mc_set_slab_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.mc_get_slab_tile.argtypes = [vp, ip, ip, ip]
    lib.mc_autotune.argtypes = [vp, ip, ip, f32p]
    lib.mc_batch_create.argtypes = [C.c_int, C.POINTER(McParams), C.POINTER(f32p), C.POINTER(vp)]
    lib.mc_batch_destroy.argtypes = [vp]
    lib.mc_batch_set_T.argtypes = [vp, C.c_int, C.c_float]
    lib.mc_batch_sweeps.argtypes = [vp, C.c_int]
    lib.mc_batch_sync.argtypes = [vp]
    lib.mc_batch_time_sweeps.argtypes = [vp, C.c_int, f32p]
    lib.mc_batch_measure.argtypes = [vp, dp, dp]
    lib.mc_batch_measure_series.argtypes = [vp, C.c_int, dp, dp]
    lib.mc_batch_set_barrier_limit.argtypes = [vp, C.c_uint]
    lib.mc_batch_get_directors_aos.argtypes = [vp, C.c_int, f32p]
    lib.mc_batch_get_state.argtypes = [vp, u32p, u32p, f32p]
    lib.mc_batch_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    lib.mc_batch_set_kernel.argtypes = [vp, C.c_int]
    lib.mc_batch_get_kernel.argtypes = [vp, ip]
    lib.mc_batch_autotune.argtypes = [vp, ip, ip, f32p]
    lib.mc_mlp_create.argtypes = [f32p, C.c_uint64, C.POINTER(vp)]
    lib.mc_mlp_destroy.argtypes = [vp]
    lib.mc_mlp_get_params.argtypes = [vp, f32p]
    lib.mc_mlp_set_params.argtypes = [vp, f32p]
    lib.mc_mlp_infer.argtypes = [vp, C.c_int, f32p, f32p, f32p, f32p]
    lib.mc_mlp_grad.argtypes = [vp, C.c_int, f32p, f32p, f32p, dp, f32p]
    lib.mc_mlp_dataset_add.argtypes = [vp, vp, C.c_int, C.c_uint32]
    lib.mc_mlp_dataset_size.argtypes = [vp, ip]
    lib.mc_mlp_dataset_clear.argtypes = [vp]
    lib.mc_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, f32p]
    lib.mc_set_learned_proposal.argtypes = [vp, vp, C.c_int]
    lib.mc_set_learned_reflection.argtypes = [vp, C.c_int]
    lib.mc_record_learned_sweep.argtypes = [vp, f32p, f32p, f32p, f32p, u8p, f32p, f32p, f32p, u8p, f32p]
    for name in EXPORTS:
        if name not in ("mc_last_error", "mc_device_count"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise EngineError(rc, load_library().mc_last_error().decode())


def _f32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _opt_f32(a):
    return None if a is None else _f32(a)


def _u8(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint8))


@dataclass
class Surface:
    kind: str
    W: float
    direction: Sequence[float] = (0.0, 0.0, 1.0)


@dataclass
class Config:
    """Parameters of one lattice; names follow the reference's Parameters."""

    nx: int
    ny: int
    nz: int
    bound: Sequence[int] = (PERIODIC, PERIODIC, PERIODIC)
    potential: int = LL
    A: float = 1.0
    B1: float = 0.04
    B2: float = 0.4
    C: float = 0.3
    ghrl: Sequence[float] = (0.0, 0.0, 0.0, 1.0, 0.0)  # rho, lambda, mu, nu, sigma
    elec: Sequence[float] = (0.0, 0.0, 0.0)
    elecA: float = 0.0
    elecE: float = 1.0
    T: float = 1.0
    surfaces: Sequence[Surface] = field(default_factory=list)
    seed: int = 1
    stream: int = 0
    order: int = ORDER_CHECKER
    ang_var: float = 0.5

    def to_struct(self) -> McParams:
        p = McParams()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        for a in range(3):
            p.bound[a] = int(self.bound[a])
        p.potential = self.potential
        p.A, p.B1, p.B2, p.C = self.A, self.B1, self.B2, self.C
        (p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma) = self.ghrl
        p.elecX, p.elecY, p.elecZ = self.elec
        p.elecA, p.elecE, p.T = self.elecA, self.elecE, self.T
        p.n_surf = len(self.surfaces)
        for s, sf in enumerate(self.surfaces):
            p.surf_kind[s] = ANCHORING[sf.kind]
            p.surf_W[s] = sf.W
            for c in range(3):
                p.surf_dir[s][c] = sf.direction[c]
        p.seed = self.seed
        p.stream = self.stream
        p.order = self.order
        p.ang_var = self.ang_var
        return p


class Lattice:
    """One device-resident lattice (an mc_handle)."""

    def __init__(self, cfg: Config, ni_aos: np.ndarray, pt: Optional[np.ndarray] = None,
                 ns_aos: Optional[np.ndarray] = None):
        lib = load_library()
        nt = cfg.nx * cfg.ny * cfg.nz
        self.cfg = cfg
        self.nt = nt
        ni = np.ascontiguousarray(ni_aos, dtype=np.float32).reshape(3 * nt)
        ptv = np.ones(nt, np.int32) if pt is None else np.ascontiguousarray(pt, dtype=np.int32).reshape(nt)
        ns = None if ns_aos is None else np.ascontiguousarray(ns_aos, dtype=np.float32).reshape(3 * nt)
        self._params = cfg.to_struct()
        h = C.c_void_p()
        _check(lib.mc_create(C.byref(self._params), _f32(ni), ptv.ctypes.data_as(C.POINTER(C.c_int32)),
                             _opt_f32(ns), C.byref(h)))
        self._h = h
        self._lib = lib

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- state -----------------------------------------------------------
    def set_T(self, T: float) -> None:
        _check(self._lib.mc_set_T(self._h, T))

    def set_field(self, E: float) -> None:
        _check(self._lib.mc_set_field(self._h, E))

    def set_ang_var(self, a: float) -> None:
        _check(self._lib.mc_set_ang_var(self._h, a))

    def ang_var(self) -> float:
        v = C.c_float()
        _check(self._lib.mc_get_ang_var(self._h, C.byref(v)))
        return v.value

    def acceptance(self):
        a, s = C.c_uint32(), C.c_uint32()
        _check(self._lib.mc_get_acceptance(self._h, C.byref(a), C.byref(s)))
        return a.value, s.value

    def set_sweep_index(self, sweep: int) -> None:
        """Index of the next sweep (keys the generator); with directors and
        ang_var it resumes a saved chain exactly."""
        _check(self._lib.mc_set_sweep_index(self._h, sweep))

    def checkpoint(self):
        """(directors, ang_var, sweep index): everything the chain's future depends on."""
        return self.directors(), self.ang_var(), self.acceptance()[1]

    def local_order(self) -> np.ndarray:
        """Local S of every site (NaN where pt == 0), as in the snapshot files."""
        out = np.empty(self.nt, np.float32)
        _check(self._lib.mc_local_order(self._h, _f32(out)))
        return out

    def directors(self) -> np.ndarray:
        out = np.empty(3 * self.nt, np.float32)
        _check(self._lib.mc_get_directors_aos(self._h, _f32(out)))
        return out

    def set_directors(self, ni_aos: np.ndarray) -> None:
        ni = np.ascontiguousarray(ni_aos, dtype=np.float32).reshape(3 * self.nt)
        _check(self._lib.mc_set_directors_aos(self._h, _f32(ni)))

    def set_tuning(self, grid: int, xcd: int) -> None:
        _check(self._lib.mc_set_tuning(self._h, grid, xcd))

    def set_slab_tile(self, ty: int, cz: int, xcd: int = 1) -> None:
        """Slab-kernel tile: ty full rows of cz planes per workgroup (0, 0: default)."""
        _check(self._lib.mc_set_slab_tile(self._h, ty, cz, xcd))

    def slab_tile(self):
        ty, cz, x = C.c_int(), C.c_int(), C.c_int()
        _check(self._lib.mc_get_slab_tile(self._h, C.byref(ty), C.byref(cz), C.byref(x)))
        return ty.value, cz.value, x.value

    def autotune(self):
        g, x, us = C.c_int(), C.c_int(), C.c_float()
        _check(self._lib.mc_autotune(self._h, C.byref(g), C.byref(x), C.byref(us)))
        return g.value, x.value, us.value

    def set_kernel(self, kind: int) -> None:
        _check(self._lib.mc_set_sweep_kernel(self._h, kind))

    def kernel(self) -> int:
        v = C.c_int()
        _check(self._lib.mc_get_sweep_kernel(self._h, C.byref(v)))
        return v.value

    # --- work --------------------------------------------------------------
    def sweeps(self, n: int, graph: bool = False) -> None:
        _check((self._lib.mc_sweeps_graph if graph else self._lib.mc_sweeps)(self._h, n))

    def sync(self) -> None:
        _check(self._lib.mc_sync(self._h))

    def time_sweeps(self, n: int, graph: bool = True):
        ms, launches = C.c_float(), C.c_int()
        _check(self._lib.mc_time_sweeps(self._h, n, 1 if graph else 0, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def time_passes(self, n: int, every: int = 1):
        """n eager sweeps between one event pair.  every >= 1 also brackets
        every `every`-th pass launch and averages those; every = 0 averages
        the region over its pass launches.  Returns (mean pass ms, region ms,
        pass launches)."""
        avg, tot, np_ = C.c_double(), C.c_double(), C.c_int()
        _check(self._lib.mc_time_passes(self._h, n, every, C.byref(avg), C.byref(tot), C.byref(np_)))
        return avg.value, tot.value, np_.value

    def measure(self):
        E, S = C.c_double(), C.c_double()
        q = (C.c_double * 6)()
        _check(self._lib.mc_measure(self._h, C.byref(E), C.byref(S), q))
        return E.value, S.value, np.array(q[:], np.float64)

    def measure_series(self, n: int, every: int = 1):
        """n measurements (E, S), each after `every` sweeps."""
        E = np.empty(n, np.float64)
        S = np.empty(n, np.float64)
        _check(self._lib.mc_measure_series_every(self._h, n, every, E.ctypes.data_as(C.POINTER(C.c_double)),
                                                 S.ctypes.data_as(C.POINTER(C.c_double))))
        return E, S

    def site_energies(self) -> np.ndarray:
        out = np.empty(self.nt, np.float32)
        _check(self._lib.mc_site_energies(self._h, _f32(out)))
        return out

    def pass_inject(self, set_id: int, prop_aos: np.ndarray, u: np.ndarray):
        prop = np.ascontiguousarray(prop_aos, dtype=np.float32).reshape(3 * self.nt)
        uu = np.ascontiguousarray(u, dtype=np.float32).reshape(self.nt)
        acc = np.zeros(self.nt, np.uint8)
        eo = np.zeros(self.nt, np.float32)
        en = np.zeros(self.nt, np.float32)
        _check(self._lib.mc_pass(self._h, set_id, _f32(prop), _f32(uu), _u8(acc), _f32(eo), _f32(en)))
        return acc, eo, en

    def set_learned(self, net: Optional["Network"], every: int = 1, reflect: bool = True) -> None:
        """Attach a network (every k-th sweep learned) or detach it (None).
        reflect: each vMF move is followed by the reflection through the
        network's mean direction (a second Metropolis-Hastings move)."""
        _check(self._lib.mc_set_learned_proposal(self._h, None if net is None else net._h, 0 if net is None else every))
        _check(self._lib.mc_set_learned_reflection(self._h, 1 if reflect else 0))
        self._net = net

    def record_learned_sweep(self):
        nt = self.nt
        r = {"mu": np.zeros(3 * nt, np.float32), "logk": np.zeros(nt, np.float32),
             "prop": np.zeros(3 * nt, np.float32), "u": np.zeros(nt, np.float32),
             "accept": np.zeros(nt, np.uint8), "dE": np.zeros(nt, np.float32),
             "refl": np.zeros(3 * nt, np.float32), "u2": np.zeros(nt, np.float32),
             "accept2": np.zeros(nt, np.uint8), "dE2": np.zeros(nt, np.float32)}
        _check(self._lib.mc_record_learned_sweep(self._h, _f32(r["mu"]), _f32(r["logk"]), _f32(r["prop"]),
                                                 _f32(r["u"]), _u8(r["accept"]), _f32(r["dE"]), _f32(r["refl"]),
                                                 _f32(r["u2"]), _u8(r["accept2"]), _f32(r["dE2"])))
        return r

    def record_sweep(self):
        nt = self.nt
        r = {k: np.zeros(nt, np.float32) for k in ("u0", "u1", "u2", "e_old", "e_new")}
        r["prop"] = np.zeros(3 * nt, np.float32)
        r["accept"] = np.zeros(nt, np.uint8)
        _check(self._lib.mc_record_sweep(self._h, _f32(r["u0"]), _f32(r["u1"]), _f32(r["u2"]), _f32(r["prop"]),
                                         _u8(r["accept"]), _f32(r["e_old"]), _f32(r["e_new"])))
        return r


class Batch:
    """Many independent bulk lattices swept together (an mc_batch)."""

    def __init__(self, cfgs: Sequence[Config], ni_list: Sequence[np.ndarray]):
        lib = load_library()
        self.cfgs = list(cfgs)
        self.nrep = len(cfgs)
        self._arr = (McParams * self.nrep)(*[c.to_struct() for c in cfgs])
        self._ni = [np.ascontiguousarray(n, dtype=np.float32).reshape(-1) for n in ni_list]
        ptrs = (C.POINTER(C.c_float) * self.nrep)(*[_f32(n) for n in self._ni])
        h = C.c_void_p()
        _check(lib.mc_batch_create(self.nrep, self._arr, ptrs, C.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_T(self, rep: int, T: float):
        _check(self._lib.mc_batch_set_T(self._h, rep, T))

    def sweeps(self, n: int):
        _check(self._lib.mc_batch_sweeps(self._h, n))

    def sync(self):
        _check(self._lib.mc_batch_sync(self._h))

    def time_sweeps(self, n: int) -> float:
        ms = C.c_float()
        _check(self._lib.mc_batch_time_sweeps(self._h, n, C.byref(ms)))
        return ms.value

    def measure(self):
        E = np.empty(self.nrep)
        S = np.empty(self.nrep)
        _check(self._lib.mc_batch_measure(self._h, E.ctypes.data_as(C.POINTER(C.c_double)),
                                          S.ctypes.data_as(C.POINTER(C.c_double))))
        return E, S

    def measure_series(self, n: int):
        E = np.empty((n, self.nrep))
        S = np.empty((n, self.nrep))
        _check(self._lib.mc_batch_measure_series(self._h, n, E.ctypes.data_as(C.POINTER(C.c_double)),
                                                 S.ctypes.data_as(C.POINTER(C.c_double))))
        return E, S

    def directors(self, rep: int) -> np.ndarray:
        c = self.cfgs[rep]
        out = np.empty(3 * c.nx * c.ny * c.nz, np.float32)
        _check(self._lib.mc_batch_get_directors_aos(self._h, rep, _f32(out)))
        return out

    def state(self):
        a = np.empty(self.nrep, np.uint32)
        s = np.empty(self.nrep, np.uint32)
        v = np.empty(self.nrep, np.float32)
        _check(self._lib.mc_batch_get_state(self._h, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            s.ctypes.data_as(C.POINTER(C.c_uint32)), _f32(v)))
        return a, s, v

    def set_tuning(self, grid: int, xcd: int):
        _check(self._lib.mc_batch_set_tuning(self._h, grid, xcd))

    def set_barrier_limit(self, polls: int):
        _check(self._lib.mc_batch_set_barrier_limit(self._h, polls))

    def set_kernel(self, kind: int) -> None:
        _check(self._lib.mc_batch_set_kernel(self._h, kind))

    def kernel(self) -> int:
        v = C.c_int()
        _check(self._lib.mc_batch_get_kernel(self._h, C.byref(v)))
        return v.value

    def autotune(self):
        g, x, us = C.c_int(), C.c_int(), C.c_float()
        _check(self._lib.mc_batch_autotune(self._h, C.byref(g), C.byref(x), C.byref(us)))
        return g.value, x.value, us.value


MLP_NPARAM = 5892


class Network:
    """The learned-proposal network (an mc_mlp)."""

    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 1):
        lib = load_library()
        h = C.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        _check(lib.mc_mlp_create(None if p is None else _f32(p), seed, C.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self) -> np.ndarray:
        out = np.empty(MLP_NPARAM, np.float32)
        _check(self._lib.mc_mlp_get_params(self._h, _f32(out)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        pp = np.ascontiguousarray(p, np.float32)
        _check(self._lib.mc_mlp_set_params(self._h, _f32(pp)))

    def infer(self, X: np.ndarray, base: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        base = np.ascontiguousarray(base, np.float32)
        n = len(X)
        mu = np.empty((n, 3), np.float32)
        lk = np.empty(n, np.float32)
        _check(self._lib.mc_mlp_infer(self._h, n, _f32(X), _f32(base), _f32(mu), _f32(lk)))
        return mu, lk

    def grad(self, X, base, dirs):
        X = np.ascontiguousarray(X, np.float32)
        base = np.ascontiguousarray(base, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        g = np.empty(MLP_NPARAM, np.float32)
        loss = C.c_double()
        _check(self._lib.mc_mlp_grad(self._h, len(X), _f32(X), _f32(base), _f32(dirs), C.byref(loss), _f32(g)))
        return loss.value, g

    def add_samples(self, lattice: "Lattice", n: int, seed: int) -> None:
        _check(self._lib.mc_mlp_dataset_add(self._h, lattice._h, n, seed))

    def dataset_size(self) -> int:
        v = C.c_int()
        _check(self._lib.mc_mlp_dataset_size(self._h, C.byref(v)))
        return v.value

    def clear_samples(self) -> None:
        _check(self._lib.mc_mlp_dataset_clear(self._h))

    def train(self, epochs: int, batch: int = 1024, lr: float = 1e-3) -> np.ndarray:
        out = np.empty(epochs, np.float32)
        _check(self._lib.mc_mlp_train(self._h, epochs, batch, lr, _f32(out)))
        return out


def device_count() -> int:
    return int(load_library().mc_device_count())
