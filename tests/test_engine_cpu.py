"""CPU-side checks of the engine library: it builds, loads, exports every
symbol the public header declares, and rejects bad input before touching a
device.  No kernel runs here."""
import os
import re
import subprocess

import numpy as np
import pytest

import mclics
from mclics import engine

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "mclics_hip.h")


def _ensure_built():
    if not os.path.exists(engine.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(REPO, "mclics-backend_amd")], check=True)


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mc_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    _ensure_built()
    lib = mclics.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_functions():
        assert name in syms, name


def test_library_is_gfx950_only():
    # The offload bundle of our own build names its code-object targets.
    _ensure_built()
    blob = open(engine.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"hipv4-amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", blob))
    assert targets == {b"gfx950"}


def test_params_struct_layout_matches_header():
    # sizeof(mc_params) as the C compiler sees it
    src = '#include "mclics_hip.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(){printf("%zu %zu %zu",' \
          'sizeof(mc_params), offsetof(mc_params, seed), offsetof(mc_params, ang_var));return 0;}\n'
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "t.c")
        exe = os.path.join(td, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", exe, c], check=True)
        size, off_seed, off_av = map(int, subprocess.run([exe], capture_output=True, text=True).stdout.split())
    import ctypes
    assert ctypes.sizeof(engine.McParams) == size
    assert engine.McParams.seed.offset == off_seed
    assert engine.McParams.ang_var.offset == off_av


@pytest.mark.parametrize("bad", ["odd_periodic", "non_unit", "nan", "bad_pt", "zero_T", "no_active", "surf_dir_angles"])
def test_create_validates_before_device(bad):
    _ensure_built()
    nx, ny, nz = 4, 4, 4
    nt = 64
    n = np.tile(np.array([0, 0, 1], np.float32), nt)
    pt = np.ones(nt, np.int32)
    cfg = mclics.Config(nx, ny, nz)
    if bad == "odd_periodic":
        cfg = mclics.Config(5, 4, 4)
        n = np.tile(np.array([0, 0, 1], np.float32), 80)
        pt = np.ones(80, np.int32)
    elif bad == "non_unit":
        n[5] = 2.0
    elif bad == "nan":
        n[7] = np.nan
    elif bad == "bad_pt":
        pt[3] = 2  # names surface 0, none defined
    elif bad == "zero_T":
        cfg.T = 0.0
    elif bad == "no_active":
        pt[:] = 0
    elif bad == "surf_dir_angles":
        # an rp direction handed in as (theta_s, phi_s) degrees instead of the unit vector
        cfg.surfaces = [mclics.Surface("rp", 2.0, (90.0, 0.0, 0.0))]
    with pytest.raises(mclics.EngineError) as ei:
        mclics.Lattice(cfg, n, pt)
    assert ei.value.code == -1
    if bad == "surf_dir_angles":
        assert "unit vector" in str(ei.value)


@pytest.mark.parametrize("xi,xf,dx", [(1.2, 0.0, -0.02), (0.0, 1.0, 0.05), (1.2, 1.0, -0.1), (0.5, 0.5, 0.1),
                                      (2.0, 0.1, -0.3)])
def test_bench_float_loop_matches_the_oracle(xi, xf, dx):
    # bench.py's temperature list for the 30^3 thermal run follows the
    # reference's FP32 float loop; the oracle's C restatement is the check.
    import _oracle as O
    import bench
    rows, ref = O.float_loop(xi, xf, dx)
    got = bench.float_loop(xi, xf, dx)
    assert len(got) == rows
    assert np.array_equal(np.asarray(got, np.float32), ref)


def test_bench_refuses_more_than_one_gpu_and_too_few_steps():
    # one GPU only: the multi-GPU decomposition is a design, not a code path;
    # the timed sweeps are split over the four temperatures
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for args, text in ((["--gpus", "2"], "one GPU"), (["--steps", "3"], "--steps")):
        p = subprocess.run([sys.executable, os.path.join(root, "bench.py")] + args, capture_output=True, text=True,
                           timeout=120)
        assert p.returncode == 2, (args, p.returncode, p.stderr)
        assert text in p.stderr, (args, p.stderr)
        assert p.stdout == "", args
