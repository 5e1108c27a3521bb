"""End-to-end runs of the drop-in executable (mc_sim) on the GPU: the *GPU
modes on small lattices, checked for the reference's po.dat columns and row
counts (float loops, src/evolve_thermal.cu:79, src/evolve_electric.cu:81),
snapshot names ((int)(100*(T+1e-5)), (int)(1000*(E+1e-5)), block index) and
format, and the snapshot's S column against the oracle's local order
parameter of the snapshot's own directors; plus the device local order
parameter (mc_local_order) against the oracle on every site."""
import math
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import mclics

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(REPO, "mclics-backend_amd", "bin", "mc_sim")


def run(tmp_path, text):
    (tmp_path / "p.txt").write_text(text)
    print(f"mc_sim in {tmp_path}", flush=True)
    r = subprocess.run([SIM, "p.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    return r.returncode, r.stdout, r.stderr


def po(tmp_path):
    lines = (tmp_path / "po.dat").read_text().splitlines()
    return lines[0], [ln.split() for ln in lines[1:]]


def snapshot(path):
    lines = path.read_text().splitlines()
    rows = [ln.split(",") for ln in lines[1:]]
    ijk = np.array([[int(v) for v in r[:3]] for r in rows])
    n = np.array([[float(v) for v in r[3:6]] for r in rows], np.float64)
    S = np.array([float(r[6]) for r in rows])
    pt = np.array([int(r[7]) for r in rows], np.int32)
    return lines[0], ijk, n, S, pt


def check_snapshot(path, dims, bound):
    nx, ny, nz = dims
    hdr, ijk, n, S, pt = snapshot(path)
    assert hdr == "x,y,z,nx,ny,nz,S,pt"
    assert len(S) == nx * ny * nz
    s = ijk[:, 0] + nx * (ijk[:, 1] + ny * ijk[:, 2])
    assert np.array_equal(s, np.arange(nx * ny * nz))
    act = pt != 0
    assert np.all(np.isnan(S[~act])) and np.all(np.isfinite(S[act]))
    # the printed directors (6 significant digits) are unit vectors
    assert np.abs(np.linalg.norm(n[act], axis=1) - 1).max() < 2e-5
    # S from the oracle on the printed directors
    po_ = O.params(nx, ny, nz, bound=bound)
    flat = n.astype(np.float32).reshape(-1)
    rng = np.random.default_rng(0)
    for t in rng.choice(np.flatnonzero(act), min(200, act.sum()), replace=False):
        i, j, k = ijk[t]
        want = O.local_S(po_, flat, pt, int(i), int(j), int(k))
        assert abs(S[t] - want) < 2e-4, (path.name, t, S[t], want)
    return pt


def test_thermal_gpu_bulk(tmp_path):
    text = """Nx 8
Ny 8
Nz 8
MCS 20
MCT 30
Ti 1.2
Tf 1.0
dT -0.1
potential LL
evol thermalGPU
ic random
seed 3
geometry bulk
xbound periodic
ybound periodic
zbound periodic
"""
    rc, out, err = run(tmp_path, text)
    assert rc == 0, (out[-2000:], err[-2000:])
    hdr, rows = po(tmp_path)
    assert hdr == "T S varS E varE"
    nrow, Ts = O.float_loop(1.2, 1.0, -0.1)
    assert len(rows) == nrow
    for r, T in zip(rows, Ts):
        assert float(r[0]) == pytest.approx(float(T), rel=1e-5)
        S, varS, E, varE = (float(v) for v in r[1:])
        assert 0.0 <= S <= 1.0 and varS >= -1e-12 and -9.0 <= E <= 0.0 and varE >= -1e-12
        name = tmp_path / f"director_field_{int(100 * (float(T) + 1e-5))}.csv"
        assert name.exists(), name
        check_snapshot(name, (8, 8, 8), (1, 1, 1))


def test_electric_gpu_slab_ghrl(tmp_path):
    text = """Nx 8
Ny 6
Nz 6
MCS 10
MCT 20
Ti 0.5
initial_E 0.0
final_E 0.2
E_variation 0.1
EF_z 1
Aniso_E 1
potential ghrl
k11 1.0
k22 0.7
k33 1.3
p0 0
evol electricGPU
ic homogeneous
theta_0 90
phi_0 0
geometry slab
xbound periodic
ybound periodic
anchoring_type 0 homeotropic
W 0 4
anchoring_type 1 homeotropic
W 1 4
"""
    rc, out, err = run(tmp_path, text)
    assert rc == 0, (out[-2000:], err[-2000:])
    hdr, rows = po(tmp_path)
    assert hdr == "T S varS E varE"
    nrow, Es = O.float_loop(0.0, 0.2, 0.1)
    assert len(rows) == nrow
    for E in Es:
        name = tmp_path / f"director_field_{int(1000 * (float(E) + 1e-5))}.csv"
        assert name.exists(), name
        pt = check_snapshot(name, (8, 6, 6), (1, 1, 0))
        # slab faces carry the surface point types (src/geometry_slab.cpp:47-67)
        pt3 = pt.reshape(6, 6, 8)
        assert np.all(pt3[0] == 2) and np.all(pt3[-1] == 3) and np.all(pt3[1:-1] == 1)


def test_thermal_gpu_pear_sphere(tmp_path):
    """Pear in the sphere geometry (sites outside the droplet inactive, a
    homeotropic shell): po.dat carries the polarization column P
    (src/evolve_thermal.cu:72,103-105), |mean director| over all Nt sites
    (src/parameter_order.cpp:142-155), which must match the snapshot."""
    text = """Nx 12
Ny 12
Nz 12
MCS 10
MCT 30
Ti 1.0
Tf 0.8
dT -0.1
potential pear
A 1.0
B1 0.3
C 0.3
evol thermalGPU
ic random
seed 9
geometry sphere
xbound periodic
ybound periodic
zbound periodic
anchoring_type 0 homeotropic
W 0 2
"""
    rc, out, err = run(tmp_path, text)
    assert rc == 0, (out[-2000:], err[-2000:])
    hdr, rows = po(tmp_path)
    assert hdr == "T S varS E varE P"
    nrow, Ts = O.float_loop(1.0, 0.8, -0.1)
    assert len(rows) == nrow
    for r, T in zip(rows, Ts):
        assert len(r) == 6
        name = tmp_path / f"director_field_{int(100 * (float(T) + 1e-5))}.csv"
        _, ijk, n, S, pt = snapshot(name)
        assert (pt == 0).any() and (pt == 2).any() and (pt == 1).any()
        # inactive sites keep their directors in the snapshot; P averages all Nt sites
        P = np.linalg.norm(n.astype(np.float64).mean(axis=0))
        assert float(r[5]) == pytest.approx(P, abs=2e-5)


@pytest.mark.parametrize("mode", ["stepGPU", "quenchGPU"])
def test_step_and_quench_gpu(tmp_path, mode):
    text = f"""Nx 8
Ny 4
Nz 6
MCS 10
MCT 20
Ti 1.3
Tf 0.6
dT 0.5
fn 3
initial_output 2
potential LL
evol {mode}
ic random
seed 5
geometry bulk
xbound periodic
ybound periodic
zbound periodic
"""
    rc, out, err = run(tmp_path, text)
    assert rc == 0, (out[-2000:], err[-2000:])
    hdr, rows = po(tmp_path)
    assert hdr == "ii S varS E varE"
    assert [r[0] for r in rows] == ["2", "3", "4"]
    for ii in (2, 3, 4):
        check_snapshot(tmp_path / f"director_field_{ii}.csv", (8, 4, 6), (1, 1, 1))


@pytest.mark.parametrize("case", ["bulk", "holes", "free"])
def test_local_order_matches_oracle(case):
    rng = np.random.default_rng(21)
    nx, ny, nz = 8, 6, 10
    nt = nx * ny * nz
    v = rng.normal(size=(nt, 3))
    if case == "free":
        v[:, 2] += 2.0  # partly ordered, so S spans a range
    n = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)
    bound = (0, 1, 0) if case == "free" else (1, 1, 1)
    pt = np.ones(nt, np.int32)
    if case == "holes":
        pt[rng.random(nt) < 0.25] = 0
    cfg = mclics.Config(nx, ny, nz, bound=bound, T=1.0, seed=1)
    with mclics.Lattice(cfg, n, pt=pt) as lat:
        S = lat.local_order()
    po_ = O.params(nx, ny, nz, bound=bound)
    for s in range(nt):
        i, j, k = s % nx, (s // nx) % ny, s // (nx * ny)
        if pt[s] == 0:
            assert math.isnan(S[s])
            continue
        want = np.float32(O.local_S(po_, n, pt, i, j, k))
        assert abs(S[s] - want) <= 2 * np.spacing(want), (s, S[s], want)


RESUME_CASES = {
    # bulk LL on the slab kernel, a float temperature loop of 5 rows
    "thermal": """Nx 8
Ny 8
Nz 8
MCS 7
MCT 33
Ti 1.2
Tf 0.8
dT -0.1
potential LL
evol thermalGPU
ic random
seed 5
xbound periodic
ybound periodic
zbound periodic
""",
    # GHRL slab with anchoring faces and a field loop, on the segment kernel
    "electric": """Nx 8
Ny 6
Nz 6
MCS 5
MCT 21
Ti 0.5
initial_E 0.0
final_E 0.3
E_variation 0.1
EF_z 1
Aniso_E 1
potential ghrl
k11 1.0
k22 0.7
k33 1.3
p0 16
evol electricGPU
ic homogeneous
theta_0 90
phi_0 0
geometry slab
xbound periodic
ybound periodic
anchoring_type 0 homeotropic_ghrl
W 0 4
anchoring_type 1 homeotropic_ghrl
W 1 4
seed 9
""",
    # the CPU-named quench mode: reference sublattice order on the site kernel
    "quench": """Nx 6
Ny 6
Nz 4
MCS 4
MCT 10
Ti 1.3
Tf 0.7
dT 0.5
fn 3
potential LL
evol quench
ic random
seed 6
""",
}


@pytest.mark.parametrize("case", sorted(RESUME_CASES))
def test_checkpoint_resume_is_exact(tmp_path, case):
    # A run stopped after two rows with --checkpoint and continued with
    # --resume writes the same po.dat, byte for byte, and the same snapshots
    # as the run that was never interrupted.
    text = RESUME_CASES[case]
    full, part = tmp_path / "full", tmp_path / "part"
    full.mkdir()
    part.mkdir()
    for d in (full, part):
        (d / "p.txt").write_text(text)

    def sim(cwd, *args):
        print(f"mc_sim {case} {' '.join(args)} in {cwd}", flush=True)
        r = subprocess.run([SIM, *args, "p.txt"], cwd=cwd, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
        return r.stdout

    sim(full)
    sim(part, "--checkpoint", "ck.bin", "--max-rows", "2")
    assert len((part / "po.dat").read_text().splitlines()) == 3   # header + 2 rows
    out = sim(part, "--checkpoint", "ck.bin", "--resume", "ck.bin")
    assert "Resuming from ck.bin: 2 row(s) done" in out
    assert (part / "po.dat").read_bytes() == (full / "po.dat").read_bytes()
    snaps = sorted(p.name for p in full.glob("director_field_*.csv"))
    assert len(snaps) >= 3
    assert snaps == sorted(p.name for p in part.glob("director_field_*.csv"))
    for name in snaps:
        assert (part / name).read_bytes() == (full / name).read_bytes(), name
    # a resume of a finished run adds nothing
    before = (part / "po.dat").read_bytes()
    sim(part, "--checkpoint", "ck.bin", "--resume", "ck.bin")
    assert (part / "po.dat").read_bytes() == before
