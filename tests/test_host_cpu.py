"""CPU tests of the C++17 host drop-in (mc_sim): parameter parsing with the
reference's keys, aliases and exit statuses, geometry set-up, initial
conditions and output formats.  Runs `mc_sim --dry-run`, which never touches a
device."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(REPO, "mclics-backend_amd", "bin", "mc_sim")
PARAMS = os.path.join(REPO, "tests", "golden", "ref_params")


def _ensure_built():
    if not os.path.exists(SIM):
        subprocess.run(["make", "-C", os.path.join(REPO, "mclics-backend_amd")], check=True)


def dry(tmp_path, text=None, path=None):
    _ensure_built()
    if text is not None:
        path = tmp_path / "p.txt"
        path.write_text(text)
    r = subprocess.run([SIM, "--dry-run", str(path)], cwd=tmp_path, capture_output=True, text=True)
    rep = None
    for line in r.stdout.splitlines():
        if line.startswith("DRYRUN "):
            rep = json.loads(line[7:])
    return r.returncode, rep, r.stderr


def read_csv(path):
    rows = open(path).read().splitlines()
    return rows[0], [r.split(",") for r in rows[1:]]


def sphere_mask(N):
    # src/geometry_sphere.cpp:28-54 restated with numpy float32 arithmetic
    H = np.float32(N // 2)
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    R = np.sqrt(((i - H) ** 2 + (j - H) ** 2 + (k - H) ** 2).astype(np.float32))
    return np.where(R < H - 1, 1, np.where(R < H + 1, 2, 0))


@pytest.mark.parametrize("name,dims,geo,valid", [
    ("test1.txt", (24, 24, 24), "Bulk", 24**3),
    ("test2.txt", (24, 24, 24), "Bulk", 24**3),
    ("test3.txt", (20, 20, 30), "Slab", 20 * 20 * 30),
    ("param.txt", (20, 20, 20), "Sphere", None),
    ("paramSphere.txt", (20, 20, 20), "Sphere", None),
    ("paramPear.txt", (20, 20, 20), "Sphere", None),
    ("test4.txt", None, "Sphere", None),
    ("input_parameters.txt", (61, 61, 61), "Sphere", None),
])
def test_reference_parameter_files(tmp_path, name, dims, geo, valid):
    rc, rep, err = dry(tmp_path, path=os.path.join(PARAMS, name))
    assert rc == 0, err
    if dims:
        assert (rep["nx"], rep["ny"], rep["nz"]) == dims
    assert rep["geometry"] == geo
    if geo == "Sphere":
        m = sphere_mask(rep["nz"]) if rep["nx"] == rep["ny"] == rep["nz"] else None
        if m is not None:
            assert rep["valid"] == int((m > 0).sum())
        assert rep["bound"] == [0, 0, 0]
    elif valid:
        assert rep["valid"] == valid
    if geo == "Slab":
        assert rep["bound"][2] == 0 and rep["n_surf"] == 2


def test_input_parameters_values(tmp_path):
    # trailing /* */ and ## comments, ghrl with E7 constants and p0 = 40
    rc, rep, err = dry(tmp_path, path=os.path.join(PARAMS, "input_parameters.txt"))
    assert rc == 0, err
    assert rep["MCS"] == 1000 and rep["MCT"] == 40000 and rep["fn"] == 30
    assert rep["Ti"] == pytest.approx(0.84) and rep["dT"] == pytest.approx(-0.01)
    po = O.params(2, 2, 2, potential=O.GHRL, k=(1.48, 1.0, 2.28, 40.0))
    assert rep["ghrl"] == pytest.approx(list(O.ghrl_tuple(po)), rel=1e-7)
    assert rep["surfaces"][0]["type"] == "homeotropic" and rep["surfaces"][0]["W"] == 4


def test_aliases_case_and_comments(tmp_path):
    text = """NX 6  # trailing words are ignored
    ny 4
    Nz 8 extra tokens
    # a comment line with keys: nx 99
    MODE thermalGPU
    potential LL
    EF_x 0.5
    Electric_field_y 0.25
    elecz 1
    Aniso_E 2
    initial_E 0.1
    final_E 0.9
    E_variation 0.2
    initial_output 7
    neighbour 1
    xbound periodic
    ybound PERIODIC
    zbound periodic
    seed 42
    ic homogeneous
    theta_0 90
    phi_0 0
    """
    rc, rep, err = dry(tmp_path, text)
    assert rc == 0, err
    assert (rep["nx"], rep["ny"], rep["nz"]) == (6, 4, 8)
    assert rep["evol"] == "thermalGPU" and rep["seed"] == 42 and rep["first_file"] == 7
    assert rep["elec"] == pytest.approx([0.5, 0.25, 1.0, 2.0])
    assert rep["E"] == pytest.approx([0.1, 0.9, 0.2])
    assert rep["bound"] == [1, 1, 1]
    hdr, rows = read_csv(tmp_path / "ic.csv")
    assert hdr == "x,y,z,nx,ny,nz,s,pt"
    assert len(rows) == 6 * 4 * 8
    v = np.array([[float(x) for x in r[3:6]] for r in rows])
    # homogeneous along (theta0=90, phi0=0) = +x  (src/ic.cpp:61-86)
    np.testing.assert_allclose(v, np.tile([1.0, 0.0, 0.0], (len(rows), 1)), atol=1e-7)
    assert all(r[6] == "1" and r[7] == "1" for r in rows)


@pytest.mark.parametrize("text,code", [
    ("nx 4\nbogus 3\n", 3),                       # unknown key
    ("nk 2\n", 1),                                # only nearest neighbours
    ("nx abc\n", 1),                              # non-numeric value
    ("evol sideways\n", 2),                       # unknown mode
    ("geometry torus\n", 2),                      # unknown geometry
    ("potential xyz\n", 2),                       # unknown potential
    ("xbound wrap\n", 2),                         # unknown boundary
    ("geometry slab\nanchoring_type 0 homeotropic\nW 0 1\n", 0),     # second surface undefined: status 0
    ("geometry slab\nanchoring_type 0 homeotropic\nanchoring_type 1 homeotropic\nW 0 1\n", 1),  # missing W
    ("geometry slab\nanchoring_type 0 weird\nanchoring_type 1 homeotropic\nW 0 1\nW 1 1\n", 2),
    ("geometry slab\nanchoring_type 0 rp\nanchoring_type 1 homeotropic\nW 0 1\nW 1 1\n", 1),  # rp needs angles
    ("ic ic_file\nic_file nothere.csv\n", 2),
    ("ic spiral\n", 2),
])
def test_exit_statuses(tmp_path, text, code):
    rc, rep, err = dry(tmp_path, text)
    assert rc == code, (rc, err)
    assert rep is None


def test_missing_parameter_file_is_status_5(tmp_path):
    rc, rep, err = dry(tmp_path, path=tmp_path / "absent.txt")
    assert rc == 5


def test_malformed_ic_file_is_status_3(tmp_path):
    (tmp_path / "ic_in.csv").write_text("x,y,z,nx,ny,nz\n0,0,0,1,0,0\n1,0,0,garbage\n")
    rc, rep, err = dry(tmp_path, "nx 2\nny 1\nnz 1\nic ic_file\nic_file ic_in.csv\n")
    assert rc == 3, err


def test_ic_file_roundtrip_from_snapshot(tmp_path):
    rc, rep, err = dry(tmp_path, "nx 4\nny 4\nnz 4\nseed 9\n")
    assert rc == 0
    _, rows = read_csv(tmp_path / "director_field_dryrun.csv")
    (tmp_path / "start.csv").write_text(open(tmp_path / "director_field_dryrun.csv").read())
    rc, rep, err = dry(tmp_path, "nx 4\nny 4\nnz 4\nic ic_file\nic_file start.csv\n")
    assert rc == 0, err
    _, rows2 = read_csv(tmp_path / "ic.csv")
    assert [r[:6] for r in rows] == [r[:6] for r in rows2]


def test_random_ic_matches_python_sampler(tmp_path):
    import mclics
    rc, rep, err = dry(tmp_path, "nx 10\nny 6\nnz 4\nseed 17\n")
    assert rc == 0
    _, rows = read_csv(tmp_path / "ic.csv")
    got = np.array([[float(x) for x in r[3:6]] for r in rows], np.float64)
    want = mclics.random_ic(10, 6, 4, seed=17).reshape(-1, 3).astype(np.float64)
    np.testing.assert_allclose(got, want, atol=6e-6)  # %g keeps 6 significant digits
    np.testing.assert_allclose(np.linalg.norm(want, axis=1), 1.0, atol=2e-7)


def test_snapshot_format_and_local_order_parameter(tmp_path):
    # free boundaries and a sphere (NaN outside, local S from active neighbours only)
    rc, rep, err = dry(tmp_path, "nx 8\nny 8\nnz 8\ngeometry sphere\nanchoring_type 0 homeotropic\nW 0 2\nseed 3\n")
    assert rc == 0, err
    hdr, rows = read_csv(tmp_path / "director_field_dryrun.csv")
    assert hdr == "x,y,z,nx,ny,nz,S,pt"
    assert len(rows) == 512
    n = np.array([[float(x) for x in r[3:6]] for r in rows], np.float32).reshape(-1)
    pt = np.array([int(r[7]) for r in rows], np.int32)
    p = O.params(8, 8, 8, bound=(0, 0, 0))
    for s, r in enumerate(rows):
        i, j, k = (int(x) for x in r[:3])
        assert s == i + 8 * (j + 8 * k)
        if pt[s] == 0:
            assert r[6] == "nan"
        else:
            # the CSV holds 6 digits of each director; recompute from those
            assert float(r[6]) == pytest.approx(O.local_S(p, n, pt, i, j, k), abs=2e-5)


def test_slab_point_types_and_normals(tmp_path):
    rc, rep, err = dry(tmp_path, "nx 4\nny 3\nnz 5\ngeometry slab\nanchoring_type 0 homeotropic\n"
                                 "anchoring_type 1 rp_ghrl\nW 0 4\nW 1 2.5\ntheta_s 1 90\nphi_s 1 30\n"
                                 "potential ghrl\nk11 1\nk22 0.7\nk33 1.3\np0 16\nzbound periodic\n")
    assert rc == 0, err
    assert rep["bound"][2] == 0  # slab forces z free (src/geometry_slab.cpp:23)
    s1 = rep["surfaces"][1]
    assert s1["type"] == "rp_ghrl" and s1["W"] == 2.5
    t, ph = math.radians(90), math.radians(30)
    assert s1["n_s"] == pytest.approx([math.cos(ph) * math.sin(t), math.sin(ph) * math.sin(t), math.cos(t)],
                                      abs=1e-6)
    assert rep["ghrl"][4] == pytest.approx(-1.030835, rel=1e-6)
    _, rows = read_csv(tmp_path / "ic.csv")
    pts = [int(r[7]) for r in rows]
    for r, pv in zip(rows, pts):
        z = int(r[2])
        assert pv == (2 if z == 0 else 3 if z == 4 else 1)


def test_custom_geometry_file(tmp_path):
    lines = ["x,y,z,nx,ny,nz,pt"]
    lines.append("0,0,0,0,0,1,0")
    lines.append("1,0,0,0,0,1,2")
    lines.append("2,0,0,1,0,0,3")
    (tmp_path / "b.csv").write_text("\n".join(lines) + "\n")
    rc, rep, err = dry(tmp_path, "nx 3\nny 2\nnz 2\ngeometry custom\nboundary_file b.csv\n"
                                 "anchoring_type 0 homeotropic\nanchoring_type 1 fg\nW 0 1\nW 1 2\n")
    assert rc == 0, err
    assert rep["n_surf"] == 2 and rep["valid"] == 11
    rc, rep, err = dry(tmp_path, "nx 2\nny 2\nnz 2\ngeometry custom\nboundary_file b.csv\n")
    assert rc == 10  # point out of the box


# --- checkpoints ------------------------------------------------------------
# File layout (little-endian): "MCLCKPT1", u32 version 1, i32 Nx Ny Nz,
# char evol[20], u64 fingerprint, u32 next sweep, f32 ang_var, i32 rows done,
# f32 next loop value, f32 directors[3*Nt], u64 FNV-1a of all preceding bytes.
CKPT_PARAMS = "nx 4\nny 4\nnz 4\nevol thermalGPU\nTi 1.0\nTf 0.8\ndT -0.1\nMCT 2\nMCS 2\nseed 4\n"


def fnv1a(b):
    h = 0xcbf29ce484222325
    for c in b:
        h = ((h ^ c) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def ckpt_bytes(nt, magic=b"MCLCKPT1", fingerprint=0, good_sum=True):
    import struct
    body = magic + struct.pack("<I3i", 1, 4, 4, 4) + b"thermalGPU".ljust(20, b"\0")
    body += struct.pack("<QIfif", fingerprint, 40, 0.5, 1, 0.9) + np.tile(np.float32([0, 0, 1]), nt).tobytes()
    s = fnv1a(body) ^ (0 if good_sum else 1)
    return body + struct.pack("<Q", s)


def run_sim(tmp_path, *args):
    _ensure_built()
    (tmp_path / "p.txt").write_text(CKPT_PARAMS)
    r = subprocess.run([SIM, *args, "p.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("case,needle", [
    ("missing", "cannot open checkpoint"),
    ("short", "does not fit a 4x4x4 lattice"),
    ("magic", "not a checkpoint file"),
    ("checksum", "checksum mismatch"),
    ("fingerprint", "other parameters"),
])
def test_resume_rejects_bad_checkpoints_before_the_device(tmp_path, case, needle):
    # Every check runs on the host before the device is opened, so each of
    # these ends with the configuration status 2 even without a GPU.
    f = tmp_path / "ck.bin"
    if case == "short":
        f.write_bytes(ckpt_bytes(64)[:-4])
    elif case == "magic":
        f.write_bytes(ckpt_bytes(64, magic=b"NOTACKPT"))
    elif case == "checksum":
        f.write_bytes(ckpt_bytes(64, good_sum=False))
    elif case == "fingerprint":
        f.write_bytes(ckpt_bytes(64, fingerprint=12345))
    (tmp_path / "po.dat").write_text("T S varS E varE\nkept\n")
    rc, out = run_sim(tmp_path, "--resume", "ck.bin")
    assert rc == 2, out
    assert needle in out
    # a resumed run keeps the previous outputs
    assert (tmp_path / "po.dat").read_text() == "T S varS E varE\nkept\n"


@pytest.mark.parametrize("args", [["--max-rows", "0", "p.txt"], ["--bogus", "p.txt"], ["p.txt", "--checkpoint"]])
def test_bad_checkpoint_options(tmp_path, args):
    _ensure_built()
    (tmp_path / "p.txt").write_text(CKPT_PARAMS)
    r = subprocess.run([SIM, *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stdout + r.stderr
