"""CPU builds of the oracle and of the host drop-in under AddressSanitizer and
UndefinedBehaviorSanitizer (gcc/g++, host code only; no GPU code is built
with them), run over the paths the CPU suite can reach: the oracle's sweeps,
passes, energies, order parameters, float loop, expf, Philox and MLP
gradient (tests/native/oracle_san.c); the drop-in's parser, geometries,
initial conditions, snapshot writer and checkpoint checks through
`mc_sim --dry-run` on every reference parameter file and through rejected
resumes.  Any sanitizer report fails the run (-fno-sanitize-recover)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PKG = os.path.join(REPO, "mclics-backend_amd")
PARAMS = os.path.join(HERE, "golden", "ref_params")
SAN = ["-O1", "-g", "-fopenmp", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
       "-fno-sanitize-recover=undefined"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

needs_gcc = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="needs gcc/g++")


def _clean(r):
    text = r.stdout + r.stderr
    assert "runtime error" not in text, text[-4000:]
    assert "AddressSanitizer" not in text and "LeakSanitizer" not in text, text[-4000:]
    return text


@needs_gcc
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_san")
    subprocess.run(["gcc", "-std=c11", *SAN, "-o", exe, os.path.join(HERE, "native", "oracle_san.c"),
                    os.path.join(REPO, "oracle", "mc_oracle.c"), os.path.join(REPO, "oracle", "mlp_oracle.c"), "-lm"],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    _clean(r)
    assert r.returncode == 0 and "oracle sanitizer run ok" in r.stdout, r.stdout + r.stderr


@pytest.fixture(scope="module")
def sim_san(tmp_path_factory):
    lib = os.path.join(PKG, "lib", "libmclics_hip.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", PKG, "lib/libmclics_hip.so"], check=True)
    exe = str(tmp_path_factory.mktemp("san") / "mc_sim_san")
    host = [os.path.join(PKG, "host", f) for f in ("main.cpp", "io.cpp", "geometry.cpp", "ic.cpp", "evolve.cpp",
                                                     "simulator.cpp")]
    subprocess.run(["g++", "-std=c++17", *SAN, "-o", exe, *host, "-L" + os.path.join(PKG, "lib"), "-lmclics_hip",
                    "-Wl,-rpath," + os.path.join(PKG, "lib"), "-Wl,-rpath,/opt/rocm/lib", "-L/opt/rocm/lib",
                    "-lamdhip64"], check=True)
    return exe


@needs_gcc
@pytest.mark.parametrize("name", sorted(f for f in os.listdir(PARAMS) if f.endswith(".txt")))
def test_dry_run_under_sanitizers(sim_san, tmp_path, name):
    r = subprocess.run([sim_san, "--dry-run", os.path.join(PARAMS, name)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300, env=ENV)
    _clean(r)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (tmp_path / "ic.csv").exists() and (tmp_path / "director_field_dryrun.csv").exists()


@needs_gcc
def test_rejected_resume_under_sanitizers(sim_san, tmp_path):
    (tmp_path / "p.txt").write_text("nx 4\nny 4\nnz 4\nevol thermalGPU\nTi 1.0\nTf 0.8\ndT -0.1\n")
    for content in (b"", b"MCLCKPT1" + b"\0" * 100, b"x" * (8 + 4 + 12 + 20 + 8 + 16 + 12 * 64 + 8)):
        (tmp_path / "ck.bin").write_bytes(content)
        r = subprocess.run([sim_san, "--resume", "ck.bin", "p.txt"], cwd=tmp_path, capture_output=True, text=True,
                           timeout=300, env=ENV)
        _clean(r)
        assert r.returncode == 2, r.stdout[-2000:] + r.stderr[-2000:]
