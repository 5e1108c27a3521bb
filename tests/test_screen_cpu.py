"""The GHRL screen's FP32 energy difference (mclics-backend_amd/csrc/
mc_ghrl_screen.h) against the reference's arithmetic, on the host.

The sweep kernels decide a GHRL move from that FP32 dE unless the uniform
lies within a band around exp(-dE/T) whose width assumes the screen is
within 920 u M of the reference's dE (u = 2^-24, M the bond bound).  The
native checker compares the header's function, compiled for the host with
the same FP32 operations, with the oracle's bonds (the reference's
float/double mix, src/potential.cpp:22-38) over two million random moves,
neighbourhoods, free-face bond masks and elastic constants, chiral and not;
a wrong term or sign shows up as a distance of order 1/u."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the native checker")
def test_ghrl_screen_within_its_bound(tmp_path):
    exe = str(tmp_path / "ghrl_screen_check")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "native", "ghrl_screen_check.cpp"),
                    "-x", "c", os.path.join(REPO, "oracle", "mc_oracle.c"), "-lm"], check=True)
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "violations 0" in r.stdout
    worst = float(r.stdout.split("worst ")[1].split()[0])
    assert worst < 100.0  # the proven bound is 920; typical distances are a few u M
