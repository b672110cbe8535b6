"""The quad adjoint's split form (forward mode in the in-plane coordinates and
properties, reverse mode for the frame, chain rule through the frame's Jacobian)
against forward mode through the whole element per coordinate, on the CPU: the
element routines in jaxsso_amd/csrc/element_math.hpp are host+device code, so a
small host program built with hipcc runs both forms on 200 random warped and
square elements (1-3 load cases, lambda = u/2 and a general lambda).  The
per-coordinate form is the one the GPU gradient tests pin against the oracle and
finite differences."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "jaxsso_amd", "csrc")
INCLUDE = os.path.join(HERE, "..", "include")


@pytest.fixture(scope="module")
def check_exe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("adj") / "adjoint_split_check")
    r = subprocess.run([hipcc, "-O2", "-std=c++17", "-I", CSRC, "-I", INCLUDE,
                        os.path.join(HERE, "native", "adjoint_split_check.cpp"), "-o", out],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def _run(exe, plant):
    r = subprocess.run([exe, str(plant)], capture_output=True, text=True, timeout=300)
    worst = float(r.stdout.strip().splitlines()[-1].split()[-1])
    return r.returncode, worst


def test_split_adjoint_matches_per_coordinate_form(check_exe):
    rc, worst = _run(check_exe, 0)
    assert rc == 0 and worst < 1e-12, worst


@pytest.mark.parametrize("plant", [1, 2])
def test_split_adjoint_check_catches_planted_mistakes(check_exe, plant):
    rc, worst = _run(check_exe, plant)
    assert rc != 0 and worst > 1e-8, worst
