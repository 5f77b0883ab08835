"""The adjoints of the constitutive laws in the differentiable run
(mpm_diff.hip: radial return of the perfectly plastic and J2-hardening laws,
NeoHookean total stress) against central finite differences of the same
forward functions, compiled for the host: 4000 random radial returns (both
yield branches, with and without hardening) and 2000 random deformation
gradients.  The errors are the finite differences' own (they fall 100-fold
when the step falls 10-fold)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_law_adjoints_match_central_differences(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "law_adjoints"
    subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "law_adjoints.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, env=env).stdout
    vals = dict(line.split()[:2] for line in out.strip().splitlines())
    print(out, flush=True)
    assert int(vals["plastic"]) > 1000 and int(vals["harden"]) > 500
    assert float(vals["radial_return"]) <= 1e-5
    assert float(vals["neohookean"]) <= 1e-6
