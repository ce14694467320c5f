"""The reciprocal quotient of the fused HS kernels (hindemith_amd/csrc/hs_div.h)
against the host's IEEE fp32 division.

fast_div_check.cpp compiles the same header the kernels use with g++ and
checks, on ~370 million operand pairs (every significand of b at four binades,
random pairs over and beyond the accepted ranges with zeros, subnormals,
infinities and NaNs, and the hardest near-midpoint quotients), that every
quotient the guard accepts equals a / b bit for bit, and that the magnitudes of
the Horn-Schunck update are all accepted.  The GPU side (the same code on
gfx950, in every variant) is covered by the bit-exact Jacobi tests, including
test_hs_jacobi_division_edge_cases.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_reciprocal_quotient_is_correctly_rounded(tmp_path):
    exe = tmp_path / "fast_div_check"
    flags = ["-O2", "-std=c++17", "-ffp-contract=off", "-Wall"]
    # a hardware fma when the host has one (glibc's fmaf is exact either way)
    probe = subprocess.run(["g++", "-mfma", "-E", "-x", "c++", "/dev/null"],
                           capture_output=True)
    if probe.returncode == 0 and "fma" in open("/proc/cpuinfo").read():
        flags.append("-mfma")
    subprocess.run(["g++", *flags, "-o", str(exe), os.path.join(HERE, "fast_div_check.cpp")],
                   check=True)
    r = subprocess.run([str(exe), "1"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout and "hs-range-rejected 0" in r.stdout
