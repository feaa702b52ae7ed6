"""The turbulence term divides by 500 (wind.hpp:83) without a divide
instruction sequence (swe_fdiv.h: product with RN(1/500) plus one exact
FMA correction, division kept for tiny, infinite and NaN inputs).  This
checks all 2^32 float32 inputs against the IEEE quotient on the CPU, with
the same header the kernels include."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to build the checker")
def test_div500_is_correctly_rounded_for_every_float(tmp_path):
    exe = str(tmp_path / "fdiv_check")
    # no contraction: x * c and the quotient must stay separate roundings;
    # -mfma only where the host has it (fmaf is exact either way)
    flags = ["-O3", "-fopenmp", "-ffp-contract=off"]
    probe = subprocess.run(["gcc", "-mfma", "-E", "-x", "c", "-", "-o", os.devnull], input="", text=True,
                           capture_output=True)
    if probe.returncode == 0 and "fma" in open("/proc/cpuinfo").read():
        flags.append("-mfma")
    subprocess.run(["gcc"] + flags + ["-o", exe, os.path.join(HERE, "fdiv_check.c"), "-lm"], check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("mismatches 0 "), res.stdout
