"""The short x / 12 of the nonlinear momentum kernels, against IEEE division.

gpu-comcot_amd/csrc/fastdiv.h computes x / 12 as RN(1/12) * x with one
residual correction.  Every float is checked against the IEEE quotient on
the host (the same IEEE operations the device performs): it must be equal
for all |x| >= 2^-124 and for both zeros, and the only differences left are
quotients in the denormal range.  The device form of both short divisions
is checked on the GPU in tests/test_gpu_fastdiv.py.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gpu-comcot_amd", "csrc")


def test_div12_equals_ieee_division_for_every_float(tmp_path):
    exe = tmp_path / "fastdiv_check"
    flags = ["-O2", "-ffp-contract=off", "-fopenmp"]
    if "fma" in open("/proc/cpuinfo").read().split():
        flags.append("-mfma")
    subprocess.run(["g++", *flags, "-I", CSRC, os.path.join(HERE, "fastdiv_check.cpp"), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True,
                         env=dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1))))
    bad, bad_any, worst = out.stdout.split()
    assert int(bad) == 0, f"{bad} floats with |x| >= 2^-124 or x = 0 differ"
    assert float.fromhex(worst) < 2.0 ** -124
    assert int(bad_any) > 0   # the denormal quotients do differ: the check has teeth
