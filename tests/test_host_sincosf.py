"""The device library's restatement of the host sinf / cosf is bit-exact.

The reference's FP32 MESE preparation and prediction coding call the host C
library's sinf / cosf; the GPU kernels reproduce them with
mray_amd/csrc/host_sincosf.hpp. This compiles that header for the host and
compares it with the C library on every float with |x| < 120 (2.2e9
arguments per function, a few seconds on 8 cores)."""
import pathlib
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_sincosf_bit_exact(tmp_path):
    exe = tmp_path / "chk"
    subprocess.run(["g++", "-O2", "-fopenmp", "-ffp-contract=off", "-o", str(exe),
                    str(HERE / "host_sincosf_check.cpp"), "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=600).stdout.split()
    n, ms, mc = map(int, out)
    assert n == 2 * 0x42f00000
    assert ms == 0 and mc == 0, f"sinf mismatches {ms}, cosf mismatches {mc}"
