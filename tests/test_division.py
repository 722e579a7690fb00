"""Division by a precomputed reciprocal is the IEEE division.

The mixed generate kernel divides sample positions by the domain extent,
densities by the data range and density differences by the phase step with
mray_amd/csrc/div_rn.hpp (q = a y, r = fma(-q, b, a), fma(r, y, q) with
y = RN(1/b)) instead of the IEEE division sequence. Bounds and sample
values must stay bit-identical to the reference's, so this compiles that
header for the host and compares it with float division for every
dividend significand against 64 random and 64 adversarial divisor
significands, at three scales (2^23 * 128 * 3 = 3.2e9 quotients). The
result is invariant under power-of-two scaling inside the range the header
keeps, so the significands cover every exponent there. A quotient that
differs in one bit fails the test."""
import pathlib
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_div_rn_matches_ieee_division(tmp_path):
    exe = tmp_path / "chk"
    subprocess.run(["g++", "-O2", "-fopenmp", "-ffp-contract=off", "-o", str(exe),
                    str(HERE / "div_rn_check.cpp")], check=True)
    out = subprocess.run([str(exe), "64"], check=True, capture_output=True, text=True, timeout=600).stdout.split()
    pairs, bad, rejected, guards = map(int, out)
    assert pairs == (64 + 64) * 3 * (1 << 23)
    assert bad == 0, f"{bad} quotients differ from IEEE division"
    assert rejected == 0, "in-range operands must take the three-operation path"
    assert guards == 1, "divisors outside [2^-30, 2^30] and zero dividends must take the IEEE division"

