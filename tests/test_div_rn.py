"""The FMA-corrected division of the device helpers (div_rn in
fairsim-vigor_amd/csrc/simmath.h) returns the IEEE quotient bit for bit.

The same operation sequence is compiled for the host (gcc, FMA from libm,
contraction off) and compared with IEEE division over random operands of
every exponent, the divisors the kernels divide by, operands next to
multiples of the divisor (where a faulty rounding step would show) and the
special values the guard sends to the plain division."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("cc") is None, reason="no host C compiler")
def test_div_rn_matches_ieee_division(tmp_path):
    exe = tmp_path / "div_rn_check"
    subprocess.run(["cc", "-O2", "-ffp-contract=off", "-o", str(exe), str(HERE / "div_rn_check.c"), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert " 0 mismatches" in out.stdout
