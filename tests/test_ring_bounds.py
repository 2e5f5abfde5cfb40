"""The batched estimation's DC-ring bounds (eb_ring_pair, run by the step and
set-up kernels) agree with commonRegion's per-pixel ring test for every squared radius.

The ratio sqrt(d2) / |k| is a non-decreasing function of the integer d2, so
the set the ring cuts is d2 < lo or d2 > hi; the host restatement of the
kernel's two binary searches is compared with the direct test at every d2 of
sizes 16..4096, including integral |k| (ratios exactly at the thresholds)."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("cc") is None, reason="no host C compiler")
def test_ring_bounds_match_per_pixel_test(tmp_path):
    exe = tmp_path / "ring_bounds_check"
    subprocess.run(["cc", "-O2", "-ffp-contract=off", "-o", str(exe), str(HERE / "ring_bounds_check.c"), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "120"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert " 0 mismatches" in out.stdout
