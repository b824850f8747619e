"""The integrate kernel replaces two of its three IEEE divisions per voxel by
a correctly rounded reciprocal and Markstein's FMA correction
(div_by_rcp, recon_common.hpp).  This runs tests/div_correction_check.c
(exhaustive over the weight-update table's divisors, random elsewhere) with
the host's correctly rounded fmaf; the GPU parity tests then compare the
kernel's D and W bit for bit with the oracle's IEEE divisions."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_quotient_correction_equals_ieee_division(tmp_path):
    exe = tmp_path / "div_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math",
                    str(HERE / "div_correction_check.c"), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split() == ["table", "0", "pixel", "0", "wide", "0"]
