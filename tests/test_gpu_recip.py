"""The FP32 density reciprocal of the collision (d3q19.h recip_rn<float>:
the hardware estimate and one Newton-Raphson correction inside
[2^-125, 2^126), the quotient elsewhere) against the IEEE quotient 1.0f / x,
for all 2^32 float bit patterns (bgk3d_hit_amd/bin/recip_check, built with
the library).  The exhaustive comparison is what makes the short form a
bit-exact stand-in for the reference's rhoinv = uno / rho; the parity tests
then check whole lattices against the oracle's IEEE division."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "bgk3d_hit_amd", "bin", "recip_check")


def test_recip_rn_equals_quotient_for_every_float(gpu):
    print("[gpu-case] recip_check over 2^32 floats", flush=True)
    sys.stdout.flush()
    assert os.path.exists(TOOL), "bin/recip_check missing: build the library first"
    r = subprocess.run([TOOL], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
