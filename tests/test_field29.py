"""The reduced-radix field (9 limbs of 29 bits, Montgomery R = 2^261, lazily
reduced: nkr-zkp_amd/csrc/field29.h) and its mixed XYZZ addition
(curve29.h), compiled as host code and checked against the canonical
32-bit-limb arithmetic of field.h / curve.h, whose products the golden
vectors and the oracle pin: conversions both ways, products at every input
bound the addition formulas use, lazy sums and differences, and chains of
mixed additions through doublings and cancellations with bit-identical
canonical coordinates after every step."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


def test_field29_against_canonical_arithmetic(tmp_path):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("needs hipcc")
    exe = str(tmp_path / "field29_check")
    subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950",
                    "-I" + os.path.join(ROOT, "nkr-zkp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "field29_check.hip"), "-o", exe], check=True)
    res = subprocess.run([exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("ok ")
    checks, doublings, cancels = [int(w) for w in res.stdout.replace("(", " ").split() if w.isdigit()][:3]
    assert checks > 300000 and doublings > 100 and cancels > 100
