"""Plan reuse on the CPU: a plan made for one program, with its coefficients
refilled from a program of the same structure and other angles, equals the
plan made directly for that program, bit for bit (tests/jit_emu/plan_cache_check.cpp:
one and two states, with and without split rotations, generator dots, and a
growth bound on scaled rotations that some refreshes cross, which must then
be refused so that the caller plans afresh)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pennylane-lightning_amd")
LIB = os.path.join(PKG, "lightning_amd", "_lib")


def test_refreshed_plans_equal_fresh_plans(tmp_path):
    if not os.path.exists(os.path.join(LIB, "liblq_amd.so")):
        pytest.skip("liblq_amd.so not built")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("no host compiler")
    obj, exe = str(tmp_path / "pcc.o"), str(tmp_path / "pcc")
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I", os.path.join(PKG, "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-c", os.path.join(ROOT, "tests", "jit_emu", "plan_cache_check.cpp"), "-o", obj], check=True,
                   capture_output=True)
    subprocess.run([hipcc, obj, "-L", LIB, "-llq_amd", f"-Wl,-rpath,{LIB}", "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "36 cases" in r.stdout and ", 0 failures" in r.stdout
