"""Generated pass kernels run on the CPU under emulation (tests/jit_emu):
the engine's planner and kernel generator in liblq_amd.so produce each pass
of three programs (hardware-efficient layers with dots; plus controlled,
parity, phase and SWAP gates; the same on two, three and four states with dots between them,
as in the adjoint sweep with one to three observables; and layers followed by
named expectation values),
the kernels are compiled for the host, run with one
host thread per GPU thread, and compared with a naive application of the same
ops (complex128 within 1e-10, complex64 within 1e-3 relative)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "jit_emu")
PKG = os.path.join(ROOT, "pennylane-lightning_amd")
LIB = os.path.join(PKG, "lightning_amd", "_lib")


@pytest.fixture(scope="module")
def emu_binary(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB, "liblq_amd.so")):
        pytest.skip("liblq_amd.so not built")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) or not shutil.which("g++"):
        pytest.skip("no host compiler")
    d = tmp_path_factory.mktemp("emu")
    obj, exe = str(d / "emu_check.o"), str(d / "emu_check")
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I", os.path.join(PKG, "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-c", os.path.join(EMU, "emu_check.cpp"), "-o", obj], check=True, capture_output=True)
    subprocess.run([hipcc, obj, "-L", LIB, "-llq_amd", f"-Wl,-rpath,{LIB}", "-ldl", "-o", exe], check=True,
                   capture_output=True)
    return exe, str(d)


@pytest.mark.parametrize("case,dtype,lds32,look,n", [
    (0, 1, 0, 1, 14), (1, 1, 0, 1, 14), (2, 1, 0, 1, 14), (1, 0, 0, 1, 14), (2, 0, 0, 1, 14),
    (0, 1, 1, 1, 14), (2, 1, 1, 1, 14), (2, 0, 1, 1, 14), (0, 1, 0, 0, 14), (1, 0, 0, 0, 14), (2, 1, 0, 0, 14),
    (0, 1, 1, 2, 16), (1, 0, 1, 2, 16), (2, 1, 1, 2, 16), (0, 1, 1, 3, 14), (1, 0, 1, 3, 14),
    (3, 0, 1, 1, 17), (3, 0, 1, 3, 17), (3, 1, 1, 3, 17), (2, 1, 1, 4, 14), (1, 0, 1, 4, 14), (0, 1, 1, 5, 14),
    (0, 1, 1, 6, 14), (2, 1, 1, 6, 14), (1, 0, 1, 6, 14), (1, 0, 1, 7, 14), (3, 0, 1, 7, 17),
    (4, 1, 1, 1, 13), (4, 0, 1, 1, 13), (5, 1, 1, 1, 13), (5, 0, 1, 1, 13)])
def test_generated_kernels_match_naive_application(emu_binary, case, dtype, lds32, look, n):
    exe, work = emu_binary
    # lds32: transposes through 32 KiB of LDS (emulated at exactly the size the kernel declares);
    # look 0: tile bits from the first ready op instead of the planner's lookahead;
    # look 2: lookahead, and loads/stores in the first/last stage's mapping wherever they coalesce
    # (16 qubits: passes whose first and last stages leave the low bits to the lanes)
    # look 3: lookahead, and the state a basis state not yet written (deferred reset): the leading
    # one-qubit gates are folded into product factors and the first pass forms the product state
    # case 3 at 17 qubits and complex64: the program of the complex64 expval GPU test
    # cases 4 and 5: three and four states (the adjoint sweep with two and three observables at once)
    env = dict(os.environ, LQ_JIT_LDS32=str(lds32), EMU_LOOKAHEAD=str(min(look, 1)),
               LQ_JIT_DIRECT="2" if look == 2 else "1")
    # look 4: bit positions read from the table after the coefficients (LQ_JIT_LIFT); look 5: that and
    # the deferred reset; look 6: register bits swapped with any lane bit by shuffles (LQ_JIT_LANES=2);
    # look 7: only lane bits 4 and 5, by row swaps (LQ_JIT_LANES=1)
    if look in (4, 5):
        env["EMU_LIFT"] = "1"
    if look in (3, 5):
        env["EMU_BASIS"] = "5"
    if look == 6:
        env["LQ_JIT_LANES"] = "2"
    if look == 7:
        env["LQ_JIT_LANES"] = "1"
    r = subprocess.run([exe, str(case), str(dtype), str(n), "3", EMU, work], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
