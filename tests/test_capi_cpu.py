"""The C-ABI library loads and exports every entry point include/lq_amd.h declares
(CPU only: nothing here calls into the device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lq_amd.h")
LIB = os.path.join(ROOT, "pennylane-lightning_amd", "lightning_amd", "_lib", "liblq_amd.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(lq_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "lq_sv_create" in names and "lq_adjoint_jacobian" in names
    assert len(names) >= 35


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblq_amd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblq_amd.so not built")
def test_python_mirror_surface():
    from lightning_amd import lightning_qubit_ops as L
    for cls in ("StateVectorC64", "StateVectorC128", "MeasurementsC64", "MeasurementsC128"):
        assert hasattr(L, cls)
    for g in L.GATE_NAMES:
        assert callable(getattr(L.StateVectorC128, g))
    for name in ("NamedObsC128", "HermitianObsC128", "TensorProdObsC128", "HamiltonianC128",
                 "NamedObsC64", "HamiltonianC64"):
        assert hasattr(L.observables, name)
    for name in ("OpsStructC128", "AdjointJacobianC128", "create_ops_listC128", "OpsStructC64",
                 "AdjointJacobianC64", "create_ops_listC64"):
        assert hasattr(L.algorithms, name)


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblq_amd.so not built")
def test_one_hip_runtime_with_torch():
    """The binding imports torch before the engine, so the engine binds to the
    HIP runtime libraries torch bundles (same sonames): one runtime per process."""
    import subprocess
    import sys
    code = ("import re, lightning_amd._capi\n"
            "m = {l.split()[-1] for l in open('/proc/self/maps') if re.search(r'libamdhip64|libhsa-runtime64', l)}\n"
            "print(len([x for x in m if 'amdhip64' in x]), len([x for x in m if 'hsa-runtime64' in x]))\n")
    env = dict(os.environ, PYTHONPATH=os.path.join(ROOT, "pennylane-lightning_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.split() == ["1", "1"], r.stdout
