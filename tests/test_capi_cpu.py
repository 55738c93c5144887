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
