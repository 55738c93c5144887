"""MI355X (gfx950) state-vector engine for the lightning.qubit hot path.

``lightning_amd.lightning_qubit_ops`` mirrors ``pennylane_lightning.lightning_qubit_ops``;
the compute runs in ``_lib/liblq_amd.so`` (HIP kernels behind a C-ABI).
"""
from . import lightning_qubit_ops  # noqa: F401
from .lightning_qubit_ops import (  # noqa: F401
    LightningException,
    MeasurementsC64,
    MeasurementsC128,
    StateVectorC64,
    StateVectorC128,
    algorithms,
    observables,
)

__version__ = "0.1.0"
