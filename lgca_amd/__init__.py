"""MI355X lattice-gas engine: HPP / FHP-I / FHP-II / FHP-III on gfx950 bit-planes.

The compute path is liblgca_hip.so (hand-written HIP kernels behind the C ABI of
include/lgca.h); this package is its Python front end.  Importing it does not
load the library; the first call that needs it does, and fails loudly when the
library has not been built.
"""
from ._capi import LgcaError, lib
from .batch import Batch, Job, job_bytes
from .lattice import Lattice, derive_parameters, disc

__all__ = ["Batch", "Job", "Lattice", "LgcaError", "derive_parameters", "disc", "job_bytes", "lib"]
