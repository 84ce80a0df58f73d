"""MI355X-native hot path of the lunar-orbit interferometer simulator
(SSshuishui/viss_recon_shade): orbits -> uvw, Moon occultation mask,
shaded direct-DFT vis-gen and recon, as gfx950 HIP kernels behind the C-ABI
of include/vrs.h (libvrs.so)."""
from ._lib import (VRS_MFMA_SPLIT16, VRS_MIXED_F64PHASE, VRS_PLAIN_F32, Knobs, VrsError,
                   load)
from .api import Context, Geometry, knobs
from . import workload

__all__ = ["Context", "Geometry", "Knobs", "VrsError", "knobs", "load", "workload",
           "VRS_PLAIN_F32", "VRS_MIXED_F64PHASE", "VRS_MFMA_SPLIT16"]
