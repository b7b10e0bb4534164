"""MI355X solver core for the NEMOH Solver stage.

The compute path is the HIP library libnemoh_core.so (C ABI in
include/nemoh_core.h); ``nemoh_amd.core`` binds it, ``nemoh_amd.solver``
mirrors the reference Solver's interface, ``nemoh_amd.mesh`` and
``nemoh_amd.io`` handle meshes and the boundary file formats.
"""
from .core import FP64, MIXED, Core, NemohError  # noqa: F401
