"""gfx950 (MI355X) LBMForce platform: the Python side of the C-ABI in include/lbm_amd.h.

Mirrors the reference's openmmlbm module (python/openmmlbm.i) for the one-GPU path: LBMForce with its
enums and setters, and a minimal System/VerletIntegrator/Context host that drives the kernel the way
OpenMM's ContextImpl does.
"""
from ._lib import LBMError, LBMParams, LBMForceSettings, load_library, LIBRARY_PATH
from .force import LBMForce
from .context import Context, System, VerletIntegrator

__all__ = ["LBMError", "LBMParams", "LBMForceSettings", "load_library", "LIBRARY_PATH", "LBMForce", "Context",
           "System", "VerletIntegrator"]
