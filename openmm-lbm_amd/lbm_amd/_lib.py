"""ctypes declarations of the C-ABI in include/lbm_amd.h.

The structures mirror lbm_params (the converted LBMLatticeParameters, openmmapi/include/LBMKernels.h:30-92)
and lbm_force_settings (the user settings of LBMForce, openmmapi/include/LBMForce.h:41-691).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "liblbm_amd.so")

LBM_OK = 0
LBM_ERROR_INVALID = -1
LBM_ERROR_RUNTIME = -2
LBM_ERROR_MACH = -3
LBM_ERROR_UNSUPPORTED = -4

STORAGE_FP32, STORAGE_FP16, STORAGE_BF16, STORAGE_FP64 = 0, 1, 2, 3

_c_double3 = ctypes.c_double * 3


class LBMParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("density", ctypes.c_double),
        ("tau", ctypes.c_double), ("omega", ctypes.c_double),
        ("initial_velocity", _c_double3),
        ("body_acceleration", _c_double3),
        ("friction", ctypes.c_double),
        ("kT", ctypes.c_double),
        ("drag_scheme", ctypes.c_int),
        ("interpolation_stencil", ctypes.c_int),
        ("fluid_fluctuations", ctypes.c_int),
        ("fluid_kT", ctypes.c_double),
        ("random_number_seed", ctypes.c_int),
        ("momentum_removal_frequency", ctypes.c_int),
        ("mach_check_frequency", ctypes.c_int),
        ("mach_number_limit", ctypes.c_double),
        ("wall_scheme", ctypes.c_int),
        ("face_boundary", ctypes.c_int * 6),
        ("face_velocity", _c_double3 * 6),
        ("face_density", ctypes.c_double * 6),
        ("storage", ctypes.c_int),
        ("density_halo_exchange", ctypes.c_int),
        ("velocity_halo_exchange", ctypes.c_int),
    ]


class LBMForceSettings(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("fluid_density", ctypes.c_double),
        ("kinematic_viscosity", ctypes.c_double),
        ("friction", ctypes.c_double),
        ("temperature", ctypes.c_double),
        ("coupling_scheme", ctypes.c_int), ("drag_scheme", ctypes.c_int),
        ("interpolation_stencil", ctypes.c_int), ("wall_scheme", ctypes.c_int),
        ("fluid_fluctuations", ctypes.c_int),
        ("random_number_seed", ctypes.c_int), ("momentum_removal_frequency", ctypes.c_int),
        ("mach_check_frequency", ctypes.c_int),
        ("mach_number_limit", ctypes.c_double),
        ("body_acceleration", _c_double3),
        ("initial_velocity", _c_double3),
        ("face_boundary", ctypes.c_int * 6),
        ("face_velocity", _c_double3 * 6),
        ("face_density", ctypes.c_double * 6),
        ("storage", ctypes.c_int),
        ("density_halo_exchange", ctypes.c_int),
        ("velocity_halo_exchange", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

# The kernel-level entry points that every backend library (the gfx950 library, and the CPU oracle
# used by the tests) exports.
KERNEL_FUNCTIONS = {
    "lbm_last_error": (ctypes.c_char_p, []),
    "lbm_create": (ctypes.c_int, [ctypes.POINTER(LBMParams), _I, ctypes.c_int, _I, _D, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(_P)]),
    "lbm_destroy": (ctypes.c_int, [_P]),
    "lbm_begin_step": (ctypes.c_int, [_P, ctypes.c_int64, _D, _D]),
    "lbm_execute": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int, _D, _D, _D]),
    "lbm_begin_step_device": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P]),
    "lbm_execute_device": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int, _P, _P, _P]),
    "lbm_copy_parameters": (ctypes.c_int, [_P, ctypes.POINTER(LBMParams)]),
    "lbm_get_fluid_fields": (ctypes.c_int, [_P, _D, _D, ctypes.c_int]),
    "lbm_get_fluid_mach_number": (ctypes.c_int, [_P, _D]),
    "lbm_get_particle_nodes": (ctypes.c_int, [_P, _D, _I]),
    "lbm_get_wall_force": (ctypes.c_int, [_P, _D]),
    "lbm_get_fluid_state": (ctypes.c_int, [_P, _D, ctypes.c_size_t]),
    "lbm_set_fluid_state": (ctypes.c_int, [_P, _D, ctypes.c_size_t]),
    "lbm_checkpoint_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "lbm_create_checkpoint": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t]),
    "lbm_load_checkpoint": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t]),
    "lbm_get_fluid_totals": (ctypes.c_int, [_P, _D, _D]),
    "lbm_synchronize": (ctypes.c_int, [_P]),
    "lbm_profile_fused": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.POINTER(ctypes.c_int)]),
}

# Entry points of the gfx950 library only.
PRODUCT_FUNCTIONS = {
    "lbm_force_defaults": (None, [ctypes.POINTER(LBMForceSettings)]),
    "lbm_compute_lattice_parameters": (ctypes.c_int, [ctypes.POINTER(LBMForceSettings), _D, ctypes.c_double, _I,
                                                      ctypes.c_int, _I, ctypes.c_int, _D, ctypes.c_int,
                                                      ctypes.POINTER(LBMParams)]),
    "lbm_profile_fused_launches": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_int)]),
}


class LBMError(RuntimeError):
    """An error reported by the C-ABI (the OpenMMException of the reference)."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def declare(lib, functions):
    for name, (restype, argtypes) in functions.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def check(lib, status):
    if status != LBM_OK:
        raise LBMError(status, lib.lbm_last_error().decode())


_product = None


def load_library(path=None):
    """Load the gfx950 library (built by openmm-lbm_amd/Makefile).  Raises if it is missing: there is no
    fallback path."""
    global _product
    if path is None and _product is not None:
        return _product
    p = path or LIBRARY_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"the gfx950 LBMForce library {p} is not built: run make -C openmm-lbm_amd")
    lib = ctypes.CDLL(p)
    declare(lib, KERNEL_FUNCTIONS)
    declare(lib, PRODUCT_FUNCTIONS)
    if path is None:
        _product = lib
    return lib
