"""Test infrastructure: loads the FP64 CPU oracle (oracle/liblbm_oracle.so), which exports the same kernel C-ABI as
the gfx950 library, plus a few test hooks.  Only tests, the smoke check at the repository root and bench.py's CPU
baseline use it."""
import ctypes
import os

from lbm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liblbm_oracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_d3q19.so")

_D = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint32)
_I = ctypes.POINTER(ctypes.c_int)

HOOKS = {
    "oracle_philox4x32": (None, [_U, _U, _U]),
    "oracle_normals": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int, _D]),
    "oracle_axis_weights": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _I, _D]),
    "oracle_equilibrium_deviation": (None, [ctypes.c_double]*4 + [_D]),
    "oracle_pi_neq": (None, [_D] + [ctypes.c_double]*4 + [_D]),
    "oracle_regularized_neq": (None, [_D, _D]),
    "oracle_guo": (None, [ctypes.c_double]*6 + [_D]),
    "oracle_mode": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
    "oracle_collision_noise": (None, [_D, ctypes.c_double, ctypes.c_double, _D]),
    "oracle_equilibrium_noise": (None, [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]),
    "oracle_num_boundary_nodes": (ctypes.c_int, [ctypes.c_void_p]),
    "oracle_seed": (ctypes.c_uint32, [ctypes.c_void_p]),
    "oracle_boundary_node": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _I, _I]),
}

REF_FUNCTIONS = {
    "ref_equilibrium_deviation": (None, [ctypes.c_double]*4 + [_D]),
    "ref_pi_neq": (None, [_D] + [ctypes.c_double]*4 + [_D]),
    "ref_regularized_neq": (None, [_D, _D]),
    "ref_guo": (None, [ctypes.c_double]*6 + [_D]),
    "ref_mode": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
    "ref_collision_noise": (None, [_D, ctypes.c_double, ctypes.c_double, _D]),
    "ref_equilibrium_noise": (None, [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]),
    "ref_boundary_stress_weight": (ctypes.c_double, [ctypes.c_double]),
}

_oracle = None


def load_oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_PATH)
        _lib.declare(lib, _lib.KERNEL_FUNCTIONS)
        _lib.declare(lib, HOOKS)
        _oracle = lib
    return _oracle


def load_ref():
    """The reference's own D3Q19.h compiled into oracle/_ref (only where the reference tree was present)."""
    if not os.path.exists(REF_PATH):
        return None
    lib = ctypes.CDLL(REF_PATH)
    _lib.declare(lib, REF_FUNCTIONS)
    return lib
