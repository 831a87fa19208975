"""Device-resident particles (include/lbm_amd_harness.h): the Verlet integrator and the extra forces of the
benchmark configurations on the GPU, around a Context's kernel."""
import ctypes

import numpy as np

from . import _lib

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)

HARNESS_FUNCTIONS = {
    "lbm_md_last_error": (ctypes.c_char_p, []),
    "lbm_md_create": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, ctypes.c_double, ctypes.POINTER(_P)]),
    "lbm_md_set_wall_potential": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double]),
    "lbm_md_set_bonds": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _D]),
    "lbm_md_step": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.c_int]),
    "lbm_md_get": (ctypes.c_int, [_P, _P, _D, _D]),
    "lbm_md_destroy": (ctypes.c_int, [_P]),
}


def _d(a):
    return a.ctypes.data_as(_D)


class DeviceParticles:
    def __init__(self, context):
        self.ctx = context
        self.lib = _lib.declare(context.lib, HARNESS_FUNCTIONS)
        masses = np.ascontiguousarray(context.masses, dtype=np.float64)
        pos = np.ascontiguousarray(context.positions, dtype=np.float64)
        vel = np.ascontiguousarray(context.velocities, dtype=np.float64)
        handle = ctypes.c_void_p()
        self._check(self.lib.lbm_md_create(len(masses), _d(masses), _d(pos), _d(vel), context.integrator.stepSize,
                                           ctypes.byref(handle)))
        self.handle = handle

    def _check(self, status):
        if status != 0:
            raise _lib.LBMError(status, self.lib.lbm_md_last_error().decode())

    def setWallPotential(self, eps, sigma, zlo, zhi):
        self._check(self.lib.lbm_md_set_wall_potential(self.handle, eps, sigma, zlo, zhi))

    def setBonds(self, pairs, k, r0):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        box = np.array(self.ctx.system.box, dtype=np.float64)
        self._check(self.lib.lbm_md_set_bonds(self.handle, pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                              len(pairs)//2, k, r0, _d(box)))

    def step(self, numSteps):
        self._check(self.lib.lbm_md_step(self.handle, self.ctx.handle, self.ctx.stepCount, int(numSteps)))
        self.ctx.stepCount += int(numSteps)

    def pull(self):
        pos = np.zeros_like(self.ctx.positions)
        vel = np.zeros_like(self.ctx.velocities)
        self._check(self.lib.lbm_md_get(self.handle, self.ctx.handle, _d(pos), _d(vel)))
        self.ctx.positions = pos
        self.ctx.velocities = vel
        return pos, vel

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.lbm_md_destroy(self.handle)
            self.handle = None
