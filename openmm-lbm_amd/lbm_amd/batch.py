"""Batched replicas (include/lbm_amd_batch.h): many independent periodic lattices advanced together on one GPU.

The reference runs the replicas of a seed-averaged study one after another (examples/cocomo/kick.py:30,105); a
Batch steps them all at once, three launches per step captured in a hipGraph.  The particles of a replica feel the
coupling force only and are integrated with leapfrog Verlet on the device.
"""
import ctypes

import numpy as np

from . import _lib
from .context import System

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_I64 = ctypes.POINTER(ctypes.c_int64)

BATCH_FUNCTIONS = {
    "lbm_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_lib.LBMParams), _I, _D, _D, _D,
                                        ctypes.POINTER(_P)]),
    "lbm_batch_autotune": (ctypes.c_int, [_P, ctypes.c_int, _D, _I, _I]),
    "lbm_batch_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "lbm_batch_synchronize": (ctypes.c_int, [_P]),
    "lbm_batch_profile_fused": (ctypes.c_int, [_P, ctypes.c_int, _D, _I]),
    "lbm_batch_get_fluid_fields": (ctypes.c_int, [_P, ctypes.c_int, _D, _D]),
    "lbm_batch_get_particles": (ctypes.c_int, [_P, _D, _D]),
    "lbm_batch_info": (ctypes.c_int, [_P, _I64, _I64]),
    "lbm_batch_destroy": (ctypes.c_int, [_P]),
}

WORKGROUP_SIZES = (256, 512, 1024)          # nodes per work-group of the lattice pass
COUPLING_SIZES = (64, 128, 256)              # threads per work-group of the coupling


def _d(a):
    return a.ctypes.data_as(_D)


class Replica:
    """One replica: an LBMForce (periodic, FP32, explicit drag) with its box, step size and particles."""

    def __init__(self, force, box, stepSize, masses, positions, velocities=None):
        self.force = force
        self.box = tuple(box)
        self.stepSize = float(stepSize)
        self.masses = np.ascontiguousarray(masses, dtype=np.float64).reshape(-1)
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        if velocities is None:
            velocities = np.zeros_like(self.positions)
        self.velocities = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)

    def latticeParameters(self):
        system = System(self.box)
        for m in self.masses:
            system.addParticle(m)
        return self.force.computeLatticeParameters(system.box, self.stepSize, system.masses)


class Batch:
    def __init__(self, replicas, library=None):
        self.lib = _lib.declare(library if library is not None else _lib.load_library(), BATCH_FUNCTIONS)
        self.replicas = list(replicas)
        R = len(self.replicas)
        params = (_lib.LBMParams*R)(*[r.latticeParameters() for r in self.replicas])
        self.params = params
        counts = np.array([len(r.masses) for r in self.replicas], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        masses = np.ascontiguousarray(np.concatenate([r.masses for r in self.replicas] + [np.zeros(0)]))
        pos = np.ascontiguousarray(np.concatenate([r.positions for r in self.replicas] + [np.zeros((0, 3))]))
        vel = np.ascontiguousarray(np.concatenate([r.velocities for r in self.replicas] + [np.zeros((0, 3))]))
        masses = masses if masses.size else np.zeros(1)
        pos = pos if pos.size else np.zeros((1, 3))
        vel = vel if vel.size else np.zeros((1, 3))
        handle = ctypes.c_void_p()
        self._check(self.lib.lbm_batch_create(R, params, counts.ctypes.data_as(_I), _d(masses), _d(pos), _d(vel),
                                              ctypes.byref(handle)))
        self.handle = handle

    def _check(self, status):
        _lib.check(self.lib, status)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.lbm_batch_destroy(h)
            self.handle = None

    def autotune(self, stepsPerTrial=20):
        """Times each tile shape (nodes per lattice-pass work-group x threads per coupling work-group; the replicas
        advance meanwhile); returns ({(fused, couple): ms per step}, (fused, couple) chosen)."""
        times = np.zeros(9)
        fused, couple = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.lbm_batch_autotune(self.handle, int(stepsPerTrial), _d(times), ctypes.byref(fused),
                                                ctypes.byref(couple)))
        shapes = [(f, c) for f in WORKGROUP_SIZES for c in COUPLING_SIZES]
        return dict(zip(shapes, times.tolist())), (fused.value, couple.value)

    def step(self, numSteps=1):
        self._check(self.lib.lbm_batch_step(self.handle, int(numSteps)))

    def synchronize(self):
        self._check(self.lib.lbm_batch_synchronize(self.handle))

    def profileFused(self, enable):
        """While enabled, steps run eagerly with events around each fused pass; disabling returns
        (total ms, number of fused passes)."""
        total = ctypes.c_double()
        count = ctypes.c_int()
        self._check(self.lib.lbm_batch_profile_fused(self.handle, 1 if enable else 0, ctypes.byref(total),
                                                     ctypes.byref(count)))
        return total.value, count.value

    def getFluidFields(self, replica):
        p = self.params[replica]
        n = p.nx*p.ny*p.nz
        rho = np.zeros(n)
        u = np.zeros((n, 3))
        self._check(self.lib.lbm_batch_get_fluid_fields(self.handle, int(replica), _d(rho), _d(u)))
        return rho, u

    def getParticles(self):
        """Positions and velocities of every replica's particles, as a list of (positions, velocities)."""
        total = int(self.offsets[-1])
        pos = np.zeros((max(total, 1), 3))
        vel = np.zeros((max(total, 1), 3))
        self._check(self.lib.lbm_batch_get_particles(self.handle, _d(pos), _d(vel)))
        return [(pos[a:b], vel[a:b]) for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def info(self):
        nodes = ctypes.c_int64()
        step = ctypes.c_int64()
        self._check(self.lib.lbm_batch_info(self.handle, ctypes.byref(nodes), ctypes.byref(step)))
        return nodes.value, step.value
