"""A minimal host for the LBMForce kernel: a System (masses and a rectangular periodic box), OpenMM's leapfrog
VerletIntegrator and a Context that calls the kernel at the points where OpenMM's ContextImpl calls
CalcLBMForceKernel (LBMForceImpl.cpp:366-380): beginStep() before the forces of a step, execute() in the force
evaluation, then the integrator update v(t+dt/2) = v(t-dt/2) + dt F(t)/m, x(t+dt) = x(t) + dt v(t+dt/2).

Particles are held on the host in FP64; the kernel library is the gfx950 library unless another library with
the same C-ABI is given (the tests pass the CPU oracle)."""
import ctypes

import numpy as np

from . import _lib


class System:
    def __init__(self, box=(1.0, 1.0, 1.0)):
        self.masses = []
        self.box = tuple(float(b) for b in box)

    def addParticle(self, mass):
        self.masses.append(float(mass))
        return len(self.masses) - 1

    def getNumParticles(self):
        return len(self.masses)

    def setDefaultPeriodicBoxVectors(self, a, b, c):
        if a[1] != 0 or a[2] != 0 or b[0] != 0 or b[2] != 0 or c[0] != 0 or c[1] != 0:
            raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "LBMForce: the periodic box must be rectangular")
        self.box = (float(a[0]), float(b[1]), float(c[2]))


class VerletIntegrator:
    def __init__(self, stepSize):
        self.stepSize = float(stepSize)


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """Drives one LBMForce.  extraForces: callables f(positions) -> forces (N x 3, kJ/mol/nm) of the other forces of
    the System, evaluated before the LBMForce as OpenMM does when LBMForce is the last force."""

    def __init__(self, system, integrator, force, library=None, extraForces=()):
        self.system = system
        self.integrator = integrator
        self.force = force
        self.lib = library if library is not None else _lib.load_library()
        self.params = force.computeLatticeParameters(system.box, integrator.stepSize, system.masses)
        n = system.getNumParticles()
        self.masses = np.array(system.masses, dtype=np.float64)
        self.positions = np.zeros((n, 3))
        self.velocities = np.zeros((n, 3))
        self.stepCount = 0
        self.extraForces = list(extraForces)
        parts = [force.getParticle(i) for i in range(force.getNumParticles())]
        self.coupled = parts
        solid = sorted(force.getSolidNodes())
        self.numNodes = self.params.nx*self.params.ny*self.params.nz
        solidArr = (ctypes.c_int*max(1, len(solid)))(*solid)
        partArr = (ctypes.c_int*max(1, len(parts)))(*parts)
        massArr = (ctypes.c_double*max(1, len(parts)))(*[system.masses[p] for p in parts])
        handle = ctypes.c_void_p()
        self._check(self.lib.lbm_create(ctypes.byref(self.params), solidArr, len(solid), partArr, massArr, len(parts), n,
                                        ctypes.byref(handle)))
        self.handle = handle

    def _check(self, status):
        _lib.check(self.lib, status)

    def __del__(self):
        handle = getattr(self, "handle", None)
        if handle:
            self.lib.lbm_destroy(handle)
            self.handle = None

    # Particles
    def setPositions(self, positions):
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3).copy()

    def setVelocities(self, velocities):
        self.velocities = np.array(velocities, dtype=np.float64).reshape(-1, 3).copy()

    def getStepCount(self):
        return self.stepCount

    def setStepCount(self, count):
        self.stepCount = int(count)

    def _otherForces(self):
        forces = np.zeros_like(self.positions)
        for f in self.extraForces:
            forces += f(self.positions)
        return forces

    def step(self, numSteps=1):
        dt = self.integrator.stepSize
        invMass = np.where(self.masses > 0, 1.0/np.where(self.masses > 0, self.masses, 1.0), 0.0)[:, None]
        for _ in range(int(numSteps)):
            pos = np.ascontiguousarray(self.positions)
            vel = np.ascontiguousarray(self.velocities)
            self._check(self.lib.lbm_begin_step(self.handle, self.stepCount, _dptr(pos), _dptr(vel)))
            forces = np.ascontiguousarray(self._otherForces())
            self._check(self.lib.lbm_execute(self.handle, self.stepCount, 1, _dptr(pos), _dptr(vel), _dptr(forces)))
            vel = vel + dt*forces*invMass
            self.positions = pos + dt*vel
            self.velocities = vel
            self.stepCount += 1

    def getForces(self):
        """The forces of a State (a force evaluation between steps): those of the next step."""
        pos = np.ascontiguousarray(self.positions)
        vel = np.ascontiguousarray(self.velocities)
        forces = np.ascontiguousarray(self._otherForces())
        self._check(self.lib.lbm_execute(self.handle, self.stepCount, 1, _dptr(pos), _dptr(vel), _dptr(forces)))
        return forces

    # Fluid
    def getFluidFields(self):
        rho = np.zeros(self.numNodes)
        u = np.zeros((self.numNodes, 3))
        self._check(self.lib.lbm_get_fluid_fields(self.handle, _dptr(rho), _dptr(u)))
        return rho, u

    def getFluidState(self):
        state = np.zeros(19*self.numNodes)
        self._check(self.lib.lbm_get_fluid_state(self.handle, _dptr(state), state.size))
        return state

    def setFluidState(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        self._check(self.lib.lbm_set_fluid_state(self.handle, _dptr(state), state.size))

    def getFluidMachNumber(self):
        m = ctypes.c_double()
        self._check(self.lib.lbm_get_fluid_mach_number(self.handle, ctypes.byref(m)))
        return m.value

    def getWallForce(self):
        f = np.zeros(3)
        self._check(self.lib.lbm_get_wall_force(self.handle, _dptr(f)))
        return f

    def getFluidTotals(self):
        mass = ctypes.c_double()
        p = np.zeros(3)
        self._check(self.lib.lbm_get_fluid_totals(self.handle, ctypes.byref(mass), _dptr(p)))
        return mass.value, p

    def createCheckpoint(self):
        """LBMForce::createCheckpoint: LBMForceImpl's header (LBMForceImpl.cpp:516-529), then the kernel's data."""
        from .serialization import checkpoint_header
        size = ctypes.c_size_t()
        self._check(self.lib.lbm_checkpoint_size(self.handle, ctypes.byref(size)))
        buf = ctypes.create_string_buffer(size.value)
        self._check(self.lib.lbm_create_checkpoint(self.handle, buf, size.value))
        return checkpoint_header(self.force, len(self.coupled)) + buf.raw

    def loadCheckpoint(self, data):
        from .serialization import read_checkpoint_header
        data = bytes(data)
        offset = read_checkpoint_header(data, self.force, len(self.coupled))
        body = data[offset:]
        buf = ctypes.create_string_buffer(body, len(body))
        self._check(self.lib.lbm_load_checkpoint(self.handle, buf, len(body)))

    def updateParametersInContext(self):
        """LBMForceImpl::updateParametersInContext (LBMForceImpl.cpp:388-416)."""
        updated = self.force.computeLatticeParameters(self.system.box, self.integrator.stepSize, self.system.masses)
        old = self.params
        checks = [
            ((updated.nx, updated.ny, updated.nz) != (old.nx, old.ny, old.nz), "the grid size cannot be changed"),
            (updated.density != old.density or updated.tau != old.tau, "the fluid density and viscosity cannot be changed"),
            (updated.wall_scheme != old.wall_scheme, "the wall scheme cannot be changed"),
            (list(updated.face_boundary) != list(old.face_boundary), "the boundary types of the faces cannot be changed"),
            (updated.drag_scheme != old.drag_scheme, "the drag scheme cannot be changed"),
            (updated.interpolation_stencil != old.interpolation_stencil, "the interpolation stencil cannot be changed"),
            (updated.fluid_fluctuations != old.fluid_fluctuations, "the fluid fluctuations cannot be switched on or off"),
        ]
        for bad, msg in checks:
            if bad:
                raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "updateParametersInContext: " + msg)
        self._check(self.lib.lbm_copy_parameters(self.handle, ctypes.byref(updated)))
        self.params = updated

    def getLatticeParameters(self):
        return self.params.dx, self.params.dt, self.params.tau
