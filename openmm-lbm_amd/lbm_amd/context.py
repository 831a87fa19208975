"""A minimal host for the LBMForce kernel: a System (masses and a rectangular periodic box), OpenMM's leapfrog
VerletIntegrator and a Context that calls the kernel at the points where OpenMM's ContextImpl calls
CalcLBMForceKernel (LBMForceImpl.cpp:366-380): beginStep() before the forces of a step, execute() in the force
evaluation, then the integrator update v(t+dt/2) = v(t-dt/2) + dt F(t)/m, x(t+dt) = x(t) + dt v(t+dt/2).

Particles are held on the host in FP64; the kernel library is the gfx950 library unless another library with
the same C-ABI is given (the tests pass the CPU oracle)."""
import ctypes
import math
import sys

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
        if not isinstance(integrator, VerletIntegrator):
            raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "LBMForce requires a VerletIntegrator: drag and random forces "
                                                        "are part of the force")
        self.system = system
        self.integrator = integrator
        self.force = force
        self.lib = library if library is not None else _lib.load_library()
        self.params = force.computeLatticeParameters(system.box, integrator.stepSize, system.masses)
        self._warn()
        n = system.getNumParticles()
        self.masses = np.array(system.masses, dtype=np.float64)
        self.positions = np.zeros((n, 3))
        self.velocities = np.zeros((n, 3))
        self.stepCount = 0
        self.extraForces = list(extraForces)
        parts = [force.getParticle(i) for i in range(force.getNumParticles())]
        self.coupled = parts
        solid = sorted(force.getSolidNodes())
        self.solid = solid
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

    def _warn(self):
        """The warnings of LBMForceImpl::initialize (LBMForceImpl.cpp:256-355), printed on stderr when the Context is
        created: a relaxation time outside [0.505, 2]; tau > 1.7 with the explicit drag (the hydrodynamic
        self-mobility of a particle at a node is small, and negative above 1.79), except with the three-point stencil;
        friction*dt > 1 with the explicit drag; the heating of the explicit drag in a fluctuating fluid,
        friction*dt*m*K/(2 m_c) for the heaviest coupled particle with K = 1, 8/27, 1/8, (57/70)^3 the self weight of
        the stencil averaged over a cell; the size of the fluctuations in lattice units (validated up to
        kT = 1/3000); and regularized walls or DensityVelocity faces with fluctuations outside 0.586 <= tau <= 3.41,
        where the stress amplitude sqrt(1 - c), c = (1 - omega)^2/(omega (2 - omega)), of their noise is not real."""
        p, f = self.params, self.force
        out = []
        coupled = f.getNumParticles() > 0
        explicit = p.drag_scheme == f.Explicit
        if p.tau < 0.505 or p.tau > 2.0:
            out.append(f"Warning: LBMForce: the relaxation time tau = {p.tau:.6g} is outside the range [0.505, 2] in "
                       "which the lattice Boltzmann model is accurate. tau = 3 nu dt/dx^2 + 1/2: change the viscosity, "
                       "the time step or the lattice spacing.")
        if coupled and explicit and p.tau > 1.7 and p.interpolation_stencil != f.ThreePoint:
            out.append(f"Warning: LBMForce: tau = {p.tau:.6g} > 1.7: with the explicit drag at the nearest node the "
                       "hydrodynamic self-mobility of a coupled particle is small, and negative above tau = 1.79, so "
                       "particles move less than they should. Reduce the viscosity or the time step, or use a coarser "
                       "lattice.")
        gammaDt = p.friction*p.dt
        if coupled and explicit and gammaDt > 1.0:
            grows = ", and grows without bound since friction*dt >= 2" if gammaDt >= 2.0 else ""
            out.append(f"Warning: LBMForce: friction*dt = {gammaDt:.6g} > 1: with the explicit drag the velocity of a "
                       f"particle relative to the fluid changes sign at every step{grows}. Reduce the friction or the "
                       "time step.")
        cellMass = p.density*p.dx**3
        if coupled and explicit and p.fluid_fluctuations and p.kT > 0 and gammaDt > 0:
            heaviest = max(self.system.masses[i] for i in
                           (f.getParticle(k) for k in range(f.getNumParticles())))
            self_weight = {f.Trilinear: 8.0/27.0, f.ThreePoint: 0.125,
                           f.Keys: (57.0/70.0)**3}.get(p.interpolation_stencil, 1.0)
            text = ("Warning: LBMForce: with fluid fluctuations the explicit drag makes the coupled particles hotter "
                    f"than the set temperature, by about friction*dt*m{'*K' if self_weight < 1 else ''}/(2 m_c) = "
                    f"{100.0*gammaDt*heaviest*self_weight/(2.0*cellMass):.6g}% for the heaviest one "
                    f"(m_c = {cellMass:.6g} Da is the mass of fluid in a cell")
            if self_weight < 1:
                text += f", K = {self_weight:.6g} the self weight of the interpolation stencil averaged over a cell"
            out.append(text + "). Use the Centered drag scheme with fluid fluctuations.")
        if p.fluid_fluctuations and p.fluid_kT > 0:
            latticeKT = p.fluid_kT*p.dt**2/(cellMass*p.dx**2)
            validated = 1.0/3000.0
            out.append(f"LBMForce: fluctuating fluid: kT/(m_c (dx/dt)^2) = {latticeKT:.2g} in lattice units, thermal "
                       f"Mach number sqrt(3 kT) = {math.sqrt(3*latticeKT):.2g} (validated up to kT = 1/3000, Mach number "
                       f"{math.sqrt(3*validated):.2g})")
            if latticeKT > validated*(1.0 + 1e-6):
                out.append(f"Warning: LBMForce: the fluctuations of the fluid, kT = {latticeKT:.2g} in lattice units, "
                           "are larger than those validated, kT = 1/3000: the fluctuating fluid may become unstable "
                           "(at 1/3000 it is for tau <= 0.501). A larger lattice spacing or a smaller time step makes "
                           "them smaller (kT in lattice units goes as dt^2/dx^5).")
            noisy = (len(f.getSolidNodes()) > 0 and p.wall_scheme == f.Regularized) or \
                any(b == f.DensityVelocity for b in p.face_boundary)
            w = p.omega
            if noisy and (1.0 - w)**2/(w*(2.0 - w)) > 1.0:
                out.append("Warning: LBMForce: with fluid fluctuations, regularized walls and DensityVelocity faces "
                           "keep the fluid next to them exactly at equilibrium only for 0.586 <= tau <= 3.41; "
                           f"tau = {p.tau:.6g} is outside, and the fluid next to them is slightly away from "
                           "equilibrium.")
        for line in out:
            print(line, file=sys.stderr)

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

    def _checkStepSize(self):
        """LBMForceImpl::calcForcesAndEnergy (LBMForceImpl.cpp:373-380): the lattice time step is the step size of
        the integrator when the Context was created."""
        if self.integrator.stepSize != self.params.dt:
            raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "LBMForce: the integrator step size changed after the Context "
                                                        "was created; reinitialize the Context")

    def step(self, numSteps=1):
        self._checkStepSize()
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
        self._checkStepSize()
        pos = np.ascontiguousarray(self.positions)
        vel = np.ascontiguousarray(self.velocities)
        forces = np.ascontiguousarray(self._otherForces())
        self._check(self.lib.lbm_execute(self.handle, self.stepCount, 1, _dptr(pos), _dptr(vel), _dptr(forces)))
        return forces

    # Fluid
    def getFluidFields(self, halo=False):
        """Density and velocity of every node (CalcLBMForceKernel::getFluidFields); with halo, of the lattice with
        a layer one node thick around it, (nx + 2)(ny + 2)(nz + 2) nodes."""
        nx, ny, nz = self.force.getGridSize()
        count = (nx + 2)*(ny + 2)*(nz + 2) if halo else self.numNodes
        rho = np.zeros(count)
        u = np.zeros((count, 3))
        self._check(self.lib.lbm_get_fluid_fields(self.handle, _dptr(rho), _dptr(u), 1 if halo else 0))
        return rho, u

    def getFluidState(self):
        state = np.zeros(19*self.numNodes)
        self._check(self.lib.lbm_get_fluid_state(self.handle, _dptr(state), state.size))
        return state

    def setFluidState(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        self._check(self.lib.lbm_set_fluid_state(self.handle, _dptr(state), state.size))

    def getParticleNodes(self, positions=None):
        """The nearest node of every coupled particle (lbm_get_particle_nodes), for the current positions or the given
        ones."""
        pos = np.ascontiguousarray(self.positions if positions is None else positions, dtype=np.float64)
        count = len(self.coupled)
        nodes = np.zeros(max(count, 1), dtype=np.int32)
        self._check(self.lib.lbm_get_particle_nodes(self.handle, _dptr(pos),
                                                    nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return nodes[:count]

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
        force = self.force
        updated = force.computeLatticeParameters(self.system.box, self.params.dt, self.system.masses)
        old = self.params
        checks = [
            ((updated.nx, updated.ny, updated.nz) != (old.nx, old.ny, old.nz), "the grid size cannot be changed"),
            (updated.density != old.density or updated.tau != old.tau, "the fluid density and viscosity cannot be changed"),
            ([force.getParticle(i) for i in range(force.getNumParticles())] != self.coupled,
             "the set of coupled particles cannot be changed"),
            (sorted(force.getSolidNodes()) != self.solid, "the solid nodes cannot be changed"),
            (updated.wall_scheme != old.wall_scheme, "the wall scheme cannot be changed"),
            ((updated.density_halo_exchange, updated.velocity_halo_exchange) !=
             (old.density_halo_exchange, old.velocity_halo_exchange),
             "the exchange of the halo cannot be switched on or off"),
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
