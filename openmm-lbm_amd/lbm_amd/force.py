"""LBMForce: the user-facing settings of the lattice Boltzmann force, with the names, enums, defaults and
checks of the reference's LBMForce (openmmapi/include/LBMForce.h:41-691, openmmapi/src/LBMForce.cpp).
The MPI-only settings (domain decomposition, particle-copy check) are kept for serialization; a decomposition other
than one domain is refused.  The halo-exchange settings select the fields of the halo layer of getFluidFields(), which
with one domain holds the lattice's own nodes across the periodic boundaries.  Units are OpenMM's: nm, ps, Da, K,
kJ/mol."""
import ctypes

from . import _lib


class LBMForce:
    # CouplingScheme (LBMForce.h:46-57)
    EulerMaruyama, NVE = 0, 1
    # DragScheme (LBMForce.h:61-77)
    Explicit, Centered = 0, 1
    # InterpolationStencil (LBMForce.h:84-103)
    NearestNode, Trilinear, ThreePoint, Keys = 0, 1, 2, 3
    # WallScheme (LBMForce.h:107-122)
    BounceBack, Regularized = 0, 1
    # Face (LBMForce.h:127-134)
    XMin, XMax, YMin, YMax, ZMin, ZMax = 0, 1, 2, 3, 4, 5
    # BoundaryType (LBMForce.h:138-158)
    Periodic, Velocity, Density, DensityVelocity = 0, 1, 2, 3
    # Storage of the populations on the device (no counterpart in the reference)
    FP32, FP16, BF16, FP64 = _lib.STORAGE_FP32, _lib.STORAGE_FP16, _lib.STORAGE_BF16, _lib.STORAGE_FP64

    def __init__(self):
        self._s = _lib.LBMForceSettings()
        _lib.load_library().lbm_force_defaults(ctypes.byref(self._s))
        self._particles = []
        self._solid = []
        self._group = 0
        self._name = "LBMForce"
        self._decomposition = [1, 1, 1]
        self._copiesCheck = True

    # Grid
    def getGridSize(self):
        return self._s.nx, self._s.ny, self._s.nz

    def setGridSize(self, nx, ny, nz):
        self._s.nx, self._s.ny, self._s.nz = int(nx), int(ny), int(nz)

    # Fluid
    def getFluidDensity(self):
        return self._s.fluid_density

    def setFluidDensity(self, density):
        self._s.fluid_density = float(density)

    def getKinematicViscosity(self):
        return self._s.kinematic_viscosity

    def setKinematicViscosity(self, viscosity):
        self._s.kinematic_viscosity = float(viscosity)

    # Coupling
    def getFriction(self):
        return self._s.friction

    def setFriction(self, friction):
        self._s.friction = float(friction)

    def getTemperature(self):
        return self._s.temperature

    def setTemperature(self, temperature):
        self._s.temperature = float(temperature)

    def getCouplingScheme(self):
        return self._s.coupling_scheme

    def setCouplingScheme(self, scheme):
        self._s.coupling_scheme = int(scheme)

    def getDragScheme(self):
        return self._s.drag_scheme

    def setDragScheme(self, scheme):
        self._s.drag_scheme = int(scheme)

    def getInterpolationStencil(self):
        return self._s.interpolation_stencil

    def setInterpolationStencil(self, stencil):
        self._s.interpolation_stencil = int(stencil)

    def getFluidFluctuations(self):
        return bool(self._s.fluid_fluctuations)

    def setFluidFluctuations(self, fluctuations):
        self._s.fluid_fluctuations = 1 if fluctuations else 0

    # Walls and faces
    def getWallScheme(self):
        return self._s.wall_scheme

    def setWallScheme(self, scheme):
        self._s.wall_scheme = int(scheme)

    @staticmethod
    def _checkFace(face):
        if face < 0 or face > 5:
            raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "LBMForce: unknown face of the box")

    def getFaceBoundary(self, face):
        self._checkFace(face)
        return self._s.face_boundary[face]

    def setFaceBoundary(self, face, boundary):
        self._checkFace(face)
        self._s.face_boundary[face] = int(boundary)

    def getFaceVelocity(self, face):
        self._checkFace(face)
        return tuple(self._s.face_velocity[face])

    def setFaceVelocity(self, face, velocity):
        self._checkFace(face)
        for k in range(3):
            self._s.face_velocity[face][k] = float(velocity[k])

    def getFaceDensity(self, face):
        self._checkFace(face)
        return self._s.face_density[face]

    def setFaceDensity(self, face, density):
        self._checkFace(face)
        self._s.face_density[face] = float(density)

    def getSolidNodes(self):
        return list(self._solid)

    def setSolidNodes(self, nodes):
        self._solid = [int(n) for n in nodes]

    # Body force, initial state
    def getBodyAcceleration(self):
        return tuple(self._s.body_acceleration)

    def setBodyAcceleration(self, acceleration):
        for k in range(3):
            self._s.body_acceleration[k] = float(acceleration[k])

    def getInitialFluidVelocity(self):
        return tuple(self._s.initial_velocity)

    def setInitialFluidVelocity(self, velocity):
        for k in range(3):
            self._s.initial_velocity[k] = float(velocity[k])

    # Run control
    def getRandomNumberSeed(self):
        return self._s.random_number_seed

    def setRandomNumberSeed(self, seed):
        self._s.random_number_seed = int(seed)

    def getFluidMomentumRemovalFrequency(self):
        return self._s.momentum_removal_frequency

    def setFluidMomentumRemovalFrequency(self, frequency):
        self._s.momentum_removal_frequency = int(frequency)

    def getMachCheckFrequency(self):
        return self._s.mach_check_frequency

    def setMachCheckFrequency(self, frequency):
        self._s.mach_check_frequency = int(frequency)

    def getMachNumberLimit(self):
        return self._s.mach_number_limit

    def setMachNumberLimit(self, limit):
        self._s.mach_number_limit = float(limit)

    # Precision form (no counterpart in the reference)
    def getStorage(self):
        return self._s.storage

    def setStorage(self, storage):
        self._s.storage = int(storage)

    # Particles
    def getNumParticles(self):
        return len(self._particles)

    def addParticle(self, particle):
        self._particles.append(int(particle))
        return len(self._particles) - 1

    def getParticle(self, index):
        return self._particles[index]

    def setParticle(self, index, particle):
        self._particles[index] = int(particle)

    def usesPeriodicBoundaryConditions(self):
        return True

    # OpenMM Force settings (Force.h) and the MPI settings of the reference (LBMForce.h), kept so that a force
    # serializes and deserializes like the reference's; this platform runs one domain, so any decomposition other
    # than one domain is refused when a Context is created.
    def getForceGroup(self):
        return self._group

    def setForceGroup(self, group):
        if not 0 <= int(group) <= 31:
            raise ValueError("Force group must be between 0 and 31")
        self._group = int(group)

    def getName(self):
        return self._name

    def setName(self, name):
        self._name = str(name)

    def getDomainDecomposition(self):
        return tuple(self._decomposition)

    def setDomainDecomposition(self, px, py, pz):
        self._decomposition = [int(px), int(py), int(pz)]

    def getParticleCopiesCheck(self):
        return self._copiesCheck

    def setParticleCopiesCheck(self, check):
        self._copiesCheck = bool(check)

    def getDensityHaloExchange(self):
        return bool(self._s.density_halo_exchange)

    def setDensityHaloExchange(self, exchange):
        self._s.density_halo_exchange = int(bool(exchange))

    def getVelocityHaloExchange(self):
        return bool(self._s.velocity_halo_exchange)

    def setVelocityHaloExchange(self, exchange):
        self._s.velocity_halo_exchange = int(bool(exchange))

    def getFluidFields(self, context, gather=False, halo=False):
        """The density (Da/nm^3) and velocity (nm/ps) of the fluid at the nodes (LBMForce.h:517-536,
        openmmlbm.i:449-460), as arrays: with one domain the whole lattice, node (i, j, k) at index i + nx*(j + ny*k);
        gather changes nothing.  halo=True adds the layer one node thick around the lattice, node (i, j, k) for i from
        -1 to nx and so on at index (i + 1) + (nx + 2)*((j + 1) + (ny + 2)*(k + 1)), holding the fields across the
        periodic boundaries whose halo is exchanged and NaN for the others and beyond the open faces."""
        if gather and halo:
            raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "LBMForce: getFluidFields() cannot gather the fields and add "
                                                        "the halo in the same call")
        return context.getFluidFields(halo=halo)

    def computeLatticeParameters(self, box, stepSize, masses):
        """LBMForceImpl::computeLatticeParameters (LBMForceImpl.cpp:44-214): the lattice parameters of this force
        for a rectangular box (three lengths, nm), an integrator step (ps) and the masses of all the particles."""
        if any(p not in (0, 1) for p in self._decomposition):     # 0: chosen automatically, one domain here
            raise _lib.LBMError(_lib.LBM_ERROR_UNSUPPORTED, "LBMForce: this platform runs the whole lattice in one "
                                                            "domain on one GPU (setDomainDecomposition(1, 1, 1))")
        lib = _lib.load_library()
        out = _lib.LBMParams()
        boxArr = (ctypes.c_double * 3)(*[float(b) for b in box])
        solid = (ctypes.c_int * max(1, len(self._solid)))(*self._solid)
        parts = (ctypes.c_int * max(1, len(self._particles)))(*self._particles)
        m = (ctypes.c_double * max(1, len(masses)))(*[float(x) for x in masses])
        _lib.check(lib, lib.lbm_compute_lattice_parameters(ctypes.byref(self._s), boxArr, float(stepSize), solid,
                                                           len(self._solid), parts, len(self._particles), m,
                                                           len(masses), ctypes.byref(out)))
        return out
