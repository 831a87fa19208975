"""Lockstep comparison of the gfx950 library with the FP64 CPU oracle on the same setup (test infrastructure)."""
import numpy as np

from lbm_amd import Context, LBMForce, System, VerletIntegrator
import oracle_lib

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1])
W = np.array([1/3] + [1/18]*6 + [1/36]*12)


def equilibrium_state(num_nodes, rho, u):
    """Equilibrium deviations f_q - w_q at every node (TestLBMFluid.h:57-75): u is (num_nodes, 3) or (3,)."""
    u = np.broadcast_to(np.asarray(u, dtype=np.float64), (num_nodes, 3))
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), (num_nodes,))
    cu = CX[:, None]*u[None, :, 0] + CY[:, None]*u[None, :, 1] + CZ[:, None]*u[None, :, 2]
    uu = (u*u).sum(axis=1)[None, :]
    return (W[:, None]*((rho - 1.0)[None, :] + rho[None, :]*(3*cu + 4.5*cu*cu - 1.5*uu))).reshape(-1)


def total_moments(state):
    s = state.reshape(19, -1)
    n = s.shape[1]
    return n + s.sum(), np.array([(CX[:, None]*s).sum(), (CY[:, None]*s).sum(), (CZ[:, None]*s).sum()])


class Setup:
    """A System + LBMForce description that can be instantiated on any backend."""

    def __init__(self, grid, dx=1.0, dt=0.01, tau=0.8, density=602.214):
        self.grid = tuple(grid)
        self.dx, self.dt, self.tau, self.density = dx, dt, tau, density
        self.masses = []
        self.positions = []
        self.velocities = []
        self.coupled = []
        self.configure = []
        self.state = None
        self.extra = []

    def add_particles(self, positions, mass=100.0, velocities=None, coupled=True):
        positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        for i, x in enumerate(positions):
            idx = len(self.masses)
            self.masses.append(mass)
            self.positions.append(x)
            self.velocities.append(np.zeros(3) if velocities is None else np.asarray(velocities[i], dtype=np.float64))
            if coupled:
                self.coupled.append(idx)

    def build(self, library=None, storage=LBMForce.FP32):
        nx, ny, nz = self.grid
        if not self.masses:
            self.add_particles([[0.1, 0.2, 0.3]], mass=1.0, coupled=False)
        s = System((nx*self.dx, ny*self.dx, nz*self.dx))
        for m in self.masses:
            s.addParticle(m)
        f = LBMForce()
        f.setGridSize(nx, ny, nz)
        f.setFluidDensity(self.density)
        f.setKinematicViscosity((self.tau - 0.5)/3.0*self.dx*self.dx/self.dt)
        f.setFluidMomentumRemovalFrequency(0)
        f.setStorage(storage)
        for p in self.coupled:
            f.addParticle(p)
        for fn in self.configure:
            fn(f)
        c = Context(s, VerletIntegrator(self.dt), f, library=library, extraForces=self.extra)
        c.setPositions(np.array(self.positions))
        c.setVelocities(np.array(self.velocities))
        if self.state is not None:
            c.setFluidState(self.state)
        return c


def rel_err(a, b, scale=None):
    a = np.asarray(a)
    b = np.asarray(b)
    s = scale if scale is not None else max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max()/s


def lockstep(setup, steps, check_every=1, storage=LBMForce.FP32, fields=True, forces=True, callback=None):
    """Run setup on the oracle and on the gfx950 library; return the worst relative errors of density, velocity
    and coupled-particle forces over the checked steps.  Relative errors are max |a - b| / max |b|."""
    ref = setup.build(oracle_lib.load_oracle())
    gpu = setup.build(None, storage=storage)
    worst = {"density": 0.0, "velocity": 0.0, "force": 0.0, "position": 0.0}
    for step in range(1, steps + 1):
        ref.step(1)
        gpu.step(1)
        if step % check_every and step != steps:
            continue
        if fields:
            r0, u0 = ref.getFluidFields()
            r1, u1 = gpu.getFluidFields()
            worst["density"] = max(worst["density"], rel_err(r1, r0))
            worst["velocity"] = max(worst["velocity"], rel_err(u1, u0))
        if forces and setup.coupled:
            f0 = ref.getForces()[setup.coupled]
            f1 = gpu.getForces()[setup.coupled]
            worst["force"] = max(worst["force"], rel_err(f1, f0))
            worst["position"] = max(worst["position"], rel_err(gpu.positions, ref.positions, scale=setup.dx))
        if callback:
            callback(step, ref, gpu)
    return worst, ref, gpu
