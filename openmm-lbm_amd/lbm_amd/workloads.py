"""The benchmark workloads, with every parameter that the configuration statements leave open fixed here.

 config1: 32^3 periodic D3Q19, dx = 1 nm, dt = 0.01 ps, rho0 = 602.214 Da/nm^3, tau = 0.8; shear wave
          u_x = 0.01 sin(2 pi z/32); Guo body force 1e-5 (lattice) along x; 64 beads of 100 Da at seeded
          uniform positions, v = 0, gamma = 10/ps, trilinear; T = 0 (no noise); removal off; 1000 steps.
 config4: 64 periodic replicas of uneven size, relaxation time, temperature, seed and bead count (batched).
 config5: 128^3 periodic, 2,000 harmonic chains of 50 beads, fluctuating; the learned closure against Keys.
 config2: 256^3, solid planes z = 0 and z = 255 with bounce-back (walls at z = 0.5 and 254.5 nm), periodic in
          x and y; Guo body force 1e-6 (lattice) along x; tau = 0.6; dx = 1 nm, dt = 0.01 ps, rho0 = 602.214;
          20,000 beads of 100 Da, gamma = 10/ps, uniform in 1.5 <= z <= 253.5 nm, Keys stencil, repulsive walls
          eps = 2.494 kJ/mol, sigma = 0.5 nm at z = 0.5 and 254.5 nm; fluctuating fluid at T = 300 K
          (kT = 2.494 kJ/mol) with Euler-Maruyama particle noise, seed 20240607; removal off; 2000 steps.
"""
import numpy as np

from .context import Context, System, VerletIntegrator
from .force import LBMForce

BOLTZ = 1.380649e-23*6.02214076e23/1000.0


def _equilibrium(num_nodes, u):
    cx = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
    cy = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0])
    cz = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1])
    w = np.array([1/3] + [1/18]*6 + [1/36]*12)
    cu = cx[:, None]*u[None, :, 0] + cy[:, None]*u[None, :, 1] + cz[:, None]*u[None, :, 2]
    uu = (u*u).sum(axis=1)[None, :]
    return (w[:, None]*(3*cu + 4.5*cu*cu - 1.5*uu)).reshape(-1)


def wall_force_function(eps, sigma, zlo, zhi):
    def f(pos):
        out = np.zeros_like(pos)
        d1 = pos[:, 2] - zlo
        d2 = zhi - pos[:, 2]
        out[:, 2] = 12*eps*((sigma/d1)**12/d1 - (sigma/d2)**12/d2)
        return out
    return f


class Workload:
    def __init__(self, name, grid, tau, dx=1.0, dt=0.01):
        self.name = name
        self.grid = grid
        self.tau = tau
        self.dx = dx
        self.dt = dt
        self.masses = []
        self.positions = np.zeros((0, 3))
        self.velocities = np.zeros((0, 3))
        self.configure = []
        self.state = None
        self.wall = None          # (eps, sigma, zlo, zhi)
        self.bonds = None         # (pairs, k, r0)
        self.steps = 0
        self.description = {}

    @property
    def num_nodes(self):
        return self.grid[0]*self.grid[1]*self.grid[2]

    def build(self, library=None, storage=LBMForce.FP32):
        nx, ny, nz = self.grid
        system = System((nx*self.dx, ny*self.dx, nz*self.dx))
        masses = self.masses if len(self.masses) else [1.0]
        for m in masses:
            system.addParticle(m)
        force = LBMForce()
        force.setGridSize(nx, ny, nz)
        force.setFluidDensity(602.214)
        force.setKinematicViscosity((self.tau - 0.5)/3.0*self.dx*self.dx/self.dt)
        force.setFluidMomentumRemovalFrequency(0)
        force.setStorage(storage)
        for i in range(len(self.masses)):
            force.addParticle(i)
        for fn in self.configure:
            fn(force)
        extra = [wall_force_function(*self.wall)] if self.wall is not None else []
        ctx = Context(system, VerletIntegrator(self.dt), force, library=library, extraForces=extra)
        pos = self.positions if len(self.masses) else np.array([[0.1, 0.2, 0.3]])
        vel = self.velocities if len(self.masses) else np.zeros((1, 3))
        ctx.setPositions(pos)
        ctx.setVelocities(vel)
        if self.state is not None:
            ctx.setFluidState(self.state)
        return ctx

    def device_particles(self, ctx):
        from .harness import DeviceParticles
        dp = DeviceParticles(ctx)
        if self.wall is not None:
            dp.setWallPotential(*self.wall)
        if self.bonds is not None:
            dp.setBonds(*self.bonds)
        return dp


def config1(stencil=LBMForce.Trilinear, num_beads=64, seed=1234):
    n = 32
    w = Workload("config1", (n, n, n), 0.8)
    z = np.arange(n**3)//(n*n)
    u = np.zeros((n**3, 3))
    u[:, 0] = 0.01*np.sin(2*np.pi*z/n)
    w.state = _equilibrium(n**3, u)
    rng = np.random.default_rng(seed)
    w.masses = [100.0]*num_beads
    w.positions = rng.uniform(0, n, size=(num_beads, 3))
    w.velocities = np.zeros((num_beads, 3))

    def conf(f):
        f.setBodyAcceleration((1e-5*w.dx/w.dt**2, 0, 0))
        f.setFriction(10.0)
        f.setTemperature(0.0)
        f.setInterpolationStencil(stencil)
    w.configure.append(conf)
    w.steps = 1000
    w.description = {"workload": "config1 32^3 periodic D3Q19 tau=0.8 shear wave + Guo 1e-5, 64 beads trilinear, T=0",
                     "lattice": [n, n, n], "beads": num_beads, "stencil": "trilinear", "fluctuations": False}
    return w


def config2(n=256, num_beads=20000, fluctuations=True, stencil=LBMForce.Keys, seed=20240607):
    w = Workload("config2", (n, n, n), 0.6)
    solid = [i + n*(j + n*k) for k in (0, n - 1) for j in range(n) for i in range(n)]
    rng = np.random.default_rng(seed)
    w.masses = [100.0]*num_beads
    pos = rng.uniform(0, n, size=(num_beads, 3))
    pos[:, 2] = rng.uniform(1.5, n - 2.5, size=num_beads)
    w.positions = pos*w.dx
    w.velocities = np.zeros((num_beads, 3))
    w.wall = (BOLTZ*300.0, 0.5*w.dx, 0.5*w.dx, (n - 1.5)*w.dx) if num_beads > 0 else None

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(LBMForce.BounceBack)
        f.setBodyAcceleration((1e-6*w.dx/w.dt**2, 0, 0))
        f.setFriction(10.0)
        f.setTemperature(300.0)
        f.setFluidFluctuations(fluctuations)
        f.setRandomNumberSeed(seed)
        f.setInterpolationStencil(stencil)
    w.configure.append(conf)
    w.steps = 2000
    w.description = {"workload": f"config2 {n}^3 Poiseuille, bounce-back walls z=0,{n-1}, Guo 1e-6, tau=0.6, "
                                 f"{num_beads} beads Keys, fluctuating kT=2.494",
                     "lattice": [n, n, n], "beads": num_beads, "stencil": "keys", "fluctuations": bool(fluctuations),
                     "walls": "bounce-back z=0,255", "tau": 0.6, "dx_nm": w.dx, "dt_ps": w.dt}
    return w


def config4(num_replicas=64, steps=1000, seed=4242, edges=(32, 48, 64, 96), max_beads=2000):
    """64 independent periodic cubes: edge drawn from {32, 48, 64, 96}, tau uniform in [0.55, 1.0], 50-2000 beads of
    100 Da (uniform positions, v = 0, gamma = 10/ps, trilinear), fluctuating fluid with kT uniform in [1.5, 3.0]
    kJ/mol (particles at the same temperature) and a Philox seed per replica; dx = 1 nm, dt = 0.01 ps,
    rho0 = 602.214 Da/nm^3; FP32."""
    from .batch import Replica
    rng = np.random.default_rng(seed)
    dx, dt = 1.0, 0.01
    replicas = []
    rows = []
    for r in range(num_replicas):
        n = int(rng.choice(edges))
        tau = float(rng.uniform(0.55, 1.0))
        beads = int(rng.integers(50, max_beads + 1))
        kT = float(rng.uniform(1.5, 3.0))
        rseed = int(rng.integers(1, 2**31 - 1))
        force = LBMForce()
        force.setGridSize(n, n, n)
        force.setFluidDensity(602.214)
        force.setKinematicViscosity((tau - 0.5)/3.0*dx*dx/dt)
        force.setFluidMomentumRemovalFrequency(0)
        force.setFriction(10.0)
        force.setTemperature(kT/BOLTZ)
        force.setFluidFluctuations(True)
        force.setRandomNumberSeed(rseed)
        force.setInterpolationStencil(LBMForce.Trilinear)
        for i in range(beads):
            force.addParticle(i)
        pos = rng.uniform(0, n*dx, size=(beads, 3))
        replicas.append(Replica(force, (n*dx, n*dx, n*dx), dt, np.full(beads, 100.0), pos))
        rows.append({"edge": n, "tau": tau, "beads": beads, "kT": kT, "seed": rseed})
    nodes = sum(row["edge"]**3 for row in rows)
    beads = sum(row["beads"] for row in rows)
    description = {"workload": f"config4 {num_replicas} periodic replicas, edges {sorted(set(edges))}, tau U[0.55,1.0], "
                               f"50-{max_beads} beads trilinear, fluctuating kT U[1.5,3.0], per-replica seeds",
                   "replicas": num_replicas, "total_nodes": nodes, "total_beads": beads, "steps": steps,
                   "dx_nm": dx, "dt_ps": dt}
    return replicas, rows, description


def config5(n=128, num_chains=2000, chain_length=50, seed=5150, temperature=300.0):
    """128^3 periodic, tau = 0.7, dx = 1 nm, dt = 0.01 ps, rho0 = 602.214 Da/nm^3; 2,000 chains of 50 beads (100 Da,
    gamma = 10/ps) bonded by harmonic springs k = 1000 kJ/mol/nm^2, r0 = 0.38 nm, started as seeded random walks of
    step r0 from uniform random origins; fluctuating fluid and particle noise at 300 K; Keys stencil (the yardstick and
    fallback of the closure); removal off."""
    w = Workload("config5", (n, n, n), 0.7)
    rng = np.random.default_rng(seed)
    count = num_chains*chain_length
    pos = np.zeros((count, 3))
    bonds = []
    r0 = 0.38
    for c in range(num_chains):
        start = c*chain_length
        pos[start] = rng.uniform(0, n*w.dx, size=3)
        steps = rng.normal(size=(chain_length - 1, 3))
        steps *= r0/np.linalg.norm(steps, axis=1)[:, None]
        pos[start + 1:start + chain_length] = pos[start] + np.cumsum(steps, axis=0)
        bonds.extend((start + i, start + i + 1) for i in range(chain_length - 1))
    w.positions = pos % (n*w.dx)
    w.masses = [100.0]*count
    w.velocities = np.zeros((count, 3))
    w.bonds = (np.array(bonds, dtype=np.int32), 1000.0, r0)

    def conf(f):
        f.setFriction(10.0)
        f.setTemperature(temperature)
        f.setFluidFluctuations(True)
        f.setRandomNumberSeed(seed)
        f.setInterpolationStencil(LBMForce.Keys)
    w.configure.append(conf)
    w.steps = 1000
    w.description = {"workload": f"config5 {n}^3 periodic, tau=0.7, {num_chains} harmonic chains x {chain_length} beads "
                                 f"(k=1000, r0=0.38), fluctuating kT={BOLTZ*temperature:.3f}, Keys yardstick",
                     "lattice": [n, n, n], "beads": count, "chains": num_chains, "chain_length": chain_length,
                     "dx_nm": w.dx, "dt_ps": w.dt, "tau": 0.7}
    return w
