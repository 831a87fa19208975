"""The device-resident step (include/lbm_amd_harness.h: lbm_begin_step_device / lbm_execute_device with the device
Verlet integrator), the path the benchmark times, against the host path of the same library (lbm_begin_step /
lbm_execute with the Python integrator, the path the oracle parity tests use) on a small config2-like setup: bounce-back
walls, the repulsive wall potential, bonded chains, the Keys stencil, fluctuations and particle noise.  The other forces
are computed on the device in one kernel and on the host with numpy, so the paths agree to rounding, not bit for bit."""
import numpy as np
import pytest

from lbm_amd import LBMForce
from lbm_amd.harness import DeviceParticles
from lbm_amd.workloads import BOLTZ, wall_force_function
from parity import Setup, rel_err

EPS, SIGMA, ZLO, ZHI = BOLTZ*300.0, 0.5, 0.5, 18.5
BOND_K, BOND_R0 = 1000.0, 0.38


def bond_force_function(pairs, k, r0, box):
    box = np.asarray(box, dtype=np.float64)

    def f(pos):
        out = np.zeros_like(pos)
        for i, j in pairs:
            d = pos[j] - pos[i]
            d -= box*np.rint(d/box)
            r = np.sqrt(d @ d)
            s = k*(r - r0)/r
            out[i] += s*d
            out[j] -= s*d
        return out
    return f


def chain_setup(chains=6, length=5):
    nx, ny, nz = 16, 16, 20
    s = Setup((nx, ny, nz), dx=1.0, dt=0.01, tau=0.6)
    rng = np.random.default_rng(8)
    pos, pairs = [], []
    for c in range(chains):
        x = rng.uniform([0, 0, 5], [nx, ny, nz - 6])
        for b in range(length):
            if b:
                step = rng.normal(size=3)
                x = x + BOND_R0*step/np.linalg.norm(step)
                pairs.append((c*length + b - 1, c*length + b))
            pos.append(x)
    s.add_particles(np.array(pos), mass=100.0, velocities=rng.normal(scale=0.2, size=(len(pos), 3)))
    solid = [i + nx*(j + ny*k) for k in (0, nz - 1) for j in range(ny) for i in range(nx)]

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(LBMForce.BounceBack)
        f.setBodyAcceleration((2.0, 0, 0))
        f.setFriction(10.0)
        f.setTemperature(300.0)
        f.setFluidFluctuations(True)
        f.setRandomNumberSeed(5)
        f.setInterpolationStencil(LBMForce.Keys)
    s.configure.append(conf)
    s.extra = [wall_force_function(EPS, SIGMA, ZLO, ZHI),
               bond_force_function(pairs, BOND_K, BOND_R0, (nx*s.dx, ny*s.dx, nz*s.dx))]
    return s, np.array(pairs, dtype=np.int32)


@pytest.mark.gpu
def test_device_step_matches_host_step():
    s, pairs = chain_setup()
    host = s.build()
    dev = s.build()
    dp = DeviceParticles(dev)
    dp.setWallPotential(EPS, SIGMA, ZLO, ZHI)
    dp.setBonds(pairs, BOND_K, BOND_R0)
    for block in range(4):
        host.step(10)
        dp.step(10)
        pos, vel = dp.pull()
        rho0, u0 = host.getFluidFields()
        rho1, u1 = dev.getFluidFields()
        errs = {"position": rel_err(pos, host.positions, scale=s.dx), "velocity": rel_err(vel, host.velocities),
                "density": rel_err(rho1, rho0), "fluid": rel_err(u1, u0),
                "wall": rel_err(dev.getWallForce(), host.getWallForce())}
        print(f"after {10*(block + 1)} steps: {errs}", flush=True)
        assert errs["position"] < 1e-10 and errs["velocity"] < 1e-8, errs
        assert errs["density"] < 1e-9 and errs["fluid"] < 1e-8 and errs["wall"] < 1e-8, errs
    # The next step's forces as a State would report them: the coupling force added once to the other forces.
    f_host = host.getForces()
    f_dev = dev.getForces()
    assert rel_err(f_dev, f_host) < 1e-8
