"""Reduced-precision storage of the populations (FP16 and BF16 with stochastic rounding; FP32 math), the precision form
of the build (no counterpart in the reference).  The bar is the velocity field: the steady channel profile within 1e-3
of its peak against FP64 storage, which the closed-form tests pin to the oracle at 1e-9.  Stochastic rounding must
also keep mass and momentum unbiased: their drift is a random walk, not a trend."""
import numpy as np
import pytest

from lbm_amd import LBMForce
from parity import Setup, equilibrium_state, total_moments

pytestmark = pytest.mark.gpu

REDUCED = [pytest.param(LBMForce.FP16, id="fp16"), pytest.param(LBMForce.BF16, id="bf16")]


def channel(storage, nx=64, ny=18, nz=64, tau=1.0, umax=0.02):
    """Poiseuille flow along x between the walls of the solid plane j = 0 (bounce-back), driven by a body force that
    gives the peak velocity umax (lattice units)."""
    dx, dt = 1.0, 0.01
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    nu = (tau - 0.5)/3.0
    h = ny - 1
    g = 8.0*nu*umax/h**2
    solid = [i + nx*ny*k for k in range(nz) for i in range(nx)]
    s.configure.append(lambda f: (f.setSolidNodes(solid), f.setBodyAcceleration((g*dx/dt**2, 0.0, 0.0)),
                                  f.setRandomNumberSeed(11)))          # the stochastic rounding's bits
    return s.build(None, storage=storage)


def mean_profile(c, steps, every, shape):
    """u_x(y) averaged over x, z and over samples taken every `every` steps during `steps` steps."""
    nx, ny, nz = shape
    acc = np.zeros(ny)
    count = 0
    for _ in range(steps//every):
        c.step(every)
        rho, u = c.getFluidFields()
        acc += u[:, 0].reshape(nz, ny, nx).mean(axis=(0, 2))
        count += 1
    return acc/count


@pytest.mark.parametrize("storage", REDUCED)
def test_reduced_storage_channel_profile(storage):
    shape = (64, 18, 64)
    profiles = {}
    for st in (LBMForce.FP64, storage):
        c = channel(st)
        c.step(2000)                       # the slowest mode decays as exp(-t/176): converged to 1e-5
        profiles[st] = mean_profile(c, 400, 20, shape)
    ref = profiles[LBMForce.FP64]
    err = np.abs(profiles[storage] - ref).max()/np.abs(ref).max()
    print(f"u_x(y) profile error against FP64 storage: {err:.2e}")      # measured: FP16 1.8e-6, BF16 5.6e-5
    assert err < 1e-3


@pytest.mark.parametrize("storage", REDUCED)
def test_reduced_storage_conservation(storage):
    """A periodic 16^3 fluid with a shear wave and density bumps, no forces: 1000 steps with stochastic rounding keep
    the mass and the momentum to the size of a random walk of the roundings.  Measured with these seeds: FP16 1.4e-7
    (mass) and 2.2e-5 (momentum, relative to 0.02 N), BF16 1.7e-6 and 6.6e-5 (other seeds: up to 3.6e-7, 1.5e-5,
    2e-6 and 1.8e-4); the bounds are 5-70 times that.  A rounding biased by half an
    ulp (truncation) would add up to 19 N half-ulps a step, of the order of 1e-2 of the mass over 1000 steps at these
    magnitudes (an estimate), about a thousand times the bound."""
    n = 16
    s = Setup((n, n, n), dx=1.0, dt=0.01, tau=0.8)
    node = np.arange(n**3)
    j = (node//n) % n
    u = np.zeros((n**3, 3))
    u[:, 0] = 0.02*np.sin(2*np.pi*j/n)
    u[:, 2] = 0.01
    rho = 1.0 + 0.01*np.cos(2*np.pi*(node % n)/n)
    s.state = equilibrium_state(n**3, rho, u)
    s.configure.append(lambda f: f.setRandomNumberSeed(7))
    c = s.build(None, storage=storage)
    m0, p0 = total_moments(c.getFluidState())
    c.step(1000)
    m1, p1 = total_moments(c.getFluidState())
    scale = 0.02*n**3
    print(f"mass drift {abs(m1 - m0)/m0:.2e}, momentum drift {np.abs(p1 - p0).max()/scale:.2e}")
    bound = {LBMForce.FP16: (1e-5, 2e-4), LBMForce.BF16: (5e-5, 1e-3)}[storage]
    assert abs(m1 - m0)/m0 < bound[0]
    assert np.abs(p1 - p0).max()/scale < bound[1]
