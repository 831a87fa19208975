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


# ------------------------------------------------------------------------------------------------------------------
# The precision form at the shape of its configuration (config2 without fluctuations: bounce-back planes, a body
# force, Keys beads under a repulsive wall potential), against the FP64 oracle.
# ------------------------------------------------------------------------------------------------------------------
C3_N, C3_STEPS, C3_CHECK = 32, 400, 100


def config3_setup(n=C3_N, num_beads=200, g=1e-5, seed=2024):
    """n^3 lattice, solid planes z = 0 and z = n - 1 (bounce-back), tau = 0.6, Guo body force g (lattice units) along
    x; num_beads Keys beads of 100 Da (gamma = 10/ps) uniform in 1.5 <= z <= n - 2.5 under the wall potential of
    config2 (eps = kT at 300 K, sigma = 0.5 nm); T = 0, no fluctuations."""
    from lbm_amd.workloads import BOLTZ, wall_force_function
    s = Setup((n, n, n), dx=1.0, dt=0.01, tau=0.6)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, n, size=(num_beads, 3))
    pos[:, 2] = rng.uniform(1.5, n - 2.5, size=num_beads)
    s.add_particles(pos, mass=100.0)
    solid = [i + n*(j + n*k) for k in (0, n - 1) for j in range(n) for i in range(n)]
    s.extra.append(wall_force_function(BOLTZ*300.0, 0.5, 0.5, n - 1.5))

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(LBMForce.BounceBack)
        f.setBodyAcceleration((g/1e-4, 0, 0))
        f.setFriction(10.0)
        f.setTemperature(0.0)
        f.setFluidFluctuations(False)
        f.setRandomNumberSeed(77)
        f.setInterpolationStencil(LBMForce.Keys)
    s.configure.append(conf)
    return s


def config3_samples(ctx, setup):
    """Every C3_CHECK steps: the mean profile u_x(z) (plane averages) and the coupling forces of the beads (the total
    force minus the wall potential's)."""
    n = setup.grid[0]
    out = []
    for _ in range(C3_STEPS//C3_CHECK):
        ctx.step(C3_CHECK)
        rho, u = ctx.getFluidFields()
        profile = u[:, 0].reshape(n, n*n).mean(axis=1)
        coupling = ctx.getForces() - ctx._otherForces()
        out.append((profile, coupling[setup.coupled]))
    return out


@pytest.fixture(scope="module")
def config3_oracle():
    import oracle_lib
    s = config3_setup()
    return s, config3_samples(s.build(oracle_lib.load_oracle()), s)


@pytest.mark.parametrize("storage, profile_bound, force_bound", [
    pytest.param(LBMForce.FP32, 1e-5, 1e-5, id="fp32"),
    pytest.param(LBMForce.FP16, 1e-3, 3e-3, id="fp16"),
    pytest.param(LBMForce.BF16, 1e-3, 2e-2, id="bf16")])
def test_config3_against_fp64_oracle(config3_oracle, storage, profile_bound, force_bound):
    """FP32, FP16 and BF16 storage against the FP64 oracle on the same run, every 100 steps over 400: the mean
    velocity profile u_x(z) within 1e-3 of its peak (the precision form's bar; FP32 within 1e-5), and the beads'
    coupling forces within the stated fraction of gamma m max|u_x|, the drag on a bead held still in the flow (FP32
    1e-5, FP16 3e-3, BF16 2e-2; measured 2.9e-8, 3.2e-4 and 2.9e-3, profiles 2.3e-9, 2.7e-5 and 2.2e-4).  At T = 0 the beads move with the fluid, so their forces are small differences of
    the interpolated velocity and of their own: the interpolation of 64 stored populations per node carries the
    storage's rounding (2^-11 relative for FP16, 2^-8 for BF16) straight into them."""
    setup, ref = config3_oracle
    gpu = setup.build(None, storage=storage)
    got = config3_samples(gpu, setup)
    gamma_m = 10.0*100.0
    worst_profile = worst_force = 0.0
    for (p_ref, f_ref), (p_gpu, f_gpu) in zip(ref, got):
        peak = np.abs(p_ref).max()
        worst_profile = max(worst_profile, np.abs(p_gpu - p_ref).max()/peak)
        worst_force = max(worst_force, np.abs(f_gpu - f_ref).max()/(gamma_m*peak))
    print(f"profile error {worst_profile:.2e} (bound {profile_bound:g}), bead-force error {worst_force:.2e} of "
          f"gamma m max|u| (bound {force_bound:g}); peak u_x {np.abs(ref[-1][0]).max():.3e} nm/ps")
    assert worst_profile < profile_bound
    assert worst_force < force_bound


@pytest.mark.parametrize("storage", REDUCED)
def test_reduced_coupling_order_is_bit_identical(storage, monkeypatch):
    """With FP16/BF16 storage each XCD couples a contiguous run of the beads in their processing order
    (LBM_AMD_COUPLE_XCD, read when a context is created).  That changes which work-group couples which bead, not the
    arithmetic: after 60 steps the fields and forces match the plain order's (every 8th work-group on each XCD) bit for
    bit, since couplings are independent, the rounding is keyed by node and step and the reactions add as integers.
    200 beads make 50 work-groups, so the order is not the identity."""
    runs = []
    for env in ("0", "1"):
        monkeypatch.setenv("LBM_AMD_COUPLE_XCD", env)
        setup = config3_setup()
        ctx = setup.build(None, storage=storage)
        ctx.step(60)
        rho, u = ctx.getFluidFields()
        runs.append((rho, u, ctx.getForces()))
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_config3_full_size_fp16():
    """config3 at its full size (256^3, walls at z = 0 and 255, body force 1e-6, 20,000 Keys beads, fluctuations off,
    T = 0), 400 device-resident steps with FP16 storage against the same run with FP64 storage (which agrees with the
    FP64 oracle to 1e-12 on the small lattices of the parity tests): the mean profile u_x(z) within 1e-3 of its peak,
    the fluid mass to 1e-6, every bead between the walls."""
    from lbm_amd import workloads
    profiles, masses, beads = {}, {}, {}
    for st in (LBMForce.FP64, LBMForce.FP16):
        w = workloads.config2(fluctuations=False)
        w.configure.append(lambda f: f.setTemperature(0.0))
        c = w.build(storage=st)
        dp = w.device_particles(c)
        m0, _ = c.getFluidTotals()
        dp.step(400)
        m1, _ = c.getFluidTotals()
        pos, _ = dp.pull()
        rho, u = c.getFluidFields()
        profiles[st] = u[:, 0].reshape(256, 256*256).mean(axis=1)
        masses[st] = abs(m1 - m0)/m0
        beads[st] = pos[:, 2]
        del dp, c
    ref = profiles[LBMForce.FP64]
    err = np.abs(profiles[LBMForce.FP16] - ref).max()/np.abs(ref).max()
    print(f"FP16 profile error {err:.2e}, mass drift FP16 {masses[LBMForce.FP16]:.1e} FP64 {masses[LBMForce.FP64]:.1e}")
    assert err < 1e-3
    assert masses[LBMForce.FP16] < 1e-6
    z = beads[LBMForce.FP16]
    assert z.min() > 0.0 and z.max() < 255.5
