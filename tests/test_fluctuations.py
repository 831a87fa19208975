"""Statistics of the fluctuating fluid and of the coupled particles, on the FP64 oracle (CPU) and on the gfx950 library
(GPU), against the closed forms the reference's tests assert (TestLBMFluctuations.h, TestLBMCoupling.h):

* a single node (1x1x1 lattice, streaming is the identity): mass and momentum exact, and the moments k = 4..18 of the
  orthogonal basis uncorrelated with variance mu b_k, mu = kT/cs^2 (TestLBMFluctuations.h:78-110), 5%;
* a fluid at rest on 8^3: density, momentum and every other moment at their equilibrium variances mu, kT and mu b_k
  (TestLBMFluctuations.h:117-153), 5%;
* the velocity spectrum on 8^3: longitudinal kT/rho and transverse 2kT/rho per band of |k|, node variance
  (N-1)/N kT/rho (TestLBMFluctuations.h:570-651), 5% and 3%, at rest, in uniform flow and under a body force;
* zero temperature with fluctuations on is the run without fluctuations, bit for bit;
* particles with the explicit drag in a fluid without fluctuations: full-step temperature T and half-step temperature
  T/(1 - gamma dt/2) (TestLBMCoupling.h:612-663), 10%;
* particles in a fluctuating fluid: the bead temperature and the fluid temperature, and (GPU) the bead velocity
  autocorrelation function of the GPU against the oracle's with another seed, within their statistical error.

The tolerances are those of the reference's tests; the sampling is the same or longer.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from lbm_amd.workloads import BOLTZ
from parity import Setup

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]
DT, DX, RHO0 = 0.01, 1.0, 602.214
B_NORM = np.array([1, 1/3, 1/3, 1/3, 2/9, 2/9, 2/9, 1/9, 1/9, 1/9, 2/27, 2/27, 2/27, 1/18, 1/18, 1/18, 7/162, 5/126,
                   1/30])


def library(backend):
    return oracle_lib.load_oracle() if backend == "oracle" else None


def lattice_kT(temperature):
    return BOLTZ*temperature*DT*DT/(RHO0*DX**5)


def mode_matrix():
    """E[k][q] = e_k(c_q) of the orthogonal basis, from the oracle's table (pinned by test_fluctuation_basis)."""
    lib = oracle_lib.load_oracle()
    return np.array([[lib.oracle_mode(k, q) for q in range(19)] for k in range(19)])


def fluid_setup(grid, tau, temperature=300.0, seed=5, fluct=True):
    s = Setup(grid, tau=tau)

    def conf(f):
        f.setFluidFluctuations(fluct)
        f.setTemperature(temperature)
        f.setRandomNumberSeed(seed)
    s.configure.append(conf)
    return s


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("tau", [1.0, 0.8])
def test_single_node_mode_variances(backend, tau):
    ctx = fluid_setup((1, 1, 1), tau).build(library(backend))
    E = mode_matrix()
    mu = 3.0*lattice_kT(300.0)
    warmup, samples = 50, 20000
    cov = np.zeros((19, 19))
    conserved = 0.0
    for step in range(warmup + samples):
        ctx.step(1)
        m = E @ ctx.getFluidState()
        conserved = max(conserved, np.abs(m[:4]).max())
        if step >= warmup:
            cov += np.outer(m, m)
    # FP32 storage rounds the 19 deviations (~1e-3) of every step by ~1e-10 each: over 20,000 steps the conserved moments
    # random-walk to ~sqrt(20000) x 19 x 1e-10 ~ 3e-7 at most (4.7e-9 measured).
    assert conserved < (1e-15 if backend == "oracle" else 3e-7)
    ratio = cov[4:, 4:]/(samples*mu*np.sqrt(np.outer(B_NORM[4:], B_NORM[4:])))
    print("diagonal", np.round(np.diag(ratio), 3), "max off-diagonal", np.abs(ratio - np.diag(np.diag(ratio))).max())
    assert np.abs(ratio - np.eye(15)).max() < 0.05


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("tau", [0.8, 2.5])
def test_equilibrium_mode_variances(backend, tau):
    n = 8
    ctx = fluid_setup((n, n, n), tau, seed=11).build(library(backend))
    E = mode_matrix()
    kT = lattice_kT(300.0)
    mu = 3.0*kT
    ctx.step(300)
    sums = np.zeros(19)
    samples = 300
    for _ in range(samples):
        ctx.step(5)
        m = E @ ctx.getFluidState().reshape(19, -1)
        sums += (m*m).sum(axis=1)
    mass, momentum = ctx.getFluidTotals()      # lattice units
    tol = 1e-13 if backend == "oracle" else 1e-5
    assert abs(mass - n**3) <= tol*n**3
    assert np.abs(momentum).max() <= tol*n**3
    count = samples*n**3
    ratios = np.concatenate([[sums[0]/(count*mu)], sums[1:4]/(count*kT), sums[4:]/(count*mu*B_NORM[4:])])
    print("variance ratios", np.round(ratios, 3))
    assert np.abs(ratios - 1).max() < 0.05


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("flow", ["rest", "uniform", "accelerated"])
def test_velocity_spectrum(backend, flow):
    n = 8
    u0 = np.array([0.05, 0.02, -0.03]) if flow == "uniform" else np.zeros(3)
    g = np.array([1e-5, 0, 0]) if flow == "accelerated" else np.zeros(3)
    s = fluid_setup((n, n, n), 1.0, seed=13)
    s.configure.append(lambda f: (f.setInitialFluidVelocity(tuple(u0*DX/DT)), f.setBodyAcceleration(tuple(g*DX/DT**2))))
    ctx = s.build(library(backend))
    kT = lattice_kT(300.0)
    m = np.fft.fftfreq(n, 1.0/n)
    mz, my, mx = np.meshgrid(m, m, m, indexing="ij")
    m2 = mx**2 + my**2 + mz**2
    band = np.where(m2 <= 3, 0, np.where(m2 <= 12, 1, 2))
    band[0, 0, 0] = -1
    khat = np.stack([mx, my, mz], axis=-1)/np.sqrt(np.where(m2 > 0, m2, 1))[..., None]
    warmup, samples, interval = 300, 400, 5
    ctx.step(warmup)
    L, T, count = np.zeros(3), np.zeros(3), np.zeros(3)
    node_var = 0.0
    for _ in range(samples):
        ctx.step(interval)
        _, u = ctx.getFluidFields()
        mean = u.mean(axis=0)
        du = (u - mean)*(DT/DX)
        node_var += (du*du).sum()
        F = np.fft.fftn(du.reshape(n, n, n, 3), axes=(0, 1, 2))/np.sqrt(n**3)
        power = (np.abs(F)**2).sum(axis=-1)
        powerL = np.abs((F*khat).sum(axis=-1))**2
        for b in range(3):
            sel = band == b
            L[b] += powerL[sel].sum()
            T[b] += (power - powerL)[sel].sum()
            count[b] += sel.sum()
    t = warmup + samples*interval
    assert np.abs(mean*(DT/DX) - (u0 + g*t)).max() < 2e-4
    print("longitudinal", np.round(L/(count*kT), 3), "transverse", np.round(T/(count*2*kT), 3),
          "node", round(node_var/(samples*(n**3 - 1)*3*kT), 4))
    assert np.abs(L/(count*kT) - 1).max() < 0.05
    assert np.abs(T/(count*2*kT) - 1).max() < 0.05
    assert abs(node_var/(samples*(n**3 - 1)*3*kT) - 1) < 0.03


@pytest.mark.parametrize("backend", BACKENDS)
def test_zero_temperature_is_no_fluctuations(backend):
    n = 8
    rng = np.random.default_rng(4)
    runs = []
    for fluct in (True, False):
        s = fluid_setup((n, n, n), 0.8, temperature=0.0, fluct=fluct)
        s.add_particles(rng.uniform(0, n, size=(10, 3)) if fluct else runs[0][2], mass=100.0)
        s.configure.append(lambda f: (f.setFriction(10.0), f.setInterpolationStencil(LBMForce.Trilinear)))
        ctx = s.build(library(backend))
        ctx.step(30)
        runs.append((ctx.getFluidState(), ctx.positions.copy(), np.array(s.positions)))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def coupled_setup(num_particles, temperature, fluid_fluct, seed=1, stencil=LBMForce.Trilinear):
    n = 8
    s = fluid_setup((n, n, n), 0.8, temperature=temperature, seed=seed, fluct=fluid_fluct)
    i = np.arange(num_particles)
    pos = np.stack([np.fmod(0.37*i, 4.0), np.fmod(0.53*i + 0.1, 4.0), np.fmod(0.71*i + 0.2, 4.0)], axis=1)*2.0
    s.add_particles(pos, mass=100.0)
    s.configure.append(lambda f: (f.setFriction(10.0), f.setInterpolationStencil(stencil)))
    return s


def bead_temperatures(ctx, samples):
    """Full-step (mean of consecutive half steps) and half-step temperatures of the beads, in K."""
    prev = ctx.velocities.copy()
    full = half = 0.0
    trace = []
    for _ in range(samples):
        ctx.step(1)
        v = ctx.velocities
        mid = 0.5*(v + prev)
        full += 100.0*(mid*mid).sum()
        half += 100.0*(v*v).sum()
        trace.append(v.copy())
        prev = v.copy()
    scale = 1.0/(3*len(prev)*samples*BOLTZ)
    return full*scale, half*scale, np.array(trace)


@pytest.mark.parametrize("backend", BACKENDS)
def test_particle_equipartition(backend):
    """Free particles, explicit drag, fluid without fluctuations (TestLBMCoupling.h:612-663)."""
    T, gamma = 300.0, 10.0
    ctx = coupled_setup(100, T, False).build(library(backend))
    ctx.step(500)
    full, half, _ = bead_temperatures(ctx, 2000)
    print(f"full-step {full:.1f} K, half-step {half:.1f} K (expected {T} and {T/(1 - 0.5*gamma*DT):.1f})")
    assert abs(full/T - 1) < 0.1
    assert abs(half/(T/(1 - 0.5*gamma*DT)) - 1) < 0.1


@pytest.mark.parametrize("backend", BACKENDS)
def test_fluctuating_fluid_and_particle_temperatures(backend):
    """Fluid and particles both thermalized at 300 K: the fluid's node velocity variance is (N-1)/N kT/rho per
    component (3%) and the beads' full-step temperature is T (10%, as above)."""
    T = 300.0
    ctx = coupled_setup(100, T, True, seed=3).build(library(backend))
    ctx.step(500)
    kT = lattice_kT(T)
    fluid = 0.0
    prev = ctx.velocities.copy()
    full = 0.0
    samples = 1500
    for k in range(samples):
        ctx.step(1)
        v = ctx.velocities
        mid = 0.5*(v + prev)
        full += 100.0*(mid*mid).sum()
        prev = v.copy()
        if k % 10 == 0:
            _, u = ctx.getFluidFields()
            du = (u - u.mean(axis=0))*(DT/DX)
            fluid += (du*du).sum()/(3*(len(du) - 1))
    fluid_ratio = fluid/(samples//10)/kT
    bead_T = full/(3*100*samples*BOLTZ)
    print(f"fluid variance ratio {fluid_ratio:.4f}, bead temperature {bead_T:.1f} K")
    assert abs(fluid_ratio - 1) < 0.03
    assert abs(bead_T/T - 1) < 0.1


def vacf(trace, lags):
    v = trace - trace.mean(axis=(0, 1))
    c = np.array([np.mean(np.sum(v[:len(v) - k]*v[k:], axis=2)) for k in range(lags)])
    return c


@pytest.mark.gpu
def test_velocity_autocorrelation_gpu_against_oracle():
    """The bead velocity autocorrelation C(t)/C(0) in a fluctuating fluid, GPU (seed 21) against the oracle (seed 22):
    equal within 5 standard errors at every lag (standard errors from 10 blocks of each run)."""
    lags, samples = 40, 3000
    curves, errors = [], []
    for backend, seed in (("gpu", 21), ("oracle", 22)):
        ctx = coupled_setup(100, 300.0, True, seed=seed).build(library(backend))
        ctx.step(300)
        _, _, trace = bead_temperatures(ctx, samples)
        blocks = np.array([vacf(b, lags) for b in np.array_split(trace, 10)])
        blocks /= blocks[:, :1]
        curves.append(blocks.mean(axis=0))
        errors.append(blocks.std(axis=0, ddof=1)/np.sqrt(len(blocks)))
    diff = np.abs(curves[0] - curves[1])
    sigma = np.sqrt(errors[0]**2 + errors[1]**2)
    print("C(t)/C(0) gpu", np.round(curves[0][:10], 3), "oracle", np.round(curves[1][:10], 3))
    assert np.all(diff[1:] <= 5*sigma[1:] + 1e-3)
