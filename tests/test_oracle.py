"""Pins the FP64 CPU oracle (oracle/lbm_oracle.cpp) before it is trusted as the checker of the GPU path.

 - Philox4x32-10 against the published known-answer vectors of Random123 (kat_vectors, philox4x32_10).
 - The D3Q19 algebra against the reference's own openmmapi/include/internal/D3Q19.h, compiled into oracle/_ref.
 - Identities of the algebra and of the stencils that a dropped term, a wrong sign or index would break.
 - The closed forms of the reference's own tests (tests/TestLBMFluid.h, TestLBMCoupling.h, TestLBMStencils.h).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from parity import CX, CY, CZ, W, Setup, equilibrium_state, total_moments

_D = ctypes.POINTER(ctypes.c_double)


def dptr(a):
    return a.ctypes.data_as(_D)


@pytest.fixture(scope="module")
def lib():
    return oracle_lib.load_oracle()


# ------------------------------------------------------------------------------------------------
# Philox4x32-10 known answers (Salmon et al. SC'11; Random123 kat_vectors).
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,expected", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff]*4, [0xffffffff]*2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(lib, ctr, key, expected):
    c = (ctypes.c_uint32*4)(*ctr)
    k = (ctypes.c_uint32*2)(*key)
    out = (ctypes.c_uint32*4)()
    lib.oracle_philox4x32(c, k, out)
    assert list(out) == expected


def test_normals_are_standard(lib):
    """Box-Muller normals from Philox: mean 0, variance 1, kurtosis 3 over 2^17 samples (5 sigma bounds)."""
    count = 16
    samples = []
    buf = np.zeros(count)
    for index in range(8192):
        lib.oracle_normals(7, 0x464C5549, index, 11, count, dptr(buf))
        samples.append(buf.copy())
    x = np.concatenate(samples)
    n = x.size
    assert abs(x.mean()) < 5/math.sqrt(n)
    assert abs(x.var() - 1) < 5*math.sqrt(2/n)
    assert abs((x**4).mean() - 3) < 5*math.sqrt(96/n)
    # Different steps and indices give different numbers.
    a, b = np.zeros(4), np.zeros(4)
    lib.oracle_normals(7, 1, 0, 0, 4, dptr(a))
    lib.oracle_normals(7, 1, 0, 1, 4, dptr(b))
    assert not np.array_equal(a, b)


# ------------------------------------------------------------------------------------------------
# D3Q19 algebra against the reference's own header (compiled in oracle/_ref where the tree was present).
# ------------------------------------------------------------------------------------------------
def random_nodes(count, seed=0):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        dr = rng.uniform(-0.05, 0.05)
        u = rng.uniform(-0.1, 0.1, 3)
        df = rng.uniform(-0.01, 0.01, 19)
        f = rng.uniform(-1e-3, 1e-3, 3)
        r = rng.normal(size=19)
        yield dr, u, df, f, r


@pytest.mark.skipif(oracle_lib.load_ref() is None, reason="oracle/_ref is built only where the reference tree is present")
def test_algebra_matches_reference_header(lib):
    ref = oracle_lib.load_ref()
    a, b = np.zeros(19), np.zeros(19)
    p, q = np.zeros(6), np.zeros(6)
    for dr, u, df, f, r in random_nodes(200):
        lib.oracle_equilibrium_deviation(dr, *u, dptr(a))
        ref.ref_equilibrium_deviation(dr, *u, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-16)
        j = u*(1 + dr)
        lib.oracle_pi_neq(dptr(df), dr, *j, dptr(p))
        ref.ref_pi_neq(dptr(df), dr, *j, dptr(q))
        assert np.allclose(p, q, rtol=0, atol=1e-16)
        lib.oracle_regularized_neq(dptr(p), dptr(a))
        ref.ref_regularized_neq(dptr(p), dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-17)
        lib.oracle_guo(*u, *f, dptr(a))
        ref.ref_guo(*u, *f, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-18)
        rr = np.ascontiguousarray(r)
        lib.oracle_collision_noise(dptr(rr), 0.3, 0.7, dptr(a))
        ref.ref_collision_noise(dptr(rr), 0.3, 0.7, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-15)
        lib.oracle_equilibrium_noise(dptr(rr), 0.2, 0.3, 0.7, dptr(a))
        ref.ref_equilibrium_noise(dptr(rr), 0.2, 0.3, 0.7, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-15)
    for k in range(19):
        for qq in range(19):
            assert lib.oracle_mode(k, qq) == ref.ref_mode(k, qq)
    for omega in (0.7, 1.0, 1.6):
        assert ref.ref_boundary_stress_weight(omega) == pytest.approx((1 - omega)**2/(omega*(2 - omega)))


def test_algebra_identities(lib):
    """sum dfeq = drho, sum c dfeq = rho u; sum S = 0, sum c S = F (D3Q19.h:123-126); Pi_neq of the equilibrium is 0;
    the regularized part carries no mass or momentum and reproduces Pi (H2 projection)."""
    eq = np.zeros(19)
    s = np.zeros(19)
    pi = np.zeros(6)
    neq = np.zeros(19)
    c = np.stack([CX, CY, CZ])
    for dr, u, df, f, r in random_nodes(100, seed=1):
        lib.oracle_equilibrium_deviation(dr, *u, dptr(eq))
        assert eq.sum() == pytest.approx(dr, abs=1e-15)
        assert np.allclose(c @ eq, (1 + dr)*u, atol=1e-15)
        lib.oracle_guo(*u, *f, dptr(s))
        assert abs(s.sum()) < 1e-18
        assert np.allclose(c @ s, f, atol=1e-18)
        lib.oracle_pi_neq(dptr(eq), dr, *((1 + dr)*u), dptr(pi))
        assert np.allclose(pi, 0, atol=1e-15)
        pi_in = np.ascontiguousarray(np.random.default_rng(2).normal(size=6)*1e-3)
        lib.oracle_regularized_neq(dptr(pi_in), dptr(neq))
        assert abs(neq.sum()) < 1e-17 and np.allclose(c @ neq, 0, atol=1e-17)
        moments = np.array([(CX*CX*neq).sum(), (CY*CY*neq).sum(), (CZ*CZ*neq).sum(), (CX*CY*neq).sum(),
                            (CX*CZ*neq).sum(), (CY*CZ*neq).sum()])
        trace = pi_in[:3].sum()/3
        assert np.allclose(moments[:3] - (neq.sum())/3, pi_in[:3] - 0*trace, atol=1e-15)
        assert np.allclose(moments[3:], pi_in[3:], atol=1e-16)


def test_fluctuation_basis(lib):
    """Orthogonality sum_q w_q e_k e_l = b_k delta_kl (TestLBMFluctuations.h:58-69); the random part conserves mass and
    momentum and gives mode k the amplitude phi_k r_k."""
    e = np.array([[lib.oracle_mode(k, q) for q in range(19)] for k in range(19)])
    gram = (e*W) @ e.T
    norms = np.array([1, 1/3, 1/3, 1/3, 2/9, 2/9, 2/9, 1/9, 1/9, 1/9, 2/27, 2/27, 2/27, 1/18, 1/18, 1/18, 7/162, 5/126,
                      1/30])
    assert np.allclose(gram, np.diag(norms), atol=1e-15)
    r = np.ascontiguousarray(np.random.default_rng(3).normal(size=19))
    xi = np.zeros(19)
    lib.oracle_collision_noise(dptr(r), 0.4, 0.9, dptr(xi))
    assert abs(xi.sum()) < 1e-15 and np.allclose(np.stack([CX, CY, CZ]) @ xi, 0, atol=1e-15)
    modes = e @ xi
    for k in range(4, 19):
        amp = 0.4 if k < 10 else 0.9
        assert modes[k] == pytest.approx(amp*math.sqrt(norms[k])*r[k - 4], abs=1e-14)


# ------------------------------------------------------------------------------------------------
# Stencils (TestLBMStencils.h:27-124).
# ------------------------------------------------------------------------------------------------
def axis(lib, stencil, s, n=16, open_axis=False):
    idx = (ctypes.c_int*4)()
    w = np.zeros(4)
    lib.oracle_axis_weights(stencil, s, n, 1 if open_axis else 0, idx, dptr(w))
    width = {LBMForce.NearestNode: 1, LBMForce.Trilinear: 2, LBMForce.ThreePoint: 3, LBMForce.Keys: 4}[stencil]
    return np.array(idx[:width]), w[:width]


@pytest.mark.parametrize("stencil", [LBMForce.Trilinear, LBMForce.ThreePoint, LBMForce.Keys])
def test_stencil_moments(lib, stencil):
    for s in np.linspace(3.0, 4.0, 41):
        idx, w = axis(lib, stencil, s)
        assert w.sum() == pytest.approx(1.0, abs=1e-14)
        assert (w*(idx - s)).sum() == pytest.approx(0.0, abs=1e-14)
        if stencil == LBMForce.Keys:
            assert (w*(idx - s)**2).sum() == pytest.approx(0.0, abs=1e-13)
        if stencil == LBMForce.ThreePoint:
            assert (w*w).sum() == pytest.approx(0.5, abs=1e-13)


def test_stencil_at_node_and_wrap(lib):
    for stencil in (LBMForce.Trilinear, LBMForce.Keys):
        idx, w = axis(lib, stencil, 5.0)
        assert w[list(idx).index(5)] == pytest.approx(1.0) and np.abs(w).sum() == pytest.approx(1.0)
    idx, w = axis(lib, LBMForce.Keys, 0.3, n=16)
    assert list(idx) == [15, 0, 1, 2]
    idx, w = axis(lib, LBMForce.Keys, 0.3, n=16, open_axis=True)
    assert list(idx) == [0, 0, 1, 2] and w[1] == 0.0 and w.sum() == pytest.approx(1.0)
    idx, w = axis(lib, LBMForce.ThreePoint, 0.2, n=2)
    assert w.sum() == pytest.approx(1.0) and w[2] == 0.0


# ------------------------------------------------------------------------------------------------
# Closed forms of the reference's tests, on the oracle.
# ------------------------------------------------------------------------------------------------
def build(setup):
    return setup.build(oracle_lib.load_oracle())


def test_uniform_flow_is_steady():
    """TestLBMFluid.h:93-108."""
    s = Setup((6, 5, 4), dx=0.5, dt=0.01, tau=0.8)
    s.configure.append(lambda f: f.setInitialFluidVelocity((0.03*50, -0.02*50, 0.01*50)))
    c = build(s)
    before = c.getFluidState()
    c.step(50)
    assert np.abs(c.getFluidState() - before).max() < 1e-13


def test_conservation():
    """TestLBMFluid.h:113-135: mass and momentum of a perturbed fluid without forces."""
    s = Setup((6, 5, 4), dx=0.5, dt=0.01, tau=0.7)
    c = build(s)
    state = c.getFluidState()
    i = np.arange(state.size)
    state = state + 1e-3*np.sin(1.3*i + 0.7*(i % 11))
    c.setFluidState(state)
    m0, p0 = total_moments(state)
    c.step(100)
    m1, p1 = total_moments(c.getFluidState())
    assert m1 == pytest.approx(m0, abs=1e-13*m0)
    assert np.linalg.norm(p0) > 1e-5
    assert np.abs(p1 - p0).max() < 1e-13*m0


@pytest.mark.parametrize("rho0", [1.0, 1.02])
def test_body_force(rho0):
    """TestLBMFluid.h:141-169: momentum n rho g per node, velocity (n + 1/2) g dt."""
    dx, dt = 0.5, 0.01
    s = Setup((4, 3, 5), dx=dx, dt=dt, tau=0.9)
    g = np.array([0.5, -0.3, 0.2])
    s.configure.append(lambda f: f.setBodyAcceleration(g))
    s.state = equilibrium_state(60, rho0, np.zeros(3))
    c = build(s)
    c.step(20)
    m, p = total_moments(c.getFluidState())
    gl = g*dt*dt/dx
    assert m == pytest.approx(rho0*60, abs=1e-12)
    assert np.allclose(p, gl*20*rho0*60, atol=1e-13)
    rho, u = c.getFluidFields()
    assert np.allclose(u, g*20.5*dt, atol=1e-12)


def test_momentum_removal_timing():
    """TestLBMFluid.h:176-199: with frequency 3 the momentum after n steps is ((n-1)%3 + 1) F per node."""
    dx, dt = 0.5, 0.01
    s = Setup((4, 4, 4), dx=dx, dt=dt, tau=1.0)
    s.configure.append(lambda f: (f.setBodyAcceleration((0.5, 0, 0)), f.setFluidMomentumRemovalFrequency(3),
                                  f.setInitialFluidVelocity((0.5, 1.0, -0.5))))
    c = build(s)
    gl = 0.5*dt*dt/dx
    for n in range(1, 8):
        c.step(1)
        m, p = total_moments(c.getFluidState())
        assert np.allclose(p, [((n - 1) % 3 + 1)*gl*64, 0, 0], atol=1e-13)


@pytest.mark.parametrize("tau", [0.6, 1.0])
def test_shear_wave_viscosity(tau):
    """TestLBMFluid.h:205-251: u_x = U sin(k y) decays as exp(-nu k^2 t), nu = (tau - 1/2)/3, within 2e-3."""
    nx, ny, nz = 2, 64, 2
    s = Setup((nx, ny, nz), dx=0.5, dt=0.01, tau=tau)
    n = nx*ny*nz
    j = (np.arange(n)//nx) % ny
    k = 2*np.pi/ny
    u = np.zeros((n, 3))
    u[:, 0] = 1e-3*np.sin(k*j)
    s.state = equilibrium_state(n, 1.0, u)
    c = build(s)

    def amplitude():
        rho, vel = c.getFluidFields()
        return 2*(vel[:, 0]*np.sin(k*j)).sum()/n/(0.5/0.01)
    c.step(200)
    a1 = amplitude()
    c.step(1000)
    a2 = amplitude()
    nu = -math.log(a2/a1)/(1000*k*k)
    assert abs(nu/((tau - 0.5)/3) - 1) < 2e-3


def wall_plane(nx, ny, nz):
    return [i + nx*(0 + ny*kk) for kk in range(nz) for i in range(nx)]


@pytest.mark.parametrize("tau", [0.7, 0.875, 1.2])
def test_bounce_back_poiseuille(tau):
    """TestLBMFluid.h:486-535: exact steady profile of halfway bounce-back with the regularized collision,
    u(y) = g/(2 nu)(y - 1/2)(n - 1/2 - y) + g(16 Lambda - 3)/(24 nu), and the walls carry the body force."""
    nx, ny, nz = 2, 12, 2
    dx, dt = 0.5, 0.01
    g = 1e-5
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    s.configure.append(lambda f: (f.setBodyAcceleration((g*dx/dt**2, 0, 0)), f.setSolidNodes(wall_plane(nx, ny, nz))))
    c = build(s)
    nu, lam, h = (tau - 0.5)/3, (tau - 0.5)/2, ny - 1
    c.step(int(4*h*h/nu))
    rho, u = c.getFluidFields()
    umax = g/(2*nu)*(0.5*h)**2 + g*(16*lam - 3)/(24*nu)
    for node in range(nx*ny*nz):
        jj = (node//nx) % ny
        if jj == 0:
            assert rho[node] == 0 and np.all(u[node] == 0)
            continue
        exact = g/(2*nu)*(jj - 0.5)*(ny - 0.5 - jj) + g*(16*lam - 3)/(24*nu)
        assert u[node, 0]*dt/dx/umax == pytest.approx(exact/umax, abs=1e-9)
    mass = rho.sum()*dx**3
    assert np.allclose(c.getWallForce(), [g*dx/dt**2*mass, 0, 0], rtol=1e-8)


@pytest.mark.parametrize("tau", [0.8, 1.0])
def test_regularized_poiseuille(tau):
    """TestLBMFluid.h:590-627: u(y) = g/(2 nu) y (n - y) + 3 g (tau - 1)/(tau - 1/2), density of the fluid at rest."""
    nx, ny, nz = 2, 12, 2
    dx, dt = 0.5, 0.01
    g = 1e-5
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    s.configure.append(lambda f: (f.setBodyAcceleration((g*dx/dt**2, 0, 0)), f.setSolidNodes(wall_plane(nx, ny, nz)),
                                  f.setWallScheme(LBMForce.Regularized)))
    c = build(s)
    nu, h = (tau - 0.5)/3, ny
    c.step(int(4*h*h/nu))
    rho, u = c.getFluidFields()
    umax = g/(2*nu)*(0.5*h)**2
    for node in range(nx*ny*nz):
        jj = (node//nx) % ny
        if jj == 0:
            continue
        assert rho[node]/602.214 == pytest.approx(1.0, abs=1e-9)
        exact = g/(2*nu)*jj*(ny - jj) + 3*g*(tau - 1)/(tau - 0.5)
        assert u[node, 0]*dt/dx/umax == pytest.approx(exact/umax, abs=1e-9)
    mass = rho.sum()*dx**3
    assert np.allclose(c.getWallForce(), [g*dx/dt**2*mass, 0, 0], rtol=1e-8)


def wall_test_nodes(nx, ny, nz):
    return [i + nx*(j + ny*k) for k in range(nz) for j in range(ny) for i in range(nx)
            if i == 0 or (2 <= i < 4 and 1 <= j < 3 and 1 <= k < 3)]


@pytest.mark.parametrize("scheme", [LBMForce.BounceBack, LBMForce.Regularized])
def test_wall_momentum_balance(scheme):
    """TestLBMFluid.h:648-685: mass conserved; P(t+dt) - P(t) = (M g - F_wall) dt at every step."""
    nx, ny, nz = 6, 5, 4
    dx, dt = 0.5, 0.01
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=0.7)
    g = np.array([1e-5, 2e-5, -1e-5])*dx/dt**2
    s.configure.append(lambda f: (f.setSolidNodes(wall_test_nodes(nx, ny, nz)), f.setWallScheme(scheme),
                                  f.setInitialFluidVelocity((1.0, -0.5, 0.25)), f.setBodyAcceleration(g)))
    c = build(s)

    def moments():
        rho, u = c.getFluidFields()
        return (rho*dx**3).sum(), ((rho*dx**3)[:, None]*u).sum(axis=0)
    m0, p0 = moments()
    scale = m0*0.02*dx/dt
    for _ in range(30):
        c.step(1)
        m, p = moments()
        assert m == pytest.approx(m0, rel=1e-13)
        balance = p - p0 - (g*m - c.getWallForce())*dt
        assert np.linalg.norm(balance)/scale < 1e-12
        p0 = p


@pytest.mark.parametrize("boundary", [LBMForce.Velocity, LBMForce.DensityVelocity])
def test_couette(boundary):
    """TestLBMFluid.h:694-716: linear profile u_x(z) = U (z + 1)/(nz + 1), uniform density."""
    nx, ny, nz = 2, 2, 10
    dx, dt, tau, U = 0.5, 0.01, 0.9, 0.01
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    s.configure.append(lambda f: (f.setFaceBoundary(LBMForce.ZMin, boundary), f.setFaceBoundary(LBMForce.ZMax, boundary),
                                  f.setFaceVelocity(LBMForce.ZMax, (U*dx/dt, 0, 0))))
    c = build(s)
    nu, h = (tau - 0.5)/3, nz + 1
    c.step(int(6*h*h/nu))
    rho, u = c.getFluidFields()
    k = np.arange(nx*ny*nz)//(nx*ny)
    assert np.allclose(u[:, 0]*dt/(dx*U), (k + 1)/h, atol=1e-9)
    assert np.allclose(rho/602.214, 1.0, atol=1e-9)


def test_first_step_drag():
    """TestLBMCoupling.h:84-103: at T = 0 a particle in fluid at rest has v1 = v0 (1 - gamma dt) after one step."""
    s = Setup((8, 8, 8), dx=0.5, dt=0.01, tau=0.8)
    v0 = np.array([[0.5, -0.3, 0.2]])
    s.add_particles([[1.0, 1.2, 2.1]], mass=100.0, velocities=v0)
    s.configure.append(lambda f: (f.setFriction(10.0), f.setTemperature(0.0)))
    c = build(s)
    c.step(1)
    assert np.allclose(c.velocities[0], v0[0]*(1 - 10.0*0.01), atol=1e-13)


@pytest.mark.parametrize("stencil", [LBMForce.NearestNode, LBMForce.Keys])
def test_total_momentum_conserved_with_particles(stencil):
    """TestLBMCoupling.h:146-172: sum m v + j m_c dx/dt is conserved with drag and random forces, two particles
    sharing a node and one crossing the periodic boundary."""
    dx, dt = 0.5, 0.01
    s = Setup((8, 8, 8), dx=dx, dt=dt, tau=0.8)
    pos = [[1.02, 2.01, 0.98], [0.97, 1.99, 1.03], [3.98, 0.02, 3.96], [2.3, 1.1, 0.6]]
    vel = [[0.5, -0.2, 0.3], [-0.4, 0.1, 0.2], [2.0, -1.5, 1.0], [0.0, 0.3, -0.6]]
    s.add_particles(pos, mass=100.0, velocities=vel)
    s.configure.append(lambda f: (f.setFriction(10.0), f.setTemperature(300.0), f.setRandomNumberSeed(3),
                                  f.setInterpolationStencil(stencil)))
    c = build(s)
    cell = 602.214*dx**3

    def momentum():
        m, j = total_moments(c.getFluidState())
        return (100.0*c.velocities).sum(axis=0) + j*cell*dx/dt
    scale = sum(100.0*np.linalg.norm(v) for v in vel)
    p0 = momentum()
    c.step(50)
    p1 = momentum()
    assert np.linalg.norm(p1 - p0)/scale < 1e-11
    assert c.positions[2, 0] > 4.0


def test_comoving_particle_feels_no_force():
    """TestLBMCoupling.h:178-193: a particle moving with a uniform flow keeps its velocity."""
    for u in ([1.0, 0, 0], [0, 1.0, 0], [0.6, -0.5, 0.6]):
        s = Setup((8, 8, 8), dx=0.5, dt=0.01, tau=0.8)
        s.add_particles([[1.1, 1.3, 1.7]], mass=100.0, velocities=[u])
        s.configure.append(lambda f, u=u: (f.setFriction(10.0), f.setTemperature(0.0), f.setInitialFluidVelocity(u)))
        c = build(s)
        c.step(100)
        assert np.allclose(c.velocities[0], u, atol=1e-13)
