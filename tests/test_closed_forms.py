"""The closed forms of the reference's own tests (tests/TestLBMFluid.h, TestLBMCoupling.h, TestLBMStencils.h), run as
the reference runs them on every platform: on the FP64 oracle, and on the gfx950 library with FP64 and with FP32
storage.  Tolerances follow the reference's: the double-precision value for the oracle and FP64 storage; for FP32
storage the storage resolution 2e-6 (getStorageTolerance, TestLBMForce.h:37-42), or the test's own single-precision
value (getFluidTolerance, TestLBMFluid.h:30-35), and at least 2e-6 for coupled particles (getCouplingTolerance,
TestLBMCoupling.h:27-31)."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from parity import Setup, equilibrium_state, total_moments

BACKENDS = [pytest.param("oracle", id="oracle"),
            pytest.param("fp64", id="gpu-fp64", marks=pytest.mark.gpu),
            pytest.param("fp32", id="gpu-fp32", marks=pytest.mark.gpu)]


def build(setup, backend):
    if backend == "oracle":
        return setup.build(oracle_lib.load_oracle())
    return setup.build(None, storage=LBMForce.FP64 if backend == "fp64" else LBMForce.FP32)


def tol(backend, double, single=2e-6):
    """getFluidTolerance: the double-precision tolerance, or the single-precision one with FP32 storage."""
    return single if backend == "fp32" else double


def ctol(backend, double):
    """getCouplingTolerance: at least 2e-6 with FP32 storage."""
    return max(double, 2e-6) if backend == "fp32" else double


@pytest.mark.parametrize("backend", BACKENDS)
def test_uniform_flow_is_steady(backend):
    """TestLBMFluid.h:93-108."""
    s = Setup((6, 5, 4), dx=0.5, dt=0.01, tau=0.8)
    s.configure.append(lambda f: f.setInitialFluidVelocity((0.03*50, -0.02*50, 0.01*50)))
    c = build(s, backend)
    before = c.getFluidState()
    c.step(50)
    assert np.abs(c.getFluidState() - before).max() < tol(backend, 1e-13)


@pytest.mark.parametrize("backend", BACKENDS)
def test_conservation(backend):
    """TestLBMFluid.h:113-135: mass and momentum of a perturbed fluid without forces."""
    s = Setup((6, 5, 4), dx=0.5, dt=0.01, tau=0.7)
    c = build(s, backend)
    state = c.getFluidState()
    i = np.arange(state.size)
    state = state + 1e-3*np.sin(1.3*i + 0.7*(i % 11))
    c.setFluidState(state)
    m0, p0 = total_moments(state)
    c.step(100)
    m1, p1 = total_moments(c.getFluidState())
    assert m1 == pytest.approx(m0, abs=tol(backend, 1e-13)*m0)
    assert np.linalg.norm(p0) > 1e-5
    assert np.abs(p1 - p0).max() < tol(backend, 1e-13)*m0


@pytest.mark.parametrize("rho0", [1.0, 1.02])
@pytest.mark.parametrize("backend", BACKENDS)
def test_body_force(backend, rho0):
    """TestLBMFluid.h:141-169: momentum n rho g per node, velocity (n + 1/2) g dt."""
    dx, dt = 0.5, 0.01
    s = Setup((4, 3, 5), dx=dx, dt=dt, tau=0.9)
    g = np.array([0.5, -0.3, 0.2])
    s.configure.append(lambda f: f.setBodyAcceleration(g))
    s.state = equilibrium_state(60, rho0, np.zeros(3))
    c = build(s, backend)
    c.step(20)
    m, p = total_moments(c.getFluidState())
    gl = g*dt*dt/dx
    assert m == pytest.approx(rho0*60, abs=tol(backend, 1e-12)*60)
    assert np.allclose(p, gl*20*rho0*60, atol=tol(backend, 1e-13), rtol=0)
    rho, u = c.getFluidFields()
    assert np.allclose(u, g*20.5*dt, atol=tol(backend, 1e-12), rtol=0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_momentum_removal_timing(backend):
    """TestLBMFluid.h:176-199: with frequency 3 the momentum after n steps is ((n-1)%3 + 1) F per node."""
    dx, dt = 0.5, 0.01
    s = Setup((4, 4, 4), dx=dx, dt=dt, tau=1.0)
    s.configure.append(lambda f: (f.setBodyAcceleration((0.5, 0, 0)), f.setFluidMomentumRemovalFrequency(3),
                                  f.setInitialFluidVelocity((0.5, 1.0, -0.5))))
    c = build(s, backend)
    gl = 0.5*dt*dt/dx
    for n in range(1, 8):
        c.step(1)
        m, p = total_moments(c.getFluidState())
        assert np.allclose(p, [((n - 1) % 3 + 1)*gl*64, 0, 0], atol=tol(backend, 1e-13), rtol=0)


@pytest.mark.parametrize("tau", [0.6, 1.0])
@pytest.mark.parametrize("backend", BACKENDS)
def test_shear_wave_viscosity(backend, tau):
    """TestLBMFluid.h:205-251: u_x = U sin(k y) decays as exp(-nu k^2 t), nu = (tau - 1/2)/3, within 2e-3."""
    nx, ny, nz = 2, 64, 2
    s = Setup((nx, ny, nz), dx=0.5, dt=0.01, tau=tau)
    n = nx*ny*nz
    j = (np.arange(n)//nx) % ny
    k = 2*np.pi/ny
    u = np.zeros((n, 3))
    u[:, 0] = 1e-3*np.sin(k*j)
    s.state = equilibrium_state(n, 1.0, u)
    c = build(s, backend)

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
@pytest.mark.parametrize("backend", BACKENDS)
def test_bounce_back_poiseuille(backend, tau):
    """TestLBMFluid.h:486-535: exact steady profile of halfway bounce-back with the regularized collision,
    u(y) = g/(2 nu)(y - 1/2)(n - 1/2 - y) + g(16 Lambda - 3)/(24 nu), and the walls carry the body force."""
    nx, ny, nz = 2, 12, 2
    dx, dt = 0.5, 0.01
    g = 1e-5
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    s.configure.append(lambda f: (f.setBodyAcceleration((g*dx/dt**2, 0, 0)), f.setSolidNodes(wall_plane(nx, ny, nz))))
    c = build(s, backend)
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
        assert u[node, 0]*dt/dx/umax == pytest.approx(exact/umax, abs=tol(backend, 1e-9, 5e-5))
    mass = rho.sum()*dx**3
    # Relative to the size of the force, as ASSERT_EQUAL_VEC compares vectors.
    assert np.allclose(c.getWallForce(), [g*dx/dt**2*mass, 0, 0], rtol=0, atol=tol(backend, 1e-8, 5e-5)*g*dx/dt**2*mass)


@pytest.mark.parametrize("tau", [0.8, 1.0])
@pytest.mark.parametrize("backend", BACKENDS)
def test_regularized_poiseuille(backend, tau):
    """TestLBMFluid.h:590-627: u(y) = g/(2 nu) y (n - y) + 3 g (tau - 1)/(tau - 1/2), density of the fluid at rest."""
    nx, ny, nz = 2, 12, 2
    dx, dt = 0.5, 0.01
    g = 1e-5
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    s.configure.append(lambda f: (f.setBodyAcceleration((g*dx/dt**2, 0, 0)), f.setSolidNodes(wall_plane(nx, ny, nz)),
                                  f.setWallScheme(LBMForce.Regularized)))
    c = build(s, backend)
    nu, h = (tau - 0.5)/3, ny
    c.step(int(4*h*h/nu))
    rho, u = c.getFluidFields()
    umax = g/(2*nu)*(0.5*h)**2
    for node in range(nx*ny*nz):
        jj = (node//nx) % ny
        if jj == 0:
            continue
        assert rho[node]/602.214 == pytest.approx(1.0, abs=tol(backend, 1e-9, 5e-5))
        exact = g/(2*nu)*jj*(ny - jj) + 3*g*(tau - 1)/(tau - 0.5)
        assert u[node, 0]*dt/dx/umax == pytest.approx(exact/umax, abs=tol(backend, 1e-9, 5e-5))
    mass = rho.sum()*dx**3
    # Relative to the size of the force, as ASSERT_EQUAL_VEC compares vectors.
    assert np.allclose(c.getWallForce(), [g*dx/dt**2*mass, 0, 0], rtol=0, atol=tol(backend, 1e-8, 5e-5)*g*dx/dt**2*mass)


def wall_test_nodes(nx, ny, nz):
    return [i + nx*(j + ny*k) for k in range(nz) for j in range(ny) for i in range(nx)
            if i == 0 or (2 <= i < 4 and 1 <= j < 3 and 1 <= k < 3)]


@pytest.mark.parametrize("scheme", [LBMForce.BounceBack, LBMForce.Regularized])
@pytest.mark.parametrize("backend", BACKENDS)
def test_wall_momentum_balance(backend, scheme):
    """TestLBMFluid.h:648-685: mass conserved; P(t+dt) - P(t) = (M g - F_wall) dt at every step."""
    nx, ny, nz = 6, 5, 4
    dx, dt = 0.5, 0.01
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=0.7)
    g = np.array([1e-5, 2e-5, -1e-5])*dx/dt**2
    s.configure.append(lambda f: (f.setSolidNodes(wall_test_nodes(nx, ny, nz)), f.setWallScheme(scheme),
                                  f.setInitialFluidVelocity((1.0, -0.5, 0.25)), f.setBodyAcceleration(g)))
    c = build(s, backend)

    def moments():
        rho, u = c.getFluidFields()
        return (rho*dx**3).sum(), ((rho*dx**3)[:, None]*u).sum(axis=0)
    m0, p0 = moments()
    scale = m0*0.02*dx/dt
    for _ in range(30):
        c.step(1)
        m, p = moments()
        assert m == pytest.approx(m0, rel=tol(backend, 1e-13))
        balance = p - p0 - (g*m - c.getWallForce())*dt
        assert np.linalg.norm(balance)/scale < tol(backend, 1e-12)
        p0 = p


@pytest.mark.parametrize("boundary", [LBMForce.Velocity, LBMForce.DensityVelocity])
@pytest.mark.parametrize("backend", BACKENDS)
def test_couette(backend, boundary):
    """TestLBMFluid.h:694-716: linear profile u_x(z) = U (z + 1)/(nz + 1), uniform density."""
    nx, ny, nz = 2, 2, 10
    dx, dt, tau, U = 0.5, 0.01, 0.9, 0.01
    s = Setup((nx, ny, nz), dx=dx, dt=dt, tau=tau)
    s.configure.append(lambda f: (f.setFaceBoundary(LBMForce.ZMin, boundary), f.setFaceBoundary(LBMForce.ZMax, boundary),
                                  f.setFaceVelocity(LBMForce.ZMax, (U*dx/dt, 0, 0))))
    c = build(s, backend)
    nu, h = (tau - 0.5)/3, nz + 1
    c.step(int(6*h*h/nu))
    rho, u = c.getFluidFields()
    k = np.arange(nx*ny*nz)//(nx*ny)
    assert np.allclose(u[:, 0]*dt/(dx*U), (k + 1)/h, atol=tol(backend, 1e-9, 5e-5), rtol=0)
    assert np.allclose(rho/602.214, 1.0, atol=tol(backend, 1e-9, 5e-5), rtol=0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_first_step_drag(backend):
    """TestLBMCoupling.h:84-103: at T = 0 a particle in fluid at rest has v1 = v0 (1 - gamma dt) after one step."""
    s = Setup((8, 8, 8), dx=0.5, dt=0.01, tau=0.8)
    v0 = np.array([[0.5, -0.3, 0.2]])
    s.add_particles([[1.0, 1.2, 2.1]], mass=100.0, velocities=v0)
    s.configure.append(lambda f: (f.setFriction(10.0), f.setTemperature(0.0)))
    c = build(s, backend)
    c.step(1)
    assert np.allclose(c.velocities[0], v0[0]*(1 - 10.0*0.01), atol=ctol(backend, 1e-13), rtol=0)


DRAGS = [pytest.param(LBMForce.Explicit, id="explicit"), pytest.param(LBMForce.Centered, id="centred")]


@pytest.mark.parametrize("rho0", [1.0, 0.98, 1.02])
@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("stencil", [LBMForce.NearestNode, LBMForce.Keys])
@pytest.mark.parametrize("backend", BACKENDS)
def test_total_momentum_conserved_with_particles(backend, stencil, drag, rho0):
    """TestLBMCoupling.h:146-172: sum m v + j m_c dx/dt is conserved with drag and random forces, two particles
    sharing a node and one crossing the periodic boundary, in a fluid at rest of lattice density rho0."""
    dx, dt = 0.5, 0.01
    s = Setup((8, 8, 8), dx=dx, dt=dt, tau=0.8)
    pos = [[1.02, 2.01, 0.98], [0.97, 1.99, 1.03], [3.98, 0.02, 3.96], [2.3, 1.1, 0.6]]
    vel = [[0.5, -0.2, 0.3], [-0.4, 0.1, 0.2], [2.0, -1.5, 1.0], [0.0, 0.3, -0.6]]
    s.add_particles(pos, mass=100.0, velocities=vel)
    s.configure.append(lambda f: (f.setFriction(10.0), f.setTemperature(300.0), f.setRandomNumberSeed(3),
                                  f.setInterpolationStencil(stencil), f.setDragScheme(drag)))
    s.state = equilibrium_state(512, rho0, np.zeros(3))
    c = build(s, backend)
    cell = 602.214*dx**3

    def momentum():
        m, j = total_moments(c.getFluidState())
        return (100.0*c.velocities).sum(axis=0) + j*cell*dx/dt
    scale = sum(100.0*np.linalg.norm(v) for v in vel)
    p0 = momentum()
    c.step(50)
    p1 = momentum()
    assert np.linalg.norm(p1 - p0)/scale < ctol(backend, 1e-11)
    assert c.positions[2, 0] > 4.0
    assert np.any(c.velocities[3] != vel[3])


@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_comoving_particle_feels_no_force(backend, drag):
    """TestLBMCoupling.h:178-193: a particle moving with a uniform flow keeps its velocity."""
    for u in ([1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [0.6, -0.5, 0.6]):
        s = Setup((8, 8, 8), dx=0.5, dt=0.01, tau=0.8)
        s.add_particles([[1.1, 1.3, 1.7]], mass=100.0, velocities=[u])
        s.configure.append(lambda f, u=u: (f.setFriction(10.0), f.setTemperature(0.0), f.setInitialFluidVelocity(u),
                                           f.setDragScheme(drag)))
        c = build(s, backend)
        c.step(100)
        assert np.allclose(c.velocities[0], u, atol=ctol(backend, 1e-13), rtol=0)


def coupling_system(positions, velocities, friction, temperature, coupled=None, configure=()):
    """createCoupledSystem (TestLBMCoupling.h:40-61): 8^3 nodes, tau = 0.8, seed 1, particles of 100 Da."""
    s = Setup((N8, N8, N8), dx=DX, dt=DT, tau=0.8, density=602.2)
    for i, (x, v) in enumerate(zip(positions, velocities)):
        s.add_particles([x], mass=100.0, velocities=[v], coupled=coupled is None or i in coupled)

    def conf(f):
        f.setFriction(friction)
        f.setTemperature(temperature)
        f.setRandomNumberSeed(1)
        for fn in configure:
            fn(f)
    s.configure.append(conf)
    return s


def wall_plane_y0():
    """wallPlane(8, 8, 8) (TestLBMFluid.h:452-458): the plane j = 0."""
    return [i + N8*N8*k for k in range(N8) for i in range(N8)]


@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_partial_coupling(backend, drag):
    """TestLBMCoupling.h:198-220: only the coupled particles feel the fluid; the others keep their velocity, no force."""
    positions = [[0.3, 0.4, 0.5], [1.3, 1.4, 1.5], [2.3, 2.4, 2.5], [3.3, 3.4, 3.5]]
    v0 = np.array([0.2, 0.1, -0.3])
    c = build(coupling_system(positions, [v0]*4, 5.0, 0.0, coupled={1, 3},
                              configure=[lambda f: f.setDragScheme(drag)]), backend)
    c.step(10)
    forces = c.getForces()
    for i in range(4):
        v = c.velocities[i]
        if i % 2 == 0:
            assert np.allclose(v, v0, atol=ctol(backend, 1e-12), rtol=0)
            assert np.all(forces[i] == 0.0)
        else:
            assert v @ v < 0.9*(v0 @ v0)


@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_force_evaluations_and_seeds(backend, drag):
    """TestLBMCoupling.h:226-254: force evaluations between the steps change nothing; the same seed gives the same
    run, seed 0 a different one each time."""
    def run(seed, queries):
        s = coupling_system([[0.3, 0.4, 0.5], [1.3, 1.4, 1.5], [2.3, 2.4, 2.5]], [[0, 0, 0]]*3, 5.0, 300.0,
                            configure=[lambda f: (f.setDragScheme(drag), f.setRandomNumberSeed(seed))])
        c = build(s, backend)
        for _ in range(20):
            if queries:
                c.getForces()
            c.step(1)
            if queries:
                c.getForces()
        return c.velocities.copy()
    assert np.array_equal(run(1234, False), run(1234, True))
    assert run(0, False)[0, 0] != run(0, False)[0, 0]


@pytest.mark.parametrize("backend", BACKENDS)
def test_nve_scheme(backend):
    """TestLBMCoupling.h:336-357: the NVE scheme has no random force at any temperature."""
    friction = 5.0

    def run(seed, steps):
        s = coupling_system([[1.1, 1.3, 1.7], [2.6, 0.4, 3.1]], [[0.3, -0.2, 0.1], [-0.1, 0.4, 0.2]], friction, 300.0,
                            configure=[lambda f: (f.setCouplingScheme(LBMForce.NVE), f.setRandomNumberSeed(seed))])
        c = build(s, backend)
        c.step(steps)
        return c.velocities.copy()
    v1 = run(1, 1)
    assert np.allclose(v1[0], np.array([0.3, -0.2, 0.1])*(1 - friction*DT), atol=ctol(backend, 1e-14), rtol=0)
    assert np.array_equal(run(1, 30), run(2, 30))


@pytest.mark.parametrize("backend", BACKENDS)
def test_wall_reflection(backend):
    """TestLBMCoupling.h:363-385: a coupled particle reaching a solid node while moving into the wall is reversed and
    slowed by the drag of the wall at rest; an uncoupled one does not see the wall."""
    friction = 1.0
    v0 = np.array([0.5, -10.0, 0.2])
    c = build(coupling_system([[1.1, 0.30, 1.7]]*2, [v0]*2, friction, 0.0, coupled={0},
                              configure=[lambda f: f.setSolidNodes(wall_plane_y0())]), backend)
    decay = 1.0 - friction*DT
    c.step(1)
    assert np.allclose(c.velocities[0], v0*decay, atol=ctol(backend, 1e-14), rtol=0)
    assert c.positions[0, 1] < 0.25
    c.step(1)
    assert np.allclose(c.velocities[0], -v0*decay*decay, atol=ctol(backend, 1e-14), rtol=0)
    assert c.positions[0, 1] > 0.25
    assert np.allclose(c.velocities[1], v0, atol=ctol(backend, 1e-12), rtol=0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_wall_reflection_direction(backend):
    """TestLBMCoupling.h:391-409: reversed only when moving into the wall (the plane j = 0 has fluid on both sides)."""
    friction = 1.0
    v0 = np.array([[0.3, 5.0, 0.1], [0.3, 5.0, 0.1], [0.3, -5.0, 0.1]])
    c = build(coupling_system([[1.1, 0.20, 1.7], [2.1, -0.10, 1.7], [3.1, -0.10, 1.7]], v0, friction, 0.0,
                              configure=[lambda f: f.setSolidNodes(wall_plane_y0())]), backend)
    c.step(1)
    decay = 1.0 - friction*DT
    assert np.allclose(c.velocities[0], v0[0]*decay, atol=ctol(backend, 1e-14), rtol=0)
    assert np.allclose(c.velocities[1], -v0[1]*decay, atol=ctol(backend, 1e-14), rtol=0)
    assert np.allclose(c.velocities[2], v0[2]*decay, atol=ctol(backend, 1e-14), rtol=0)


@pytest.mark.parametrize("scheme", [pytest.param(LBMForce.BounceBack, id="bounce-back"),
                                    pytest.param(LBMForce.Regularized, id="regularized")])
@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_wall_momentum_balance_with_particles(backend, drag, scheme):
    """TestLBMCoupling.h:416-476: particles, fluid and walls conserve momentum, the walls' share being the sum of
    getWallForce() dt; one particle is reflected by the wall j = 0."""
    walls = wall_plane_y0() + [node + N8*4 for node in wall_plane_y0()]
    positions = [[1.1, 0.30, 1.7], [2.3, 1.2, 0.6], [0.7, 3.1, 2.9], [3.3, 1.6, 3.6]]
    velocities = [[0.5, -10.0, 0.2], [0.4, 0.3, -0.2], [-0.6, 0.2, 0.5], [0.1, -0.4, 0.3]]
    c = build(coupling_system(positions, velocities, 10.0, 300.0, configure=[lambda f: (
        f.setDragScheme(drag), f.setWallScheme(scheme), f.setSolidNodes(walls),
        f.setInitialFluidVelocity((0.3, 0.4, -0.2)))]), backend)

    def momentum():
        p = (100.0*c.velocities).sum(axis=0)
        scale = (100.0*np.linalg.norm(c.velocities, axis=1)).sum()
        rho, u = c.getFluidFields()
        nodes = u*(rho*DX**3)[:, None]
        return p + nodes.sum(axis=0), scale + np.linalg.norm(nodes, axis=1).sum()
    p0, _ = momentum()
    assert np.all(np.asarray(c.getWallForce()) == 0.0)
    wall = np.zeros(3)
    for n in range(100):
        c.step(1)
        wall += np.asarray(c.getWallForce())*DT
        if n == 1:
            assert c.velocities[0, 1] > 0
    p1, scale = momentum()
    assert wall @ wall > 1.0
    tolerance = ctol(backend, 1e-11 if scheme == LBMForce.Regularized else 1e-12)
    if scheme == LBMForce.Regularized and backend == "fp32":
        tolerance = 1e-5
    assert np.linalg.norm(p1 + wall - p0)/scale < tolerance


# ------------------------------------------------------------------------------------------------
# Stencils (TestLBMStencils.h), on the 8^3 coupled system of the reference's coupling tests (dx = 0.5 nm, dt = 0.01 ps,
# rho0 = 602.2 Da/nm^3, tau = 0.8, particles of 100 Da).
# ------------------------------------------------------------------------------------------------
STENCILS = [LBMForce.Trilinear, LBMForce.ThreePoint, LBMForce.Keys]
DX, DT, N8 = 0.5, 0.01, 8
U0, X0 = np.array([0.01, -0.02, 0.015]), np.array([2.0, 2.0, 2.0])
GRAD = np.array([[0.002, -0.001, 0.0005], [0.0015, 0.001, -0.002], [-0.001, 0.0005, 0.002]])


def coupled_setup(positions, velocities, friction, stencil, state=None, faces=()):
    s = Setup((N8, N8, N8), dx=DX, dt=DT, tau=0.8, density=602.2)
    s.add_particles(positions, mass=100.0, velocities=velocities)
    s.state = state

    def conf(f):
        f.setFriction(friction)
        f.setTemperature(0.0)
        f.setRandomNumberSeed(1)
        f.setInterpolationStencil(stencil)
        for face, kind in faces:
            f.setFaceBoundary(face, kind)
    s.configure.append(conf)
    return s


def linear_velocity(x):
    return U0 + (np.asarray(x) - X0) @ GRAD


def linear_state():
    """Equilibrium with a density and a velocity linear in space (TestLBMStencils.h:131-148)."""
    node = np.arange(N8**3)
    i, j, k = node % N8, (node//N8) % N8, node//(N8*N8)
    x = np.stack([i, j, k], axis=1)*DX
    rho = 1.0 + 0.01*(i - 4) - 0.005*(j - 4) + 0.008*(k - 4)
    return equilibrium_state(N8**3, rho, linear_velocity(x)*(DT/DX))


def varied_state(rho_varies):
    node = np.arange(N8**3)
    u = np.stack([0.01*np.sin(node), 0.01*np.cos(node), 0.005*np.sin(3.0*node)], axis=1)
    rho = 1.0 + 0.001*(node % 7) if rho_varies else np.ones(N8**3)
    return equilibrium_state(N8**3, rho, u)


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_stencil_linear_field(backend, stencil):
    """TestLBMStencils.h:130-164: in the first step a particle at rest feels gamma m u(X), u the linear field."""
    friction, position = 5.0, np.array([2.13, 1.87, 2.21])
    c = build(coupled_setup([position], [[0, 0, 0]], friction, stencil, linear_state()), backend)
    c.step(1)
    assert np.allclose(c.velocities[0], linear_velocity(position)*(friction*DT), atol=ctol(backend, 1e-12), rtol=0)


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_stencil_open_faces(backend, stencil):
    """TestLBMStencils.h:172-223: along z the stencil stops at the faces; the particle feels the linear field at the
    mean node sum_c w_c z_c of its stencil (its own z away from the faces; node n - 1 for trilinear past it)."""
    lib = oracle_lib.load_oracle()
    width = {LBMForce.Trilinear: 2, LBMForce.ThreePoint: 3, LBMForce.Keys: 4}[stencil]
    friction = 5.0
    for z in (0.1, 0.35, 0.6, 1.3, 2.2, 3.3, 3.62, 3.9):
        position = np.array([2.13, 1.87, z])
        faces = ((LBMForce.ZMin, LBMForce.Velocity), (LBMForce.ZMax, LBMForce.Density))
        c = build(coupled_setup([position], [[0, 0, 0]], friction, stencil, linear_state(), faces), backend)
        c.step(1)
        idx = (ctypes.c_int*4)()
        w = np.zeros(4)
        lib.oracle_axis_weights(stencil, z/DX, N8, 1, idx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        mean = sum(w[a]*idx[a] for a in range(width))
        if width/2 < z/DX < N8 - 1 - width/2:
            assert mean == pytest.approx(z/DX, abs=1e-13)
        if stencil == LBMForce.Trilinear and z/DX > N8 - 1:
            assert mean == pytest.approx(N8 - 1.0, abs=1e-15)
        seen = linear_velocity([position[0], position[1], mean*DX])
        assert np.allclose(c.velocities[0], seen*(friction*DT), atol=ctol(backend, 1e-12), rtol=0), z


@pytest.mark.parametrize("backend", BACKENDS)
def test_stencil_at_node(backend):
    """TestLBMStencils.h:229-259: on a node the trilinear and Keys stencils weigh only that node, so the first step is
    the nearest node's, bit for bit, for the particle and for the fluid."""
    results = []
    for stencil in (LBMForce.NearestNode, LBMForce.Trilinear, LBMForce.Keys):
        c = build(coupled_setup([[1.0, 1.5, 2.0]], [[0.3, -0.2, 0.1]], 5.0, stencil, varied_state(True)), backend)
        c.step(1)
        results.append((c.velocities[0].copy(), c.getFluidState()))
    for v, fluid in results[1:]:
        assert np.array_equal(v, results[0][0])
        assert np.array_equal(fluid, results[0][1])


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_stencil_translation(backend, stencil):
    """TestLBMStencils.h:496-541: moving the particles and the fluid by one node along each axis moves the whole run."""
    positions = np.array([[1.13, 2.07, 0.91], [1.31, 1.88, 1.17], [2.6, 3.2, 0.4]])
    velocities = [[0.5, -0.2, 0.3], [-0.4, 0.1, 0.2], [0.2, -0.5, 0.1]]

    def shifted(state):
        s = state.reshape(19, N8, N8, N8)          # [q][k][j][i]
        return np.roll(s, (1, 1, 1), axis=(1, 2, 3)).reshape(-1)
    state = varied_state(False)
    runs = []
    for shift in (0, 1):
        c = build(coupled_setup(positions + DX*shift, velocities, 10.0, stencil, shifted(state) if shift else state),
                  backend)
        c.step(30)
        runs.append((c.getFluidState(), c.positions.copy()))
    assert np.allclose(shifted(runs[0][0]), runs[1][0], atol=ctol(backend, 1e-12), rtol=0)
    assert np.allclose(runs[0][1] + DX, runs[1][1], atol=ctol(backend, 1e-12), rtol=0)


# ------------------------------------------------------------------------------------------------
# Walls and open faces (TestLBMFluid.h:541-884), on the reference's fluid systems (dx = 0.5 nm, dt = 0.01 ps,
# rho0 = 602.2 Da/nm^3).
# ------------------------------------------------------------------------------------------------
RHO_F = 602.2


def fluid_setup(grid, tau):
    return Setup(grid, dx=DX, dt=DT, tau=tau, density=RHO_F)


@pytest.mark.parametrize("backend", BACKENDS)
def test_wall_conservation(backend):
    """TestLBMFluid.h:541-580: a solid block conserves the fluid mass; with removal the momentum is removed."""
    nx, ny, nz = 6, 5, 4
    block = [i + nx*(j + ny*k) for k in (1, 2) for j in (1, 2) for i in (2, 3)]
    for removal in (0, 1):
        s = fluid_setup((nx, ny, nz), 0.7)
        s.configure.append(lambda f, removal=removal: (
            f.setSolidNodes(block), f.setInitialFluidVelocity((0.02*DX/DT, -0.01*DX/DT, 0.0)),
            f.setFluidMomentumRemovalFrequency(removal)))
        c = build(s, backend)

        def mass():
            rho, u = c.getFluidFields()
            return (rho/RHO_F).sum(), rho, u
        m0, _, _ = mass()
        assert m0 == pytest.approx(nx*ny*nz - len(block), abs=tol(backend, 1e-13))
        c.step(100 if removal == 0 else 1)
        m1, rho, u = mass()
        assert m1 == pytest.approx(m0, abs=tol(backend, 1e-13))
        if removal:
            p = (u*(rho/RHO_F*DT/DX)[:, None]).sum(axis=0)
            assert np.allclose(p, 0, atol=tol(backend, 1e-13))


@pytest.mark.parametrize("inlet,outlet", [(LBMForce.Velocity, LBMForce.Density),
                                          (LBMForce.DensityVelocity, LBMForce.Density),
                                          (LBMForce.DensityVelocity, LBMForce.DensityVelocity)])
@pytest.mark.parametrize("backend", BACKENDS)
def test_uniform_flow_through_faces(backend, inlet, outlet):
    """TestLBMFluid.h:724-762: a uniform flow through an inlet and an outlet face stays uniform."""
    nx, ny, nz = 3, 8, 3
    u = np.array([0.0, 0.02, 0.0])*(DX/DT)
    s = fluid_setup((nx, ny, nz), 0.8)
    s.configure.append(lambda f: (f.setFaceBoundary(LBMForce.YMin, inlet), f.setFaceBoundary(LBMForce.YMax, outlet),
                                  f.setFaceVelocity(LBMForce.YMin, u), f.setFaceVelocity(LBMForce.YMax, u),
                                  f.setInitialFluidVelocity(u)))
    c = build(s, backend)
    c.step(50)
    rho, vel = c.getFluidFields()
    assert np.allclose(vel*(DT/DX), u*(DT/DX), atol=tol(backend, 1e-13), rtol=0)
    assert np.allclose(rho/RHO_F, 1.0, atol=tol(backend, 1e-13), rtol=0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_density_velocity_reservoir(backend):
    """TestLBMFluid.h:820-851: fluid at rest between two equal DensityVelocity faces reaches their uniform flow."""
    nx, ny, nz = 3, 12, 3
    rho, u = 1.01, np.array([0.01, 0.02, 0.0])
    s = fluid_setup((nx, ny, nz), 1.0)

    def conf(f):
        for face in (LBMForce.YMin, LBMForce.YMax):
            f.setFaceBoundary(face, LBMForce.DensityVelocity)
            f.setFaceDensity(face, rho*RHO_F)
            f.setFaceVelocity(face, u*(DX/DT))
    s.configure.append(conf)
    c = build(s, backend)
    c.step(2500)
    density, vel = c.getFluidFields()
    assert np.allclose(vel*(DT/DX), u, atol=tol(backend, 1e-9, 1e-5), rtol=0)
    assert np.allclose(density/RHO_F, rho, atol=tol(backend, 1e-9, 1e-5), rtol=0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_density_velocity_inlet(backend):
    """TestLBMFluid.h:853-884: a DensityVelocity inlet at 1.01 and a Density outlet at 1 give the uniform flow whose
    entering populations rho (1 + 3u + 3u^2)/6 are the inlet's, at the outlet's density."""
    nx, ny, nz = 3, 16, 3
    rho_in, ub = 1.01, 0.01
    s = fluid_setup((nx, ny, nz), 0.8)
    s.configure.append(lambda f: (
        f.setFaceBoundary(LBMForce.YMin, LBMForce.DensityVelocity), f.setFaceBoundary(LBMForce.YMax, LBMForce.Density),
        f.setFaceDensity(LBMForce.YMin, rho_in*RHO_F), f.setFaceVelocity(LBMForce.YMin, (0.0, ub*DX/DT, 0.0))))
    c = build(s, backend)
    c.step(4000)
    density, vel = c.getFluidFields()
    a = rho_in*(1.0 + 3.0*ub + 3.0*ub*ub) - 1.0
    u = (-3.0 + math.sqrt(9.0 + 12.0*a))/6.0
    assert np.allclose(vel*(DT/DX), [0.0, u, 0.0], atol=tol(backend, 1e-10, 1e-5), rtol=0)
    assert np.allclose(density/RHO_F, 1.0, atol=tol(backend, 1e-10, 1e-5), rtol=0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_pressure_driven_duct(backend):
    """TestLBMFluid.h:764-818: a square duct between two Density faces at 1.01 and 1: the density at the middle on the
    line between the faces' (held one node beyond them), and the series solution of the rectangular duct with the
    gradient measured there, within 2% of the peak; the flow is steady."""
    n, ny, tau = 8, 16, 1.0
    walls = [i + n*(j + ny*k) for k in range(n) for j in range(ny) for i in range(n) if i == 0 or k == 0]
    s = fluid_setup((n, ny, n), tau)
    s.configure.append(lambda f: (
        f.setSolidNodes(walls), f.setFaceBoundary(LBMForce.YMin, LBMForce.Density),
        f.setFaceBoundary(LBMForce.YMax, LBMForce.Density), f.setFaceDensity(LBMForce.YMin, 1.01*RHO_F),
        f.setFaceDensity(LBMForce.YMax, RHO_F)))
    c = build(s, backend)
    nu = (tau - 0.5)/3.0
    c.step(int(3*ny*ny/nu))
    density, velocity = [], []
    for _ in range(2):
        rho, u = c.getFluidFields()
        density.append(rho)
        velocity.append(u)
        c.step(1)
    middle = ny//2
    H = n - 1

    def at(i, j, k):
        return i + n*(j + ny*k)
    rho = density[0][at(n//2, middle, n//2)]/RHO_F
    assert rho == pytest.approx(1.0 + 0.01*(ny - middle)/(ny + 1), abs=2e-4)
    gradient = (density[0][at(n//2, middle - 2, n//2)] - density[0][at(n//2, middle + 2, n//2)])/(4*RHO_F)
    assert 0.01/(ny + 1) < gradient < 0.01/(ny - 3)
    K = (gradient/3.0)/(rho*nu)
    umax = error = change = 0.0
    for k in range(1, n):
        for i in range(1, n):
            node = at(i, middle, k)
            x, z = i - 0.5*n, k - 0.5*n
            exact = 0.0
            for m in range(1, 200, 2):
                a = m*math.pi/H
                exact += 4*K*H*H/(m*math.pi)**3*(1 if m % 4 == 1 else -1)*(1 - math.cosh(a*z)/math.cosh(a*H/2))*math.cos(a*x)
            u = velocity[0][node][1]*DT/DX
            umax = max(umax, exact)
            error = max(error, abs(u - exact))
            change = max(change, abs(velocity[1][node][1] - velocity[0][node][1])*DT/DX)
    assert error < 0.02*umax
    assert change < tol(backend, 1e-9, 1e-4)*umax


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_stencil_momentum_conservation(backend, drag, stencil):
    """TestLBMStencils.h:265-335: with every stencil the momentum of particles and fluid is conserved (overlapping
    stencils, random force, a particle crossing the periodic boundary); with a wall the momentum given to the solid
    nodes of the stencils closes the balance."""
    positions = [[1.02, 2.01, 0.98], [1.27, 1.79, 1.13], [3.98, 0.02, 3.96], [2.3, 1.1, 0.6]]
    velocities = [[0.5, -0.2, 0.3], [-0.4, 0.1, 0.2], [2.0, -1.5, 1.0], [0.0, 0.3, -0.6]]
    s = coupling_system(positions, velocities, 10.0, 300.0,
                        configure=[lambda f: (f.setInterpolationStencil(stencil), f.setDragScheme(drag))])
    s.state = equilibrium_state(512, 1.01, np.zeros(3))
    c = build(s, backend)
    cell = 602.2*DX**3

    def momentum():
        m, j = total_moments(c.getFluidState())
        return (100.0*c.velocities).sum(axis=0) + j*cell*DX/DT
    scale = sum(100.0*np.linalg.norm(v) for v in velocities)
    p0 = momentum()
    c.step(50)
    assert np.linalg.norm(momentum() - p0)/scale < ctol(backend, 1e-11)
    assert c.positions[2, 0] > 4.0

    # Near the wall j = 0, whose solid nodes the stencils cover; the fluid's momentum from its fields.
    c = build(coupling_system([[1.1, 0.40, 1.7], [2.3, 0.61, 0.6], [0.7, 3.55, 2.9]],
                              [[0.5, -1.0, 0.2], [0.4, 0.3, -0.2], [-0.6, 0.2, 0.5]], 10.0, 300.0,
                              configure=[lambda f: (f.setInterpolationStencil(stencil), f.setDragScheme(drag),
                                                    f.setSolidNodes(wall_plane_y0()))]), backend)

    def fields_momentum():
        p = (100.0*c.velocities).sum(axis=0)
        scale = (100.0*np.linalg.norm(c.velocities, axis=1)).sum()
        rho, u = c.getFluidFields()
        return p + (u*(rho*DX**3)[:, None]).sum(axis=0), scale
    p0, _ = fields_momentum()
    wall = np.zeros(3)
    for _ in range(50):
        c.step(1)
        wall += np.asarray(c.getWallForce())*DT
    p1, scale = fields_momentum()
    assert wall @ wall > 0
    assert np.linalg.norm(p1 + wall - p0)/scale < (1e-5 if backend == "fp32" else 1e-11)


# ------------------------------------------------------------------------------------------------
# Checkpoints (TestLBMCoupling.h:479-617).  The host's own checkpoint (positions, velocities, step count) is copied
# from one Context to the other, as OpenMM's Context checkpoint would.
# ------------------------------------------------------------------------------------------------
def host_state(c):
    return c.positions.copy(), c.velocities.copy(), c.getStepCount()


def set_host_state(c, state):
    positions, velocities, step = state
    c.setPositions(positions)
    c.setVelocities(velocities)
    c.setStepCount(step)


@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_checkpoint_with_random_force(backend, drag):
    """TestLBMCoupling.h:479-527: a run restarted from checkpoints (with the random force, a wall and the removal, a
    force evaluation just before them) is bit-identical to the uninterrupted run; the wall force is restored."""
    num_steps, split = 23, 9

    def system():
        s = coupling_system([[0.3, 0.8, 0.5], [1.3, 1.4, 1.5], [2.3, 2.4, 2.5], [3.1, 0.9, 0.2]],
                            [[0.1, -0.2, 0.3], [-0.3, 0.2, 0.1], [0.2, 0.2, -0.2], [0.0, -0.4, 0.1]], 10.0, 300.0,
                            configure=[lambda f: (f.setDragScheme(drag), f.setRandomNumberSeed(3),
                                                  f.setFluidMomentumRemovalFrequency(4),
                                                  f.setSolidNodes(wall_plane_y0()))])
        return build(s, backend)

    def finish(c):
        for step in range(split, num_steps):
            c.step(1)
            if step % 3 == 0:
                c.getForces()
        return c.positions.copy(), c.velocities.copy(), c.getFluidState()
    first = system()
    first.step(split)
    first.getForces()
    data, host, wall = first.createCheckpoint(), host_state(first), np.asarray(first.getWallForce()).copy()
    uninterrupted = finish(first)
    second = system()
    set_host_state(second, host)
    second.loadCheckpoint(data)
    assert wall @ wall > 0
    assert np.array_equal(np.asarray(second.getWallForce()), wall)
    for a, b in zip(uninterrupted, finish(second)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("backend", BACKENDS)
def test_checkpoint_mismatch(backend):
    """TestLBMCoupling.h:536-567: a checkpoint is refused with another number of coupled particles, and data that are
    not a checkpoint are refused."""
    from lbm_amd import LBMError
    c2 = build(coupling_system([[0.3, 0.4, 0.5], [1.3, 1.4, 1.5]], [[0, 0, 0]]*2, 5.0, 0.0), backend)
    data = c2.createCheckpoint()
    c3 = build(coupling_system([[0.3, 0.4, 0.5], [1.3, 1.4, 1.5], [2.3, 2.4, 2.5]], [[0, 0, 0]]*3, 5.0, 0.0), backend)
    with pytest.raises(LBMError):
        c3.loadCheckpoint(data)
    with pytest.raises(LBMError):
        c2.loadCheckpoint(b"this is not a checkpoint")


@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_restart_with_particles(backend, drag):
    """TestLBMCoupling.h:574-617: a run restarted with the fluid restored by setFluidState() is identical to the
    uninterrupted run (T = 0: the random state is not part of a fluid state)."""
    num_steps, split = 35, 13

    def system():
        s = coupling_system([[0.3, 0.4, 0.5], [1.3, 1.4, 1.5], [2.3, 2.4, 2.5]],
                            [[0.5, -0.2, 0.3], [-0.4, 0.1, 0.2], [1.0, -0.5, 0.0]], 5.0, 0.0,
                            configure=[lambda f: (f.setDragScheme(drag), f.setFluidMomentumRemovalFrequency(5))])
        return build(s, backend)
    a = system()
    a.step(num_steps)
    b = system()
    b.step(split)
    saved = b.getFluidState()
    c = system()
    set_host_state(c, host_state(b))
    c.setFluidState(saved)
    c.step(num_steps - split)
    assert np.array_equal(a.velocities, c.velocities)
    assert np.array_equal(a.getFluidState(), c.getFluidState())


# ------------------------------------------------------------------------------------------------
# Parameters, the Mach check and fluid restarts (TestLBMFluid.h:287-405, TestLBMFluctuations.h:427-469).
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("backend", BACKENDS)
def test_restart_from_fluid_state(backend):
    """TestLBMFluid.h:287-320: a run restarted with the fluid of setFluidState() is bit-identical (body force, initial
    velocity, removal every 5 steps)."""
    num_steps, split = 35, 13

    def system():
        s = fluid_setup((4, 4, 4), 0.8)
        s.configure.append(lambda f: (f.setBodyAcceleration((0.5, -0.2, 0.1)),
                                      f.setInitialFluidVelocity((0.01*DX/DT, 0.0, 0.0)),
                                      f.setFluidMomentumRemovalFrequency(5)))
        return build(s, backend)
    a = system()
    a.step(num_steps)
    b = system()
    b.step(split)
    saved = b.getFluidState()
    c = system()
    c.setStepCount(split)
    c.setFluidState(saved)
    c.step(num_steps - split)
    assert c.getStepCount() == num_steps
    assert np.array_equal(a.getFluidState(), c.getFluidState())


@pytest.mark.parametrize("backend", BACKENDS)
def test_update_parameters(backend):
    """TestLBMFluid.h:322-377: updateParametersInContext changes the body force, the removal and the Mach check, and
    refuses a change of the viscosity."""
    from lbm_amd import LBMError
    s = fluid_setup((4, 4, 4), 0.8)
    s.configure.append(lambda f: f.setBodyAcceleration((0.5, 0, 0)))
    c = build(s, backend)
    force = c.force
    c.step(5)
    assert total_moments(c.getFluidState())[1][0] > 1e-6
    force.setBodyAcceleration((0, 0, 0))
    force.setFluidMomentumRemovalFrequency(1)
    c.updateParametersInContext()
    c.step(1)
    assert np.allclose(total_moments(c.getFluidState())[1], 0, atol=tol(backend, 1e-13), rtol=0)
    force.setBodyAcceleration((0.5, 0, 0))
    force.setFluidMomentumRemovalFrequency(0)
    c.updateParametersInContext()
    c.step(5)
    force.setMachCheckFrequency(1)
    force.setMachNumberLimit(0.5*c.getFluidMachNumber())
    c.updateParametersInContext()
    with pytest.raises(LBMError, match="Mach number"):
        c.step(1)
    force.setKinematicViscosity(2.0*force.getKinematicViscosity())
    with pytest.raises(LBMError, match="cannot be changed"):
        c.updateParametersInContext()


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("backend", BACKENDS)
def test_mach_number_check(backend, variant):
    """TestLBMFluid.h:379-405: the Mach number is reported and checked every N steps: at Ma = 0.35 the step that checks
    throws with the default limit, not without the check or with the limit 0.5."""
    from lbm_amd import LBMError
    mach = 0.35
    speed = mach/math.sqrt(3.0)
    s = fluid_setup((4, 4, 4), 0.8)

    def conf(f):
        f.setInitialFluidVelocity((speed*DX/DT, 0, 0))
        f.setMachCheckFrequency(0 if variant == 1 else 10)
        if variant == 2:
            f.setMachNumberLimit(0.5)
    s.configure.append(conf)
    c = build(s, backend)
    assert c.getFluidMachNumber() == pytest.approx(mach, abs=tol(backend, 1e-12))
    c.step(9)
    if variant == 0:
        with pytest.raises(LBMError, match="Mach number"):
            c.step(1)
    else:
        c.step(1)


@pytest.mark.parametrize("backend", BACKENDS)
def test_fluctuation_parameters(backend):
    """TestLBMFluctuations.h:427-469: updateParametersInContext changes the temperature but cannot switch the
    fluctuations; a checkpoint is refused by a Context with the fluctuations switched differently."""
    from lbm_amd import LBMError
    positions = [[0.3, 0.8, 0.5], [1.3, 1.4, 1.5]]
    c = build(coupling_system(positions, [[0, 0, 0]]*2, 5.0, 300.0,
                              configure=[lambda f: f.setFluidFluctuations(True)]), backend)
    c.step(2)
    c.force.setTemperature(310.0)
    c.updateParametersInContext()
    c.step(1)
    c.force.setFluidFluctuations(False)
    with pytest.raises(LBMError):
        c.updateParametersInContext()
    c.force.setFluidFluctuations(True)
    data = c.createCheckpoint()
    other = build(coupling_system(positions, [[0, 0, 0]]*2, 5.0, 300.0), backend)
    with pytest.raises(LBMError):
        other.loadCheckpoint(data)


# ------------------------------------------------------------------------------------------------
# The force's own suite (TestLBMForce.h:126-218) and the remaining fluid, boundary and coupling cases.
# ------------------------------------------------------------------------------------------------
def stol(backend):
    """getStorageTolerance (TestLBMForce.h:37-42)."""
    return 2e-6 if backend == "fp32" else 1e-12


def force_system(configure=()):
    """createSystem and createPositions (TestLBMForce.h:47-65): 8^3 nodes in a 4 nm box, ten coupled particles of
    100 Da, the default density, viscosity and removal of LBMForce."""
    d = LBMForce()
    s = Setup((8, 8, 8), dx=0.5, dt=0.01, density=d.getFluidDensity())
    s.add_particles([[0.37*i, 0.53*i + 0.1, 0.71*i + 0.2] for i in range(10)], mass=100.0)

    def conf(f):
        f.setKinematicViscosity(d.getKinematicViscosity())
        f.setFluidMomentumRemovalFrequency(d.getFluidMomentumRemovalFrequency())
        for fn in configure:
            fn(f)
    s.configure.append(conf)
    return s


@pytest.mark.parametrize("backend", BACKENDS)
def test_zero_force(backend):
    """TestLBMForce.h:126-138: no coupling force before the first step."""
    c = build(force_system(), backend)
    assert np.all(c.getForces() == 0.0)
    c.step(5)


@pytest.mark.parametrize("backend", BACKENDS)
def test_initial_fluid_fields(backend):
    """TestLBMForce.h:143-163: the fluid starts at equilibrium with the density and velocity of the force."""
    u0 = np.array([0.1, -0.05, 0.02])
    s = force_system([lambda f: (f.setFluidDensity(500.0), f.setInitialFluidVelocity(u0))])
    s.density = 500.0
    c = build(s, backend)
    rho, u = c.getFluidFields()
    assert rho.shape == (512,) and u.shape == (512, 3)
    assert np.abs(rho/500.0 - 1).max() <= 5*stol(backend)
    assert np.abs(u - u0).max() <= 1e3*stol(backend)


@pytest.mark.parametrize("backend", BACKENDS)
def test_fluid_state_round_trip(backend):
    """TestLBMForce.h:168-218: the state (deviations from the rest equilibrium) reads zero at rest; moving 0.01 of
    the rest population of node 5 to population 1 keeps its density and moves it along +x at 0.01 dx/dt; a state
    of the wrong size is refused."""
    from lbm_amd import LBMError
    c = build(force_system(), backend)
    t = stol(backend)
    state = c.getFluidState()
    n = 512
    assert state.size == 19*n
    assert np.abs(state).max() <= t
    delta = 0.01
    state[5] -= delta
    state[n + 5] += delta
    c.setFluidState(state)
    assert np.abs(c.getFluidState() - state).max() <= t
    rho, u = c.getFluidFields()
    density = c.force.getFluidDensity()
    assert abs(rho[5]/density - 1) <= 5*t
    assert np.abs(u[5] - [delta*4.0/8/0.01, 0, 0]).max() <= 1e3*t
    assert np.abs(u[6]).max() <= 1e3*t
    with pytest.raises(LBMError):
        c.setFluidState(state[:10])


@pytest.mark.parametrize("backend", BACKENDS)
def test_lattice_parameters(backend):
    """TestLBMFluid.h:412-424."""
    dx, dt, tau = build(fluid_setup((4, 6, 8), 0.9), backend).getLatticeParameters()
    assert dx == pytest.approx(DX, rel=1e-14) and dt == pytest.approx(DT, rel=1e-14)
    assert tau == pytest.approx(0.9, abs=1e-12)


def boundary_setup(scheme, y_faces):
    """TestLBMFluid.h:965-975: 4x5x4, the wall plane j = 0, open faces perpendicular to y."""
    s = fluid_setup((4, 5, 4), 0.8)
    s.positions = []
    s.add_particles([[0.1, 1.2, 0.3]], mass=1.0, coupled=False)
    s.configure.append(lambda f: (f.setSolidNodes(wall_plane(4, 5, 4)), f.setWallScheme(scheme),
                                  f.setFaceBoundary(LBMForce.YMin, y_faces), f.setFaceBoundary(LBMForce.YMax, y_faces)))
    return s


@pytest.mark.parametrize("backend", BACKENDS)
def test_checkpoint_boundaries(backend):
    """TestLBMFluid.h:965-1008: a checkpoint is loaded with the same wall scheme and faces, and refused with another
    wall scheme or other faces."""
    from lbm_amd import LBMError
    s = boundary_setup(LBMForce.Regularized, LBMForce.Density)
    s.masses, s.positions, s.velocities, s.coupled = [], [], [], []
    s.add_particles([[0.1, 1.2, 0.3]], mass=1.0, coupled=False)
    c = build(s, backend)
    c.step(3)
    data = c.createCheckpoint()
    for scheme, faces, message in [(LBMForce.Regularized, LBMForce.Density, None),
                                   (LBMForce.BounceBack, LBMForce.Density, "wall scheme"),
                                   (LBMForce.Regularized, LBMForce.Velocity, "boundary types of the faces")]:
        other = build(boundary_setup(scheme, faces), backend)
        if message is None:
            other.loadCheckpoint(data)
            assert np.array_equal(other.getFluidState(), c.getFluidState())
        else:
            with pytest.raises(LBMError, match=message):
                other.loadCheckpoint(data)


@pytest.mark.parametrize("scheme", [LBMForce.BounceBack, LBMForce.Regularized], ids=["bounce-back", "regularized"])
@pytest.mark.parametrize("backend", BACKENDS)
def test_fluid_state_restart_with_boundaries(backend, scheme):
    """TestLBMFluid.h:1015-1043: with walls, a Velocity inlet, a Density outlet (whose filter reads the previous
    velocity from the populations) and a body force, a run restarted from the fluid state is bit-identical."""
    num_steps, split = 30, 11

    def system():
        s = fluid_setup((6, 5, 4), 0.8)
        s.configure.append(lambda f: (
            f.setSolidNodes(wall_plane(6, 5, 4)), f.setWallScheme(scheme),
            f.setFaceBoundary(LBMForce.XMin, LBMForce.Velocity), f.setFaceBoundary(LBMForce.XMax, LBMForce.Density),
            f.setFaceVelocity(LBMForce.XMin, (0.4, 0.1, 0.0)), f.setFaceDensity(LBMForce.XMax, 1.003*RHO_F),
            f.setBodyAcceleration((0.0, 0.0, 2.0))))
        return build(s, backend)
    a = system()
    a.step(split)
    saved = a.getFluidState()
    a.step(num_steps - split)
    b = system()
    b.setFluidState(saved)
    b.step(num_steps - split)
    assert np.array_equal(a.getFluidState(), b.getFluidState())


SIX_FACES = [LBMForce.Density, LBMForce.Velocity, LBMForce.DensityVelocity, LBMForce.Density, LBMForce.Density,
             LBMForce.DensityVelocity]


def six_face_setup():
    """TestLBMFluid.h:885-895: 4^3 with every face open, of the types Density, Velocity, DensityVelocity, Density,
    Density, DensityVelocity (XMin ... ZMax), here with a velocity and a density of its own on every face."""
    s = fluid_setup((4, 4, 4), 0.8)
    velocity = {1: (0.01, 0.002, -0.003), 2: (-0.004, 0.012, 0.001), 5: (0.003, -0.002, 0.008)}
    density = {0: 1.02, 2: 1.01, 3: 0.98, 4: 1.03, 5: 0.99}

    def conf(f):
        for face, kind in enumerate(SIX_FACES):
            f.setFaceBoundary(face, kind)
        for face, v in velocity.items():
            f.setFaceVelocity(face, tuple(x*DX/DT for x in v))
        for face, r in density.items():
            f.setFaceDensity(face, r*RHO_F)
    s.configure.append(conf)
    return s


def test_boundary_kinds():
    """TestLBMFluid.h:885-915: on nodes shared by several faces the first face (XMin ... ZMax) that sets the velocity
    gives the velocity, and the density too if it is a DensityVelocity face; nodes of Density faces take the density
    of their face, or, on several Density faces, of the first one with zero velocity (the oracle's classification)."""
    WALL, VELOCITY, DENSITY, AT_REST, DENSITY_VELOCITY = range(5)
    c = build(six_face_setup(), "oracle")
    lib = c.lib
    assert lib.oracle_num_boundary_nodes(c.handle) == 4**3 - 2**3
    expected = [((0, 1, 1), DENSITY, 0), ((3, 1, 1), VELOCITY, 1), ((1, 0, 1), DENSITY_VELOCITY, 2),
                ((0, 0, 1), DENSITY_VELOCITY, 2), ((3, 0, 1), VELOCITY, 1), ((1, 3, 0), AT_REST, 3),
                ((0, 3, 0), AT_REST, 0), ((0, 3, 3), DENSITY_VELOCITY, 5), ((3, 0, 3), VELOCITY, 1)]
    for (i, j, k), kind, face in expected:
        got_kind, got_face = ctypes.c_int(), ctypes.c_int()
        assert lib.oracle_boundary_node(c.handle, i + 4*(j + 4*k), ctypes.byref(got_kind), ctypes.byref(got_face))
        assert (got_kind.value, got_face.value) == (kind, face), (i, j, k)
    assert not lib.oracle_boundary_node(c.handle, 1 + 4*(1 + 4*1), ctypes.byref(got_kind), ctypes.byref(got_face))


@pytest.mark.gpu
def test_boundary_kinds_gpu():
    """The six open faces of TestLBMFluid.h:885-915 on the gfx950 library (FP64 storage) against the oracle, step by
    step: a node classified differently (kind or face) takes another density or velocity at once."""
    from parity import lockstep
    worst, _, _ = lockstep(six_face_setup(), 20, storage=LBMForce.FP64, forces=False)
    assert worst["density"] < 1e-12 and worst["velocity"] < 1e-11, worst


@pytest.mark.parametrize("drag", DRAGS)
@pytest.mark.parametrize("backend", BACKENDS)
def test_full_step_kinetic_energy(backend, drag):
    """TestLBMCoupling.h:111-140: OpenMM's kinetic energy of VerletIntegrator, from v(t - dt/2) + dt F(t)/(2m) with the
    forces of a State, is that of the full step, sum m ((v(t - dt/2) + v(t + dt/2))/2)^2/2: the force between steps is
    the one of the next step (random force and removal every 2 steps included)."""
    n = 6
    positions = [[0.6*i, 0.4*i + 0.3, 3.7 - 0.5*i] for i in range(n)]
    velocities = [[0.2*math.cos(i), -0.3*math.sin(i), 0.1*i - 0.25] for i in range(n)]
    c = build(coupling_system(positions, velocities, 10.0, 300.0, configure=[
        lambda f: (f.setDragScheme(drag), f.setRandomNumberSeed(5), f.setFluidMomentumRemovalFrequency(2))]), backend)
    c.step(5)
    for _ in range(4):
        v = c.velocities.copy()
        f = c.getForces()
        reported = 0.5*100.0*((v + 0.5*DT*f/100.0)**2).sum()
        c.step(1)
        expected = 0.5*100.0*((0.5*(v + c.velocities))**2).sum()
        assert reported == pytest.approx(expected, rel=ctol(backend, 1e-10))


@pytest.mark.parametrize("backend", BACKENDS)
def test_repeated_force_evaluation(backend):
    """TestLBMCoupling.h:309-330 and CommonLBMKernels.cpp:1365-1374: OpenMM repeats every force evaluation of a step
    whose neighbour list overflowed, before the step count advances.  The repeated evaluation returns the coupling
    forces of the step again, bit for bit, and leaves the fluid alone: the step conserves the total momentum and the
    run is the same as without the repeat."""
    positions = [[0.3 + 0.45*i, 0.4 + 0.3*i, 0.5 + 0.35*i] for i in range(8)]
    velocities = [[0.2*math.cos(i), -0.3*math.sin(i), 0.1*i - 0.25] for i in range(8)]

    def system():
        c = build(coupling_system(positions, velocities, 10.0, 300.0), backend)
        c.step(5)
        return c
    a, b = system(), system()
    cell = RHO_F*DX**3

    def momentum(c):
        m, j = total_moments(c.getFluidState())
        return (100.0*c.velocities).sum(axis=0) + j*cell*DX/DT
    scale = (100.0*np.linalg.norm(a.velocities, axis=1)).sum()
    p0 = momentum(a)
    ptr = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    pos, vel = a.positions.copy(), a.velocities.copy()
    a._check(a.lib.lbm_begin_step(a.handle, a.stepCount, ptr(pos), ptr(vel)))
    first = np.zeros_like(pos)
    a._check(a.lib.lbm_execute(a.handle, a.stepCount, 1, ptr(pos), ptr(vel), ptr(first)))
    repeated = np.zeros_like(pos)
    a._check(a.lib.lbm_execute(a.handle, a.stepCount, 1, ptr(pos), ptr(vel), ptr(repeated)))
    assert np.array_equal(first, repeated)
    vel = vel + DT*repeated*(1.0/100.0)        # Context.step's update
    a.positions, a.velocities = pos + DT*vel, vel
    a.stepCount += 1
    assert np.linalg.norm(momentum(a) - p0)/scale < ctol(backend, 1e-11)
    b.step(1)
    assert np.array_equal(a.velocities, b.velocities)
    assert np.array_equal(a.getFluidState(), b.getFluidState())


@pytest.mark.parametrize("backend", BACKENDS)
def test_queries_do_not_advance_fluid(backend):
    """TestLBMFluid.h:257-280: force evaluations and new particle velocities between steps leave the fluid of a run
    with a body force and an initial flow bit-identical."""
    states = []
    for queried in (False, True):
        s = fluid_setup((4, 4, 4), 0.8)
        s.configure.append(lambda f: (f.setBodyAcceleration((0.5, -0.2, 0.1)),
                                      f.setInitialFluidVelocity((0.01*DX/DT, 0.0, 0.0))))
        c = build(s, backend)
        rng = np.random.default_rng(1234)
        for _ in range(10):
            c.step(1)
            if queried:
                c.getForces()
                c.getForces()
                c.setVelocities(rng.normal(size=(1, 3)))
        states.append(c.getFluidState())
    assert np.array_equal(states[0], states[1])
