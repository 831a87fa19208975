"""The kernel interface at the C-ABI: getFluidFields with the halo layer (LBMKernels.h:136-145,
ReferenceLBMKernels.cpp:1403-1436, node order of LBMDecomposition.cpp:183-204 with one domain) and the checks of
copyParametersToContext (LBMForceImpl.cpp:388-414), on the FP64 oracle and on the gfx950 library.

The halo is pinned independently of both implementations: numpy's periodic padding (np.pad, mode "wrap") of the
fields without halo gives the layer, NaN replaces it where the halo of the field is not exchanged and beyond the open
faces.  Inside the lattice the halo call must return the fields of the plain call bit for bit."""
import ctypes

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce, _lib
from parity import Setup, equilibrium_state

BACKENDS = [pytest.param("oracle", id="oracle"),
            pytest.param("fp64", id="gpu-fp64", marks=pytest.mark.gpu),
            pytest.param("fp32", id="gpu-fp32", marks=pytest.mark.gpu)]


def build(setup, backend):
    if backend == "oracle":
        return setup.build(oracle_lib.load_oracle())
    return setup.build(None, storage=LBMForce.FP64 if backend == "fp64" else LBMForce.FP32)


def expected_halo(field, grid, open_axes, exchanged):
    """The halo layout from the plain fields: (nz+2, ny+2, nx+2[, 3]) padded periodically, NaN where not exchanged or
    beyond an open face."""
    nx, ny, nz = grid
    f = field.reshape((nz, ny, nx) + field.shape[1:])
    widths = [(1, 1)]*3 + [(0, 0)]*(f.ndim - 3)
    out = np.pad(f, widths, mode="wrap").astype(np.float64)
    inside = np.zeros((nz + 2, ny + 2, nx + 2), dtype=bool)
    inside[1:-1, 1:-1, 1:-1] = True
    if not exchanged:
        out[~inside] = np.nan
    # open axes: the layer across those faces is NaN (axis 0 = x is the last array axis)
    for axis, is_open in enumerate(open_axes):
        if is_open:
            sl = [slice(None)]*3
            for end in (0, -1):
                sl[2 - axis] = end
                out[tuple(sl)] = np.nan
    return out.reshape((-1,) + field.shape[1:])


def fluid_setup(grid, faces=(), solid=()):
    s = Setup(grid, tau=0.8)
    n = grid[0]*grid[1]*grid[2]
    rng = np.random.default_rng(3)
    u = rng.normal(scale=0.01, size=(n, 3))
    rho = 1 + rng.normal(scale=0.01, size=n)
    s.state = equilibrium_state(n, rho, u)

    def conf(f):
        f.setBodyAcceleration((2e-5/1e-4, 0, 0))
        for face, kind in faces:
            f.setFaceBoundary(face, kind)
        if solid:
            f.setSolidNodes(list(solid))
    s.configure.append(conf)
    return s


CASES = [
    pytest.param(((5, 4, 3), (), ()), id="periodic"),
    pytest.param(((5, 4, 6), ((LBMForce.ZMin, LBMForce.Velocity), (LBMForce.ZMax, LBMForce.Velocity)), (7, 30)),
                 id="open-z-solid"),
    pytest.param(((4, 5, 3), ((LBMForce.XMin, LBMForce.Velocity), (LBMForce.XMax, LBMForce.Density)), ()),
                 id="open-x"),
]


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("exchange", [(False, False), (True, False), (False, True), (True, True)],
                         ids=["none", "density", "velocity", "both"])
def test_fluid_fields_with_halo(backend, case, exchange):
    grid, faces, solid = case
    s = fluid_setup(grid, faces, solid)
    s.configure.append(lambda f: (f.setDensityHaloExchange(exchange[0]), f.setVelocityHaloExchange(exchange[1])))
    c = build(s, backend)
    c.step(3)
    rho, u = c.getFluidFields()
    hrho, hu = c.getFluidFields(halo=True)
    nx, ny, nz = grid
    assert hrho.shape == ((nx + 2)*(ny + 2)*(nz + 2),) and hu.shape == (hrho.size, 3)
    open_axes = [c.force.getFaceBoundary(2*a) != LBMForce.Periodic for a in range(3)]
    erho = expected_halo(rho, grid, open_axes, exchange[0])
    eu = expected_halo(u, grid, open_axes, exchange[1])
    assert np.array_equal(np.isnan(hrho), np.isnan(erho))
    assert np.array_equal(np.isnan(hu), np.isnan(eu))
    assert np.array_equal(hrho[~np.isnan(erho)], erho[~np.isnan(erho)])
    assert np.array_equal(hu[~np.isnan(eu)], eu[~np.isnan(eu)])
    if solid:
        assert (rho[list(solid)] == 0).all()
    # The force-level call: gather changes nothing with one domain; gather with halo is refused (LBMForceImpl.cpp:426).
    frho, fu = c.force.getFluidFields(c, gather=True)
    assert np.array_equal(frho, rho) and np.array_equal(fu, u)
    with pytest.raises(_lib.LBMError, match="cannot gather the fields and add the halo"):
        c.force.getFluidFields(c, gather=True, halo=True)


@pytest.mark.gpu
def test_fluid_fields_with_halo_match_oracle():
    """GPU halo fields against the oracle's on a run with open faces, solid nodes, a body force and particles."""
    grid = (6, 5, 7)
    s = fluid_setup(grid, ((LBMForce.ZMin, LBMForce.Velocity), (LBMForce.ZMax, LBMForce.Density)), (8, 40, 41))
    s.add_particles(np.random.default_rng(4).uniform(1, 4, size=(6, 3)), mass=100.0)
    s.configure.append(lambda f: (f.setDensityHaloExchange(True), f.setVelocityHaloExchange(True),
                                  f.setFriction(10.0), f.setTemperature(0.0)))
    ref = s.build(oracle_lib.load_oracle())
    gpu = s.build(None, storage=LBMForce.FP64)
    ref.step(20)
    gpu.step(20)
    r0, u0 = ref.getFluidFields(halo=True)
    r1, u1 = gpu.getFluidFields(halo=True)
    assert np.array_equal(np.isnan(r0), np.isnan(r1)) and np.array_equal(np.isnan(u0), np.isnan(u1))
    ok = ~np.isnan(r0)
    assert np.abs(r1[ok] - r0[ok]).max() <= 1e-12*np.abs(r0[ok]).max()
    ok = ~np.isnan(u0)
    assert np.abs(u1[ok] - u0[ok]).max() <= 1e-12*np.abs(u0[ok]).max()


REFUSED = [
    ("grid", lambda q: setattr(q, "nx", q.nx + 1), "the grid size cannot be changed"),
    ("density", lambda q: setattr(q, "density", q.density*1.01), "the fluid density and viscosity cannot be changed"),
    ("tau", lambda q: setattr(q, "tau", q.tau + 0.1), "the fluid density and viscosity cannot be changed"),
    ("wall", lambda q: setattr(q, "wall_scheme", 1 - q.wall_scheme), "the wall scheme cannot be changed"),
    ("halo", lambda q: setattr(q, "velocity_halo_exchange", 1), "the exchange of the halo cannot be switched on or off"),
    ("face", lambda q: q.face_boundary.__setitem__(4, LBMForce.Velocity),
     "the boundary types of the faces cannot be changed"),
    ("drag", lambda q: setattr(q, "drag_scheme", LBMForce.Centered), "the drag scheme cannot be changed"),
    ("stencil", lambda q: setattr(q, "interpolation_stencil", LBMForce.Keys),
     "the interpolation stencil cannot be changed"),
    ("fluctuations", lambda q: setattr(q, "fluid_fluctuations", 1),
     "the fluid fluctuations cannot be switched on or off"),
]


@pytest.mark.parametrize("backend", BACKENDS)
def test_copy_parameters_refuses_fixed_settings(backend):
    """lbm_copy_parameters called directly (without the Python-level checks of updateParametersInContext) refuses a
    change of every setting fixed at creation, with the reference's message, and leaves the context as it was; a
    change of the body force is accepted."""
    s = Setup((6, 6, 6), tau=0.8)
    s.add_particles([[1.2, 2.3, 3.4], [4.5, 0.5, 2.5]], mass=100.0)
    s.configure.append(lambda f: (f.setSolidNodes([0, 1, 2]), f.setFriction(10.0), f.setTemperature(0.0),
                                  f.setInterpolationStencil(LBMForce.NearestNode)))
    c = build(s, backend)
    twin = build(s, backend)
    p = c.params
    for name, change, message in REFUSED:
        q = _lib.LBMParams()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
        change(q)
        status = c.lib.lbm_copy_parameters(c.handle, ctypes.byref(q))
        assert status == _lib.LBM_ERROR_INVALID, name
        assert c.lib.lbm_last_error().decode() == "updateParametersInContext: " + message, name
    c.step(5)
    twin.step(5)
    assert np.array_equal(c.getFluidState(), twin.getFluidState())
    assert np.array_equal(c.getForces(), twin.getForces())
    q = _lib.LBMParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    q.body_acceleration[0] = 1e-5
    _lib.check(c.lib, c.lib.lbm_copy_parameters(c.handle, ctypes.byref(q)))
    c.step(1)
    twin.step(1)
    assert not np.array_equal(c.getFluidState(), twin.getFluidState())


# ------------------------------------------------------------------------------------------------------------------
# The time filter of the Density faces (ReferenceLBMKernels.cpp:1264-1272; docs/theory.md:366-382), pinned on one step
# away from steady state, where the filter weight shows.
# ------------------------------------------------------------------------------------------------------------------
CXA = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
CYA = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0])
CZA = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1])
WA = np.array([1/3] + [1/18]*6 + [1/36]*12)
OPPOSITE = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])


def equilibrium_deviation(drho, u):
    """feq_q - w_q = w_q [drho + rho (3 c.u + 9/2 (c.u)^2 - 3/2 u.u)] (the second-order equilibrium, cs^2 = 1/3)."""
    cu = CXA*u[0] + CYA*u[1] + CZA*u[2]
    return WA*(drho + (1 + drho)*(3*cu + 4.5*cu*cu - 1.5*np.dot(u, u)))


def hermite2():
    """H2(c_q) = c c - I/3 as xx, yy, zz, xy, xz, yz (rows q)."""
    return np.stack([CXA*CXA - 1/3, CYA*CYA - 1/3, CZA*CZA - 1/3, CXA*CYA, CXA*CZA, CYA*CZA], axis=1)


def test_density_filter_helpers_match_golden_vectors():
    """The helpers of the filter test against the reference's D3Q19 algebra (golden vectors from its header)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "d3q19_vectors.npz"))
    H = hermite2()
    for k in range(len(g["dr"])):
        assert np.abs(equilibrium_deviation(g["dr"][k], g["u"][k]) - g["equilibrium"][k]).max() < 1e-15
        # Pi_neq of the deviations df at the momentum (1 + dr) u, then the regularized non-equilibrium part
        pi = H.T @ (g["df"][k] - equilibrium_deviation(g["dr"][k], g["u"][k]))
        assert np.abs(pi - g["pi_neq"][k]).max() < 1e-15
        fneq = 4.5*WA*(H[:, :3] @ g["pi_neq"][k][:3] + 2*(H[:, 3:] @ g["pi_neq"][k][3:]))
        assert np.abs(fneq - g["regularized_neq"][k]).max() < 1e-15


@pytest.mark.parametrize("backend", BACKENDS[:2])
def test_density_face_time_filter(backend):
    """One step from a state away from steady flow (random velocities and non-equilibrium parts): the populations that
    a node of a Density face (ZMin, density 1.004) rebuilds, its unknown directions c_z = +1, are the equilibrium at
    the face density and the velocity u_b, plus (1 - omega) times the regularized non-equilibrium part of the node.
    u_b keeps the node's velocity v_x = j/rho along the face; across it, u_b is (v_ZH + v_x)/2, where v_ZH is the
    velocity that gives the node the density of the face from the populations that arrived, the unknown ones replaced
    by the bounce-back of their opposites (Zou and He 1997: rho_face = rho_0 + 2 rho_out + rho_face v_ZH).  The test
    takes the post-streaming populations and the node's moments from the states before and after the step, and checks
    that the weight 0 or 1 (v_ZH alone, v_x alone) would miss by more than 1e4 times the tolerance."""
    nx, ny, nz = 3, 3, 4
    n = nx*ny*nz
    rng = np.random.default_rng(41)
    s = Setup((nx, ny, nz), dx=1.0, dt=0.01, tau=0.8)
    u = rng.normal(scale=0.02, size=(n, 3))
    rho = 1.0 + rng.normal(scale=0.003, size=n)
    state = equilibrium_state(n, rho, u).reshape(19, n) + rng.normal(scale=4e-4, size=(19, n))*WA[:, None]
    s.state = state.reshape(-1)
    face_density = 1.004

    def conf(f):
        f.setFaceBoundary(LBMForce.ZMin, LBMForce.Density)
        f.setFaceBoundary(LBMForce.ZMax, LBMForce.Density)
        f.setFaceDensity(LBMForce.ZMin, face_density*602.214)
        f.setFaceDensity(LBMForce.ZMax, 602.214)
    s.configure.append(conf)
    c = build(s, backend)
    c.step(1)
    after = c.getFluidState().reshape(19, n)
    omega = 1/0.8
    node = 1 + nx*(1 + ny*0)
    f0 = state[:, node]
    drho = f0.sum()
    j = np.array([(CXA*f0).sum(), (CYA*f0).sum(), (CZA*f0).sum()])
    v_x = j/(1 + drho)
    pi = hermite2().T @ (f0 - equilibrium_deviation(drho, v_x))
    H = hermite2()
    fneq = 4.5*WA*(H[:, :3] @ pi[:3] + 2*(H[:, 3:] @ pi[3:]))
    unknown = CZA == 1
    arrived = np.where(unknown, after[OPPOSITE, node], after[:, node])
    v_zh = ((face_density - 1.0) - arrived.sum())/face_density          # inward = +z on ZMin

    def rebuilt(beta):
        u_b = np.array([v_x[0], v_x[1], beta*v_zh + (1 - beta)*v_x[2]])
        return (equilibrium_deviation(face_density - 1.0, u_b) + (1 - omega)*fneq)[unknown]

    got = after[unknown, node]
    err = np.abs(got - rebuilt(0.5)).max()
    miss = min(np.abs(got - rebuilt(0.0)).max(), np.abs(got - rebuilt(1.0)).max())
    print(f"v_ZH {v_zh:.3e} v_x {v_x[2]:.3e}: rebuilt populations within {err:.1e} of the weight 1/2; "
          f"weights 0 and 1 miss by {miss:.1e}")
    assert err < 1e-13
    assert miss > 1e4*1e-13
