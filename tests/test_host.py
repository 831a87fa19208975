"""The host side of the force that sits above the kernel (openmmapi/src/LBMForce.cpp, LBMForceImpl.cpp), as the
reference's tests check it: the settings and their defaults, the setups refused when a Context is created, the
refusals of updateParametersInContext() and of checkpoints, and the warnings printed on stderr.  The Contexts here run
on the oracle library: these checks come before any kernel call (the same host code drives the gfx950 library)."""
import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMError, LBMForce
from lbm_amd.context import Context, System, VerletIntegrator
from parity import Setup

DX, DT, RHO_F = 0.5, 0.01, 602.2


def fluid_setup(grid, tau):
    """createFluidSystem (TestLBMFluid.h:40-58): one uncoupled particle, no removal."""
    return Setup(grid, dx=DX, dt=DT, tau=tau, density=RHO_F)


def coupled_setup(num, friction, temperature, configure=()):
    """createCoupledSystem (TestLBMCoupling.h:40-61): 8^3 nodes, tau = 0.8, seed 1, all particles of 100 Da coupled."""
    s = Setup((8, 8, 8), dx=DX, dt=DT, tau=0.8, density=RHO_F)
    s.add_particles([[0.3 + i, 0.4 + i, 0.5 + i] for i in range(num)], mass=100.0)

    def conf(f):
        f.setFriction(friction)
        f.setTemperature(temperature)
        f.setRandomNumberSeed(1)
        for fn in configure:
            fn(f)
    s.configure.append(conf)
    return s


def build(setup):
    return setup.build(oracle_lib.load_oracle())


def warnings_of(setup, capsys):
    capsys.readouterr()
    build(setup)
    return capsys.readouterr().err


def test_parameters():
    """TestLBMForce.h:67-120: setters, getters and defaults."""
    f = LBMForce()
    assert f.getGridSize()[0] == 0
    f.setGridSize(10, 12, 14)
    assert f.getGridSize() == (10, 12, 14)
    f.setFluidDensity(500.0)
    f.setKinematicViscosity(2.0)
    f.setFriction(5.0)
    f.setTemperature(310.0)
    f.setRandomNumberSeed(123)
    f.setBodyAcceleration((0.1, 0.2, 0.3))
    f.setInitialFluidVelocity((-0.1, 0.0, 0.05))
    f.setFluidMomentumRemovalFrequency(10)
    assert f.getFluidDensity() == 500.0 and f.getKinematicViscosity() == 2.0
    assert f.getFriction() == 5.0 and f.getTemperature() == 310.0 and f.getRandomNumberSeed() == 123
    assert tuple(f.getBodyAcceleration()) == (0.1, 0.2, 0.3)
    assert tuple(f.getInitialFluidVelocity()) == (-0.1, 0.0, 0.05)
    assert f.getFluidMomentumRemovalFrequency() == 10
    assert f.getMachCheckFrequency() == 100
    assert f.getCouplingScheme() == LBMForce.EulerMaruyama
    f.setCouplingScheme(LBMForce.NVE)
    assert f.getCouplingScheme() == LBMForce.NVE
    assert f.getDragScheme() == LBMForce.Explicit
    f.setDragScheme(LBMForce.Centered)
    assert f.getDragScheme() == LBMForce.Centered
    assert f.getWallScheme() == LBMForce.BounceBack
    f.setWallScheme(LBMForce.Regularized)
    assert f.getWallScheme() == LBMForce.Regularized
    assert f.getMachNumberLimit() == 0.3
    f.setMachCheckFrequency(50)
    f.setMachNumberLimit(0.25)
    assert f.getMachCheckFrequency() == 50 and f.getMachNumberLimit() == 0.25
    assert list(f.getSolidNodes()) == []
    f.setSolidNodes([5, 2, 9])
    assert list(f.getSolidNodes()) == [5, 2, 9]
    assert f.addParticle(3) == 0
    assert f.addParticle(7) == 1
    assert f.getNumParticles() == 2
    f.setParticle(1, 8)
    assert f.getParticle(1) == 8
    assert f.usesPeriodicBoundaryConditions()


class LangevinMiddleIntegrator:
    """An integrator that is not a VerletIntegrator."""

    def __init__(self, temperature, friction, stepSize):
        self.stepSize = stepSize


@pytest.mark.parametrize("case", ["integrator", "viscosity", "cells", "particle", "duplicate"])
def test_invalid_setup(case):
    """TestLBMForce.h:223-255: a Langevin integrator, zero viscosity, non-cubic cells, a coupled index out of range
    and a particle coupled twice are refused when the Context is created."""
    s = System((4.0, 4.0, 4.0))
    f = LBMForce()
    f.setGridSize(8, 8, 8)
    for i in range(10):
        s.addParticle(100.0)
        f.addParticle(i)
    integrator = VerletIntegrator(0.01)
    if case == "integrator":
        integrator = LangevinMiddleIntegrator(300.0, 1.0, 0.01)
    elif case == "viscosity":
        f.setKinematicViscosity(0.0)
    elif case == "cells":
        f.setGridSize(8, 8, 10)
    elif case == "particle":
        f.addParticle(10)
    else:
        f.addParticle(0)
    with pytest.raises(LBMError):
        Context(s, integrator, f, library=oracle_lib.load_oracle())


@pytest.mark.parametrize("nodes", [[-1], [64], [3, 5, 3], list(range(64))], ids=["negative", "beyond", "twice", "all"])
def test_solid_node_checks(nodes):
    """TestLBMFluid.h:463-484."""
    s = fluid_setup((4, 4, 4), 0.8)
    s.configure.append(lambda f: f.setSolidNodes(nodes))
    with pytest.raises(LBMError):
        build(s)


@pytest.mark.parametrize("case", [0, 1, 2])
def test_face_checks(case):
    """TestLBMFluid.h:921-941: one open face of an axis, open faces with the removal, two nodes along an open axis."""
    s = fluid_setup((4, 2 if case == 2 else 4, 4), 0.8)

    def conf(f):
        f.setFaceBoundary(LBMForce.YMin, LBMForce.Velocity)
        if case != 0:
            f.setFaceBoundary(LBMForce.YMax, LBMForce.Density)
        if case == 1:
            f.setFluidMomentumRemovalFrequency(1)
    s.configure.append(conf)
    with pytest.raises(LBMError):
        build(s)


def test_stencil_errors():
    """TestLBMStencils.h:547-604: the centred drag and open faces are accepted with a stencil, an unknown stencil is
    refused; the stencil cannot be changed in a Context, and a checkpoint is refused by another stencil."""
    def setup(extra):
        return coupled_setup(1, 5.0, 0.0, [lambda f: f.setInterpolationStencil(LBMForce.Trilinear), extra])
    build(setup(lambda f: None))
    build(setup(lambda f: f.setDragScheme(LBMForce.Centered)))
    build(setup(lambda f: (f.setFaceBoundary(LBMForce.XMin, LBMForce.Velocity),
                           f.setFaceBoundary(LBMForce.XMax, LBMForce.Velocity))))
    with pytest.raises(LBMError):
        build(setup(lambda f: f.setInterpolationStencil(7)))
    c = build(coupled_setup(1, 5.0, 0.0, [lambda f: f.setInterpolationStencil(LBMForce.ThreePoint)]))
    c.force.setInterpolationStencil(LBMForce.Keys)
    with pytest.raises(LBMError):
        c.updateParametersInContext()
    c.force.setInterpolationStencil(LBMForce.ThreePoint)
    data = c.createCheckpoint()
    other = build(coupled_setup(1, 5.0, 0.0))
    with pytest.raises(LBMError, match="interpolation stencil"):
        other.loadCheckpoint(data)


def test_centred_drag_cannot_be_changed():
    """TestLBMCentered.h:297-307: updateParametersInContext() refuses another drag scheme.  (The System of this host
    has no virtual sites, and the other forces are evaluated before the LBMForce, which is therefore the last force:
    the other requirements of the centred drag, TestLBMCentered.h:280-295, hold by construction.)"""
    c = build(coupled_setup(2, 5.0, 0.0, [lambda f: f.setDragScheme(LBMForce.Centered)]))
    c.force.setDragScheme(LBMForce.Explicit)
    with pytest.raises(LBMError, match="drag scheme cannot be changed"):
        c.updateParametersInContext()


@pytest.mark.parametrize("change,message", [
    (lambda f: f.addParticle(2), "set of coupled particles"),
    (lambda f: f.setSolidNodes([0, 1]), "solid nodes"),
    (lambda f: f.setWallScheme(LBMForce.Regularized), "wall scheme"),
    (lambda f: (f.setFaceBoundary(LBMForce.XMin, LBMForce.Velocity), f.setFaceBoundary(LBMForce.XMax, LBMForce.Velocity)),
     "boundary types of the faces"),
    (lambda f: f.setGridSize(4, 4, 4), "grid size"),
])
def test_update_refusals(change, message):
    """LBMForceImpl.cpp:388-414: what updateParametersInContext() cannot change."""
    s = coupled_setup(2, 5.0, 0.0)
    s.add_particles([[2.0, 2.0, 2.0]], coupled=False)
    c = build(s)
    change(c.force)
    with pytest.raises(LBMError, match=message):
        c.updateParametersInContext()


def test_step_size_fixed():
    """LBMForceImpl.cpp:373-378: the lattice time step is the integrator step of the Context's creation."""
    c = build(coupled_setup(1, 5.0, 0.0))
    c.step(1)
    c.integrator.stepSize = 0.02
    with pytest.raises(LBMError, match="step size changed"):
        c.step(1)
    with pytest.raises(LBMError, match="step size changed"):
        c.getForces()


@pytest.mark.parametrize("tau", [0.503, 1.0, 2.2])
def test_relaxation_time_warning(tau, capsys):
    """TestLBMFluid.h:429-449."""
    err = warnings_of(fluid_setup((4, 4, 4), tau), capsys)
    assert ("relaxation time" in err) == (tau < 0.505 or tau > 2.0)


@pytest.mark.parametrize("coupling", [0, 1, 2], ids=["uncoupled", "explicit", "centred"])
@pytest.mark.parametrize("tau", [1.6, 1.75])
def test_self_mobility_warning(tau, coupling, capsys):
    """TestLBMCoupling.h:670-701: tau > 1.7 warns with the explicit drag only."""
    if coupling == 0:
        s = fluid_setup((8, 8, 8), tau)
    else:
        nu = (tau - 0.5)/3.0*DX*DX/DT
        s = coupled_setup(1, 1.0, 0.0, [lambda f: f.setKinematicViscosity(nu)])
        if coupling == 2:
            s.configure.append(lambda f: f.setDragScheme(LBMForce.Centered))
    err = warnings_of(s, capsys)
    assert ("> 1.7" in err) == (tau > 1.7 and coupling == 1)


@pytest.mark.parametrize("drag", [LBMForce.Explicit, LBMForce.Centered], ids=["explicit", "centred"])
@pytest.mark.parametrize("friction", [50.0, 150.0])
def test_friction_warning(friction, drag, capsys):
    """TestLBMCoupling.h:707-729: friction*dt > 1 warns with the explicit drag only."""
    err = warnings_of(coupled_setup(1, friction, 0.0, [lambda f: f.setDragScheme(drag)]), capsys)
    assert ("friction*dt" in err) == (friction*DT > 1.0 and drag == LBMForce.Explicit)


@pytest.mark.parametrize("scheme", [LBMForce.EulerMaruyama, LBMForce.NVE], ids=["em", "nve"])
@pytest.mark.parametrize("drag", [LBMForce.Explicit, LBMForce.Centered], ids=["explicit", "centred"])
@pytest.mark.parametrize("fluctuations", [False, True], ids=["still", "fluctuating"])
def test_fluctuation_drag_warning(fluctuations, drag, scheme, capsys):
    """TestLBMFluctuations.h:507-533: the heating of the explicit drag, 6.64% for 100 Da at friction*dt = 0.1."""
    err = warnings_of(coupled_setup(2, 10.0, 300.0, [lambda f: (f.setFluidFluctuations(fluctuations),
                                                                f.setDragScheme(drag), f.setCouplingScheme(scheme))]),
                      capsys)
    warned = "with fluid fluctuations the explicit drag" in err
    assert warned == (fluctuations and drag == LBMForce.Explicit and scheme == LBMForce.EulerMaruyama)
    if warned:
        assert "= 6.64" in err


@pytest.mark.parametrize("boundary", range(5),
                         ids=["regularized", "bounce-back", "density-velocity", "velocity", "still-regularized"])
@pytest.mark.parametrize("tau", [0.55, 1.0])
def test_boundary_noise_warning(tau, boundary, capsys):
    """TestLBMFluctuations.h:540-569: regularized walls and DensityVelocity faces warn outside 0.586 <= tau <= 3.41."""
    s = fluid_setup((6, 5, 4), tau)

    def conf(f):
        f.setTemperature(300.0)
        f.setFluidFluctuations(boundary != 4)
        if boundary in (0, 1, 4):
            f.setSolidNodes([i + 6*5*k for k in range(4) for i in range(6)])
            f.setWallScheme(LBMForce.BounceBack if boundary == 1 else LBMForce.Regularized)
        else:
            for face in (LBMForce.ZMin, LBMForce.ZMax):
                f.setFaceBoundary(face, LBMForce.DensityVelocity if boundary == 2 else LBMForce.Velocity)
    s.configure.append(conf)
    err = warnings_of(s, capsys)
    warned = "exactly at equilibrium only for 0.586 <= tau <= 3.41" in err
    assert warned == (tau < 0.586 and boundary in (0, 2))


def test_stencil_warnings(capsys):
    """TestLBMStencils.h:613-656: the heating warning carries the self weight K of the stencil (1.968%, 0.830%,
    3.586% for trilinear, three-point and Keys); tau > 1.7 warns with every stencil but the three-point one."""
    heating = {LBMForce.Trilinear: "= 1.968", LBMForce.ThreePoint: "= 0.830", LBMForce.Keys: "= 3.586"}
    for stencil, text in heating.items():
        err = warnings_of(coupled_setup(1, 10.0, 300.0, [lambda f: (f.setFluidFluctuations(True),
                                                                    f.setInterpolationStencil(stencil))]), capsys)
        assert "the explicit drag makes the coupled particles hotter" in err
        assert text in err and "averaged over a cell" in err
    nu = (1.75 - 0.5)/3.0*DX*DX/DT
    for stencil in (LBMForce.NearestNode, LBMForce.Trilinear, LBMForce.ThreePoint, LBMForce.Keys):
        err = warnings_of(coupled_setup(1, 1.0, 0.0, [lambda f: (f.setKinematicViscosity(nu),
                                                                 f.setInterpolationStencil(stencil))]), capsys)
        assert ("> 1.7" in err) == (stencil != LBMForce.ThreePoint)


def test_fluctuation_size_message(capsys):
    """LBMForceImpl.cpp:318-340: the size of the fluctuations in lattice units; above kT = 1/3000 a warning."""
    err = warnings_of(coupled_setup(1, 10.0, 300.0, [lambda f: f.setFluidFluctuations(True)]), capsys)
    cell = RHO_F*DX**3
    kT = 0.00831446261815324*300.0*DT**2/(cell*DX**2)
    assert f"kT/(m_c (dx/dt)^2) = {kT:.2g}" in err
    assert ("larger than those validated" in err) == (kT > 1/3000)
    err = warnings_of(coupled_setup(1, 10.0, 300.0), capsys)
    assert "fluctuating fluid" not in err
