"""The C-ABI libraries load on a machine without a GPU and export every entry point declared in include/*.h."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(lbm_\w+)\s*\(", text, flags=re.M)))


def test_headers_declare_entry_points():
    names = declared_functions("lbm_amd.h")
    for expected in ("lbm_create", "lbm_begin_step", "lbm_execute", "lbm_get_fluid_fields", "lbm_get_fluid_state",
                     "lbm_set_fluid_state", "lbm_get_wall_force", "lbm_get_fluid_mach_number", "lbm_destroy",
                     "lbm_compute_lattice_parameters", "lbm_create_checkpoint", "lbm_load_checkpoint"):
        assert expected in names


def test_gpu_library_exports_everything_declared():
    import lbm_amd
    lib = ctypes.CDLL(lbm_amd.LIBRARY_PATH)
    headers = sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h"))
    assert {"lbm_amd.h", "lbm_amd_harness.h", "lbm_amd_batch.h"} <= set(headers)
    for header in headers:
        for name in declared_functions(header):
            assert hasattr(lib, name), f"{name} declared in {header} is not exported"


def test_oracle_exports_the_kernel_abi():
    import oracle_lib
    from lbm_amd import _lib
    lib = ctypes.CDLL(oracle_lib.ORACLE_PATH)
    for name in _lib.KERNEL_FUNCTIONS:
        assert hasattr(lib, name), name


def test_lattice_parameters_conversion():
    """LBMForceImpl.cpp:94-111: tau = 3 nu dt/dx^2 + 1/2, velocities and accelerations in lattice units."""
    from lbm_amd import LBMForce
    f = LBMForce()
    f.setGridSize(10, 20, 5)
    f.setKinematicViscosity(0.5)
    f.setBodyAcceleration((2.0, 0, -1.0))
    f.setInitialFluidVelocity((0.3, 0, 0))
    f.setTemperature(250.0)
    p = f.computeLatticeParameters((5.0, 10.0, 2.5), 0.02, [1.0])
    dx, dt = 0.5, 0.02
    assert p.dx == pytest.approx(dx)
    assert p.tau == pytest.approx(3*0.5*dt/dx**2 + 0.5)
    assert p.omega == pytest.approx(1/p.tau)
    assert p.body_acceleration[0] == pytest.approx(2.0*dt*dt/dx)
    assert p.initial_velocity[0] == pytest.approx(0.3*dt/dx)
    assert p.kT == pytest.approx(0.00831446261815324*250.0)
    assert p.fluid_kT == 0.0
    assert p.momentum_removal_frequency == 1 and p.mach_check_frequency == 100


@pytest.mark.parametrize("change,message", [
    (lambda f: f.setGridSize(0, 4, 4), "grid size must be set"),
    (lambda f: f.setKinematicViscosity(-1), "kinematic viscosity must be positive"),
    (lambda f: f.setFriction(-1), "friction must not be negative"),
    (lambda f: f.setFaceBoundary(0, 1), "must be both periodic or both open"),
    (lambda f: (f.setFaceBoundary(0, 1), f.setFaceBoundary(1, 2)), "momentum cannot be removed"),
    (lambda f: f.setSolidNodes([3, 3]), "listed more than once"),
    (lambda f: f.setSolidNodes([1000]), "out of range"),
    (lambda f: (f.setFluidMomentumRemovalFrequency(0), f.setFaceBoundary(4, 2), f.setFaceBoundary(5, 2),
                f.setFluidFluctuations(True)), "Density faces cannot be used with fluid fluctuations"),
])
def test_invalid_setups(change, message):
    """The validation of LBMForceImpl::computeLatticeParameters, with its messages (TestLBMForce.h:223-255)."""
    from lbm_amd import LBMForce, LBMError
    f = LBMForce()
    f.setGridSize(4, 4, 4)
    change(f)
    with pytest.raises(LBMError, match=message):
        f.computeLatticeParameters((4.0, 4.0, 4.0), 0.01, [1.0])


def test_non_cubic_cells_rejected():
    from lbm_amd import LBMForce, LBMError
    f = LBMForce()
    f.setGridSize(4, 4, 4)
    with pytest.raises(LBMError, match="cubic"):
        f.computeLatticeParameters((4.0, 5.0, 4.0), 0.01, [1.0])
