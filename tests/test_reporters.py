"""The run reporters (lbm_amd/reporters.py), on the CPU oracle: the temperature file (python/openmmlbm.i:619-680) and
the VTK files (:682-860), read back and compared with the Context's own fields and particles."""
import re
import struct

import numpy as np

import oracle_lib
from lbm_amd import LBMForce
from lbm_amd.reporters import LBMTemperatureReporter, LBMVTKReporter, run
from lbm_amd.workloads import BOLTZ
from parity import Setup


def small_context(drag=LBMForce.Explicit):
    n = 6
    s = Setup((n, n, n), tau=0.8)
    rng = np.random.default_rng(2)
    s.add_particles(rng.uniform(-2, 8, size=(5, 3)), mass=100.0, velocities=rng.normal(scale=0.3, size=(5, 3)))
    solid = [i + n*n*0 for i in range(n)]

    def conf(f):
        f.setFriction(5.0)
        f.setTemperature(0.0)
        f.setDragScheme(drag)
        f.setSolidNodes(solid)
        f.setBodyAcceleration((0.3, 0, 0))
    s.configure.append(conf)
    return s.build(oracle_lib.load_oracle()), s


def read_vtk(path):
    data = open(path, "rb").read()
    head_end = data.index(b'<AppendedData encoding="raw">\n_') + len(b'<AppendedData encoding="raw">\n_')
    head = data[:head_end].decode()
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)"(?: Name="(\w+)")?(?: NumberOfComponents="3")? format="appended" '
                         r'offset="(\d+)"/>', head):
        kind, name, off = m.group(1), m.group(2) or "Points", int(m.group(3))
        size = struct.unpack_from("<Q", data, head_end + off)[0]
        dtype = {"Float32": "<f4", "Float64": "<f8", "UInt8": "u1", "Int32": "<i4", "Int64": "<i8"}[kind]
        arrays[name] = np.frombuffer(data, dtype=dtype, count=size//np.dtype(dtype).itemsize, offset=head_end + off + 8)
    return head, arrays


def test_vtk_files(tmp_path):
    ctx, s = small_context()
    prefix = str(tmp_path/"run")
    rep = LBMVTKReporter(prefix, 4, ctx.force, double=True)
    run(ctx, 8, [rep])
    head, a = read_vtk(f"{prefix}_velocity_{8:010d}.vti")
    rho, u = ctx.getFluidFields()
    assert 'Spacing="1.0 1.0 1.0"' in head and 'WholeExtent="0 5 0 5 0 5"' in head
    assert np.array_equal(a["velocity"], u.reshape(-1))
    assert a["solid"].sum() == 6 and a["solid"][:6].all()
    _, d = read_vtk(f"{prefix}_density_{8:010d}.vti")
    assert np.array_equal(d["density"], rho)
    _, p = read_vtk(f"{prefix}_particles_{8:010d}.vtp")
    pos = p["Points"].reshape(-1, 3)
    box = np.array(s.grid, dtype=float)*s.dx
    assert np.allclose(pos, ctx.positions - np.floor(ctx.positions/box)*box)
    assert np.all((pos >= 0) & (pos < box))
    assert np.array_equal(p["velocity"].reshape(-1, 3), ctx.velocities)
    assert list(p["coupled"]) == [1]*5 and list(p["index"]) == list(range(5))
    pvd = open(prefix + ".pvd").read()
    assert pvd.count("<DataSet ") == 6 and 'timestep="0.04"' in pvd and 'timestep="0.08"' in pvd


def test_temperature_file(tmp_path):
    for drag in (LBMForce.Explicit, LBMForce.Centered):
        ctx, s = small_context(drag)
        path = str(tmp_path/f"temperature_{drag}.txt")
        rep = LBMTemperatureReporter(path, 5, ctx.force)
        run(ctx, 10, [rep])
        lines = open(path).read().splitlines()
        assert lines[0] == '#"Step","Time (ps)","Temperature (K)"' and len(lines) == 3
        step, time, temperature = lines[-1].split()
        assert int(step) == 10 and abs(float(time) - 0.1) < 1e-12
        v = ctx.velocities.copy()
        if drag == LBMForce.Explicit:
            v = v + ctx.getForces()*(0.5*0.01/100.0)
        expected = 100.0*np.sum(v*v)/(3*5*BOLTZ)
        assert abs(float(temperature) - expected) <= 1e-3
