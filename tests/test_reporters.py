"""The run reporters (lbm_amd/reporters.py), on the CPU oracle: the temperature file (python/openmmlbm.i:619-680) and
the VTK files (:682-860), read back and compared with the Context's own fields and particles; on the GPU library, the
same files against the oracle's."""
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from lbm_amd.reporters import LBMTemperatureReporter, LBMVTKReporter, run
from lbm_amd.workloads import BOLTZ
from parity import Setup


def small_context(drag=LBMForce.Explicit, backend="oracle"):
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
    if backend == "oracle":
        return s.build(oracle_lib.load_oracle()), s
    return s.build(None, storage=LBMForce.FP64 if backend == "fp64" else LBMForce.FP32), s


def read_vtk(path):
    """The XML part of a VTK file (parsed) and its appended arrays by name ("Points" for the unnamed one)."""
    data = open(path, "rb").read()
    mark = b'<AppendedData encoding="raw">\n_'
    start = data.index(mark)
    root = ET.fromstring(data[:start].decode("ascii") + "</VTKFile>")
    base = start + len(mark)
    arrays = {}
    for el in root.iter("DataArray"):
        if el.get("format") != "appended":
            continue
        dtype = np.dtype({"Float32": "<f4", "Float64": "<f8", "UInt8": "u1", "Int32": "<i4", "Int64": "<i8"}[el.get("type")])
        off = base + int(el.get("offset"))
        size = struct.unpack_from("<Q", data, off)[0]
        arrays[el.get("Name", "Points")] = np.frombuffer(data, dtype=dtype, count=size//dtype.itemsize, offset=off + 8)
    assert data.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
    return root, arrays


def units(root):
    codes = [int(c) for c in root.find(".//FieldData/DataArray").text.split()]
    return bytes(codes).decode("ascii").split("\0")[:-1]


def test_vtk_files(tmp_path):
    ctx, s = small_context()
    prefix = str(tmp_path/"run")
    rep = LBMVTKReporter(prefix, 4, ctx.force, double=True)
    run(ctx, 8, [rep])
    root, a = read_vtk(f"{prefix}_velocity_{8:010d}.vti")
    rho, u = ctx.getFluidFields()
    image = root.find("ImageData")
    assert image.get("Spacing") == "1.0 1.0 1.0" and image.get("WholeExtent") == "0 5 0 5 0 5"
    assert units(root) == ["Origin, Spacing: nm", "velocity: nm/ps", "solid: 1 for a solid node, 0 for a fluid node"]
    assert np.array_equal(a["velocity"], u.reshape(-1))
    assert a["solid"].sum() == 6 and a["solid"][:6].all()
    _, d = read_vtk(f"{prefix}_density_{8:010d}.vti")
    assert np.array_equal(d["density"], rho)
    root, p = read_vtk(f"{prefix}_particles_{8:010d}.vtp")
    assert root.find("PolyData/Piece").get("NumberOfVerts") == "5"
    pos = p["Points"].reshape(-1, 3)
    box = np.array(s.grid, dtype=float)*s.dx
    assert np.allclose(pos, ctx.positions - np.floor(ctx.positions/box)*box)
    assert np.all((pos >= 0) & (pos < box))
    assert np.array_equal(p["velocity"].reshape(-1, 3), ctx.velocities)
    assert list(p["coupled"]) == [1]*5 and list(p["index"]) == list(range(5))
    assert list(p["connectivity"]) == list(range(5)) and list(p["offsets"]) == list(range(1, 6))
    sets = ET.parse(prefix + ".pvd").getroot().findall("Collection/DataSet")
    assert [(d.get("timestep"), d.get("part")) for d in sets] == [(t, str(k)) for t in ("0.04", "0.08") for k in range(3)]
    # A restart with append=True keeps the listed datasets; single precision by default.
    more = LBMVTKReporter(prefix, 4, ctx.force, append=True, particles=False)
    run(ctx, 4, [more])
    _, f = read_vtk(f"{prefix}_density_{12:010d}.vti")
    assert f["density"].dtype == np.float32
    assert len(ET.parse(prefix + ".pvd").getroot().findall("Collection/DataSet")) == 8


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


def test_temperature_continues_an_appended_file(tmp_path):
    ctx, _ = small_context()
    path = tmp_path/"t.txt"
    run(ctx, 4, [LBMTemperatureReporter(str(path), 2, ctx.force)])
    with open(path, "a") as stream:
        run(ctx, 4, [LBMTemperatureReporter(stream, 2, ctx.force)])
    lines = path.read_text().splitlines()
    assert [line.split()[0] for line in lines] == ['#"Step","Time', "2", "4", "6", "8"]


@pytest.mark.gpu
@pytest.mark.parametrize("backend, tol", [("fp64", 1e-12), ("fp32", 1e-5)], ids=["gpu-fp64", "gpu-fp32"])
def test_reporters_on_the_gpu_library(tmp_path, backend, tol):
    """Both reporters over the gfx950 library write the files that they write over the oracle: the same names, the
    same XML, and arrays and temperatures that agree to 1e-12 (FP64 storage) or 1e-5 (FP32) relative to their
    largest value."""
    files = {}
    for name in ("oracle", backend):
        ctx, _ = small_context(LBMForce.Explicit, name)
        prefix = str(tmp_path/f"{name}_run")
        temperature = str(tmp_path/f"{name}_temperature.txt")
        run(ctx, 8, [LBMVTKReporter(prefix, 4, ctx.force, double=True), LBMTemperatureReporter(temperature, 2, ctx.force)])
        files[name] = (prefix, temperature)
    (p0, t0), (p1, t1) = files["oracle"], files[backend]
    for kind, ext in (("velocity", "vti"), ("density", "vti"), ("particles", "vtp")):
        for step in (4, 8):
            r0, a0 = read_vtk(f"{p0}_{kind}_{step:010d}.{ext}")
            r1, a1 = read_vtk(f"{p1}_{kind}_{step:010d}.{ext}")
            assert sorted(a0) == sorted(a1)
            assert ET.tostring(r0.find(".//FieldData")) == ET.tostring(r1.find(".//FieldData"))
            for name in a0:
                x0, x1 = a0[name].astype(np.float64), a1[name].astype(np.float64)
                scale = max(np.abs(x0).max(), 1e-300)
                assert np.abs(x1 - x0).max() <= tol*scale, (kind, step, name)
    rows0 = [line.split() for line in open(t0).read().splitlines()[1:]]
    rows1 = [line.split() for line in open(t1).read().splitlines()[1:]]
    assert [r[:2] for r in rows0] == [r[:2] for r in rows1]
    for a, b in zip(rows0, rows1):
        assert abs(float(b[2]) - float(a[2])) <= tol*max(abs(float(a[2])), 1e-3)
