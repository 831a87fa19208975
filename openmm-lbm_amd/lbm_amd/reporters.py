"""Reporters of a run, for this package's Context: the counterparts of the reference's LBMTemperatureReporter and
LBMVTKReporter (python/openmmlbm.i:619-860), which serve openmm.app.Simulation.

A reporter has describeNextReport(context) -> steps until its next report, and report(context); run(context, steps,
reporters) steps a Context and calls them when due.

* LBMTemperatureReporter: the temperature of the coupled particles, from the velocity that has the right temperature
  for the drag scheme: the full-step velocity v + dt F/(2m) with the explicit drag, the state's velocity with the
  centred drag; three degrees of freedom per coupled particle; lines "step time temperature" after a header written
  only at the start of a file.
* LBMVTKReporter: VTK XML files (ParaView, VisIt) in OpenMM units: <prefix>_density_<step>.vti and
  <prefix>_velocity_<step>.vti (ImageData, point (i, j, k) at (i, j, k) dx, with a "solid" array when the lattice has
  solid nodes), <prefix>_particles_<step>.vtp (PolyData: positions wrapped into the box unless wrap=False,
  "velocity", "mass", "index", "coupled"), and <prefix>.pvd listing them with their times.  Arrays are appended in
  binary (raw, little endian, each after its UInt64 byte count), single precision unless double=True.
"""
import os
import struct

import numpy as np

from .force import LBMForce
from .workloads import BOLTZ


def run(context, steps, reporters):
    """Advance context by steps, calling each reporter whenever its interval divides the step count."""
    done = 0
    while done < steps:
        todo = min([r.describeNextReport(context) for r in reporters] + [steps - done])
        context.step(todo)
        done += todo
        for r in reporters:
            if r.describeNextReport(context) == r.interval:
                r.report(context)


class LBMTemperatureReporter:
    def __init__(self, file, reportInterval, force):
        self.interval = int(reportInterval)
        self.force = force
        self._opened = isinstance(file, str)
        self._out = open(file, "a" if self._opened and os.path.exists(file) else "w") if self._opened else file
        self._particles = None

    def __del__(self):
        if getattr(self, "_opened", False):
            self._out.close()

    def describeNextReport(self, context):
        return self.interval - context.getStepCount() % self.interval

    def report(self, context):
        if self._particles is None:
            self._particles = [self.force.getParticle(i) for i in range(self.force.getNumParticles())]
            self._masses = np.array([context.system.masses[i] for i in self._particles])
            try:
                new = self._out.tell() == 0
            except (AttributeError, OSError, ValueError):
                new = True
            if new:
                print('#"Step","Time (ps)","Temperature (K)"', file=self._out)
        v = np.asarray(context.velocities)[self._particles]
        if self.force.getDragScheme() != LBMForce.Centered:
            f = np.asarray(context.getForces())[self._particles]
            v = v + f*(0.5*context.integrator.stepSize/self._masses[:, None])
        energy = 0.5*np.sum(self._masses[:, None]*v*v)
        temperature = 2*energy/(3*len(self._particles)*BOLTZ)
        step = context.getStepCount()
        print("%d %.4f %.3f" % (step, step*context.integrator.stepSize, temperature), file=self._out)
        self._out.flush()


def _units_field(lines):
    """Field data with the string array "units" (one "array: unit" line each), ASCII: character codes, each string
    ended by 0."""
    codes = " ".join(" ".join(str(ord(c)) for c in line) + " 0" for line in lines)
    return ('    <FieldData>\n'
            f'      <DataArray type="String" Name="units" NumberOfTuples="{len(lines)}" format="ascii">\n'
            f'        {codes}\n'
            '      </DataArray>\n'
            '    </FieldData>\n')


def _write_vtk(path, kind, opening, closing, arrays):
    """A VTK XML file whose data arrays are appended in binary; opening holds {offset<i>} placeholders."""
    blocks = [np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes() for a in arrays]
    offsets, pos = {}, 0
    for i, b in enumerate(blocks):
        offsets[f"offset{i}"] = pos
        pos += 8 + len(b)
    head = ('<?xml version="1.0"?>\n'
            f'<VTKFile type="{kind}" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n' +
            opening.format(**offsets) + closing + '  <AppendedData encoding="raw">\n_')
    with open(path, "wb") as out:
        out.write(head.encode("ascii"))
        for b in blocks:
            out.write(struct.pack("<Q", len(b)))
            out.write(b)
        out.write(b"\n  </AppendedData>\n</VTKFile>\n")


class LBMVTKReporter:
    def __init__(self, prefix, reportInterval, force, fluid=True, particles=True, double=False, wrap=True, append=False,
                 density=True, velocity=True):
        self.prefix = prefix
        self.interval = int(reportInterval)
        self.force = force
        self._density = fluid and density
        self._velocity = fluid and velocity
        self._particles = particles
        self._double = double
        self._wrap = wrap
        self._datasets = []
        if append and os.path.exists(prefix + ".pvd"):
            with open(prefix + ".pvd") as pvd:
                self._datasets = [line for line in pvd if "<DataSet " in line]

    def describeNextReport(self, context):
        return self.interval - context.getStepCount() % self.interval

    def report(self, context):
        real = np.float64 if self._double else np.float32
        tname = "Float64" if self._double else "Float32"
        step = context.getStepCount()
        time = step*context.integrator.stepSize
        files = []
        nx, ny, nz = self.force.getGridSize()
        solid = self.force.getSolidNodes()
        dx = context.getLatticeParameters()[0]
        if self._density or self._velocity:
            rho, u = context.getFluidFields()
        mask = None
        if solid:
            mask = np.zeros(nx*ny*nz, dtype=np.uint8)
            mask[np.asarray(solid)] = 1
        for field, comps, unit, values in (("density", 1, "Da/nm^3", rho if self._density else None),
                                           ("velocity", 3, "nm/ps", u if self._velocity else None)):
            if values is None:
                continue
            extent = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
            units = ["Origin, Spacing: nm", f"{field}: {unit}"]
            arrays = [np.asarray(values, dtype=real).reshape(-1)]
            if mask is not None:
                units.append("solid: 1 for a solid node, 0 for a fluid node")
                arrays.append(mask)
            kind = "Scalars" if comps == 1 else "Vectors"
            ncomp = ' NumberOfComponents="3"' if comps == 3 else ""
            opening = (f'  <ImageData WholeExtent="{extent}" Origin="0 0 0" Spacing="{dx!r} {dx!r} {dx!r}">\n' +
                       _units_field(units) +
                       f'    <Piece Extent="{extent}">\n'
                       f'      <PointData {kind}="{field}">\n'
                       f'        <DataArray type="{tname}" Name="{field}"{ncomp} format="appended" offset="{{offset0}}"/>\n')
            if mask is not None:
                opening += '        <DataArray type="UInt8" Name="solid" format="appended" offset="{offset1}"/>\n'
            closing = "      </PointData>\n    </Piece>\n  </ImageData>\n"
            path = f"{self.prefix}_{field}_{step:010d}.vti"
            _write_vtk(path, "ImageData", opening, closing, arrays)
            files.append(path)
        if self._particles:
            n = context.system.getNumParticles()
            pos = np.asarray(context.positions, dtype=np.float64)
            if self._wrap:
                box = np.asarray(context.system.box, dtype=np.float64)
                pos = pos - np.floor(pos/box)*box
            coupled = np.zeros(n, dtype=np.uint8)
            coupled[[self.force.getParticle(i) for i in range(self.force.getNumParticles())]] = 1
            arrays = [pos.astype(real).reshape(-1), np.asarray(context.velocities, dtype=real).reshape(-1),
                      np.asarray(context.system.masses, dtype=real), np.arange(n, dtype=np.int32), coupled,
                      np.arange(n, dtype=np.int64), np.arange(1, n + 1, dtype=np.int64)]
            units = ["Points: nm", "velocity: nm/ps, at the half step of the leapfrog", "mass: Da",
                     "index: index of the particle in the System", "coupled: 1 for a particle coupled to the fluid"]
            opening = ("  <PolyData>\n" + _units_field(units) +
                       f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}" NumberOfLines="0" NumberOfStrips="0" '
                       'NumberOfPolys="0">\n'
                       '      <Points>\n'
                       f'        <DataArray type="{tname}" NumberOfComponents="3" format="appended" offset="{{offset0}}"/>\n'
                       '      </Points>\n'
                       '      <PointData Vectors="velocity">\n'
                       f'        <DataArray type="{tname}" Name="velocity" NumberOfComponents="3" format="appended" '
                       'offset="{offset1}"/>\n'
                       f'        <DataArray type="{tname}" Name="mass" format="appended" offset="{{offset2}}"/>\n'
                       '        <DataArray type="Int32" Name="index" format="appended" offset="{offset3}"/>\n'
                       '        <DataArray type="UInt8" Name="coupled" format="appended" offset="{offset4}"/>\n'
                       '      </PointData>\n'
                       '      <Verts>\n'
                       '        <DataArray type="Int64" Name="connectivity" format="appended" offset="{offset5}"/>\n'
                       '        <DataArray type="Int64" Name="offsets" format="appended" offset="{offset6}"/>\n'
                       '      </Verts>\n')
            closing = "    </Piece>\n  </PolyData>\n"
            path = f"{self.prefix}_particles_{step:010d}.vtp"
            _write_vtk(path, "PolyData", opening, closing, arrays)
            files.append(path)
        for path in files:
            self._datasets.append(f'    <DataSet timestep="{time!r}" part="0" file="{os.path.basename(path)}"/>\n')
        with open(self.prefix + ".pvd", "w") as pvd:
            pvd.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">\n'
                      "  <Collection>\n" + "".join(self._datasets) + "  </Collection>\n</VTKFile>\n")
