"""Run reporters for this package's Context, counterparts of the reference's LBMTemperatureReporter and
LBMVTKReporter (python/openmmlbm.i:619-872), which report from an openmm.app.Simulation.

A reporter offers describeNextReport(context), the number of steps until its next report, and report(context);
run(context, steps, reporters) advances the Context and calls each reporter when its interval divides the step count
(the part Simulation.step plays for the reference's reporters).

Temperature files hold the header '#"Step","Time (ps)","Temperature (K)"' (only when the file starts empty) and one
line "step time temperature" per report.  The temperature counts three degrees of freedom per coupled particle and
uses the velocity whose temperature is right for the drag scheme: with the explicit drag the full-step velocity
v + dt F/(2m), with the centred drag the velocity of the state.

VTK files (ParaView, VisIt) are written in OpenMM units, each array stored after the XML in an appended raw section
(little endian, every array preceded by its UInt64 byte count), FP32 unless double=True:
  <prefix>_density_<step>.vti, <prefix>_velocity_<step>.vti  ImageData, node (i, j, k) at (i, j, k) dx, plus an
                                                              array "solid" when the lattice has solid nodes;
  <prefix>_particles_<step>.vtp                               PolyData: positions (wrapped into the periodic box
                                                              unless wrap=False), "velocity", "mass", "index",
                                                              "coupled", one vertex per particle;
  <prefix>.pvd                                                the collection of every file written, with its time.
Each dataset carries a field-data string array "units" naming the unit of each of its arrays.
"""
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

from .force import LBMForce
from .workloads import BOLTZ


def run(context, steps, reporters):
    """Advance context by steps, calling each reporter at every step count that its interval divides."""
    target = context.getStepCount() + steps
    while context.getStepCount() < target:
        stride = min([r.describeNextReport(context) for r in reporters] + [target - context.getStepCount()])
        context.step(stride)
        due = [r for r in reporters if context.getStepCount() % r.interval == 0]
        for r in due:
            r.report(context)


def _steps_to_next(interval, context):
    return interval - context.getStepCount() % interval


class LBMTemperatureReporter:
    """file: a path (created or truncated) or an open text stream, such as sys.stdout or a file opened for appending
    (a restart continues it without a second header)."""

    def __init__(self, file, reportInterval, force):
        self.interval = int(reportInterval)
        self.force = force
        self._owns = isinstance(file, str)
        self._stream = open(file, "w") if self._owns else file
        self._coupled = None

    def __del__(self):
        if getattr(self, "_owns", False) and not self._stream.closed:
            self._stream.close()

    def describeNextReport(self, context):
        return _steps_to_next(self.interval, context)

    def _starts_empty(self):
        try:
            return self._stream.tell() == 0
        except (AttributeError, OSError, ValueError):
            return True          # a stream without a position

    def report(self, context):
        if self._coupled is None:
            self._coupled = np.array([self.force.getParticle(i) for i in range(self.force.getNumParticles())],
                                     dtype=np.int64)
            self._mass = np.asarray(context.masses, dtype=np.float64)[self._coupled]
            if self._starts_empty():
                self._stream.write('#"Step","Time (ps)","Temperature (K)"\n')
        dt = context.integrator.stepSize
        v = np.asarray(context.velocities, dtype=np.float64)[self._coupled]
        if self.force.getDragScheme() != LBMForce.Centered:
            # Leapfrog velocities sit half a step before the forces of this step.
            v = v + np.asarray(context.getForces())[self._coupled]*(dt/(2.0*self._mass))[:, None]
        twice_kinetic = float(np.einsum("i,ij,ij->", self._mass, v, v))
        temperature = twice_kinetic/(3.0*len(self._coupled)*BOLTZ)
        step = context.getStepCount()
        self._stream.write("%d %.4f %.3f\n" % (step, step*dt, temperature))
        self._stream.flush()


class _AppendedDataset:
    """A VTK XML file under construction: an element tree whose DataArray elements point into one appended raw
    section, filled in the order the arrays are added."""

    _TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.uint8): "UInt8",
              np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64"}

    def __init__(self, kind, units):
        self.root = ET.Element("VTKFile", type=kind, version="1.0", byte_order="LittleEndian", header_type="UInt64")
        self.data = ET.SubElement(self.root, kind)
        self._units = units
        self._payload = []
        self._offset = 0

    def array(self, parent, values, name=None, components=1):
        values = np.ascontiguousarray(values)
        attrs = {"type": self._TYPES[values.dtype]}
        if name is not None:
            attrs["Name"] = name
        if components > 1:
            attrs["NumberOfComponents"] = str(components)
        attrs.update(format="appended", offset=str(self._offset))
        ET.SubElement(parent, "DataArray", attrs)
        raw = values.astype(values.dtype.newbyteorder("<"), copy=False).tobytes()
        self._payload.append(struct.pack("<Q", len(raw)) + raw)
        self._offset += len(self._payload[-1])

    def write(self, path):
        # String arrays in ASCII are the character codes of each string followed by 0.
        units = ET.Element("FieldData")
        codes = []
        for line in self._units:
            codes.extend(str(b) for b in line.encode("ascii"))
            codes.append("0")
        strings = ET.SubElement(units, "DataArray", type="String", Name="units",
                                NumberOfTuples=str(len(self._units)), format="ascii")
        strings.text = " ".join(codes)
        self.data.insert(0, units)
        ET.indent(self.root, space="  ")
        xml = ET.tostring(self.root, encoding="unicode")
        body, tail = xml.rsplit("</VTKFile>", 1)
        with open(path, "wb") as out:
            out.write(b'<?xml version="1.0"?>\n')
            out.write(body.encode("ascii"))
            out.write(b'  <AppendedData encoding="raw">\n_')
            for block in self._payload:
                out.write(block)
            out.write(b"\n  </AppendedData>\n</VTKFile>\n")


class LBMVTKReporter:
    """Arguments as the reference's: fluid/density/velocity/particles choose the files, double the precision, wrap
    the folding of positions into the box, append keeps the datasets already listed in <prefix>.pvd (a restart)."""

    def __init__(self, prefix, reportInterval, force, fluid=True, particles=True, double=False, wrap=True, append=False,
                 density=True, velocity=True):
        self.prefix = prefix
        self.interval = int(reportInterval)
        self.force = force
        self._fields = [name for name, on in (("density", fluid and density), ("velocity", fluid and velocity)) if on]
        self._particles = particles
        self._real = np.float64 if double else np.float32
        self._wrap = wrap
        self._collection = []
        if append and os.path.exists(prefix + ".pvd"):
            listed = ET.parse(prefix + ".pvd").getroot().find("Collection")
            self._collection = [dict(d.attrib) for d in listed.iter("DataSet")]

    def describeNextReport(self, context):
        return _steps_to_next(self.interval, context)

    def _fluid_file(self, context, field, values, solid):
        nx, ny, nz = self.force.getGridSize()
        dx = context.getLatticeParameters()[0]
        extent = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
        units = ["Origin, Spacing: nm", "density: Da/nm^3" if field == "density" else "velocity: nm/ps"]
        if solid is not None:
            units.append("solid: 1 for a solid node, 0 for a fluid node")
        ds = _AppendedDataset("ImageData", units)
        ds.data.set("WholeExtent", extent)
        ds.data.set("Origin", "0 0 0")
        ds.data.set("Spacing", f"{dx!r} {dx!r} {dx!r}")
        piece = ET.SubElement(ds.data, "Piece", Extent=extent)
        point = ET.SubElement(piece, "PointData", {"Scalars" if field == "density" else "Vectors": field})
        ds.array(point, np.asarray(values, dtype=self._real).reshape(-1), field, 1 if field == "density" else 3)
        if solid is not None:
            ds.array(point, solid, "solid")
        return ds

    def _particle_file(self, context):
        n = context.system.getNumParticles()
        pos = np.asarray(context.positions, dtype=np.float64)
        if self._wrap:
            box = np.asarray(context.system.box, dtype=np.float64)
            pos = pos - box*np.floor(pos/box)
        coupled = np.zeros(n, dtype=np.uint8)
        coupled[[self.force.getParticle(i) for i in range(self.force.getNumParticles())]] = 1
        ds = _AppendedDataset("PolyData", ["Points: nm", "velocity: nm/ps, at the half step of the leapfrog",
                                           "mass: Da", "index: index of the particle in the System",
                                           "coupled: 1 for a particle coupled to the fluid"])
        piece = ET.SubElement(ds.data, "Piece", NumberOfPoints=str(n), NumberOfVerts=str(n), NumberOfLines="0",
                              NumberOfStrips="0", NumberOfPolys="0")
        ds.array(ET.SubElement(piece, "Points"), pos.astype(self._real).reshape(-1), components=3)
        point = ET.SubElement(piece, "PointData", Vectors="velocity")
        ds.array(point, np.asarray(context.velocities, dtype=self._real).reshape(-1), "velocity", 3)
        ds.array(point, np.asarray(context.system.masses, dtype=self._real), "mass")
        ds.array(point, np.arange(n, dtype=np.int32), "index")
        ds.array(point, coupled, "coupled")
        verts = ET.SubElement(piece, "Verts")
        ds.array(verts, np.arange(n, dtype=np.int64), "connectivity")
        ds.array(verts, np.arange(1, n + 1, dtype=np.int64), "offsets")
        return ds

    def report(self, context):
        step = context.getStepCount()
        written = []
        if self._fields:
            rho, u = context.getFluidFields()
            nodes = self.force.getSolidNodes()
            solid = None
            if len(nodes):
                nx, ny, nz = self.force.getGridSize()
                solid = np.zeros(nx*ny*nz, dtype=np.uint8)
                solid[np.asarray(nodes, dtype=np.int64)] = 1
            for field in self._fields:
                path = f"{self.prefix}_{field}_{step:010d}.vti"
                self._fluid_file(context, field, rho if field == "density" else u, solid).write(path)
                written.append(path)
        if self._particles:
            path = f"{self.prefix}_particles_{step:010d}.vtp"
            self._particle_file(context).write(path)
            written.append(path)
        time = step*context.integrator.stepSize
        for part, path in enumerate(written):
            self._collection.append({"timestep": repr(time), "part": str(part), "file": os.path.basename(path)})
        pvd = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
        listed = ET.SubElement(pvd, "Collection")
        for entry in self._collection:
            ET.SubElement(listed, "DataSet", entry)
        ET.indent(pvd, space="  ")
        with open(self.prefix + ".pvd", "w") as out:
            out.write('<?xml version="1.0"?>\n' + ET.tostring(pvd, encoding="unicode") + "\n")
