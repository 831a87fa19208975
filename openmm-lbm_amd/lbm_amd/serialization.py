"""XML serialization of LBMForce and the force-level checkpoint header, in the reference's formats.

* serialize / deserialize: the LBMForce element of OpenMM's XmlSerializer as the reference's LBMForceProxy writes it
  (serialization/src/LBMForceProxy.cpp:22-147): version 9, the same property and child names, and the reading rules
  of versions 1-8 (missing settings take the values they had before they existed).
* checkpoint_header / read_checkpoint_header: the header that LBMForceImpl puts before the kernel's checkpoint
  (openmmapi/src/LBMForceImpl.cpp:502-567): tag "LBMCKPT1", format version 6, the platform name, and 14 ints (grid,
  coupled particles, drag scheme, fluctuations, wall scheme, six face types, stencil), with the same checks and
  messages on reading.
"""
import struct
import xml.etree.ElementTree as ET

from . import _lib
from .force import LBMForce

VERSION = 9
CHECKPOINT_TAG = b"LBMCKPT1"
CHECKPOINT_VERSION = 6
PLATFORM = "HIP"


def _num(x):
    return repr(float(x))


def serialize(force):
    """The XML text of a force (the <Force type="LBMForce" version="9" .../> element)."""
    nx, ny, nz = force.getGridSize()
    px, py, pz = force.getDomainDecomposition()
    props = {
        "version": str(VERSION), "forceGroup": str(force.getForceGroup()), "name": force.getName(),
        "nx": str(nx), "ny": str(ny), "nz": str(nz), "domainsX": str(px), "domainsY": str(py), "domainsZ": str(pz),
        "particleCopiesCheck": str(int(force.getParticleCopiesCheck())),
        "densityHaloExchange": str(int(force.getDensityHaloExchange())),
        "velocityHaloExchange": str(int(force.getVelocityHaloExchange())),
        "density": _num(force.getFluidDensity()), "viscosity": _num(force.getKinematicViscosity()),
        "friction": _num(force.getFriction()), "temperature": _num(force.getTemperature()),
        "randomSeed": str(force.getRandomNumberSeed()),
        "momentumRemovalFrequency": str(force.getFluidMomentumRemovalFrequency()),
        "machCheckFrequency": str(force.getMachCheckFrequency()),
        "couplingScheme": str(force.getCouplingScheme()), "dragScheme": str(force.getDragScheme()),
        "interpolationStencil": str(force.getInterpolationStencil()),
        "fluidFluctuations": str(int(force.getFluidFluctuations())), "wallScheme": str(force.getWallScheme()),
        "machNumberLimit": _num(force.getMachNumberLimit()), "type": "LBMForce",
    }
    root = ET.Element("Force", dict(sorted(props.items())))
    g = force.getBodyAcceleration()
    ET.SubElement(root, "BodyAcceleration", {"x": _num(g[0]), "y": _num(g[1]), "z": _num(g[2])})
    u = force.getInitialFluidVelocity()
    ET.SubElement(root, "InitialFluidVelocity", {"x": _num(u[0]), "y": _num(u[1]), "z": _num(u[2])})
    particles = ET.SubElement(root, "Particles")
    for i in range(force.getNumParticles()):
        ET.SubElement(particles, "Particle", {"index": str(force.getParticle(i))})
    solid = ET.SubElement(root, "SolidNodes")
    for index in force.getSolidNodes():
        ET.SubElement(solid, "Node", {"index": str(index)})
    faces = ET.SubElement(root, "Faces")
    for face in range(6):
        v = force.getFaceVelocity(face)
        ET.SubElement(faces, "Face", {"density": _num(force.getFaceDensity(face)), "type": str(force.getFaceBoundary(face)),
                                      "vx": _num(v[0]), "vy": _num(v[1]), "vz": _num(v[2])})
    ET.indent(root, space="\t")
    return '<?xml version="1.0" ?>\n' + ET.tostring(root, encoding="unicode") + "\n"


def _bool(text):
    return text not in ("0", "false", "False")


def deserialize(text):
    """An LBMForce from the XML of serialize() or of the reference's XmlSerializer, versions 1-9."""
    root = ET.fromstring(text)
    a = root.attrib
    version = int(a["version"])
    if version < 1 or version > VERSION:
        raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "Unsupported version number")
    f = LBMForce()
    f.setForceGroup(int(a.get("forceGroup", 0)))
    f.setName(a.get("name", f.getName()))
    f.setGridSize(int(a["nx"]), int(a["ny"]), int(a["nz"]))
    f.setFluidDensity(float(a["density"]))
    f.setKinematicViscosity(float(a["viscosity"]))
    f.setFriction(float(a["friction"]))
    f.setTemperature(float(a["temperature"]))
    f.setRandomNumberSeed(int(a["randomSeed"]))
    f.setFluidMomentumRemovalFrequency(int(a["momentumRemovalFrequency"]))
    if version >= 2:
        f.setMachCheckFrequency(int(a["machCheckFrequency"]))
        f.setMachNumberLimit(float(a["machNumberLimit"]))
    if version >= 3:
        f.setCouplingScheme(int(a["couplingScheme"]))
    if version >= 4:                      # before: the explicit drag
        f.setDragScheme(int(a["dragScheme"]))
    if version >= 9:                      # before: the nearest node
        f.setInterpolationStencil(int(a["interpolationStencil"]))
    if version >= 5:                      # before: no fluctuations
        f.setFluidFluctuations(_bool(a["fluidFluctuations"]))
    if version >= 8:                      # before: one domain
        f.setDomainDecomposition(int(a["domainsX"]), int(a["domainsY"]), int(a["domainsZ"]))
        f.setParticleCopiesCheck(_bool(a.get("particleCopiesCheck", "1")))
        f.setDensityHaloExchange(_bool(a.get("densityHaloExchange", "0")))
        f.setVelocityHaloExchange(_bool(a.get("velocityHaloExchange", "0")))
    if version >= 6:                      # before: bounce-back
        f.setWallScheme(int(a["wallScheme"]))
    g = root.find("BodyAcceleration").attrib
    f.setBodyAcceleration((float(g["x"]), float(g["y"]), float(g["z"])))
    u = root.find("InitialFluidVelocity").attrib
    f.setInitialFluidVelocity((float(u["x"]), float(u["y"]), float(u["z"])))
    for p in root.find("Particles"):
        f.addParticle(int(p.attrib["index"]))
    if version >= 2:
        f.setSolidNodes([int(n.attrib["index"]) for n in root.find("SolidNodes")])
    if version >= 7:                      # before: a periodic box
        faces = list(root.find("Faces"))
        if len(faces) != 6:
            raise _lib.LBMError(_lib.LBM_ERROR_INVALID, "LBMForce: the serialized force must have 6 faces")
        for face, node in enumerate(faces):
            b = node.attrib
            f.setFaceBoundary(face, int(b["type"]))
            f.setFaceVelocity(face, (float(b["vx"]), float(b["vy"]), float(b["vz"])))
            f.setFaceDensity(face, float(b["density"]))
    return f


def checkpoint_header(force, num_particles, platform=PLATFORM):
    """LBMForceImpl::writeCheckpointHeader (LBMForceImpl.cpp:516-529), little-endian ints."""
    nx, ny, nz = force.getGridSize()
    name = platform.encode()
    ints = [nx, ny, nz, num_particles, force.getDragScheme(), int(force.getFluidFluctuations()), force.getWallScheme()]
    ints += [force.getFaceBoundary(face) for face in range(6)] + [force.getInterpolationStencil()]
    return CHECKPOINT_TAG + struct.pack("<ii", CHECKPOINT_VERSION, len(name)) + name + struct.pack("<14i", *ints)


def read_checkpoint_header(data, force, num_particles, platform=PLATFORM):
    """LBMForceImpl::readCheckpointHeader (LBMForceImpl.cpp:531-567): checks the header against the force and returns
    the offset of the kernel's data."""
    def fail(msg):
        raise _lib.LBMError(_lib.LBM_ERROR_INVALID, msg)
    if len(data) < 16 or data[:8] != CHECKPOINT_TAG:
        fail("LBMForce: the data are not a checkpoint written by LBMForce::createCheckpoint()")
    version, length = struct.unpack_from("<ii", data, 8)
    if version < 1 or version > CHECKPOINT_VERSION:
        fail("LBMForce: unsupported checkpoint version")
    if length < 0 or length > 1000 or len(data) < 16 + length:
        fail("LBMForce: the checkpoint is damaged")
    name = data[16:16 + length].decode(errors="replace")
    if name != platform:
        fail(f"LBMForce: the checkpoint was written on the platform {name}, not on {platform}")
    header = [0, 0, 0, 0, LBMForce.Explicit, 0, LBMForce.BounceBack] + [LBMForce.Periodic]*6 + [LBMForce.NearestNode]
    count = version + 3 if version <= 4 else (13 if version == 5 else 14)
    offset = 16 + length
    if len(data) < offset + 4*count:
        fail("LBMForce: the checkpoint is truncated")
    header[:count] = struct.unpack_from(f"<{count}i", data, offset)
    nx, ny, nz = force.getGridSize()
    if header[:4] != [nx, ny, nz, num_particles]:
        fail("LBMForce: the checkpoint was written for a different grid size or number of coupled particles")
    if header[4] != force.getDragScheme():
        fail("LBMForce: the checkpoint was written with a different drag scheme")
    if header[5] != int(force.getFluidFluctuations()):
        fail("LBMForce: the checkpoint was written " + ("with" if header[5] else "without") +
             " fluid fluctuations, and this Context has them " + ("on" if force.getFluidFluctuations() else "off"))
    if header[6] != force.getWallScheme():
        fail("LBMForce: the checkpoint was written with a different wall scheme")
    if header[7:13] != [force.getFaceBoundary(face) for face in range(6)]:
        fail("LBMForce: the checkpoint was written with different boundary types of the faces (setFaceBoundary())")
    if header[13] != force.getInterpolationStencil():
        fail("LBMForce: the checkpoint was written with a different interpolation stencil")
    return offset + 4*count
