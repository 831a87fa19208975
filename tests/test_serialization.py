"""The reference's formats around the force (CPU): the XML serialization of LBMForce (LBMForceProxy.cpp, the reading
rules of TestSerializeLBMForce.cpp) and the force-level checkpoint header (LBMForceImpl.cpp:502-567)."""
import struct

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMError, LBMForce
from lbm_amd.serialization import CHECKPOINT_TAG, checkpoint_header, deserialize, read_checkpoint_header, serialize
from parity import Setup


def non_default_force():
    """The values of the reference's serialization test (TestSerializeLBMForce.cpp), all different from the
    defaults."""
    f = LBMForce()
    f.setForceGroup(3)
    f.setName("fluid")
    f.setGridSize(10, 12, 14)
    f.setDomainDecomposition(2, 0, 3)
    f.setParticleCopiesCheck(False)
    f.setDensityHaloExchange(True)
    f.setVelocityHaloExchange(True)
    f.setFluidDensity(500.0)
    f.setKinematicViscosity(2.5)
    f.setFriction(5.0)
    f.setTemperature(310.0)
    f.setRandomNumberSeed(123)
    f.setFluidMomentumRemovalFrequency(10)
    f.setMachCheckFrequency(25)
    f.setMachNumberLimit(0.2)
    f.setCouplingScheme(LBMForce.NVE)
    f.setDragScheme(LBMForce.Centered)
    f.setInterpolationStencil(LBMForce.ThreePoint)
    f.setFluidFluctuations(True)
    f.setWallScheme(LBMForce.Regularized)
    f.setFaceBoundary(LBMForce.YMin, LBMForce.Density)
    f.setFaceBoundary(LBMForce.YMax, LBMForce.Velocity)
    f.setFaceDensity(LBMForce.YMin, 610.0)
    f.setFaceVelocity(LBMForce.YMax, (0.1, -0.2, 0.3))
    f.setFaceBoundary(LBMForce.ZMin, LBMForce.DensityVelocity)
    f.setFaceBoundary(LBMForce.ZMax, LBMForce.DensityVelocity)
    f.setFaceDensity(LBMForce.ZMin, 590.0)
    f.setFaceVelocity(LBMForce.ZMin, (-0.2, 0.1, 0.4))
    f.setSolidNodes([0, 7, 42])
    f.setBodyAcceleration((0.1, 0.2, 0.3))
    f.setInitialFluidVelocity((-0.1, 0.0, 0.05))
    for p in (3, 7, 1):
        f.addParticle(p)
    return f


GETTERS = ["getForceGroup", "getName", "getGridSize", "getDomainDecomposition", "getParticleCopiesCheck",
           "getDensityHaloExchange", "getVelocityHaloExchange", "getFluidDensity", "getKinematicViscosity",
           "getFriction", "getTemperature", "getRandomNumberSeed", "getFluidMomentumRemovalFrequency",
           "getMachCheckFrequency", "getMachNumberLimit", "getCouplingScheme", "getDragScheme",
           "getInterpolationStencil", "getFluidFluctuations", "getWallScheme", "getSolidNodes", "getBodyAcceleration",
           "getInitialFluidVelocity", "getNumParticles"]


def test_xml_round_trip():
    f = non_default_force()
    text = serialize(f)
    assert 'type="LBMForce"' in text and 'version="9"' in text
    g = deserialize(text)
    for getter in GETTERS:
        assert getattr(f, getter)() == getattr(g, getter)(), getter
    for face in range(6):
        assert f.getFaceBoundary(face) == g.getFaceBoundary(face)
        assert f.getFaceVelocity(face) == g.getFaceVelocity(face)
        assert f.getFaceDensity(face) == g.getFaceDensity(face)
    assert [f.getParticle(i) for i in range(3)] == [g.getParticle(i) for i in range(3)]


def older(text, version, remove=()):
    for attr in remove:
        start = text.index(f' {attr}="')
        end = text.index('"', text.index('="', start) + 2) + 1
        text = text[:start] + text[end:]
    return text.replace('version="9"', f'version="{version}"')


def test_older_versions_take_the_former_defaults():
    text = serialize(non_default_force())
    assert deserialize(older(text, 8, ["interpolationStencil"])).getInterpolationStencil() == LBMForce.NearestNode
    v6 = deserialize(older(text, 6, ["interpolationStencil"]))
    assert all(v6.getFaceBoundary(face) == LBMForce.Periodic for face in range(6))
    assert v6.getDomainDecomposition() == (1, 1, 1)
    assert deserialize(older(text, 5, ["interpolationStencil"])).getWallScheme() == LBMForce.BounceBack
    assert not deserialize(older(text, 4, ["interpolationStencil"])).getFluidFluctuations()
    assert deserialize(older(text, 3, ["interpolationStencil"])).getDragScheme() == LBMForce.Explicit
    with pytest.raises(LBMError, match="Unsupported version"):
        deserialize(text.replace('version="9"', 'version="10"'))


def test_checkpoint_header_layout_and_checks():
    f = LBMForce()
    f.setGridSize(8, 8, 8)
    f.setDragScheme(LBMForce.Centered)
    f.setInterpolationStencil(LBMForce.Keys)
    h = checkpoint_header(f, 5)
    # tag, version 6, platform name, 14 ints
    assert h[:8] == CHECKPOINT_TAG and struct.unpack_from("<ii", h, 8) == (6, 3) and h[16:19] == b"HIP"
    assert list(struct.unpack_from("<14i", h, 19)) == [8, 8, 8, 5, 1, 0, 0, 0, 0, 0, 0, 0, 0, 3]
    assert read_checkpoint_header(h + b"body", f, 5) == len(h)
    other = LBMForce()
    other.setGridSize(8, 8, 8)
    other.setInterpolationStencil(LBMForce.Keys)
    with pytest.raises(LBMError, match="different drag scheme"):
        read_checkpoint_header(h, other, 5)
    with pytest.raises(LBMError, match="grid size or number of coupled particles"):
        read_checkpoint_header(h, f, 4)
    with pytest.raises(LBMError, match="not a checkpoint"):
        read_checkpoint_header(b"XXXXXXXX" + h[8:], f, 5)
    with pytest.raises(LBMError, match="written on the platform CUDA"):
        read_checkpoint_header(checkpoint_header(f, 5, platform="CUDA"), f, 5)
    # A version-1 header (grid and particles only) reads with the defaults of the later settings.
    v1 = CHECKPOINT_TAG + struct.pack("<ii", 1, 3) + b"HIP" + struct.pack("<4i", 8, 8, 8, 5)
    plain = LBMForce()
    plain.setGridSize(8, 8, 8)
    assert read_checkpoint_header(v1, plain, 5) == len(v1)


def test_context_checkpoint_round_trip_with_header():
    s = Setup((6, 6, 6), tau=0.8)
    s.add_particles([[1.0, 2.0, 3.0]], mass=100.0)
    s.configure.append(lambda f: (f.setFriction(5.0), f.setTemperature(0.0)))
    ctx = s.build(oracle_lib.load_oracle())
    ctx.step(3)
    data = ctx.createCheckpoint()
    state = ctx.getFluidState()
    ctx.step(2)
    ctx.loadCheckpoint(data)
    assert np.array_equal(ctx.getFluidState(), state)
