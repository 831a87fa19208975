"""Parity of the gfx950 library with the FP64 CPU oracle (same setups, same Philox streams), through the C-ABI.

Tolerances: FP32 storage and arithmetic, 1e-5 relative (max |gpu - oracle| / max |oracle|) for density,
velocity and coupled-particle forces at every checked step; FP64 storage 1e-11.
"""
import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from parity import Setup, equilibrium_state, lockstep, rel_err, total_moments

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def shear_wave_setup(n=32, tau=0.8, amplitude=0.01, g=1e-5, dt=0.01, dx=1.0):
    """32^3 periodic D3Q19, u_x = 0.01 sin(2 pi z/32), uniform Guo force 1e-5 (lattice units) along x."""
    s = Setup((n, n, n), dx=dx, dt=dt, tau=tau)
    z = np.arange(n*n*n)//(n*n)
    u = np.zeros((n*n*n, 3))
    u[:, 0] = amplitude*np.sin(2*np.pi*z/n)
    s.state = equilibrium_state(n*n*n, 1.0, u)
    s.configure.append(lambda f: f.setBodyAcceleration((g*dx/(dt*dt), 0, 0)))
    return s


def config1_setup(n=32, num_beads=64, stencil=LBMForce.Trilinear, seed=1234):
    """The CPU-runnable configuration: 32^3 periodic, dx = 1 nm, tau = 0.8, shear wave 0.01 sin(2 pi z/32), Guo force
    1e-5 along x, 64 free beads (100 Da) at seeded uniform positions, v = 0, gamma = 10/ps, trilinear, T = 0."""
    s = shear_wave_setup(n)
    rng = np.random.default_rng(seed)
    s.add_particles(rng.uniform(0, n, size=(num_beads, 3)), mass=100.0)

    def conf(f):
        f.setFriction(10.0)
        f.setTemperature(0.0)
        f.setInterpolationStencil(stencil)
    s.configure.append(conf)
    return s


def test_fluid_periodic_parity():
    worst, ref, gpu = lockstep(shear_wave_setup(), 100, check_every=5, forces=False)
    assert worst["density"] < TOL32 and worst["velocity"] < TOL32, worst


def test_config1_parity_1000_steps():
    """Every step for 1000 steps: fields, bead forces and bead positions agree to 1e-5."""
    worst, ref, gpu = lockstep(config1_setup(), 1000, check_every=1)
    print("config[1] worst relative errors:", worst)
    assert worst["density"] < TOL32, worst
    assert worst["velocity"] < TOL32, worst
    assert worst["force"] < TOL32, worst


@pytest.mark.parametrize("stencil", [LBMForce.NearestNode, LBMForce.Trilinear, LBMForce.ThreePoint, LBMForce.Keys])
def test_stencil_coupling_parity(stencil):
    s = config1_setup(n=16, num_beads=24, stencil=stencil, seed=7)
    worst, ref, gpu = lockstep(s, 60, check_every=3)
    assert max(worst.values()) < TOL32, worst


def test_node_assignment_bit_exact():
    """Particle-to-node assignment (nearestNode, ReferenceLBMKernels.cpp:973-983) compared as integer node indices,
    GPU against the oracle, at ties and near them: positions on half-integer planes, one ulp either side of them,
    outside the box by whole and half periods, negative, on the lattice nodes, and 2000 random ones; then the
    couplings of three steps agree."""
    s = config1_setup(n=8, num_beads=0, stencil=LBMForce.NearestNode)
    pos = [[0.5, 1.5, 2.5], [7.5, 7.4999999999, 0.5000000001], [-0.5, 8.5, 16.5], [3.0, 3.0, 3.0],
           [-8.5, -7.5, 23.5], [1e-300, 7.999999999999999, -1e-12]]
    for h in (0.5, 3.5, 7.5):
        pos.append([np.nextafter(h, 0), np.nextafter(h, 9), h])
    rng = np.random.default_rng(3)
    ties = np.floor(rng.uniform(-20, 28, size=(1000, 3))) + 0.5
    pos = np.concatenate([np.array(pos), ties, rng.uniform(-20, 28, size=(1000, 3))])
    s.add_particles(pos, mass=100.0, velocities=np.ones((len(pos), 3))*0.1)
    ref = s.build(oracle_lib.load_oracle())
    gpu = s.build(None)
    for trial in range(2):
        nodes_ref = ref.getParticleNodes()
        nodes_gpu = gpu.getParticleNodes()
        assert nodes_ref.dtype == np.int32 and len(nodes_ref) == len(pos)
        assert np.array_equal(nodes_gpu, nodes_ref)
        ref.step(1)
        gpu.step(1)
    worst, ref, gpu = lockstep(s, 3, check_every=1)
    assert worst["force"] < TOL32, worst


def test_removal_parity():
    s = shear_wave_setup(n=12)
    s.configure.append(lambda f: f.setFluidMomentumRemovalFrequency(3))
    worst, ref, gpu = lockstep(s, 20, check_every=1, forces=False)
    assert worst["density"] < TOL32 and worst["velocity"] < TOL32, worst


def wall_setup(scheme, g=1e-5, tau=0.7):
    nx, ny, nz = 6, 10, 5
    s = Setup((nx, ny, nz), dx=0.5, dt=0.01, tau=tau)
    solid = [i + nx*(j + ny*k) for k in range(nz) for j in range(ny) for i in range(nx)
             if j == 0 or (2 <= i < 4 and 3 <= j < 5 and 1 <= k < 3)]

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(scheme)
        f.setInitialFluidVelocity((0.02*0.5/0.01, -0.01*0.5/0.01, 0.005*0.5/0.01))
        f.setBodyAcceleration((g*0.5/1e-4, 2*g*0.5/1e-4, -g*0.5/1e-4))
    s.configure.append(conf)
    return s


@pytest.mark.parametrize("scheme", [LBMForce.BounceBack, LBMForce.Regularized])
def test_wall_parity(scheme):
    def wall(step, ref, gpu):
        w0, w1 = ref.getWallForce(), gpu.getWallForce()
        assert rel_err(w1, w0) < 1e-5, (step, w0, w1)
    worst, ref, gpu = lockstep(wall_setup(scheme), 40, check_every=1, forces=False, callback=wall)
    assert worst["density"] < TOL32 and worst["velocity"] < TOL32, worst


@pytest.mark.parametrize("faces", ["velocity", "density", "density_velocity", "mixed"])
def test_face_parity(faces):
    s = Setup((5, 7, 6), dx=0.5, dt=0.01, tau=0.8)
    U = 0.01*0.5/0.01

    def conf(f):
        if faces == "velocity":
            f.setFaceBoundary(LBMForce.ZMin, LBMForce.Velocity)
            f.setFaceBoundary(LBMForce.ZMax, LBMForce.Velocity)
            f.setFaceVelocity(LBMForce.ZMax, (U, 0.3*U, 0))
        elif faces == "density":
            f.setFaceBoundary(LBMForce.YMin, LBMForce.Velocity)
            f.setFaceVelocity(LBMForce.YMin, (0, U, 0))
            f.setFaceBoundary(LBMForce.YMax, LBMForce.Density)
        elif faces == "density_velocity":
            f.setFaceBoundary(LBMForce.XMin, LBMForce.DensityVelocity)
            f.setFaceBoundary(LBMForce.XMax, LBMForce.DensityVelocity)
            f.setFaceVelocity(LBMForce.XMin, (U, 0, 0))
            f.setFaceDensity(LBMForce.XMax, 602.214*1.001)
        else:
            for face in range(6):
                f.setFaceBoundary(face, [LBMForce.Velocity, LBMForce.Density, LBMForce.DensityVelocity][face % 3])
            f.setFaceVelocity(LBMForce.XMin, (U, 0, 0))
            f.setFaceDensity(LBMForce.YMin, 602.214*0.999)
    s.configure.append(conf)
    worst, ref, gpu = lockstep(s, 30, check_every=1, forces=False)
    assert worst["density"] < TOL32 and worst["velocity"] < TOL32, worst


def test_regularized_walls_with_faces_parity():
    nx, ny, nz = 6, 8, 7
    s = Setup((nx, ny, nz), dx=0.5, dt=0.01, tau=0.9)
    solid = [i + nx*(j + ny*k) for k in range(nz) for j in range(ny) for i in range(nx) if j == 0]

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(LBMForce.Regularized)
        f.setFaceBoundary(LBMForce.ZMin, LBMForce.Velocity)
        f.setFaceBoundary(LBMForce.ZMax, LBMForce.Density)
        f.setFaceVelocity(LBMForce.ZMin, (0, 0, 0.5))
    s.configure.append(conf)
    worst, ref, gpu = lockstep(s, 25, check_every=1, forces=False)
    assert worst["density"] < TOL32 and worst["velocity"] < TOL32, worst


def test_fluctuations_same_stream_parity():
    """With the same Philox stream the fluctuating fluid follows the oracle's trajectory (FP32 normals from hardware
    transcendentals agree with the FP64 ones to ~1e-6 relative)."""
    s = shear_wave_setup(n=10, amplitude=0.0, g=0.0)

    def conf(f):
        f.setFluidFluctuations(True)
        f.setTemperature(300.0)
        f.setRandomNumberSeed(4321)
    s.configure.append(conf)
    s.dx = 2.0
    worst, ref, gpu = lockstep(s, 20, check_every=1, forces=False)
    assert worst["velocity"] < 1e-4 and worst["density"] < 1e-5, worst


def test_fluctuating_walls_and_particles_parity():
    nx, ny, nz = 8, 8, 8
    s = Setup((nx, ny, nz), dx=2.0, dt=0.01, tau=0.8)
    solid = [i + nx*(j + ny*k) for k in range(nz) for j in range(ny) for i in range(nx) if k == 0]
    rng = np.random.default_rng(3)
    s.add_particles(rng.uniform([0, 0, 3], [16, 16, 13], size=(10, 3)), mass=200.0)

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(LBMForce.Regularized)
        f.setFluidFluctuations(True)
        f.setTemperature(300.0)
        f.setRandomNumberSeed(99)
        f.setFriction(5.0)
        f.setInterpolationStencil(LBMForce.ThreePoint)
    s.configure.append(conf)
    worst, ref, gpu = lockstep(s, 15, check_every=1)
    assert worst["velocity"] < 1e-4 and worst["density"] < 1e-5 and worst["force"] < 1e-4, worst


@pytest.mark.parametrize("stencil", [LBMForce.Trilinear, LBMForce.Keys])
def test_wall_force_with_particles_at_solid_nodes(stencil):
    """Particles next to bounce-back walls, whose stencils reach solid nodes: the reaction on the solid slots leaves
    the fluid and goes to the wall (ReferenceLBMKernels.cpp:959-971), so the wall force of every step includes it."""
    nx, ny, nz = 8, 8, 10
    s = Setup((nx, ny, nz), dx=1.0, dt=0.01, tau=0.8)
    solid = [i + nx*(j + ny*k) for k in (0, nz - 1) for j in range(ny) for i in range(nx)]
    rng = np.random.default_rng(17)
    pos = rng.uniform([0, 0, 0.6], [8, 8, 2.0], size=(12, 3))
    pos[6:, 2] = rng.uniform(7.0, 8.4, size=6)
    s.add_particles(pos, mass=100.0, velocities=rng.normal(scale=0.5, size=(12, 3)))

    def conf(f):
        f.setSolidNodes(solid)
        f.setWallScheme(LBMForce.BounceBack)
        f.setFriction(10.0)
        f.setTemperature(0.0)
        f.setInterpolationStencil(stencil)
        f.setBodyAcceleration((0.5, 0, 0))
    s.configure.append(conf)

    def wall(step, ref, gpu):
        w0, w1 = ref.getWallForce(), gpu.getWallForce()
        assert rel_err(w1, w0) < 1e-5, (step, w0, w1)
    worst, ref, gpu = lockstep(s, 20, check_every=1, callback=wall)
    assert worst["velocity"] < TOL32 and worst["force"] < TOL32, worst


def test_fp64_storage_parity():
    s = config1_setup(n=12, num_beads=10, stencil=LBMForce.Keys)
    worst, ref, gpu = lockstep(s, 30, check_every=1, storage=LBMForce.FP64)
    assert max(worst.values()) < 1e-11, worst


def test_state_round_trip_and_checkpoint():
    s = config1_setup(n=10, num_beads=5)
    gpu = s.build()
    gpu.step(5)
    state = gpu.getFluidState()
    gpu.setFluidState(state)
    assert np.array_equal(gpu.getFluidState(), state)
    ck = gpu.createCheckpoint()
    gpu.step(3)
    after = gpu.getFluidState()
    gpu.loadCheckpoint(ck)
    assert np.array_equal(gpu.getFluidState(), state)
    gpu.setStepCount(5)
    gpu.positions, gpu.velocities = gpu.positions, gpu.velocities


def test_checkpoint_portable_between_oracle_and_gpu():
    import oracle_lib
    s = shear_wave_setup(n=8)
    ref = s.build(oracle_lib.load_oracle())
    ref.step(4)
    gpu = s.build()
    gpu.loadCheckpoint(ref.createCheckpoint())
    for c in (ref, gpu):
        c.setStepCount(4)
        c.step(4)
    r0, u0 = ref.getFluidFields()
    r1, u1 = gpu.getFluidFields()
    assert rel_err(u1, u0) < TOL32 and rel_err(r1, r0) < TOL32


def test_queries_do_not_advance_fluid():
    s = config1_setup(n=8, num_beads=4)
    a = s.build()
    b = s.build()
    a.step(3)
    b.step(1)
    b.getForces()
    b.getFluidMachNumber()
    b.step(2)
    assert np.array_equal(a.getFluidState(), b.getFluidState())


def test_mach_check_raises():
    from lbm_amd import LBMError
    s = shear_wave_setup(n=8, amplitude=0.0, g=0.0)
    n = 8**3
    u = np.zeros((n, 3))
    u[:, 0] = 0.25
    s.state = equilibrium_state(n, 1.0, u)
    s.configure.append(lambda f: f.setMachCheckFrequency(2))
    gpu = s.build()
    gpu.step(1)
    with pytest.raises(LBMError, match="Mach number"):
        gpu.step(1)


def test_deterministic_runs():
    s = config1_setup(n=16, num_beads=40, stencil=LBMForce.Keys)
    a = s.build()
    b = s.build()
    a.step(20)
    b.step(20)
    assert np.array_equal(a.getFluidState(), b.getFluidState())
    assert np.array_equal(a.positions, b.positions)


def test_fluid_totals_fp64():
    s = shear_wave_setup(n=16, g=0.0)
    gpu = s.build()
    m0, p0 = gpu.getFluidTotals()
    gpu.step(50)
    m1, p1 = gpu.getFluidTotals()
    mass_ref, p_ref = total_moments(gpu.getFluidState())
    assert abs(m1 - mass_ref) < 1e-9*mass_ref
    assert abs(m1 - m0) < 1e-6*m0


@pytest.mark.parametrize("storage", [LBMForce.FP32, LBMForce.FP16, LBMForce.FP64], ids=["fp32", "fp16", "fp64"])
def test_wide_addressing_is_bit_identical(storage, monkeypatch):
    """The fused pass addresses the populations by 32-bit byte offsets while an array of them fits in 4 GiB and by
    64-bit indices beyond (LBM_AMD_WIDE_ADDRESSING=1 forces the 64-bit path): both paths give the same run bit for bit,
    with regularized walls, open faces, fluctuations and coupled particles."""
    nx, ny, nz = 8, 6, 7
    s = Setup((nx, ny, nz), dx=2.0, dt=0.01, tau=0.8)
    rng = np.random.default_rng(5)
    s.add_particles(rng.uniform([0, 2, 2], [16, 10, 12], size=(12, 3)), mass=200.0)

    def conf(f):
        f.setSolidNodes([i + nx*ny*k for k in range(nz) for i in range(nx)])      # the plane j = 0
        f.setWallScheme(LBMForce.Regularized)
        f.setFaceBoundary(LBMForce.ZMin, LBMForce.Velocity)
        f.setFaceBoundary(LBMForce.ZMax, LBMForce.DensityVelocity)
        f.setFaceVelocity(LBMForce.ZMin, (0.0, 0.0, 0.5))
        f.setFluidFluctuations(True)
        f.setTemperature(300.0)
        f.setRandomNumberSeed(17)
        f.setFriction(5.0)
        f.setInterpolationStencil(LBMForce.Keys)
    s.configure.append(conf)
    runs = []
    for wide in ("0", "1"):
        monkeypatch.setenv("LBM_AMD_WIDE_ADDRESSING", wide)
        c = s.build(None, storage=storage)
        c.step(12)
        runs.append((c.getFluidState(), c.positions.copy(), c.getWallForce()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2])


@pytest.mark.parametrize("storage", [LBMForce.FP32, LBMForce.FP16, LBMForce.BF16, LBMForce.FP64],
                         ids=["fp32", "fp16", "bf16", "fp64"])
@pytest.mark.parametrize("grid", [(256, 3, 4), (128, 5, 3), (128, 3, 3)], ids=["rows256", "rows128", "rows128-tail"])
def test_lds_staged_pass_is_bit_identical(storage, grid, monkeypatch):
    """The fused pass with its pulls staged through LDS in aligned 16-byte chunks (LBM_AMD_FUSED_ROWS=1; rows of 128
    nodes or a multiple of 256) and the pass with direct pulls (=0) give the same run bit for bit: walls along y and
    z with bounce-back, fluctuations, a body force and Keys beads whose stencils wrap across the periodic x faces.
    128 x 3 x 3 leaves a partial last work-group."""
    nx, ny, nz = grid
    s = Setup((nx, ny, nz), dx=1.0, dt=0.01, tau=0.7)
    rng = np.random.default_rng(8)
    pos = rng.uniform([0, 1, 1], [nx, ny - 1, nz - 1], size=(24, 3))
    pos[:4, 0] = [0.2, nx - 0.3, 1.1, nx - 1.6]
    s.add_particles(pos, mass=150.0, velocities=rng.normal(scale=0.5, size=(24, 3)))

    def conf(f):
        f.setSolidNodes([i + nx*ny*(nz - 1) for i in range(nx*ny)] + [i for i in range(nx)])
        f.setWallScheme(LBMForce.BounceBack)
        f.setBodyAcceleration((2e-5/1e-4, 0, 0))
        f.setFluidFluctuations(True)
        f.setTemperature(300.0)
        f.setRandomNumberSeed(23)
        f.setFriction(8.0)
        f.setInterpolationStencil(LBMForce.Keys)
    s.configure.append(conf)
    runs = []
    for rows in ("0", "1"):
        monkeypatch.setenv("LBM_AMD_FUSED_ROWS", rows)
        c = s.build(None, storage=storage)
        c.step(9)
        runs.append((c.getFluidState(), c.positions.copy(), c.getWallForce()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2])


def test_lattice_beyond_4_gib():
    """A 400^3 lattice in FP32 (4.9 GB per population array, so the fused pass takes 64-bit indices): beads near the
    top of the array (element offsets above 2^32 bytes) exchange momentum with the fluid, and the total momentum of
    beads and fluid is conserved over 20 steps, to FP32 storage precision."""
    n = 400
    s = Setup((n, n, n), dx=1.0, dt=0.01, tau=0.8)
    rng = np.random.default_rng(2)
    pos = np.column_stack([rng.uniform(0, n, 8), rng.uniform(0, n, 8), np.linspace(300, 399, 8)])
    vel = rng.normal(scale=1.0, size=(8, 3))
    s.add_particles(pos, mass=100.0, velocities=vel)
    s.configure.append(lambda f: (f.setFriction(10.0), f.setTemperature(0.0),
                                  f.setInterpolationStencil(LBMForce.Keys)))
    c = s.build()
    cell = 602.214

    def momentum():
        mass, j = c.getFluidTotals()
        return (100.0*c.velocities).sum(axis=0) + j*cell*1.0/0.01
    p0 = momentum()
    c.step(20)
    p1 = momentum()
    scale = (100.0*np.linalg.norm(vel, axis=1)).sum()
    moved = (100.0*np.linalg.norm(c.velocities - vel, axis=1)).sum()
    print(f"relative momentum change {np.linalg.norm(p1 - p0)/scale:.2e}, momentum given to the fluid {moved/scale:.2f}")
    assert moved > 0.5*scale
    assert np.linalg.norm(p1 - p0)/scale < 1e-5


def test_config2_full_size_properties():
    """The benchmark workload at its full size (256^3, bounce-back walls at z = 0 and 255, body force, fluctuating
    fluid, 20,000 Keys beads under a wall potential), 2000 device-resident steps: properties that hold at any size.  The
    fluid mass is conserved (walls, fluctuations and coupling move no mass) to FP32 storage rounding, the Mach number
    stays finite and far below the limit, the beads stay between the walls, and the body force has started the flow
    along +x everywhere in the mean."""
    from lbm_amd import workloads
    w = workloads.config2()
    c = w.build()
    dp = w.device_particles(c)
    m0, p0 = c.getFluidTotals()
    dp.step(2000)
    m1, p1 = c.getFluidTotals()
    pos, vel = dp.pull()
    mach = c.getFluidMachNumber()
    print(f"mass drift {abs(m1 - m0)/m0:.2e}, momentum {p1}, Mach {mach:.4f}")
    assert abs(m1 - m0)/m0 < 1e-7
    assert np.isfinite(mach) and mach < 0.1
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    z = pos[:, 2]/w.dx
    assert z.min() > 0.0 and z.max() < 255.5          # none has crossed a wall (positions are not wrapped)
    assert p1[0] > p0[0] and abs(p1[1]) < 0.1*p1[0] and abs(p1[2]) < 0.1*p1[0]


def test_reaction_beyond_fixed_point_range_is_reported():
    """A particle reaction beyond the range of the fixed-point accumulators (a diverged coupling: a 10^6 Da bead at 1
    lattice unit per step with gamma dt = 0.5) is dropped on the device and reported at the next check as
    LBM_ERROR_RUNTIME, instead of wrapping around in the sums."""
    from lbm_amd import LBMError, _lib
    s = Setup((8, 8, 8), dx=1.0, dt=0.01, tau=0.8)
    s.add_particles([[4.2, 3.7, 4.4]], mass=1e6, velocities=[[100.0, 0.0, 0.0]])
    s.configure.append(lambda f: (f.setFriction(50.0), f.setTemperature(0.0), f.setInterpolationStencil(LBMForce.Keys),
                                  f.setMachCheckFrequency(1)))
    c = s.build(None)
    with pytest.raises(LBMError, match="diverged") as err:
        c.step(1)
    assert err.value.status == _lib.LBM_ERROR_RUNTIME


@pytest.mark.parametrize("storage", [LBMForce.FP32, LBMForce.FP16], ids=["fp32", "fp16"])
@pytest.mark.parametrize("stencil", [LBMForce.Keys, LBMForce.ThreePoint, LBMForce.Trilinear, LBMForce.NearestNode],
                         ids=["keys", "threepoint", "trilinear", "nearest"])
@pytest.mark.parametrize("case", ["sparse-walls", "dense-overflow"])
def test_coupling_variants_are_bit_identical(case, stencil, storage, monkeypatch):
    """The ways of coupling the explicit drag give the same run bit for bit: the reactions spread with fixed-point
    atomics (LBM_AMD_SPREAD=atomic) or by the one-writer tile spread (=tiles: beads binned by the tile of their
    stencil's first node, each tile's sums formed in LDS and written by one thread per node), with the node moments
    pulled per stencil slot (LBM_AMD_FIELD=0) or read from a lattice-wide moment field (=1), with each particle's
    stencil formed by every lane of its group (LBM_AMD_STENCIL_SETUP=0) or once per particle (=1).  Fields, positions,
    velocities, the wall force and a force evaluation between steps are compared.  Sparse beads between bounce-back
    walls with fluctuations; and dense beads crowded into a few tiles (bins overflow, and those beads spread
    themselves with the atomics in the same step) with the momentum removal every second step."""
    nx, ny, nz = 64, 8, 12
    s = Setup((nx, ny, nz), dx=1.0, dt=0.01, tau=0.7)
    rng = np.random.default_rng(17)
    if case == "sparse-walls":
        pos = rng.uniform([0, 0, 1.5], [nx, ny, nz - 2.5], size=(60, 3))
        pos[:6, 0] = [0.1, nx - 0.2, 31.9, 32.1, 63.5, 0.5]          # stencils across tile and periodic boundaries
    else:
        pos = np.concatenate([rng.normal([20.0, 3.0, 5.0], 1.2, size=(900, 3)),
                              rng.uniform([0, 0, 0], [nx, ny, nz], size=(300, 3))])
    s.add_particles(pos, mass=120.0, velocities=rng.normal(scale=0.4, size=(len(pos), 3)))

    def conf(f):
        if case == "sparse-walls":
            f.setSolidNodes([i + nx*(j + ny*k) for k in (0, nz - 1) for j in range(ny) for i in range(nx)])
            f.setWallScheme(LBMForce.BounceBack)
            f.setFluidFluctuations(True)
            f.setTemperature(300.0)
        else:
            f.setFluidMomentumRemovalFrequency(2)
            f.setTemperature(100.0)
        f.setBodyAcceleration((1e-5/1e-4, 0, 0))
        f.setRandomNumberSeed(5)
        f.setFriction(10.0)
        f.setInterpolationStencil(stencil)
    s.configure.append(conf)
    runs = []
    for spread, field, setup in (("atomic", "0", "0"), ("tiles", "0", "1"), ("atomic", "1", "1"), ("tiles", "1", "0"),
                                 ("atomic", "0", "1")):
        monkeypatch.setenv("LBM_AMD_SPREAD", spread)
        monkeypatch.setenv("LBM_AMD_FIELD", field)
        monkeypatch.setenv("LBM_AMD_STENCIL_SETUP", setup)
        c = s.build(None, storage=storage)
        c.step(7)
        runs.append((c.getFluidState(), c.positions.copy(), c.velocities.copy(), c.getWallForce(), c.getForces()))
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("storage", [LBMForce.FP32, LBMForce.FP16, LBMForce.BF16, LBMForce.FP64],
                         ids=["fp32", "fp16", "bf16", "fp64"])
def test_row_block_layouts_are_bit_identical(storage, monkeypatch):
    """The population layout with 1, 2 or 4 lattice rows per block (LBM_AMD_ROW_BLOCK) changes where values live, not
    what is computed: the same run bit for bit, with regularized walls along y (the rows crossed by the walls' links
    span blocks), Velocity and DensityVelocity faces along z, fluctuations and Keys beads.  ny = 12 takes every block
    size; ny = 6 falls back from 4 rows to 2."""
    for ny in (12, 6):
        nx, nz = 10, 7
        s = Setup((nx, ny, nz), dx=1.0, dt=0.01, tau=0.75)
        rng = np.random.default_rng(9)
        s.add_particles(rng.uniform([0, 1.5, 1.5], [nx, ny - 1.5, nz - 1.5], size=(20, 3)), mass=150.0,
                        velocities=rng.normal(scale=0.3, size=(20, 3)))

        def conf(f):
            f.setSolidNodes([i + nx*ny*k for k in range(nz) for i in range(nx)])        # the plane j = 0
            f.setWallScheme(LBMForce.Regularized)
            f.setFaceBoundary(LBMForce.ZMin, LBMForce.Velocity)
            f.setFaceBoundary(LBMForce.ZMax, LBMForce.DensityVelocity)
            f.setFaceVelocity(LBMForce.ZMin, (0.0, 0.0, 0.3))
            f.setFaceDensity(LBMForce.ZMax, 602.214*1.002)
            f.setFluidFluctuations(True)
            f.setTemperature(300.0)
            f.setRandomNumberSeed(31)
            f.setFriction(8.0)
            f.setInterpolationStencil(LBMForce.Keys)
        s.configure.append(conf)
        runs = []
        for rows in ("1", "2", "4"):
            monkeypatch.setenv("LBM_AMD_ROW_BLOCK", rows)
            c = s.build(None, storage=storage)
            c.step(9)
            runs.append((c.getFluidState(), c.positions.copy(), c.getWallForce(), c.getFluidFields()[1]))
        for other in runs[1:]:
            for a, b in zip(runs[0], other):
                assert np.array_equal(a, b)
