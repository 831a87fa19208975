"""Batched replicas (include/lbm_amd_batch.h).

On the GPU, a batch of uneven replicas (different shapes, relaxation times, temperatures, seeds and particle counts,
fluctuations on) must reproduce, bit for bit, the same replicas run one by one through the single-lattice C-ABI
with the device Verlet integrator: both use the same node update, the same Philox streams (per-replica seed, step and
node or particle index) and fixed-point reaction sums, so any difference is a bug in the batching (tile table,
step parity, particle offsets, graph replay).  The single-lattice path is itself pinned to the oracle by the parity
tests.  On the CPU, the argument checks that run before any device work.
"""
import numpy as np
import pytest

from lbm_amd import Context, LBMError, LBMForce, System, VerletIntegrator
from lbm_amd.batch import Batch, Replica
from lbm_amd.harness import DeviceParticles
from lbm_amd.workloads import BOLTZ


def make_replica(grid, tau, beads, kT, seed, stencil=LBMForce.Trilinear, dx=1.0, dt=0.01, fluct=True,
                 rng_seed=0):
    nx, ny, nz = grid
    force = LBMForce()
    force.setGridSize(nx, ny, nz)
    force.setFluidDensity(602.214)
    force.setKinematicViscosity((tau - 0.5)/3.0*dx*dx/dt)
    force.setFluidMomentumRemovalFrequency(0)
    force.setFriction(10.0)
    force.setTemperature(kT/BOLTZ)
    force.setFluidFluctuations(fluct)
    force.setRandomNumberSeed(seed)
    force.setInterpolationStencil(stencil)
    for i in range(beads):
        force.addParticle(i)
    rng = np.random.default_rng(rng_seed)
    pos = rng.uniform(0, 1, size=(beads, 3))*np.array(grid)*dx
    vel = rng.normal(0, 0.1, size=(beads, 3))
    return Replica(force, (nx*dx, ny*dx, nz*dx), dt, np.full(beads, 100.0), pos, vel)


def uneven_replicas(stencil=LBMForce.Trilinear):
    return [
        make_replica((12, 12, 12), 0.62, 9, 2.5, 11, stencil=stencil, rng_seed=1),
        make_replica((16, 12, 10), 0.95, 30, 1.5, 12345, stencil=stencil, rng_seed=2),
        make_replica((8, 8, 8), 0.55, 0, 3.0, 777, stencil=stencil, rng_seed=3),
        make_replica((20, 9, 14), 0.8, 3, 2.0, 31, stencil=stencil, rng_seed=4),
    ]


def run_single(replica, steps):
    """The same replica through the single-lattice C-ABI and the device Verlet integrator."""
    system = System(replica.box)
    masses = replica.masses if len(replica.masses) else [1.0]
    for m in masses:
        system.addParticle(m)
    ctx = Context(system, VerletIntegrator(replica.stepSize), replica.force)
    if len(replica.masses):
        ctx.setPositions(replica.positions)
        ctx.setVelocities(replica.velocities)
    else:
        ctx.setPositions(np.array([[0.1, 0.2, 0.3]]))
        ctx.setVelocities(np.zeros((1, 3)))
    dp = DeviceParticles(ctx)
    dp.step(steps)
    dp.pull()
    rho, u = ctx.getFluidFields()
    return rho, u, ctx.positions[:len(replica.masses)].copy(), ctx.velocities[:len(replica.masses)].copy()


@pytest.mark.gpu
@pytest.mark.parametrize("stencil", [LBMForce.Trilinear, LBMForce.Keys, LBMForce.ThreePoint, LBMForce.NearestNode],
                         ids=["trilinear", "keys", "threepoint", "nearest"])
def test_batch_matches_single_replicas_bit_for_bit(stencil):
    replicas = uneven_replicas(stencil)
    batch = Batch(replicas)
    nodes, _ = batch.info()
    assert nodes == sum(int(np.prod(r.force.getGridSize())) for r in replicas)
    batch.step(27)                # two graph replays and seven single steps
    batch.synchronize()
    particles = batch.getParticles()
    for r, replica in enumerate(replicas):
        rho, u, pos, vel = run_single(replica, 27)
        brho, bu = batch.getFluidFields(r)
        bpos, bvel = particles[r]
        print(f"replica {r} grid {replica.force.getGridSize()}: max|drho| {np.abs(brho - rho).max():.3e} "
              f"max|du| {np.abs(bu - u).max():.3e}", flush=True)
        assert np.array_equal(brho, rho), r
        assert np.array_equal(bu, u), r
        assert np.array_equal(bpos, pos), r
        assert np.array_equal(bvel, vel), r


@pytest.mark.gpu
@pytest.mark.parametrize("order", ["1", "2", "3"])
def test_batch_tile_orders_are_bit_identical(order, monkeypatch):
    """The XCD tile orders (LBM_AMD_BATCH_XCD, read when a batch is created) change which work-group runs which tile,
    not the arithmetic: the batch's state after 23 steps matches the plain order's bit for bit.  A replica of 700
    beads gives the coupling 22 tiles, and with a 72^3 replica the fused pass has 770 tiles of 512 nodes (more than
    one round of eight runs of 64), so neither order is the identity."""
    replicas = uneven_replicas() + [make_replica((24, 24, 24), 0.7, 700, 2.5, 99, rng_seed=5),
                                    make_replica((72, 72, 72), 0.7, 40, 2.0, 5, rng_seed=6)]
    states = []
    for env in ("0", order):
        monkeypatch.setenv("LBM_AMD_BATCH_XCD", env)
        batch = Batch(replicas)
        batch.step(23)
        batch.synchronize()
        states.append(([batch.getFluidFields(r) for r in range(len(replicas))], batch.getParticles()))
    (fields0, parts0), (fields1, parts1) = states
    for r in range(len(replicas)):
        assert np.array_equal(fields0[r][0], fields1[r][0]) and np.array_equal(fields0[r][1], fields1[r][1]), r
        assert np.array_equal(parts0[r][0], parts1[r][0]) and np.array_equal(parts0[r][1], parts1[r][1]), r


@pytest.mark.gpu
def test_batch_autotune_keeps_the_state():
    """The autotuner tries nine tile shapes on the batch itself, which keeps stepping: the result after it is the
    single-lattice run of the same number of steps, bit for bit (tile shapes change the work split, not the
    arithmetic)."""
    replicas = uneven_replicas()[:2]
    batch = Batch(replicas)
    times, best = batch.autotune(10)      # 3 rounds x 9 tile shapes x (10 warm-up + 10 timed) steps
    print("autotune ms/step:", times, "best", best, flush=True)
    assert best in times and len(times) == 9 and all(t > 0 for t in times.values())
    batch.step(15)
    _, step = batch.info()
    assert step == 555
    for r, replica in enumerate(replicas):
        rho, u, pos, vel = run_single(replica, 555)
        brho, bu = batch.getFluidFields(r)
        assert np.array_equal(brho, rho) and np.array_equal(bu, u), r


@pytest.mark.gpu
def test_batch_fluid_temperature():
    """Equipartition of the fluctuating replicas: the velocity variance of a node, averaged over the lattice and
    over late steps, is kT/(rho dx^3) of each replica within 5% (the oracle gives 0.994 and 0.997 of it at 12^3,
    tau = 0.6 and 0.8, over the same sampling)."""
    replicas = [make_replica((16, 16, 16), 0.8, 0, 2.5, 5), make_replica((12, 12, 20), 0.6, 0, 1.5, 6)]
    batch = Batch(replicas)
    batch.step(200)
    samples = [[] for _ in replicas]
    for _ in range(20):
        batch.step(10)
        for r in range(len(replicas)):
            _, u = batch.getFluidFields(r)
            samples[r].append(np.var(u - u.mean(axis=0), axis=0).mean())
    for r, replica in enumerate(replicas):
        kT = replica.force.getTemperature()*BOLTZ
        expected = kT/(602.214*1.0)          # (nm/ps)^2: kT / cell mass
        measured = float(np.mean(samples[r]))
        print(f"replica {r}: <u^2> {measured:.4e} expected {expected:.4e}", flush=True)
        assert abs(measured/expected - 1) < 0.05


@pytest.mark.parametrize("change, message", [
    (lambda f: f.setInterpolationStencil(LBMForce.Keys), "same interpolation stencil"),
    (lambda f: f.setFluidFluctuations(False), "fluctuations on"),
    (lambda f: (f.setFaceBoundary(0, LBMForce.Velocity), f.setFaceBoundary(1, LBMForce.Velocity)), "periodic"),
    (lambda f: f.setFluidMomentumRemovalFrequency(10), "momentum removal"),
    (lambda f: f.setStorage(LBMForce.FP16), "FP32"),
])
def test_batch_rejects_what_it_does_not_support(change, message):
    replicas = uneven_replicas()[:2]
    change(replicas[1].force)
    with pytest.raises(LBMError, match=message):
        Batch(replicas)


def oracle_context(replica):
    """One replica as a Context on the FP64 oracle: the same force, box, step and particles; the coupling force is
    the only force, integrated with the same leapfrog Verlet."""
    import oracle_lib
    system = System(replica.box)
    for m in replica.masses:
        system.addParticle(m)
    ctx = Context(system, VerletIntegrator(replica.stepSize), replica.force, library=oracle_lib.load_oracle())
    ctx.setPositions(replica.positions)
    ctx.setVelocities(replica.velocities)
    return ctx


@pytest.mark.gpu
def test_config4_full_size():
    """The batching form at its full size (workloads.config4: 64 periodic replicas, edges {32, 48, 64, 96}, tau in
    [0.55, 1], 50-2000 trilinear beads, fluctuations on with per-replica seeds and kT), stepped together.  The two
    smallest replicas run in lockstep with the FP64 oracle for 10 steps: density, velocity and bead positions agree
    to 1e-5 relative (measured 2.6e-9, 1.1e-6 and 1.6e-10 nm; the fluctuations use the same Philox stream; the normals come from FP32 hardware
    transcendentals on the GPU and FP64 ones in the oracle).  After 300 steps every replica has kept its mass to
    1e-6 and its fluid temperature, the velocity variance of a node, is kT/(rho dx^3) within 5%."""
    from lbm_amd import workloads
    replicas, rows, _ = workloads.config4()
    batch = Batch(replicas)
    masses0 = []
    for r, row in enumerate(rows):
        rho, _ = batch.getFluidFields(r)
        masses0.append(rho.sum())
    order = np.argsort([row["edge"]*10000 + row["beads"] for row in rows])[:2]
    oracles = {int(r): oracle_context(replicas[r]) for r in order}
    batch.step(10)
    particles = batch.getParticles()
    for r, ctx in oracles.items():
        ctx.step(10)
        rho0, u0 = ctx.getFluidFields()
        rho1, u1 = batch.getFluidFields(r)
        pos1 = particles[r][0]
        d_rho = np.abs(rho1 - rho0).max()/np.abs(rho0).max()
        d_u = np.abs(u1 - u0).max()/np.abs(u0).max()
        d_x = np.abs(pos1 - ctx.positions[:len(pos1)]).max()
        print(f"replica {r} edge {rows[r]['edge']} beads {rows[r]['beads']}: density {d_rho:.1e} velocity {d_u:.1e} "
              f"positions {d_x:.1e} nm", flush=True)
        assert d_rho < 1e-5 and d_u < 1e-5 and d_x < 1e-5
    batch.step(290)
    for r, row in enumerate(rows):
        rho, u = batch.getFluidFields(r)
        drift = abs(rho.sum() - masses0[r])/masses0[r]
        expected = row["kT"]/602.214
        measured = np.var(u - u.mean(axis=0), axis=0).mean()
        assert drift < 1e-6, (r, drift)
        assert abs(measured/expected - 1) < 0.05, (r, measured/expected)
    print("all 64 replicas: mass kept to 1e-6 and fluid temperature within 5% of kT", flush=True)
