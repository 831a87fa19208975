"""The learned coupling closure on the GPU (include/lbm_amd_closure.h), against the closure oracle (oracle/mlp_oracle.py,
itself pinned by tests/test_closure_oracle.py) and against the library's own stencil coupling.

Tolerances: features and targets are FP32 results of FP64 sums, compared with FP64 at 1e-6 relative; the network's
outputs and gradients come from FP16 operands with FP32 accumulation and are compared with the oracle's emulation of
the same FP16 roundings at 5e-3 and 2e-2 relative (norm of the difference over norm of the oracle, per parameter
block); the couplings that must reduce to the stencil coupling are compared bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from lbm_amd.closure import MLPClosure, pairs, stats, OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, NUM_PARAMS
from lbm_amd.workloads import BOLTZ
from parity import Setup
from test_closure_oracle import random_params, random_state, symmetric_image

sys.path.insert(0, os.path.join(oracle_lib.ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (OFF_W1, OFF_B1), "b1": (OFF_B1, OFF_W2), "W2": (OFF_W2, OFF_B2), "b2": (OFF_B2, OFF_W3),
          "W3": (OFF_W3, OFF_B3), "b3": (OFF_B3, NUM_PARAMS)}


def bead_setup(n, state, pos, vel, stencil=LBMForce.Keys, temperature=0.0, fluct=False, seed=99, storage=None):
    s = Setup((n, n, n), tau=0.7)
    s.state = state
    s.add_particles(pos, mass=100.0, velocities=vel)

    def conf(f):
        f.setFriction(10.0)
        f.setTemperature(temperature)
        f.setFluidFluctuations(fluct)
        f.setRandomNumberSeed(seed)
        f.setInterpolationStencil(stencil)
    s.configure.append(conf)
    return s


def beads(rng, n, count):
    pos = rng.uniform(0, n, size=(count, 3))
    pos[:4] = np.floor(pos[:4]) + 0.5
    return pos, rng.normal(scale=0.2, size=(count, 3))


@pytest.mark.parametrize("storage", [LBMForce.FP64, LBMForce.FP32, LBMForce.FP16, LBMForce.BF16],
                         ids=["fp64", "fp32", "fp16", "bf16"])
def test_features_and_targets_match_the_oracle(storage):
    n = 12
    rng = np.random.default_rng(5)
    pos, vel = beads(rng, n, 60)
    ctx = bead_setup(n, random_state(n, rng), pos, vel).build(storage=storage)
    feats, targs = pairs(ctx)
    state = ctx.getFluidState()
    ref_f = mo.features(state, (n, n, n), 1.0, 0.01, pos, vel)
    ref_t = mo.targets(state, (n, n, n), 1.0, pos)
    print("max |feature - oracle|", np.abs(feats - ref_f).max(), "max |target - oracle|", np.abs(targs - ref_t).max())
    assert np.abs(feats - ref_f).max() <= 1e-6*np.abs(ref_f).max()
    assert np.abs(targs - ref_t).max() <= 1e-6*np.abs(ref_f[:, 3:84]).max()


@pytest.mark.parametrize("name,R", [
    ("reflect y", np.diag([1, -1, 1])),
    ("exchange y, z, reflect x", np.array([[-1, 0, 0], [0, 0, 1], [0, 1, 0]])),
])
def test_gpu_features_are_the_same_for_a_reflected_configuration(name, R):
    """The library's pairs are in each particle's frame, so a configuration mapped by a symmetry of the lattice (the
    populations, positions and velocities together) gives the same features and targets, particle by particle."""
    n = 12
    rng = np.random.default_rng(41)
    pos = rng.uniform(0, n, size=(60, 3))
    vel = rng.normal(scale=0.2, size=(60, 3))
    state = random_state(n, rng)
    f0, t0 = pairs(bead_setup(n, state, pos, vel).build())
    st, ps, vs = symmetric_image(state, n, R, pos, vel)
    f1, t1 = pairs(bead_setup(n, st, ps, vs).build())
    print(name, "max |dfeature|", np.abs(f1 - f0).max(), "max |dtarget|", np.abs(t1 - t0).max())
    assert np.abs(f1 - f0).max() <= 1e-6*np.abs(f0).max()
    assert np.abs(t1 - t0).max() <= 1e-6*np.abs(f0[:, 3:84]).max()


def network_with(rng, scale=0.3):
    net = MLPClosure(seed=3)
    net.setParameters(random_params(rng, scale).astype(np.float32))
    norm = dict(u_scale=40.0, v_scale=4.0, out_scale=60.0, u_bound=1e3, v_bound=1e3)
    net.setNormalization(**norm)
    return net, norm


def raw_batch(rng, count):
    raw = np.zeros((count, 87), dtype=np.float32)
    raw[:, 0:3] = rng.uniform(-1, 1, size=(count, 3))
    raw[:, 3:84] = rng.normal(scale=0.025, size=(count, 81))
    raw[:, 84:87] = rng.normal(scale=0.25, size=(count, 3))
    return raw


def test_forward_matches_the_oracle_emulation():
    rng = np.random.default_rng(21)
    net, norm = network_with(rng)
    raw = raw_batch(rng, 1000)
    out = net.forward(raw)
    params = net.getParameters().astype(np.float64)
    y, _ = mo.forward(params, mo.normalize(raw.astype(np.float64), norm["u_scale"], norm["v_scale"]))
    ref = y/norm["out_scale"]
    err = np.abs(out - ref).max()/np.abs(ref).max()
    print("forward relative error vs the FP16-emulating oracle:", err)
    assert err < 5e-3


def test_gradients_match_the_oracle():
    rng = np.random.default_rng(22)
    net, norm = network_with(rng)
    raw = raw_batch(rng, 300)            # not a multiple of the 256-row chunk: padding rows must not contribute
    t_raw = rng.normal(scale=0.02, size=(300, 3)).astype(np.float32)
    g, loss = net.gradients(raw, t_raw)
    params = net.getParameters().astype(np.float64)
    x = mo.normalize(raw.astype(np.float64), norm["u_scale"], norm["v_scale"])
    t = t_raw.astype(np.float64)*norm["out_scale"]
    ref = mo.backward(params, x, t, emulate16=True)
    ref_loss = mo.loss(params, x, t, emulate16=True)
    assert abs(loss - ref_loss) <= 1e-3*ref_loss
    for name, (a, b) in BLOCKS.items():
        rel = np.linalg.norm(g[a:b] - ref[a:b])/max(np.linalg.norm(ref[a:b]), 1e-30)
        print(f"gradient {name}: relative error {rel:.2e}")
        assert rel < 2e-2, name
    # Padding entries stay zero.
    pad = np.concatenate([np.arange(OFF_W1 + 87*64, OFF_B1),
                          (OFF_W3 + np.arange(64)[:, None]*16 + np.arange(3, 16)[None, :]).ravel(),
                          np.arange(OFF_B3 + 3, NUM_PARAMS)])
    assert np.all(g[pad] == 0)


def run_pair(n, rng, steps, net=None, stencil_b=LBMForce.NearestNode, count=80):
    """Two contexts from one setup with fluctuations and particle noise: A couples through the closure (Keys stencil),
    B with the stencil stencil_b; both stepped in lockstep.  With 27 count >= n^3 the closure reads the node velocities
    from a lattice-wide pass, below it pulls them per particle."""
    pos, vel = beads(rng, n, count)
    state = random_state(n, rng, amplitude=0.01)
    a = bead_setup(n, state, pos, vel, LBMForce.Keys, temperature=300.0, fluct=True).build()
    b = bead_setup(n, state, pos, vel, stencil_b, temperature=300.0, fluct=True).build()
    net.attach(a)
    for _ in range(steps):
        a.step(1)
        b.step(1)
    return a, b


@pytest.mark.parametrize("count", [80, 20])
def test_zero_network_is_the_nearest_node_coupling(count):
    rng = np.random.default_rng(8)
    net = MLPClosure(seed=1)
    net.setParameters(np.zeros(NUM_PARAMS, dtype=np.float32))
    net.setNormalization(u_scale=1.0, v_scale=1.0, out_scale=1.0, u_bound=1e30, v_bound=1e30)
    a, b = run_pair(12, rng, 15, net, LBMForce.NearestNode, count=count)
    ra, ua = a.getFluidFields()
    rb, ub = b.getFluidFields()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.velocities, b.velocities)
    assert stats(a) == (count*15, 0)


def test_out_of_distribution_particles_take_the_keys_coupling():
    rng = np.random.default_rng(9)
    net = MLPClosure(seed=1)
    net.setNormalization(u_scale=1.0, v_scale=1.0, out_scale=1.0, u_bound=0.0, v_bound=0.0)
    a, b = run_pair(12, rng, 15, net, LBMForce.Keys)
    ra, ua = a.getFluidFields()
    rb, ub = b.getFluidFields()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.velocities, b.velocities)
    assert stats(a) == (0, 80*15)


def test_closure_force_is_nearest_plus_network_correction():
    """At one state: F_closure - F_nearest = friction m dx/dt Du (in distribution) and F_closure = F_Keys (out of
    distribution), with Du the network's output for the particle's inputs, turned from the particle's frame back to
    the lattice's."""
    n = 12
    rng = np.random.default_rng(10)
    pos, vel = beads(rng, n, 64)
    state = random_state(n, rng)
    keys = bead_setup(n, state, pos, vel, LBMForce.Keys).build()
    near = bead_setup(n, state, pos, vel, LBMForce.NearestNode).build()
    feats, _ = pairs(keys)
    net, norm = network_with(rng, scale=0.1)
    # Put roughly a quarter of the particles out of distribution through the bead-velocity bound.
    vb = np.quantile(np.abs(feats[:, 84:87]*norm["v_scale"]).max(axis=1), 0.75)
    net.setNormalization(**dict(norm, v_bound=float(vb)))
    p, sg = mo.frame_of(pos, 1.0, (n, n, n))
    du = mo.from_frame(net.forward(feats).astype(np.float64), p, sg)
    ood = (np.abs(feats[:, 84:87]*np.float32(norm["v_scale"])) > np.float32(vb)).any(axis=1)
    net.attach(keys)
    for c in (keys, near):
        c.setStepCount(1)
    f_cl = keys.getForces()
    f_n = near.getForces()
    MLPClosure.detach(keys)
    f_k = keys.getForces()
    expected = f_n + 10.0*100.0*1.0/0.01*du
    scale = np.abs(f_n).max()
    assert 0 < ood.sum() < len(ood)
    assert np.abs(f_cl[~ood] - expected[~ood]).max() <= 1e-9*scale
    assert np.array_equal(f_cl[ood], f_k[ood])


def test_training_learns_the_keys_correction():
    """Pairs from fluctuating states (16^3, kT = 2.5 kJ/mol): the GPU-trained network's error on held-out pairs is well
    below that of the nearest-node coupling alone (whose error is the target itself)."""
    n = 16
    rng = np.random.default_rng(12)
    s = Setup((n, n, n), tau=0.7)

    def conf(f):
        f.setTemperature(2.5/BOLTZ)
        f.setFluidFluctuations(True)
        f.setRandomNumberSeed(77)
        f.setInterpolationStencil(LBMForce.Keys)
        f.setFriction(10.0)
    s.configure.append(conf)
    count = 3000
    s.add_particles(rng.uniform(0, n, size=(count, 3)), mass=100.0)
    ctx = s.build()
    feats, targs = [], []
    ctx.step(50)
    for snap in range(14):
        ctx.step(10)
        pos = rng.uniform(0, n, size=(count, 3))
        vel = rng.normal(scale=np.sqrt(2.5/100.0), size=(count, 3))
        f, t = pairs(ctx, pos, vel)
        feats.append(f)
        targs.append(t)
    train_f, train_t = np.concatenate(feats[:12]), np.concatenate(targs[:12])
    test_f, test_t = np.concatenate(feats[12:]), np.concatenate(targs[12:])
    net = MLPClosure(seed=5)
    losses = net.train(train_f, train_t, epochs=30, batch=512, lr=3e-3, decay=0.9, seed=2)
    pred = net.forward(test_f)
    err_net = np.sqrt(np.mean((pred - test_t)**2))
    err_nearest = np.sqrt(np.mean(test_t**2))
    print(f"losses {losses[0]:.4f} -> {losses[-1]:.4f}; held-out rms error: network {err_net:.3e}, "
          f"nearest only {err_nearest:.3e} ({err_net/err_nearest:.3f})")
    assert losses[-1] < 0.5*losses[0]
    assert err_net < 0.6*err_nearest


def test_config5_full_size_keys_run():
    """The learned form's workload at its full size (workloads.config5: 128^3, 2,000 harmonic chains of 50 beads,
    fluctuating fluid and particle noise at 300 K, Keys coupling), 300 device-resident steps: the bonds keep their
    length (mean within 1% of the thermal mean of a harmonic bond, <r> = int r^3 exp(-k (r - r0)^2/2kT) dr /
    int r^2 exp(-k (r - r0)^2/2kT) dr = 0.3929 nm for r0 = 0.38 nm, none beyond 2 r0), the beads are at the temperature of the explicit drag's
    leapfrog half-step velocities, T/(1 - gamma dt/2), within 2% (the explicit drag heats them by about 0.45%), and the fluid's node velocity variance is kT/(rho
    dx^3) within 5%.  Then the closure (a network with every particle out of distribution) takes the Keys coupling
    for all of them: 100,000 fallbacks per step and the same state bit for bit as the Keys run itself."""
    from lbm_amd import workloads
    w = workloads.config5()
    runs = []
    for with_closure in (False, True):
        c = w.build()
        dp = w.device_particles(c)
        net = None
        # The first step starts from rest (every input zero, so in distribution even with zero bounds): the closure
        # is attached after it.
        dp.step(1)
        if with_closure:
            net = MLPClosure(seed=1)
            net.setNormalization(u_scale=1.0, v_scale=1.0, out_scale=1.0, u_bound=0.0, v_bound=0.0)
            net.attach(c)
        dp.step(299)
        pos, vel = dp.pull()
        rho, u = c.getFluidFields()
        counts = stats(c) if with_closure else None
        runs.append((pos, vel, rho, u, counts))
        del dp, c
    pos, vel, rho, u, _ = runs[0]
    pairs_ = w.bonds[0].reshape(-1, 2)
    box = np.array(w.grid, dtype=np.float64)*w.dx
    d = pos[pairs_[:, 1]] - pos[pairs_[:, 0]]
    d -= box*np.round(d/box)
    bond = np.linalg.norm(d, axis=1)
    kT = BOLTZ*300.0
    T_beads = 100.0*np.mean(vel**2)/kT
    T_expected = 1.0/(1.0 - 10.0*w.dt/2)
    fluid = np.var(u - u.mean(axis=0), axis=0).mean()/(kT/602.214)
    print(f"bond length mean {bond.mean():.4f} max {bond.max():.4f} nm; bead temperature {T_beads:.4f} kT "
          f"(expected {T_expected:.4f}); fluid temperature {fluid:.4f} kT", flush=True)
    r = np.linspace(1e-4, 1.2, 200001)
    weight = r**2*np.exp(-1000.0*(r - 0.38)**2/(2*kT))
    thermal_mean = (r*weight).sum()/weight.sum()
    assert abs(bond.mean()/thermal_mean - 1) < 0.01 and bond.max() < 0.76
    assert abs(T_beads/T_expected - 1) < 0.02
    assert abs(fluid - 1) < 0.05
    c_pos, c_vel, c_rho, c_u, counts = runs[1]
    assert counts == (0, 299*len(w.masses))
    assert np.array_equal(c_pos, pos) and np.array_equal(c_vel, vel)
    assert np.array_equal(c_rho, rho) and np.array_equal(c_u, u)


def test_closure_fallback_against_the_oracle_keys():
    """Chains on a small lattice (16^3, 12 chains of 10 beads, fluctuating fluid and particles at 300 K): the closure
    with every particle out of distribution couples through its Keys fallback, in lockstep with the FP64 oracle's Keys
    coupling on the same Philox streams, for 20 steps: density, velocity, bead forces and positions within 1e-5
    (measured 1.1e-9, 3.1e-7, 1.1e-8 and 7.8e-10)."""
    from lbm_amd import workloads
    from parity import rel_err
    import oracle_lib
    w = workloads.config5(n=16, num_chains=12, chain_length=10, seed=99)
    bonds = w.bonds

    def bond_forces(pos):
        pairs_ = bonds[0].reshape(-1, 2)
        box = np.array(w.grid, dtype=np.float64)*w.dx
        d = pos[pairs_[:, 1]] - pos[pairs_[:, 0]]
        d -= box*np.round(d/box)
        r = np.linalg.norm(d, axis=1)
        f = (bonds[1]*(r - bonds[2])/r)[:, None]*d
        out = np.zeros_like(pos)
        np.add.at(out, pairs_[:, 0], f)
        np.add.at(out, pairs_[:, 1], -f)
        return out
    w.bonds = None
    ref = w.build(oracle_lib.load_oracle())
    gpu = w.build()
    for c in (ref, gpu):
        c.extraForces.append(bond_forces)
    net = MLPClosure(seed=1)
    net.setNormalization(u_scale=1.0, v_scale=1.0, out_scale=1.0, u_bound=0.0, v_bound=0.0)
    worst = {"density": 0.0, "velocity": 0.0, "force": 0.0, "position": 0.0}
    for step in range(20):
        ref.step(1)
        gpu.step(1)
        if step == 0:
            net.attach(gpu)        # after the first step from rest, whose zero inputs are in distribution
        r0, u0 = ref.getFluidFields()
        r1, u1 = gpu.getFluidFields()
        worst["density"] = max(worst["density"], rel_err(r1, r0))
        worst["velocity"] = max(worst["velocity"], rel_err(u1, u0))
        worst["force"] = max(worst["force"], rel_err(gpu.getForces(), ref.getForces()))
        worst["position"] = max(worst["position"], rel_err(gpu.positions, ref.positions, scale=w.dx))
    network, fallback = stats(gpu)
    print("closure fallback vs oracle Keys:", worst, "couplings", network, fallback)
    assert network == 0 and fallback >= 19*len(w.masses)
    assert max(worst.values()) < 1e-5
