"""Pins of the closure oracle (oracle/mlp_oracle.py), on the CPU.

* backward(): against central finite differences of loss() in FP64 (no FP16 emulation);
* the Keys interpolation: exact on quadratic fields (the stencil reproduces polynomials to degree 2) and equal to the
  node value at a node;
* the targets and the centre-node features: against the forces of the FP64 oracle library's own nearest-node and
  Keys coupling at the same state (T = 0), through F = -gamma m (v - u):
  u_Keys - u_nearest = (F_Keys - F_nearest) dt / (friction mass dx);
* the bead's frame: offsets non-negative and sorted, vectors mapped there and back; the Keys weights of the 3x3x3
  block give, applied in the bead's frame, the bead-frame image of their lattice-frame result (a wrong node, sign or
  axis of the mapping changes the weights' arguments or the values they meet); and a configuration reflected through
  a lattice plane or with two axes exchanged has the same bead-frame features and targets.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from parity import Setup, equilibrium_state

sys.path.insert(0, os.path.join(oracle_lib.ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def random_params(rng, scale=0.3):
    W1 = np.zeros((mo.N_IN_PAD, mo.N_H))
    W1[:mo.N_IN] = rng.normal(scale=scale, size=(mo.N_IN, mo.N_H))
    W3 = np.zeros((mo.N_H, mo.N_OUT_PAD))
    W3[:, :mo.N_OUT] = rng.normal(scale=scale, size=(mo.N_H, mo.N_OUT))
    b3 = np.zeros(mo.N_OUT_PAD)
    b3[:mo.N_OUT] = rng.normal(scale=0.1, size=mo.N_OUT)
    return mo.pack(W1, rng.normal(scale=0.1, size=mo.N_H), rng.normal(scale=scale, size=(mo.N_H, mo.N_H)),
                   rng.normal(scale=0.1, size=mo.N_H), W3, b3)


def test_parameter_count():
    # 87*64 + 64 + 64*64 + 64 + 64*3 + 3 real parameters, padded to 96 inputs and 16 outputs.
    assert 87*64 + 64 + 64*64 + 64 + 64*3 + 3 == 9987
    assert mo.N_PARAMS == 96*64 + 64 + 64*64 + 64 + 64*16 + 16


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_params(rng)
    x = np.zeros((20, mo.N_IN_PAD))
    x[:, :mo.N_IN] = rng.normal(size=(20, mo.N_IN))
    t = rng.normal(size=(20, 3))
    g = mo.backward(p, x, t, emulate16=False)
    real = [i for i in range(mo.N_PARAMS)
            if not (mo.OFF_W1 + mo.N_IN*mo.N_H <= i < mo.OFF_B1)
            and not (mo.OFF_W3 <= i < mo.OFF_B3 and (i - mo.OFF_W3) % mo.N_OUT_PAD >= mo.N_OUT)
            and not (i >= mo.OFF_B3 + mo.N_OUT)]
    assert len(real) == 9987
    idx = rng.choice(real, size=200, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss(pp, x, t, False) - mo.loss(pm, x, t, False))/(2*h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6*abs(fd), (i, fd, g[i])
    # Padding entries have zero gradient.
    pad = np.setdiff1d(np.arange(mo.N_PARAMS), real)
    assert np.all(g[pad] == 0)


def test_keys_interpolation_is_exact_on_quadratic_fields():
    n = 16
    grid = (n, n, n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(float)
    c0 = np.array([0.01, -0.02, 0.005])
    A = np.array([[1e-3, 2e-4, -1e-4], [-3e-4, 5e-4, 2e-4], [1e-4, -2e-4, 7e-4]])
    Q = np.array([2e-5, -1e-5, 3e-5])
    u = c0 + xyz @ A.T + (xyz**2) @ np.diag(Q)
    rng = np.random.default_rng(1)
    pos = rng.uniform(4.0, 11.0, size=(50, 3))
    exact = c0 + pos @ A.T + (pos**2) @ np.diag(Q)
    assert np.abs(mo.keys_velocity(u, grid, 1.0, pos) - exact).max() < 1e-15
    # At a node the Keys value is the node value: the target vanishes.
    state = equilibrium_state(n**3, 1.0, u)
    nodes = rng.integers(0, n, size=(10, 3)).astype(float)
    assert np.abs(mo.targets(state, grid, 1.0, nodes)).max() < 1e-15
    # node_velocities inverts the equilibrium.
    assert np.abs(mo.node_velocities(state) - u).max() < 1e-15


def random_state(n, rng, amplitude=0.02):
    """Smooth flow plus node-to-node noise, at equilibrium (a state the tests can hand to either backend)."""
    z = np.arange(n**3)//(n*n)
    y = (np.arange(n**3)//n) % n
    u = np.zeros((n**3, 3))
    u[:, 0] = amplitude*np.sin(2*np.pi*z/n)
    u[:, 1] = amplitude*np.cos(2*np.pi*y/n)
    u += rng.normal(scale=amplitude/2, size=u.shape)
    rho = 1.0 + rng.normal(scale=0.01, size=n**3)
    return equilibrium_state(n**3, rho, u)


def coupling_forces(state, n, pos, vel, stencil, mass=100.0, friction=10.0):
    s = Setup((n, n, n), tau=0.7)
    s.state = state
    s.add_particles(pos, mass=mass, velocities=vel)

    def conf(f):
        f.setFriction(friction)
        f.setTemperature(0.0)
        f.setInterpolationStencil(stencil)
    s.configure.append(conf)
    ctx = s.build(oracle_lib.load_oracle())
    ctx.setStepCount(1)          # a force evaluation between steps, from the current state
    return ctx.getForces(), s


def test_targets_and_centre_features_match_the_oracle_coupling():
    n = 12
    rng = np.random.default_rng(7)
    state = random_state(n, rng)
    pos = rng.uniform(0, n, size=(40, 3))
    pos[:5] = np.floor(pos[:5]) + 0.5           # exact ties of the nearest-node rule
    vel = rng.normal(scale=0.2, size=(40, 3))
    mass, friction, dt, dx = 100.0, 10.0, 0.01, 1.0
    fk, _ = coupling_forces(state, n, pos, vel, LBMForce.Keys, mass, friction)
    fn, _ = coupling_forces(state, n, pos, vel, LBMForce.NearestNode, mass, friction)
    expected = (fk - fn)*dt/(friction*mass*dx)
    t = mo.lattice_targets(state, (n, n, n), dx, pos)
    assert np.abs(t - expected).max() <= 1e-12*max(1.0, np.abs(expected).max())
    feats = mo.lattice_features(state, (n, n, n), dx, dt, pos, vel)
    centre = feats[:, 3 + 3*13:6 + 3*13]
    assert np.abs(centre - (vel*dt/dx + fn*dt/(friction*mass*dx))).max() < 1e-12
    assert np.all(np.abs(feats[:, 0:3]) <= 1.0)
    # In the bead's frame: the same target turned, and the centre node stays the centre slot.
    p, sg = mo.frame_of(pos, dx, (n, n, n))
    assert np.abs(mo.from_frame(mo.targets(state, (n, n, n), dx, pos), p, sg) - expected).max() \
        <= 1e-12*max(1.0, np.abs(expected).max())
    framed = mo.features(state, (n, n, n), dx, dt, pos, vel)
    assert np.abs(mo.from_frame(framed[:, 42:45], p, sg) - centre).max() < 1e-15


def test_neighbour_features_match_the_oracle_fields():
    """The 27 velocity features are the oracle library's node velocities (no body force: u = j/rho)."""
    n = 10
    rng = np.random.default_rng(11)
    state = random_state(n, rng)
    s = Setup((n, n, n), tau=0.7)
    s.state = state
    ctx = s.build(oracle_lib.load_oracle())
    _, u_fields = ctx.getFluidFields()
    u_lat = u_fields*(s.dt/s.dx)
    pos = rng.uniform(0, n, size=(15, 3))
    feats = mo.lattice_features(state, (n, n, n), s.dx, s.dt, pos, np.zeros((15, 3)))
    _, idx = mo.nearest(pos, s.dx, (n, n, n))
    slot = 0
    for c in (-1, 0, 1):
        for b in (-1, 0, 1):
            for a in (-1, 0, 1):
                nodes = mo.node_index(idx[:, 0] + a, idx[:, 1] + b, idx[:, 2] + c, (n, n, n))
                assert np.abs(feats[:, 3 + 3*slot:6 + 3*slot] - u_lat[nodes]).max() < 1e-14
                slot += 1


def test_frame_sorts_the_offsets_and_maps_vectors_both_ways():
    rng = np.random.default_rng(31)
    o = rng.uniform(-0.5, 0.5, size=(2000, 3))
    o[:100, 1] = -o[:100, 0]                     # equal magnitudes, opposite signs
    o[100:200] = -0.5                            # the rounding tie on every axis
    o[200:300, 2] = 0.0
    p, sg = mo.frame(o)
    framed = mo.to_frame(o, p, sg)
    assert np.all(framed >= 0)
    assert np.all(framed[:, 0] >= framed[:, 1]) and np.all(framed[:, 1] >= framed[:, 2])
    assert np.array_equal(np.sort(p, axis=1), np.tile([0, 1, 2], (len(o), 1)))
    # Equal magnitudes keep the axis order.
    assert np.all(p[100:200] == [0, 1, 2])
    tie = np.abs(o[:, 0]) == np.abs(o[:, 1])
    assert np.all(np.where(p[tie] == 0)[1] < np.where(p[tie] == 1)[1])
    w = rng.normal(size=o.shape)
    assert np.array_equal(mo.from_frame(mo.to_frame(w, p, sg), p, sg), w)
    # The frame is a rotation or reflection: lengths and dot products are kept.
    v = rng.normal(size=o.shape)
    assert np.allclose(np.sum(mo.to_frame(w, p, sg)*mo.to_frame(v, p, sg), axis=1), np.sum(w*v, axis=1), atol=1e-14)


def block_keys(features):
    """u_Keys - u(nearest) with the Keys weights of the 3x3x3 block only, from features of either frame."""
    off = features[:, 0:3]/2.0
    u = features[:, 3:84].reshape(-1, 27, 3)
    out = -u[:, 13, :]
    for c in range(3):
        for b in range(3):
            for a in range(3):
                w = mo.keys_kernel(off[:, 0] - (a - 1))*mo.keys_kernel(off[:, 1] - (b - 1))*mo.keys_kernel(off[:, 2] - (c - 1))
                out = out + w[:, None]*u[:, a + 3*b + 9*c, :]
    return out


def test_block_keys_weights_commute_with_the_frame():
    n = 10
    rng = np.random.default_rng(32)
    state = random_state(n, rng)
    pos = rng.uniform(0, n, size=(300, 3))
    vel = rng.normal(scale=0.2, size=(300, 3))
    lat = mo.lattice_features(state, (n, n, n), 1.0, 0.01, pos, vel)
    p, sg = mo.frame_of(pos, 1.0, (n, n, n))
    framed = mo.features(state, (n, n, n), 1.0, 0.01, pos, vel)
    ref = mo.to_frame(block_keys(lat), p, sg)
    assert np.abs(block_keys(framed) - ref).max() <= 1e-14*np.abs(ref).max()
    assert np.abs(framed[:, 84:87] - mo.to_frame(lat[:, 84:87], p, sg)).max() == 0
    # Planting a wrong sign in the block mapping is caught.
    bad = framed.copy()
    bad[:, 3:84] = bad[:, 3:84].reshape(-1, 27, 3)[:, ::-1, :].reshape(-1, 81)
    assert np.abs(block_keys(bad) - ref).max() > 1e-3*np.abs(ref).max()


def symmetric_image(state, n, R, pos, vel):
    """The configuration mapped by the lattice symmetry R (a signed permutation matrix: c -> R c)."""
    C = np.stack([mo.CX, mo.CY, mo.CZ], axis=1)
    image = C @ R.T
    qmap = [int(np.where((C == image[q]).all(axis=1))[0][0]) for q in range(19)]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    to = (xyz @ R.T) % n
    node_to = to[:, 0] + n*(to[:, 1] + n*to[:, 2])
    f = np.asarray(state, dtype=np.float64).reshape(19, -1)
    out = np.empty_like(f)
    for q in range(19):
        out[qmap[q], node_to] = f[q]
    return out.ravel(), np.mod(pos @ R.T, n), vel @ R.T


@pytest.mark.parametrize("name,R", [
    ("reflect x", np.diag([-1, 1, 1])),
    ("exchange x, y", np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1]])),
    ("reflect z, cycle axes", np.array([[0, 0, 1], [1, 0, 0], [0, -1, 0]])),
])
def test_bead_frame_inputs_are_invariant_under_lattice_symmetries(name, R):
    n = 10
    rng = np.random.default_rng(33)
    state = random_state(n, rng)
    pos = rng.uniform(0, n, size=(200, 3))
    vel = rng.normal(scale=0.2, size=(200, 3))
    f0 = mo.features(state, (n, n, n), 1.0, 0.01, pos, vel)
    t0 = mo.targets(state, (n, n, n), 1.0, pos)
    st, ps, vs = symmetric_image(state, n, R, pos, vel)
    f1 = mo.features(st, (n, n, n), 1.0, 0.01, ps, vs)
    t1 = mo.targets(st, (n, n, n), 1.0, ps)
    assert np.abs(f1 - f0).max() <= 1e-12*np.abs(f0).max(), name
    assert np.abs(t1 - t0).max() <= 1e-12*np.abs(t0).max(), name
    # The lattice-frame features do change (the test is not vacuous).
    assert np.abs(mo.lattice_features(st, (n, n, n), 1.0, 0.01, ps, vs)
                  - mo.lattice_features(state, (n, n, n), 1.0, 0.01, pos, vel)).max() > 1e-3
