"""CPU ORACLE OF THE LEARNED COUPLING CLOSURE -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the reference results for the gfx950 closure kernels; nothing in the product imports it.

The closure has no counterpart in the reference.  What it is built from does:

* the per-node fluid velocity the coupling interpolates, u = j/rho from the populations, with the momentum-removal
  shift j -= rho u_cm when removal is due (ReferenceLBMKernels.cpp:503-553 nearest node, 643-697 stencils);
* the nearest node floor(s + 1/2) of the wrapped lattice coordinate s = x/dx (ReferenceLBMKernels.cpp:973-983);
* the Keys stencil: first node floor(s) - 1, four nodes per axis, phi(r) = 1 - 2.5 r^2 + 1.5 |r|^3 for |r| <= 1,
  2 - 4|r| + 2.5 r^2 - 0.5 |r|^3 for 1 < |r| < 2 (LBMStencils.h:42-104, slots x fastest, ReferenceLBMKernels.cpp:998-1018).

Restated here in numpy, FP64, first in the lattice's frame (lattice_features, lattice_targets):

  features(x) = [ 2 (s - round(s))            3 values, the sub-node offset o of the bead, scaled to [-1, 1)
                  u(nearest + (a, b, c))      81 values: a, b, c in {-1, 0, 1}, a fastest, then b, then c; xyz each
                  v dt/dx ]                   3 values, the bead velocity in lattice units
  target(x)   = u_Keys(x) - u(nearest)        3 values, lattice units: (F_Keys - F_nearest)/(gamma m) of the explicit
                                              drag F = -gamma m (v - u) at the same state

then in the bead's frame (features, targets), which is what the library's pairs hold and its network sees: the Keys
stencil and the lattice are symmetric under the 48 reflections and axis permutations of the cube, and the frame is the
one of them that makes the offset non-negative and sorted.  Frame axis i is lattice axis p_i (|o| in decreasing order,
equal values in axis order), with sign s_i = -1 where o_(p_i) < 0; a vector w has w'_i = s_i w_(p_i), the node
nearest + d sits at d'_i = s_i d_(p_i) of the block, and the offset becomes |o_(p_i)|.  to_frame / from_frame map
vectors both ways.

and the network: 87 -> 64 -> 64 -> 3, SiLU activations, on normalized inputs (offsets as above, velocities times
u_scale, bead velocity times v_scale; inputs padded with zeros to 96) and normalized targets (times out_scale).  The
loss of a batch is the mean over beads of (1/3) |y - t|^2.  forward/backward can emulate the gfx950 kernels' FP16
roundings (weights, inputs, hidden activations and back-propagated errors stored as FP16, sums in FP32 or better).

Pinned by tests/test_closure_oracle.py: finite differences of the loss (backward), the Keys stencil's exactness on
quadratic fields and at nodes, the lattice-frame target and centre-velocity features against the forces of the FP64
oracle library's own nearest-node and Keys coupling (oracle/lbm_oracle.cpp) at the same state, and the frame through
what the symmetry fixes: the Keys weights of the 3x3x3 block applied in the bead's frame give the bead-frame image of
the lattice-frame result, and a configuration reflected or with its axes exchanged has the same bead-frame features
and targets.
"""
import numpy as np

N_IN, N_IN_PAD, N_H, N_OUT, N_OUT_PAD = 87, 96, 64, 3, 16

# Flat parameter layout, FP32 master copy: W1[i][o] (96 x 64), b1, W2[i][o] (64 x 64), b2, W3[i][o] (64 x 16), b3 (16).
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN_PAD*N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H*N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H*N_OUT_PAD
N_PARAMS = OFF_B3 + N_OUT_PAD

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1])


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_IN_PAD, N_H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(N_H, N_H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(N_H, N_OUT_PAD), p[OFF_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3])


def node_velocities(state, ucm=None):
    """u = j/rho at every node (lattice units) from the FP64 deviations state[q*N + node]."""
    f = np.asarray(state, dtype=np.float64).reshape(19, -1)
    rho = 1.0 + f.sum(axis=0)
    j = np.stack([CX @ f, CY @ f, CZ @ f], axis=1)
    if ucm is not None:
        j -= rho[:, None]*np.asarray(ucm)[None, :]
    return j/rho[:, None]


def lattice_coordinates(positions, dx, grid):
    s = np.asarray(positions, dtype=np.float64)/dx
    size = np.asarray(grid, dtype=np.float64)
    return s - np.floor(s/size)*size


def nearest(positions, dx, grid):
    s = lattice_coordinates(positions, dx, grid)
    idx = np.floor(s + 0.5).astype(np.int64) % np.asarray(grid)
    return s, idx


def node_index(ix, iy, iz, grid):
    nx, ny, nz = grid
    return (ix % nx) + nx*((iy % ny) + ny*(iz % nz))


def lattice_features(state, grid, dx, dt, positions, velocities, ucm=None):
    """Raw (unnormalized) features in the lattice's frame, FP64, [n][87]."""
    u = node_velocities(state, ucm)
    s, idx = nearest(positions, dx, grid)
    n = len(s)
    out = np.zeros((n, N_IN))
    out[:, 0:3] = 2.0*(s - np.floor(s + 0.5))
    slot = 0
    for c in (-1, 0, 1):
        for b in (-1, 0, 1):
            for a in (-1, 0, 1):
                nodes = node_index(idx[:, 0] + a, idx[:, 1] + b, idx[:, 2] + c, grid)
                out[:, 3 + 3*slot:6 + 3*slot] = u[nodes]
                slot += 1
    out[:, 84:87] = np.asarray(velocities, dtype=np.float64)*(dt/dx)
    return out


def keys_kernel(r):
    r = np.abs(r)
    return np.where(r <= 1.0, 1.0 - 2.5*r*r + 1.5*r**3, np.where(r < 2.0, 2.0 - 4.0*r + 2.5*r*r - 0.5*r**3, 0.0))


def keys_velocity(u, grid, dx, positions):
    """Keys-interpolated velocity (periodic lattice; the repeated-node merge only matters below 4 nodes per axis)."""
    s = lattice_coordinates(positions, dx, grid)
    first = np.floor(s).astype(np.int64) - 1
    out = np.zeros((len(s), 3))
    for c in range(4):
        wz = keys_kernel(s[:, 2] - (first[:, 2] + c))
        for b in range(4):
            wy = keys_kernel(s[:, 1] - (first[:, 1] + b))
            for a in range(4):
                wx = keys_kernel(s[:, 0] - (first[:, 0] + a))
                nodes = node_index(first[:, 0] + a, first[:, 1] + b, first[:, 2] + c, grid)
                out += (wx*wy*wz)[:, None]*u[nodes]
    return out


def lattice_targets(state, grid, dx, positions, ucm=None):
    """u_Keys(x) - u(nearest) in the lattice's frame, [n][3]."""
    u = node_velocities(state, ucm)
    _, idx = nearest(positions, dx, grid)
    centre = u[node_index(idx[:, 0], idx[:, 1], idx[:, 2], grid)]
    return keys_velocity(u, grid, dx, positions) - centre


def frame(offsets):
    """The bead's frame for sub-node offsets o [n][3]: (p [n][3] lattice axis of each frame axis, s [n][3] signs)."""
    o = np.asarray(offsets, dtype=np.float64)
    p = np.argsort(-np.abs(o), axis=1, kind="stable")
    s = np.where(np.take_along_axis(o, p, axis=1) < 0, -1.0, 1.0)
    return p, s


def frame_of(positions, dx, grid):
    s, _ = nearest(positions, dx, grid)
    return frame(s - np.floor(s + 0.5))


def to_frame(w, p, s):
    """Vectors [n][3] of the lattice's frame in the bead's frame: w'_i = s_i w_(p_i)."""
    return np.take_along_axis(np.asarray(w, dtype=np.float64), p, axis=1)*s


def from_frame(w, p, s):
    """Vectors [n][3] of the bead's frame back in the lattice's frame."""
    out = np.empty(np.shape(w))
    np.put_along_axis(out, p, np.asarray(w, dtype=np.float64)*s, axis=1)
    return out


def features_to_frame(raw, p, s):
    """Lattice-frame features [n][87] in the bead's frame."""
    raw = np.asarray(raw, dtype=np.float64)
    n = len(raw)
    out = np.empty_like(raw)
    out[:, 0:3] = np.abs(np.take_along_axis(raw[:, 0:3], p, axis=1))
    rows = np.arange(n)
    for c in (-1, 0, 1):
        for b in (-1, 0, 1):
            for a in (-1, 0, 1):
                d_frame = np.array([a, b, c])
                d = np.zeros((n, 3), dtype=np.int64)
                np.put_along_axis(d, p, (s*d_frame[None, :]).astype(np.int64), axis=1)
                src = (d[:, 0] + 1) + 3*(d[:, 1] + 1) + 9*(d[:, 2] + 1)
                dst = (a + 1) + 3*(b + 1) + 9*(c + 1)
                u = raw[rows[:, None], 3 + 3*src[:, None] + np.arange(3)[None, :]]
                out[:, 3 + 3*dst:6 + 3*dst] = to_frame(u, p, s)
    out[:, 84:87] = to_frame(raw[:, 84:87], p, s)
    return out


def features(state, grid, dx, dt, positions, velocities, ucm=None):
    """Raw (unnormalized) features in each bead's frame, FP64, [n][87]: the library's training features."""
    p, s = frame_of(positions, dx, grid)
    return features_to_frame(lattice_features(state, grid, dx, dt, positions, velocities, ucm), p, s)


def targets(state, grid, dx, positions, ucm=None):
    """u_Keys(x) - u(nearest) in each bead's frame, [n][3]: the library's training targets."""
    p, s = frame_of(positions, dx, grid)
    return to_frame(lattice_targets(state, grid, dx, positions, ucm), p, s)


def normalize(raw, u_scale, v_scale):
    x = np.zeros((len(raw), N_IN_PAD))
    x[:, 0:3] = raw[:, 0:3]
    x[:, 3:84] = raw[:, 3:84]*u_scale
    x[:, 84:87] = raw[:, 84:87]*v_scale
    return x


def fp16(a):
    return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


def silu(z):
    return z/(1.0 + np.exp(-z))


def dsilu(z):
    sg = 1.0/(1.0 + np.exp(-z))
    return sg*(1.0 + z*(1.0 - sg))


def forward(params, x, emulate16=True):
    """Normalized inputs x [n][96] -> normalized outputs [n][3] and the intermediate values."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    r = fp16 if emulate16 else (lambda a: np.asarray(a, dtype=np.float64))
    x16 = r(x)
    z1 = x16 @ r(W1) + b1
    h1 = r(silu(z1))
    z2 = h1 @ r(W2) + b2
    h2 = r(silu(z2))
    y = h2 @ r(W3) + b3
    return y[:, :N_OUT], dict(x=x16, z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def loss(params, x, t, emulate16=True):
    y, _ = forward(params, x, emulate16)
    return float(np.mean(np.sum((y - t)**2, axis=1)/3.0))


def backward(params, x, t, emulate16=True):
    """Gradient of loss() with respect to the flat FP32 parameters (padding entries included, zero)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    r = fp16 if emulate16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(params, x, emulate16)
    n = len(x)
    dy = np.zeros((n, N_OUT_PAD))
    dy[:, :N_OUT] = (2.0/3.0)*(y - t)
    dy = r(dy)
    dz2 = r((dy @ r(W3).T)*dsilu(c["z2"]))
    dz1 = r((dz2 @ r(W2).T)*dsilu(c["z1"]))
    gW3 = c["h2"].T @ dy/n
    gW2 = c["h1"].T @ dz2/n
    gW1 = c["x"].T @ dz1/n
    return pack(gW1, dz1.sum(0)/n, gW2, dz2.sum(0)/n, gW3, dy.sum(0)/n)
