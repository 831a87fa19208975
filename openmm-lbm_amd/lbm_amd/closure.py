"""The learned coupling closure (include/lbm_amd_closure.h): nearest-node coupling corrected toward the Keys stencil by
an 87 -> 64 -> 64 -> 3 network that the library trains and evaluates with its own FP16 MFMA kernels.

Typical use::

    features, targets = pairs(context)              # at states of a Keys-coupled run, repeated over snapshots
    net = MLPClosure(seed=1)
    losses = net.train(features, targets)           # GPU training
    net.attach(context)                             # the context now couples through the network (Keys fallback)
"""
import ctypes

import numpy as np

from . import _lib

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I64 = ctypes.POINTER(ctypes.c_int64)

NUM_INPUTS = 87
NUM_OUTPUTS = 3
NUM_PARAMS = 11408
# Flat parameter layout: W1[i][o] (96 x 64), b1, W2[i][o] (64 x 64), b2, W3[i][o] (64 x 16), b3 (16).
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 6144, 6208, 10304, 10368, 11392


class Normalization(ctypes.Structure):
    _fields_ = [("u_scale", ctypes.c_double), ("v_scale", ctypes.c_double), ("out_scale", ctypes.c_double),
                ("u_bound", ctypes.c_double), ("v_bound", ctypes.c_double)]


CLOSURE_FUNCTIONS = {
    "lbm_mlp_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    "lbm_mlp_destroy": (ctypes.c_int, [_P]),
    "lbm_mlp_get_parameters": (ctypes.c_int, [_P, _F, ctypes.c_size_t]),
    "lbm_mlp_set_parameters": (ctypes.c_int, [_P, _F, ctypes.c_size_t]),
    "lbm_mlp_get_normalization": (ctypes.c_int, [_P, ctypes.POINTER(Normalization)]),
    "lbm_mlp_set_normalization": (ctypes.c_int, [_P, ctypes.POINTER(Normalization)]),
    "lbm_mlp_train": (ctypes.c_int, [_P, _F, _F, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_uint64, _D]),
    "lbm_mlp_forward": (ctypes.c_int, [_P, _F, ctypes.c_int64, _F]),
    "lbm_mlp_forward_noalloc": (ctypes.c_int, [_P, _F, ctypes.c_int64, _F]),
    "lbm_mlp_gradients": (ctypes.c_int, [_P, _F, _F, ctypes.c_int64, _F, _D]),
    "lbm_closure_pairs": (ctypes.c_int, [_P, _D, _D, _F, _F]),
    "lbm_set_closure": (ctypes.c_int, [_P, _P]),
    "lbm_closure_stats": (ctypes.c_int, [_P, _I64, _I64]),
}


def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


def _lib_of(library):
    return _lib.declare(library if library is not None else _lib.load_library(), CLOSURE_FUNCTIONS)


def pairs(context, positions=None, velocities=None):
    """Training pairs at the context's current state: (features [n][87], targets [n][3]) of its coupled particles,
    each in the particle's frame (the reflection and axis order that make its sub-node offset non-negative and
    sorted; include/lbm_amd_closure.h)."""
    lib = _lib_of(context.lib)
    pos = np.ascontiguousarray(context.positions if positions is None else positions, dtype=np.float64)
    vel = np.ascontiguousarray(context.velocities if velocities is None else velocities, dtype=np.float64)
    n = len(context.coupled)
    feats = np.zeros((max(n, 1), NUM_INPUTS), dtype=np.float32)
    targs = np.zeros((max(n, 1), NUM_OUTPUTS), dtype=np.float32)
    _lib.check(lib, lib.lbm_closure_pairs(context.handle, pos.ctypes.data_as(_D), vel.ctypes.data_as(_D),
                                          feats.ctypes.data_as(_F), targs.ctypes.data_as(_F)))
    return feats[:n], targs[:n]


def stats(context):
    """(couplings through the network, couplings through the Keys fallback) since the last call."""
    lib = _lib_of(context.lib)
    a, b = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(lib, lib.lbm_closure_stats(context.handle, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


class MLPClosure:
    def __init__(self, seed=1, library=None):
        self.lib = _lib_of(library)
        handle = ctypes.c_void_p()
        self._check(self.lib.lbm_mlp_create(int(seed), ctypes.byref(handle)))
        self.handle = handle

    def _check(self, status):
        _lib.check(self.lib, status)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.lbm_mlp_destroy(h)
            self.handle = None

    def getParameters(self):
        p = np.zeros(NUM_PARAMS, dtype=np.float32)
        self._check(self.lib.lbm_mlp_get_parameters(self.handle, p.ctypes.data_as(_F), NUM_PARAMS))
        return p

    def setParameters(self, params):
        p = _f32(params, (NUM_PARAMS,))
        self._check(self.lib.lbm_mlp_set_parameters(self.handle, p.ctypes.data_as(_F), NUM_PARAMS))

    def getNormalization(self):
        n = Normalization()
        self._check(self.lib.lbm_mlp_get_normalization(self.handle, ctypes.byref(n)))
        return {k: getattr(n, k) for k, _ in Normalization._fields_}

    def setNormalization(self, **values):
        n = Normalization(**{k: float(values[k]) for k, _ in Normalization._fields_})
        self._check(self.lib.lbm_mlp_set_normalization(self.handle, ctypes.byref(n)))

    def train(self, features, targets, epochs=40, batch=2048, lr=2e-3, decay=0.92, seed=1):
        """Adam on the GPU; sets the normalization from the data; returns the mean loss of each epoch."""
        x = _f32(features, (-1, NUM_INPUTS))
        t = _f32(targets, (-1, NUM_OUTPUTS))
        losses = np.zeros(int(epochs))
        self._check(self.lib.lbm_mlp_train(self.handle, x.ctypes.data_as(_F), t.ctypes.data_as(_F), len(x), int(epochs),
                                           int(batch), float(lr), float(decay), int(seed), losses.ctypes.data_as(_D)))
        return losses

    def forward(self, features):
        """The network's Du (lattice units, in the frame of the inputs) for raw inputs."""
        x = _f32(features, (-1, NUM_INPUTS))
        out = np.zeros((len(x), NUM_OUTPUTS), dtype=np.float32)
        self._check(self.lib.lbm_mlp_forward(self.handle, x.ctypes.data_as(_F), len(x), out.ctypes.data_as(_F)))
        return out

    def gradients(self, features, targets):
        """(gradient of the batch loss with respect to the parameters, loss), from the training kernels."""
        x = _f32(features, (-1, NUM_INPUTS))
        t = _f32(targets, (-1, NUM_OUTPUTS))
        g = np.zeros(NUM_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.lbm_mlp_gradients(self.handle, x.ctypes.data_as(_F), t.ctypes.data_as(_F), len(x),
                                               g.ctypes.data_as(_F), ctypes.byref(loss)))
        return g, loss.value

    def attach(self, context):
        self._check(self.lib.lbm_set_closure(context.handle, self.handle))
        context._closure = self          # keeps the network alive while attached

    @staticmethod
    def detach(context):
        lib = _lib_of(context.lib)
        _lib.check(lib, lib.lbm_set_closure(context.handle, None))
        context._closure = None
