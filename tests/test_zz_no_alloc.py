"""lbm_mlp_forward_noalloc (include/lbm_amd_closure.h), the network's forward pass with no device allocation: its
outputs equal lbm_mlp_forward's bit for bit (the same kernels on the same inputs), on two calls in a row, the second
with fewer rows, and sizes beyond LBM_MLP_NOALLOC_MAX_ROWS are refused before any device work."""
import ctypes

import numpy as np
import pytest

from lbm_amd import _lib
from lbm_amd.closure import MLPClosure, NUM_INPUTS, NUM_OUTPUTS, _lib_of
from test_closure_oracle import random_params

pytestmark = pytest.mark.gpu

MAX_ROWS = 16384
_F = ctypes.POINTER(ctypes.c_float)


def raw_batch(rng, count):
    raw = np.zeros((count, NUM_INPUTS), dtype=np.float32)
    raw[:, 0:3] = rng.uniform(-1, 1, size=(count, 3))
    raw[:, 3:84] = rng.normal(scale=0.025, size=(count, 81))
    raw[:, 84:87] = rng.normal(scale=0.25, size=(count, 3))
    return raw


def forward_noalloc(net, raw):
    lib = _lib_of(None)
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    out = np.full((len(raw), NUM_OUTPUTS), np.nan, dtype=np.float32)
    _lib.check(lib, lib.lbm_mlp_forward_noalloc(net.handle, raw.ctypes.data_as(_F), len(raw), out.ctypes.data_as(_F)))
    return out


def test_noalloc_forward_matches_forward():
    rng = np.random.default_rng(31)
    net = MLPClosure(seed=4)
    net.setParameters(random_params(rng, 0.3).astype(np.float32))
    net.setNormalization(u_scale=40.0, v_scale=4.0, out_scale=60.0, u_bound=1e3, v_bound=1e3)
    for count in (MAX_ROWS - 5, 777):
        raw = raw_batch(rng, count)
        want = net.forward(raw)
        got = forward_noalloc(net, raw)
        print(f"{count} rows: max |noalloc - forward| = {np.abs(got - want).max():.3e}", flush=True)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got, want)


def test_noalloc_forward_refuses_too_many_rows():
    net = MLPClosure(seed=4)
    lib = _lib_of(None)
    raw = np.zeros((1, NUM_INPUTS), dtype=np.float32)
    out = np.zeros((1, NUM_OUTPUTS), dtype=np.float32)
    for n in (MAX_ROWS + 1, 0):
        status = lib.lbm_mlp_forward_noalloc(net.handle, raw.ctypes.data_as(_F), n, out.ctypes.data_as(_F))
        assert status == _lib.LBM_ERROR_INVALID
        assert "LBM_MLP_NOALLOC_MAX_ROWS" in lib.lbm_last_error().decode()
