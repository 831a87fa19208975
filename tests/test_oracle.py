"""Pins the FP64 CPU oracle (oracle/lbm_oracle.cpp) before it is trusted as the reference for the GPU path.

 - Philox4x32-10 against the published known-answer vectors of Random123 (kat_vectors, philox4x32_10).
 - The D3Q19 algebra against the reference's own openmmapi/include/internal/D3Q19.h, compiled into oracle/_ref.
 - Identities of the algebra and of the stencils that a dropped term, a wrong sign or index would break.
 The closed forms of the reference's own tests run on the oracle and on the GPU library in test_closed_forms.py.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib
from lbm_amd import LBMForce
from parity import CX, CY, CZ, W, Setup, equilibrium_state, total_moments

_D = ctypes.POINTER(ctypes.c_double)


def dptr(a):
    return a.ctypes.data_as(_D)


@pytest.fixture(scope="module")
def lib():
    return oracle_lib.load_oracle()


# ------------------------------------------------------------------------------------------------
# Philox4x32-10 known answers (Salmon et al. SC'11; Random123 kat_vectors).
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,expected", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff]*4, [0xffffffff]*2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(lib, ctr, key, expected):
    c = (ctypes.c_uint32*4)(*ctr)
    k = (ctypes.c_uint32*2)(*key)
    out = (ctypes.c_uint32*4)()
    lib.oracle_philox4x32(c, k, out)
    assert list(out) == expected


def test_normals_are_standard(lib):
    """Box-Muller normals from Philox: mean 0, variance 1, kurtosis 3 over 2^17 samples (5 sigma bounds)."""
    count = 16
    samples = []
    buf = np.zeros(count)
    for index in range(8192):
        lib.oracle_normals(7, 0x464C5549, index, 11, count, dptr(buf))
        samples.append(buf.copy())
    x = np.concatenate(samples)
    n = x.size
    assert abs(x.mean()) < 5/math.sqrt(n)
    assert abs(x.var() - 1) < 5*math.sqrt(2/n)
    assert abs((x**4).mean() - 3) < 5*math.sqrt(96/n)
    # Different steps and indices give different numbers.
    a, b = np.zeros(4), np.zeros(4)
    lib.oracle_normals(7, 1, 0, 0, 4, dptr(a))
    lib.oracle_normals(7, 1, 0, 1, 4, dptr(b))
    assert not np.array_equal(a, b)


# ------------------------------------------------------------------------------------------------
# D3Q19 algebra against the reference's own header (compiled in oracle/_ref where the tree was present).
# ------------------------------------------------------------------------------------------------
def random_nodes(count, seed=0):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        dr = rng.uniform(-0.05, 0.05)
        u = rng.uniform(-0.1, 0.1, 3)
        df = rng.uniform(-0.01, 0.01, 19)
        f = rng.uniform(-1e-3, 1e-3, 3)
        r = rng.normal(size=19)
        yield dr, u, df, f, r


@pytest.mark.skipif(oracle_lib.load_ref() is None, reason="oracle/_ref is built only where the reference tree is present")
def test_algebra_matches_reference_header(lib):
    ref = oracle_lib.load_ref()
    a, b = np.zeros(19), np.zeros(19)
    p, q = np.zeros(6), np.zeros(6)
    for dr, u, df, f, r in random_nodes(200):
        lib.oracle_equilibrium_deviation(dr, *u, dptr(a))
        ref.ref_equilibrium_deviation(dr, *u, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-16)
        j = u*(1 + dr)
        lib.oracle_pi_neq(dptr(df), dr, *j, dptr(p))
        ref.ref_pi_neq(dptr(df), dr, *j, dptr(q))
        assert np.allclose(p, q, rtol=0, atol=1e-16)
        lib.oracle_regularized_neq(dptr(p), dptr(a))
        ref.ref_regularized_neq(dptr(p), dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-17)
        lib.oracle_guo(*u, *f, dptr(a))
        ref.ref_guo(*u, *f, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-18)
        rr = np.ascontiguousarray(r)
        lib.oracle_collision_noise(dptr(rr), 0.3, 0.7, dptr(a))
        ref.ref_collision_noise(dptr(rr), 0.3, 0.7, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-15)
        lib.oracle_equilibrium_noise(dptr(rr), 0.2, 0.3, 0.7, dptr(a))
        ref.ref_equilibrium_noise(dptr(rr), 0.2, 0.3, 0.7, dptr(b))
        assert np.allclose(a, b, rtol=0, atol=1e-15)
    for k in range(19):
        for qq in range(19):
            assert lib.oracle_mode(k, qq) == ref.ref_mode(k, qq)
    for omega in (0.7, 1.0, 1.6):
        assert ref.ref_boundary_stress_weight(omega) == pytest.approx((1 - omega)**2/(omega*(2 - omega)))


def test_algebra_matches_reference_vectors(lib):
    """The oracle's D3Q19 algebra against tests/golden/d3q19_vectors.npz, input/output vectors of the reference's own
    D3Q19.h (equilibriumDeviation, nonEquilibriumMomentFromDeviation, regularizedNonEquilibrium, guoForcing, the
    collision and equilibrium noise, the mode basis) written by tests/golden/make_d3q19_vectors.py: the same check as
    above where the reference tree is absent."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "d3q19_vectors.npz"))
    a, p = np.zeros(19), np.zeros(6)
    for n in range(len(g["dr"])):
        dr, u, f = float(g["dr"][n]), g["u"][n], g["force"][n]
        lib.oracle_equilibrium_deviation(dr, *u, dptr(a))
        assert np.allclose(a, g["equilibrium"][n], rtol=0, atol=1e-16)
        df = np.ascontiguousarray(g["df"][n])
        lib.oracle_pi_neq(dptr(df), dr, *(u*(1 + dr)), dptr(p))
        assert np.allclose(p, g["pi_neq"][n], rtol=0, atol=1e-16)
        pp = np.ascontiguousarray(g["pi_neq"][n])
        lib.oracle_regularized_neq(dptr(pp), dptr(a))
        assert np.allclose(a, g["regularized_neq"][n], rtol=0, atol=1e-17)
        lib.oracle_guo(*u, *f, dptr(a))
        assert np.allclose(a, g["guo"][n], rtol=0, atol=1e-18)
        r = np.ascontiguousarray(g["normals"][n])
        lib.oracle_collision_noise(dptr(r), 0.3, 0.7, dptr(a))
        assert np.allclose(a, g["collision_noise"][n], rtol=0, atol=1e-15)
        lib.oracle_equilibrium_noise(dptr(r), 0.2, 0.3, 0.7, dptr(a))
        assert np.allclose(a, g["equilibrium_noise"][n], rtol=0, atol=1e-15)
    modes = np.array([[lib.oracle_mode(k, q) for q in range(19)] for k in range(19)])
    assert np.array_equal(modes, g["modes"])


def test_algebra_identities(lib):
    """sum dfeq = drho, sum c dfeq = rho u; sum S = 0, sum c S = F (D3Q19.h:123-126); Pi_neq of the equilibrium is 0;
    the regularized part carries no mass or momentum and reproduces Pi (H2 projection)."""
    eq = np.zeros(19)
    s = np.zeros(19)
    pi = np.zeros(6)
    neq = np.zeros(19)
    c = np.stack([CX, CY, CZ])
    for dr, u, df, f, r in random_nodes(100, seed=1):
        lib.oracle_equilibrium_deviation(dr, *u, dptr(eq))
        assert eq.sum() == pytest.approx(dr, abs=1e-15)
        assert np.allclose(c @ eq, (1 + dr)*u, atol=1e-15)
        lib.oracle_guo(*u, *f, dptr(s))
        assert abs(s.sum()) < 1e-18
        assert np.allclose(c @ s, f, atol=1e-18)
        lib.oracle_pi_neq(dptr(eq), dr, *((1 + dr)*u), dptr(pi))
        assert np.allclose(pi, 0, atol=1e-15)
        pi_in = np.ascontiguousarray(np.random.default_rng(2).normal(size=6)*1e-3)
        lib.oracle_regularized_neq(dptr(pi_in), dptr(neq))
        assert abs(neq.sum()) < 1e-17 and np.allclose(c @ neq, 0, atol=1e-17)
        moments = np.array([(CX*CX*neq).sum(), (CY*CY*neq).sum(), (CZ*CZ*neq).sum(), (CX*CY*neq).sum(),
                            (CX*CZ*neq).sum(), (CY*CZ*neq).sum()])
        trace = pi_in[:3].sum()/3
        assert np.allclose(moments[:3] - (neq.sum())/3, pi_in[:3] - 0*trace, atol=1e-15)
        assert np.allclose(moments[3:], pi_in[3:], atol=1e-16)


def test_fluctuation_basis(lib):
    """Orthogonality sum_q w_q e_k e_l = b_k delta_kl (TestLBMFluctuations.h:58-69); the random part conserves mass and
    momentum and gives mode k the amplitude phi_k r_k."""
    e = np.array([[lib.oracle_mode(k, q) for q in range(19)] for k in range(19)])
    gram = (e*W) @ e.T
    norms = np.array([1, 1/3, 1/3, 1/3, 2/9, 2/9, 2/9, 1/9, 1/9, 1/9, 2/27, 2/27, 2/27, 1/18, 1/18, 1/18, 7/162, 5/126,
                      1/30])
    assert np.allclose(gram, np.diag(norms), atol=1e-15)
    r = np.ascontiguousarray(np.random.default_rng(3).normal(size=19))
    xi = np.zeros(19)
    lib.oracle_collision_noise(dptr(r), 0.4, 0.9, dptr(xi))
    assert abs(xi.sum()) < 1e-15 and np.allclose(np.stack([CX, CY, CZ]) @ xi, 0, atol=1e-15)
    modes = e @ xi
    for k in range(4, 19):
        amp = 0.4 if k < 10 else 0.9
        assert modes[k] == pytest.approx(amp*math.sqrt(norms[k])*r[k - 4], abs=1e-14)


# ------------------------------------------------------------------------------------------------
# Stencils (TestLBMStencils.h:27-124).
# ------------------------------------------------------------------------------------------------
def axis(lib, stencil, s, n=16, open_axis=False):
    idx = (ctypes.c_int*4)()
    w = np.zeros(4)
    lib.oracle_axis_weights(stencil, s, n, 1 if open_axis else 0, idx, dptr(w))
    width = {LBMForce.NearestNode: 1, LBMForce.Trilinear: 2, LBMForce.ThreePoint: 3, LBMForce.Keys: 4}[stencil]
    return np.array(idx[:width]), w[:width]


@pytest.mark.parametrize("stencil", [LBMForce.Trilinear, LBMForce.ThreePoint, LBMForce.Keys])
def test_stencil_moments(lib, stencil):
    for s in np.linspace(3.0, 4.0, 41):
        idx, w = axis(lib, stencil, s)
        assert w.sum() == pytest.approx(1.0, abs=1e-14)
        assert (w*(idx - s)).sum() == pytest.approx(0.0, abs=1e-14)
        if stencil == LBMForce.Keys:
            assert (w*(idx - s)**2).sum() == pytest.approx(0.0, abs=1e-13)
        if stencil == LBMForce.ThreePoint:
            assert (w*w).sum() == pytest.approx(0.5, abs=1e-13)


def test_stencil_at_node_and_wrap(lib):
    for stencil in (LBMForce.Trilinear, LBMForce.Keys):
        idx, w = axis(lib, stencil, 5.0)
        assert w[list(idx).index(5)] == pytest.approx(1.0) and np.abs(w).sum() == pytest.approx(1.0)
    idx, w = axis(lib, LBMForce.Keys, 0.3, n=16)
    assert list(idx) == [15, 0, 1, 2]
    idx, w = axis(lib, LBMForce.Keys, 0.3, n=16, open_axis=True)
    assert list(idx) == [0, 0, 1, 2] and w[1] == 0.0 and w.sum() == pytest.approx(1.0)
    idx, w = axis(lib, LBMForce.ThreePoint, 0.2, n=2)
    assert w.sum() == pytest.approx(1.0) and w[2] == 0.0
