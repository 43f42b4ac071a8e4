"""CPU pins of the generated straight-line FMM operators (cppe_amd/csrc/device/fmm_ops.gen.hh).

The header is compiled for the host (tests/fmm_ops_host_driver.cc) and compared with
  * the reference's own generated fmmgen kernels (oracle/_ref, built from the reference
    sources where the checkout exists): same multi-index ordering, same M/L layouts;
  * the oracle's Coulomb derivatives and its FMM field (always available).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEV = os.path.join(ROOT, "cppe_amd", "csrc", "device")
_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)


def nterms(p):
    return 0 if p < 0 else (p + 1) * (p + 2) * (p + 3) // 6


@pytest.fixture(scope="module", params=[2, 3, 4, 5, 6, 7, 8])
def gen(request, tmp_path_factory):
    p = request.param
    out = tmp_path_factory.mktemp("fmmops") / f"libgen_p{p}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", f"-DCPPE_FMMOPS_ONLY_P={p}", "-I", DEV,
                    os.path.join(ROOT, "tests", "fmm_ops_host_driver.cc"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    dd = ctypes.c_double
    for f in ("gen_p2m", "gen_m2m", "gen_m2l", "gen_l2l", "gen_l2p"):
        getattr(L, f).argtypes = [ctypes.c_int, dd, dd, dd, _d, _d]
    L.gen_m2l_f32.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f, _f]
    return p, L


def sizes(mc, p):
    ms = nterms(p) if mc else nterms(p) - 1
    ls = nterms(p) if mc else nterms(p - 1)
    return ms, ls


def P(a):
    return a.ctypes.data_as(_d)


def test_ops_vs_reference_kernels(gen):
    p, G = gen
    R = O.ref_kernels()
    if R is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    if p > 7:
        pytest.skip("the reference's generated operators cover orders 2-7 (fmm/field_m*.cc dispatchers)")
    rng = np.random.default_rng(p)
    for mc, refm in ((0, 1), (1, 0)):
        ms, ls = sizes(mc, p)
        for _ in range(5):
            d = rng.normal(size=3) * 3
            M = rng.normal(size=ms)
            # M2M
            a = np.zeros(ms)
            b = np.zeros(ms)
            G.gen_m2m(mc, *d, P(M.copy()), P(a))
            R.ref_m2m(refm, p, *d, P(M.copy()), P(b))
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
            # M2L
            a = np.zeros(ls)
            b = np.zeros(ls)
            G.gen_m2l(mc, *d, P(M.copy()), P(a))
            R.ref_m2l(refm, p, *d, P(M.copy()), P(b))
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-12 * np.abs(b).max())
            # L2L
            Lp = rng.normal(size=ls)
            a = np.zeros(ls)
            b = np.zeros(ls)
            G.gen_l2l(mc, *d, P(Lp.copy()), P(a))
            R.ref_l2l(refm, p, *d, P(Lp.copy()), P(b))
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
            # L2P
            a = np.zeros(3)
            b = np.zeros(3)
            G.gen_l2p(mc, *d, P(Lp.copy()), P(a))
            R.ref_l2p(refm, p, *d, P(Lp.copy()), P(b))
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
            # P2M == M2M of a point multipole
            if mc == 0:
                S = rng.normal(size=3)
                Q = np.zeros(ms)
                Q[:3] = S
            else:
                S = np.zeros(10)
                S[0] = rng.normal()
                Q = np.zeros(ms)
                Q[0] = S[0]
            a = np.zeros(ms)
            b = np.zeros(ms)
            G.gen_p2m(mc, *d, P(S), P(a))
            R.ref_m2m(refm, p, *d, P(Q), P(b))
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-13 * np.abs(b).max())


def test_m2l_coulomb_derivatives(gen):
    """L_0 of a unit moment in slot j equals D_j (oracle recursion), for every j."""
    p, G = gen
    r = np.array([2.3, -1.1, 0.7])
    D = O.coulomb_derivs(p, r)
    ms, ls = sizes(1, p)
    for j in range(ms):
        M = np.zeros(ms)
        M[j] = 1.0
        L = np.zeros(ls)
        G.gen_m2l(1, *r, P(M), P(L))
        assert abs(L[0] - D[j]) <= 1e-12 * max(1.0, abs(D[j]))


def test_m2l_f32_close_to_f64(gen):
    p, G = gen
    rng = np.random.default_rng(11)
    ms, ls = sizes(0, p)
    d = np.array([5.0, -3.0, 2.0])
    M = rng.normal(size=ms)
    a = np.zeros(ls)
    G.gen_m2l(0, *d, P(M), P(a))
    b = np.zeros(ls, dtype=np.float32)
    G.gen_m2l_f32(0, *d.astype(np.float32), M.astype(np.float32).ctypes.data_as(_f), b.ctypes.data_as(_f))
    assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


def test_p2m_mc_sources(gen):
    """Combined sources: charge, dipole and quadrupole slots of P2M equal shifted point
    multipoles (M_j = S_j at d = 0)."""
    p, G = gen
    S = np.arange(1, 11, dtype=float)
    ms, _ = sizes(1, p)
    M = np.zeros(ms)
    G.gen_p2m(1, 0.0, 0.0, 0.0, P(S), P(M))
    np.testing.assert_array_equal(M[:10], S)
    assert not M[10:].any()
