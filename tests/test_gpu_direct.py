"""GPU parity of the direct path and the octree, through the C ABI, against the CPU oracle.

Bars: FP64 results within 1e-10 relative of the oracle (the plain form's parity tolerance), tree cells
and interaction lists identical as sets, FP32 applies within 1e-5 relative.
"""
import functools
import os
import sys

import numpy as np
import pytest

import oracle as O
from cppe_amd import capi, synthetic

from .conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@functools.lru_cache(None)
def water(nw, seed, aniso):
    return synthetic.water_box(nw, seed, anisotropic=aniso)


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@functools.lru_cache(None)
def loprop():
    return O.read_potfile(os.path.join(GOLDEN, "loprop_solvated_20.pot.gz"))


def cell_ids(geo, meta):
    return [(int(meta[c, 1]), geo[c, 0], geo[c, 1], geo[c, 2]) for c in range(len(geo))]


def compare_trees(pos, ncrit, theta, order=5):
    say("tree case n=%d ncrit=%d theta=%g" % (len(pos), ncrit, theta))
    t = O.Tree(pos, ncrit=ncrit, order=order, theta=theta)
    geo_o, meta_o, leafs_o = t.cells()
    m2l_o, p2p_o = t.lists()
    S = capi.System(pos, summation="fmm", ncrit=ncrit, order=order, theta=theta)
    ex = S.tree_export()
    st = S.tree_stats()
    ids_o = cell_ids(geo_o, meta_o)
    ids_d = cell_ids(ex["geo"], ex["meta"])
    assert len(ids_o) == len(ids_d) == st["ncells"]
    assert set(ids_o) == set(ids_d)
    # geometry bit-identical per cell identity (r, rmax)
    map_d = {k: c for c, k in enumerate(ids_d)}
    for c, k in enumerate(ids_o):
        d = map_d[k]
        assert geo_o[c, 3] == ex["geo"][d, 3] and geo_o[c, 4] == ex["geo"][d, 4]
        assert meta_o[c, 2] == ex["meta"][d, 2]          # occupied-octant mask
        if meta_o[c, 2] == 0:                               # leaf: same member sites
            mem_o = sorted(int(v) for v in leafs_o[c] if v >= 0)
            s0, cnt = int(ex["start"][d]), int(ex["meta"][d, 3])
            mem_d = sorted(int(v) for v in ex["perm"][s0:s0 + cnt])
            assert mem_o == mem_d
    m2l_so = {(ids_o[b], ids_o[a]) for b, a in m2l_o}
    m2l_sd = {(ids_d[b], ids_d[a]) for b, a in ex["m2l"]}
    p2p_so = {(ids_o[a], ids_o[b]) for a, b in p2p_o}
    p2p_sd = {(ids_d[a], ids_d[b]) for a, b in ex["p2p"]}
    assert len(m2l_o) == st["n_m2l"] and len(p2p_o) == st["n_p2p"]
    assert m2l_so == m2l_sd
    assert p2p_so == p2p_sd
    S.close()


def test_tree_bit_exact_synthetic():
    w = water(1000, 1001, False)
    compare_trees(w.pos, 64, 0.5)
    compare_trees(w.pos, 16, 0.7)


def test_tree_bit_exact_loprop():
    compare_trees(loprop().pos, 64, 0.5)
    compare_trees(loprop().pos, 64, 0.99, order=3)


@pytest.mark.parametrize("damping", [0.0, 2.1304])
def test_apply_b_direct_fp64(damping):
    w = water(1000, 1001, True)
    say("apply_b fp64 damping", damping)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=damping)
    B = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=damping)
    x = np.random.default_rng(3).normal(size=3 * w.n)
    assert relerr(S.apply_b(x), B.apply(x)) < 1e-12
    assert relerr(S.apply_alpha(x), B.apply_diagonal_inverse(x)) < 1e-14
    assert relerr(S.apply_alpha_inv(x), B.apply_diagonal(x)) < 1e-13


def test_apply_b_direct_fp32():
    w = water(1000, 1001, True)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304, precision="mixed")
    B = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=2.1304)
    x = np.random.default_rng(4).normal(size=3 * w.n)
    assert relerr(S.apply_b(x), B.apply(x)) < 1e-5


@pytest.mark.parametrize("m", [0, 1, 2])
@pytest.mark.parametrize("damping", [0.0, 2.1304])
def test_field_direct(m, damping):
    w = water(700, 1007, True)
    say("field m", m, "damping", damping)
    rng = np.random.default_rng(m)
    Ssrc = rng.normal(size=(w.n, O.ncomp(m)))
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=damping)
    F = S.field(m, Ssrc)
    ref = O.direct_field(m, w.pos, Ssrc, w.alpha_iso() if damping else None, w.excl_csr(), damping)
    assert relerr(F, ref) < 1e-11


def test_multipole_fields_golden_loprop():
    """Static multipole fields on the reference's 8,805-site file equal the golden vector."""
    pot = loprop()
    fmuls = np.load(os.path.join(GOLDEN, "loprop_fmm_ref.npz"))["fmuls"]
    S = capi.System(pot.pos, excl=pot.excl_csr())
    F = S.multipole_fields(pot.mul[:, 0], pot.mul[:, 1:4], pot.mul[:, 4:10])
    assert relerr(F, fmuls) < 1e-12
    np.testing.assert_allclose(F, fmuls, rtol=1e-10, atol=1e-14)


def test_multipole_fields_damped():
    w = water(700, 1008, True)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304)
    F = S.multipole_fields(w.charges, w.dipoles, w.quadrupoles)
    al = w.alpha_iso()
    th = w.quadrupoles.copy()
    th[:, [0, 3, 5]] *= 0.5
    ref = (O.direct_field(0, w.pos, w.charges, al, w.excl_csr(), 2.1304)
           - O.direct_field(1, w.pos, w.dipoles, al, w.excl_csr(), 2.1304)
           + O.direct_field(2, w.pos, th, al, w.excl_csr(), 2.1304))
    assert relerr(F, ref) < 1e-11


def test_nuclear_fields_and_energy():
    m = np.load(os.path.join(GOLDEN, "molecules.npz"))
    pot = O.read_potfile(os.path.join(GOLDEN, "pna_6w.pot"))
    z = m["pna_z"].astype(float)
    F = capi.nuclear_fields(pot.pos, z, m["pna_xyz"])
    np.testing.assert_allclose(F, O.nuclear_fields(pot.pos, z, m["pna_xyz"]), rtol=1e-13, atol=1e-14)
    E = capi.nuclear_multipole_energy(pot.pos, pot.kmax, pot.mul, z, m["pna_xyz"])
    np.testing.assert_almost_equal(E, -0.321349401430, decimal=9)


@pytest.mark.parametrize("damping", [0.0, 2.1304])
def test_solve_pna_vs_dense(damping):
    m = np.load(os.path.join(GOLDEN, "molecules.npz"))
    pot = O.read_potfile(os.path.join(GOLDEN, "pna_6w.pot"))
    z = m["pna_z"].astype(float)
    F = O.nuclear_fields(pot.pos, z, m["pna_xyz"]) + O.multipole_fields(pot)
    S = capi.System(pot.pos, pot.pol, pot.alpha_iso(), pot.excl_csr(), damping=damping)
    Bd = S.dense_b()
    B = O.BSystem(pot.pos, pot.pol, pot.excl_csr(), damping=damping)
    np.testing.assert_allclose(Bd, B.dense(), atol=1e-13)
    mu, it, norms = S.solve(F, thresh=1e-12)
    ref = np.linalg.inv(B.dense()) @ F
    np.testing.assert_allclose(mu, ref, atol=1e-10)


@pytest.mark.parametrize("aniso", [False, True])
def test_solve_c1_vs_oracle(aniso):
    w = water(1000, 1001, aniso)
    say("solve c1 aniso", aniso)
    rhs = O.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304)
    B = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=2.1304)
    mu_o, it_o, n_o = B.solve(rhs, thresh=1e-8)
    mu, it, norms = S.solve(rhs, thresh=1e-8)
    assert abs(it - it_o) <= 1
    assert relerr(mu, mu_o) < 1e-9
    e, e_o = -0.5 * mu @ rhs, -0.5 * mu_o @ rhs
    assert abs(e - e_o) <= 1e-10 * abs(e_o)
    # mixed precision: same FP64 stopping test
    Sm = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304, precision="mixed")
    mu_m, it_m, norms_m = Sm.solve(rhs, thresh=1e-8)
    r = rhs - B.apply(mu_m)
    assert np.linalg.norm(r) < 1e-8
    assert abs(-0.5 * mu_m @ rhs - e_o) <= 1e-8 * abs(e_o)


def test_solve_nonconvergence():
    w = water(200, 1009, False)
    rhs = O.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304)
    with pytest.raises(capi.CppeHipError, match="Failed to converge induced dipole moments."):
        S.solve(rhs, thresh=1e-14, maxiter=1)


def test_resident_solve_matches():
    w = water(1000, 1001, True)
    rhs = O.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304)
    mu, it, _ = S.solve(rhs)
    S.upload_rhs(rhs)
    it2, _, ms = S.solve_resident(0)
    assert it2 == it and ms > 0
    np.testing.assert_array_equal(S.download_mu(), mu)  # deterministic reductions


@pytest.mark.parametrize("damping", [0.0, 2.1304])
def test_field_targets_vs_oracle(damping):
    """FP64 direct dipole field at sampled targets (the measuring aid of the tree path's error)."""
    w = water(700, 1007, True)
    rng = np.random.default_rng(17)
    mu = rng.normal(size=(w.n, 3)) * 0.1
    targets = rng.choice(w.n, 97, replace=False)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=damping)
    F = S.field_targets(targets, mu=mu)
    ref = O.direct_field_targets(1, w.pos, mu, targets, w.alpha_iso() if damping else None, w.excl_csr(), damping)
    assert relerr(F.reshape(-1), ref) < 1e-12
