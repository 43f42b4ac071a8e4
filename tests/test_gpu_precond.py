"""GPU checks of the local Neumann preconditioner (cppe_hip_set_preconditioner), through the C ABI.

* the operator: on a direct-summation system the device CG's first residual norms equal those of
  a CPU PCG with the same preconditioner built independently (dense B from the oracle, pairs
  within 3 A from a k-d tree, z = alpha r + w alpha T_loc alpha r), so the pair set, the tensor
  blocks and the weight are the ones stated;
* solves: FP64 and mixed, single systems, a snapshot batch and several right-hand sides stop on
  the FP64 residual (checked with the oracle's operator) with fewer iterations than the Jacobi
  form; the host API option solver_preconditioner gives the Jacobi form's energies.
The reference preconditions with alpha r only (electric_fields.cc:191-239); this one has no
counterpart there.
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
DAMP = 2.1304
RC = capi.System.PRECOND_RADIUS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@functools.lru_cache(None)
def water(nw, seed):
    return synthetic.water_box(nw, seed, anisotropic=True)


def alpha_blocks(w):
    A6 = w.pol
    al = np.zeros((w.n, 3, 3))
    al[:, 0, 0] = A6[:, 0]
    al[:, 0, 1] = al[:, 1, 0] = A6[:, 1]
    al[:, 0, 2] = al[:, 2, 0] = A6[:, 2]
    al[:, 1, 1] = A6[:, 3]
    al[:, 1, 2] = al[:, 2, 1] = A6[:, 4]
    al[:, 2, 2] = A6[:, 5]
    return al


def test_operator_matches_independent_pcg():
    from scipy.spatial import cKDTree
    w = water(500, 31)
    n = w.n
    rhs = O.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Bd = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=DAMP).dense(np.arange(n, dtype=np.int32))
    al = alpha_blocks(w)
    pos = w.pos.reshape(-1, 3)
    T = np.zeros_like(Bd)
    for i, j in cKDTree(pos).query_pairs(RC, output_type="ndarray"):
        T[3 * i:3 * i + 3, 3 * j:3 * j + 3] = -Bd[3 * i:3 * i + 3, 3 * j:3 * j + 3]
        T[3 * j:3 * j + 3, 3 * i:3 * i + 3] = -Bd[3 * j:3 * j + 3, 3 * i:3 * i + 3]
    weight = 0.8

    def amul(v):
        return np.einsum("nij,nj->ni", al, v.reshape(-1, 3)).ravel()

    def M(r):
        a = amul(r)
        return a + weight * amul(T @ a)

    x = amul(rhs)
    r = rhs - Bd @ x
    z = M(r)
    p = z.copy()
    norms = []
    for _ in range(60):
        Ap = Bd @ p
        rz = r @ z
        a = rz / (p @ Ap)
        x += a * p
        r -= a * Ap
        norms.append(np.linalg.norm(r))
        if norms[-1] < 1e-8:
            break
        z = M(r)
        p = z + (z @ r) / rz * p
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=DAMP)
    S.set_preconditioner("local", RC, weight)
    mu, it, dn = S.solve(rhs, thresh=1e-8, maxiter=60)
    say("local PCG: device", it, "iterations, CPU", len(norms))
    assert it == len(norms)
    np.testing.assert_allclose(dn[:6], norms[:6], rtol=1e-5)
    np.testing.assert_allclose(mu, x, rtol=0, atol=1e-7 * np.abs(x).max())
    S.set_preconditioner("jacobi")
    _, it_j, _ = S.solve(rhs, thresh=1e-8, maxiter=60)
    assert it < it_j


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fmm_solves_converge_with_fewer_iterations(precision):
    w = water(2000, 32)
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr(), summation="fmm")
    rhs = Fn + Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)
    Bo = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=DAMP, scheme="fmm")
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP, precision=precision)
    S.upload_rhs(rhs)
    its = {}
    for kind in ("jacobi", "local"):
        S.set_preconditioner(kind)
        for graphs in (True, False):
            S.set_graphs(graphs)
            it, norms, _ = S.solve_resident(0, 1e-8, 100)
            mu = S.download_mu()
            res = np.linalg.norm(rhs - Bo.apply(mu))
            say(precision, kind, "graphs", graphs, "iterations", it, "oracle residual", res)
            assert res < 1e-8
            its[kind, graphs] = it
    assert its["local", True] == its["local", False]
    assert its["local", True] < its["jacobi", True]


def test_batch_and_multi_rhs():
    sizes = [150, 420, 300]
    snaps = [synthetic.water_box(nw, 7300 + k, anisotropic=True) for k, nw in enumerate(sizes)]
    direct = [False, True, False]
    rhs = np.concatenate([O.nuclear_fields(s.pos, s.atoms_z.astype(float), s.atoms_pos) for s in snaps])
    B = capi.Batch(snaps, direct, ncrit=32, damping=DAMP, precision="fp64")
    B.set_preconditioner("local")
    B.upload_rhs(rhs)
    B.solve_resident(0, 1e-8, 100)
    it, rn, conv = B.segment_results()
    assert conv.all()
    mu = B.split(B.download_mu())
    for g, s in enumerate(snaps):
        Bo = O.BSystem(s.pos, s.pol, s.excl_csr(), damping=DAMP, scheme="direct" if direct[g] else "fmm", ncrit=32)
        assert np.linalg.norm(B.split(rhs)[g] - Bo.apply(mu[g])) < 1e-8, g
        # the snapshot alone, same preconditioner: same iterations (no pair crosses snapshots)
        S = capi.System(s.pos, s.pol, s.alpha_iso(), s.excl_csr(), summation="direct" if direct[g] else "fmm",
                        ncrit=32, damping=DAMP)
        S.set_preconditioner("local")
        _, its, _ = S.solve(B.split(rhs)[g], thresh=1e-8, maxiter=100)
        assert its == it[g], g
    B.close()
    w = water(600, 33)
    R = np.zeros((3, 3 * w.n))
    for a in range(3):
        R[a, a::3] = 1.0
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP)
    S.set_preconditioner("local")
    mus, itm, _ = S.solve_multi(R, thresh=1e-8, maxiter=100)
    for a in range(3):
        mu1, it1, _ = S.solve(R[a], thresh=1e-8, maxiter=100)
        assert it1 == itm[a]
        np.testing.assert_allclose(mus[a], mu1, rtol=0, atol=1e-12 * np.abs(mu1).max())


def test_host_api_option():
    cppe = pytest.importorskip("cppe_amd.pycppe")
    m = np.load(os.path.join(GOLDEN, "molecules.npz"))
    mol = cppe.Molecule()
    for z, c in zip(m["pna_z"], m["pna_xyz"]):
        mol.append(cppe.Atom(int(z), *c))
    out = {}
    for pc in ("jacobi", "local"):
        opts = {"potfile": os.path.join(GOLDEN, "pna_6w.pot"), "damp_induced": True, "solver_preconditioner": pc}
        s = cppe.CppeState(opts, mol, lambda x: None)
        s.calculate_static_energies_and_fields()
        s.update_induced_moments(np.zeros_like(s.static_fields), False)
        out[pc] = (s.total_energy, s.last_iterations)
    assert abs(out["local"][0] - out["jacobi"][0]) <= 1e-8 * abs(out["jacobi"][0])
    with pytest.raises(ValueError, match="solver_preconditioner"):
        s = cppe.CppeState({"potfile": os.path.join(GOLDEN, "pna_6w.pot"), "solver_preconditioner": "ilu"}, mol,
                           lambda x: None)
        s.calculate_static_energies_and_fields()
        s.update_induced_moments(np.zeros_like(s.static_fields), False)
