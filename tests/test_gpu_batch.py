"""GPU checks of snapshot batches (one handle, one tree per snapshot, a CG recurrence per snapshot).

* forest: the cells and lists of every snapshot are bit-identical to the oracle's tree of that
  snapshot alone (FMM snapshots), or a single leaf with itself as its only near partner (direct)
* static fields of a batch equal each snapshot's own fields (direct: oracle direct sum, 1e-11;
  FMM: the single-system FMM, 1e-12)
* FP64 and mixed solves: every snapshot stops on its own FP64 residual, verified with the
  oracle's B of that snapshot; dipoles match the single-snapshot solves
* graph replay and kernel-by-kernel launch give bitwise identical dipoles
"""
import functools
import sys

import numpy as np
import pytest

import oracle as O
from cppe_amd import capi, synthetic

pytestmark = pytest.mark.gpu
DAMP = 2.1304


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@functools.lru_cache(None)
def snapshots():
    # uneven sizes; small ones summed directly, the rest through the tree
    sizes = [60, 230, 410, 95, 700, 330, 520]
    snaps = [synthetic.water_box(nw, 7000 + k, anisotropic=True) for k, nw in enumerate(sizes)]
    direct = [s.n < 1000 for s in snaps]
    return snaps, direct


def rhs_of(w, summation, ncrit):
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr(), summation=summation, ncrit=ncrit)
    F = Fn + Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)
    Sm.close()
    return F


def batch_rhs(snaps, direct):
    """Static fields of all snapshots at once: an undamped batch (damp_multipole off)."""
    Fn = np.concatenate([capi.nuclear_fields(s.pos, s.atoms_z.astype(float), s.atoms_pos) for s in snaps])
    q = np.concatenate([s.charges for s in snaps])
    mu = np.concatenate([s.dipoles for s in snaps])
    th = np.concatenate([s.quadrupoles for s in snaps])
    Bs = capi.Batch(snaps, direct, ncrit=32)
    F = Fn + Bs.multipole_fields(q, mu, th)
    Bs.close()
    return F


def cell_ids(geo, meta, rows):
    return [(int(meta[c, 1]), geo[c, 0], geo[c, 1], geo[c, 2]) for c in rows]


def test_forest_matches_single_trees():
    snaps, direct = snapshots()
    say("forest of", len(snaps))
    ncrit = 32
    B = capi.Batch(snaps, direct, ncrit=ncrit, damping=DAMP)
    ex = B.tree_export()
    st = B.tree_stats()
    start, meta, geo = ex["start"], ex["meta"], ex["geo"]
    n_m2l = n_p2p = 0
    for g, (s, d) in enumerate(zip(snaps, direct)):
        a, b = int(B.seg_off[g]), int(B.seg_off[g + 1])
        rows = [c for c in range(st["ncells"]) if a <= start[c] < b]
        if d:
            assert len(rows) == 1 and meta[rows[0], 1] == 0 and meta[rows[0], 3] == s.n
            (r0,) = rows
            assert [tuple(p) for p in ex["p2p"] if p[0] == r0 or p[1] == r0] == [(r0, r0)]
            n_p2p += 1
            continue
        t = O.Tree(s.pos, ncrit=ncrit)
        geo_o, meta_o, leafs_o = t.cells()
        ids_o = cell_ids(geo_o, meta_o, range(t.ncells))
        ids_d = cell_ids(geo, meta, rows)
        assert sorted(ids_o) == sorted(ids_d)
        local = {c: k for c, k in zip(rows, ids_d)}
        for c, k in zip(rows, ids_d):
            if meta[c, 2] == 0:  # leaf members, as local indices
                mem_d = sorted(int(v) - a for v in ex["perm"][start[c]:start[c] + meta[c, 3]])
                co = ids_o.index(k)
                assert mem_d == sorted(int(v) for v in leafs_o[co] if v >= 0)
        m2l_o, p2p_o = t.lists()
        m2l_d = {(local[int(u)], local[int(v)]) for u, v in ex["m2l"] if int(v) in local}
        p2p_d = {(local[int(u)], local[int(v)]) for u, v in ex["p2p"] if int(u) in local}
        assert m2l_d == {(ids_o[u], ids_o[v]) for u, v in m2l_o}
        assert p2p_d == {(ids_o[u], ids_o[v]) for u, v in p2p_o}
        n_m2l += len(m2l_o)
        n_p2p += len(p2p_o)
    assert st["n_m2l"] == n_m2l and st["n_p2p"] == n_p2p


def test_batch_static_fields():
    snaps, direct = snapshots()
    B = capi.Batch(snaps, direct, ncrit=32, damping=DAMP)
    F = B.split(batch_rhs(snaps, direct))
    for g, (s, d) in enumerate(zip(snaps, direct)):
        ref = rhs_of(s, "direct" if d else "fmm", 32)
        err = np.abs(F[g] - ref).max() / np.abs(ref).max()
        say("snapshot", g, "direct" if d else "fmm", "static field rel err %.2e" % err)
        assert err < (1e-11 if d else 1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_solve(precision):
    snaps, direct = snapshots()
    say("batch solve", precision)
    B = capi.Batch(snaps, direct, ncrit=32, damping=DAMP, precision=precision)
    rhs = batch_rhs(snaps, direct)
    B.upload_rhs(rhs)
    it, _, _ = B.solve_resident(0, 1e-8, 100)
    iters, rn, conv = B.segment_results()
    assert conv.all() and (rn < 1e-8).all() and it == iters.max()
    mu = B.split(B.download_mu())
    F = B.split(rhs)
    for g, (s, d) in enumerate(zip(snaps, direct)):
        Bo = O.BSystem(s.pos, s.pol, s.excl_csr(), damping=DAMP, scheme="direct" if d else "fmm", ncrit=32)
        res = np.linalg.norm(F[g] - Bo.apply(mu[g]))
        # the oracle's FP64 operator agrees with the device's to ~1e-13 relative
        assert res < 1e-8 * 1.01 + 1e-13 * np.linalg.norm(F[g]), (g, res)
        S1 = capi.System(s.pos, s.pol, s.alpha_iso(), s.excl_csr(), summation="direct" if d else "fmm", ncrit=32,
                         damping=DAMP, precision=precision)
        S1.upload_rhs(F[g])
        it1, _, _ = S1.solve_resident(0, 1e-8, 100)
        mu1 = S1.download_mu()
        say("snapshot", g, "iters batch", iters[g], "single", it1)
        assert np.abs(mu1 - mu[g]).max() < 1e-6 * np.abs(mu1).max()


def test_graph_replay_matches_eager():
    snaps, direct = snapshots()
    B = capi.Batch(snaps, direct, ncrit=32, damping=DAMP, precision="mixed")
    rhs = batch_rhs(snaps, direct)
    B.upload_rhs(rhs)
    B.set_graphs(False)
    B.solve_resident(0, 1e-8, 100)
    mu_eager = B.download_mu()
    it_e = B.segment_results()[0]
    B.set_graphs(True)
    for rep in range(3):  # capture on the first solve, replays after; eager work between solves
        B.solve_resident(0, 1e-8, 100)
        assert np.array_equal(B.download_mu(), mu_eager), rep
        assert np.array_equal(B.segment_results()[0], it_e)
        B.apply_b(mu_eager)


def test_single_system_graphs_match_eager():
    w = synthetic.water_box(900, 7100, anisotropic=True)
    F = rhs_of(w, "fmm", 64)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP, precision="mixed")
    S.upload_rhs(F)
    S.set_graphs(False)
    it0, n0, _ = S.solve_resident(0, 1e-8, 100)
    mu0 = S.download_mu()
    S.set_graphs(True)
    for _ in range(2):
        it1, n1, _ = S.solve_resident(0, 1e-8, 100)
        assert it1 == it0 and np.array_equal(n1, n0)
        assert np.array_equal(S.download_mu(), mu0)


def test_batch_rejects_cross_snapshot_exclusions():
    snaps, _ = snapshots()
    bad = synthetic.water_box(60, 7000, anisotropic=True)
    bad.excl[0] = [bad.n + 1]  # points into the next snapshot once concatenated
    with pytest.raises(capi.CppeHipError):
        capi.Batch([bad, snaps[1]], [True, True], damping=DAMP)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_converged_guess_needs_no_iteration(precision):
    """A guess that already meets the threshold: the plain form takes the reference's one step before
    its first test (electric_fields.cc:213-229), the FP32-refined form stops at 0 iterations once an
    FP64 true residual confirms it; a zero right-hand side with the alpha F guess (an exactly zero
    residual) converges at once in both."""
    w = synthetic.water_box(300, 7300, anisotropic=True)
    F = rhs_of(w, "fmm", 64)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP, precision=precision)
    S.upload_rhs(F)
    it0, _, _ = S.solve_resident(0, 1e-10, 100)
    assert it0 > 0
    it1, norms, _ = S.solve_resident(1, 1e-8, 100)  # restart from the converged moments
    if precision == "fp64":
        assert it1 == 1 and len(norms) == 1 and norms[0] < 1e-8
    else:
        assert it1 == 0 and len(norms) == 0
    info = S.last_solve_info()
    assert info["reliable_updates"] == (1 if precision == "mixed" else 0)
    S.upload_rhs(np.zeros_like(F))
    it2, _, _ = S.solve_resident(0, 1e-8, 100)
    assert it2 == 0 and not S.download_mu().any()
