"""Several right-hand sides solved together (cppe_hip_solve_multi): every right-hand side must
follow its own single-solve CG recurrence -- the same iteration count and the same dipoles as a
solve of that right-hand side alone (FP64: 1e-12 relative; the shared near-field launch evaluates
each pair's geometry once and applies it to every dipole vector with the single solve's
arithmetic) -- in both precision forms, for groups of three (one shared launch) and for the
remainder (one at a time), under any near-field launch shape. CppeState::induced_moments_eef
(three unit fields, reference cppe_state.cc:111-128) is checked against the dense inverse in
test_api_reference.py."""
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


def box(nw=1500, seed=7300):
    return synthetic.water_box(nw, seed, anisotropic=True)


def rhs_set(w, k, seed=3):
    rng = np.random.default_rng(seed)
    F = O.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    out = [F]
    for _ in range(k - 1):
        out.append(F * rng.uniform(0.5, 2.0) + 0.01 * rng.normal(size=F.shape))
    return np.array(out)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", [3, 4, 2])
def test_multi_matches_single_solves(precision, k):
    w = box()
    say("multi", precision, "nrhs", k, "sites", w.n)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP, precision=precision)
    S.set_tune(np.array([8, 256, 16, 1, 8, 256, 16, 1], dtype=np.int32))
    R = rhs_set(w, k)
    mu, it, norms = S.solve_multi(R, thresh=1e-8, maxiter=60)
    B = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=DAMP, scheme="fmm")
    for j in range(k):
        m1, it1, n1 = S.solve(R[j], thresh=1e-8, maxiter=60)
        assert it[j] == it1, (j, it[j], it1)
        tol = 1e-12 if precision == "fp64" else 1e-6
        assert np.abs(mu[j] - m1).max() <= tol * np.abs(m1).max(), j
        assert len(norms[j]) == it1
        # the stopping test holds for the oracle's operator too (FP64 residual of the tree operator)
        assert np.linalg.norm(R[j] - B.apply(mu[j])) < 1.01e-8


def test_multi_other_launch_shapes():
    """Thread-per-target near field (no shared launch): right-hand sides applied one at a time."""
    w = box(800, 7301)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP)
    R = rhs_set(w, 3)
    ref = [S.solve(R[j], thresh=1e-9)[0] for j in range(3)]
    S.set_tune(np.array([0, 256, 16, 0, 0, 256, 16, 0], dtype=np.int32))
    mu, it, _ = S.solve_multi(R, thresh=1e-9)
    for j in range(3):
        assert np.abs(mu[j] - ref[j]).max() <= 1e-11 * np.abs(ref[j]).max()


def test_multi_direct_summation_and_guess():
    """Direct summation (one-leaf tree) and a caller-supplied first guess."""
    w = box(600, 7302)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=DAMP)
    R = rhs_set(w, 3)
    mu0, _, _ = S.solve_multi(R, thresh=1e-10)
    D = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=DAMP)
    for j in range(3):
        assert np.linalg.norm(R[j] - D.apply(mu0[j])) < 1.01e-10
    mu1, it1, _ = S.solve_multi(R, guess=mu0, make_guess=False, thresh=1e-10)
    assert (it1 <= 1).all()
    assert np.abs(mu1 - mu0).max() <= 1e-9 * np.abs(mu0).max()


def test_multi_rejects_batches():
    ws = [box(300, 7303 + s) for s in range(2)]
    Bt = capi.Batch(ws, direct=[True, True], damping=DAMP)
    with pytest.raises(capi.CppeHipError, match="one-snapshot"):
        Bt.solve_multi(np.zeros((3, 3 * Bt.n)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("damping", [DAMP, 0.0])
@pytest.mark.parametrize("overlap", [1, 2, 3])
def test_apply_multi_matches_single_applies(precision, damping, overlap):
    """B applied to 3 and 4 vectors together equals one apply per vector (same launch shape)."""
    w = box(1500, 7304)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=damping,
                    precision=precision)
    S.set_tune(np.array([8, 256, 16, overlap, 8, 256, 16, overlap], dtype=np.int32))
    X = np.random.default_rng(21).normal(size=(4, 3 * w.n))
    single = np.array([S.apply_b(x) for x in X])
    tol = 1e-13 if precision == "fp64" else 1e-6
    for k in (3, 4):
        Y = S.apply_b_multi(X[:k])
        for j in range(k):
            assert np.abs(Y[j] - single[j]).max() <= tol * np.abs(single[j]).max(), (k, j)
