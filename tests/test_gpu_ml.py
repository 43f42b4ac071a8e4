"""GPU checks of the learned warm start against the CPU reference (oracle/mlp_ref.py).

* descriptors and bases: GPU kernels vs brute force (FP32 features 2e-5 of their scale; v1 1e-12,
  v2 / v3 from FP32 tensor blocks 2e-6 of their scale)
* forward: BF16-MFMA network vs the FP64 reference with the same bfloat16 rounding points
* gradient: GPU backward vs the FP64 reference gradient (bf16-emulated tight, pure FP64 loose)
* training + guess: a small training run lowers the loss and the learned guess needs fewer CG
  iterations than alpha F on a held-out snapshot, with the same 1e-8 FP64 stopping test (both
  training objectives: dipole error and its alpha^-1 norm)
* alpha^-1 objective: the collected samples are the dipole-objective samples mapped by L^T,
  L L^T = alpha^-1 (oracle: numpy Cholesky; |L^T e|^2 = e^T alpha^-1 e checked on its own)
"""
import numpy as np
import pytest

import oracle as O
from cppe_amd import capi, synthetic
from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu


def rhs_of(w):
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr())
    return Fn + Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)


def test_features_vs_reference():
    w = synthetic.water_box(250, 21, anisotropic=True)
    F = rhs_of(w)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304)
    S.upload_rhs(F)
    X, V = capi.ml_features(S)
    Xr, Vr = R.features(w.pos, w.pol, F.reshape(-1, 3), w.excl_csr(), 2.1304)
    # V: per site (v1, v2, v3). v1 = alpha F in FP64; v2, v3 from FP32 tensor blocks and FP32
    # neighbour vectors (sums in FP64)
    np.testing.assert_allclose(V[:, :3], Vr[:, :3], rtol=1e-12, atol=1e-14 * np.abs(Vr[:, :3]).max())
    for b in (1, 2):
        blk, ref = V[:, 3 * b:3 * b + 3], Vr[:, 3 * b:3 * b + 3]
        assert np.abs(blk - ref).max() <= 2e-6 * np.abs(ref).max()
    scale = np.maximum(np.abs(Xr).max(axis=0), 1e-3)
    assert (np.abs(X - Xr) / scale).max() < 2e-5


def _random_model(seed):
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=R.NPARAMS) * 0.15).astype(np.float32)
    mean = rng.normal(size=64).astype(np.float32) * 0.1
    istd = rng.uniform(0.5, 1.5, size=64).astype(np.float32)
    return p, mean, istd


def test_forward_vs_reference():
    p, mean, istd = _random_model(3)
    M = capi.Mlp(p, mean, istd)
    X = np.random.default_rng(4).normal(size=(300, 64)).astype(np.float32)
    y = M.forward(X)
    yr = R.forward(p, mean, istd, X, emulate_bf16=True)
    np.testing.assert_allclose(y, yr, rtol=2e-3, atol=2e-3 * np.abs(yr).max())
    y64 = R.forward(p, mean, istd, X, emulate_bf16=False)
    assert np.abs(y - y64).max() < 3e-2 * np.abs(y64).max()


def test_gradient_vs_reference():
    p, mean, istd = _random_model(5)
    M = capi.Mlp(p, mean, istd)
    rng = np.random.default_rng(6)
    ns = 203  # ragged last tile
    X = rng.normal(size=(ns, 64)).astype(np.float32)
    V = rng.normal(size=(ns, 9)).astype(np.float32)
    T = rng.normal(size=(ns, 3)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=ns).astype(np.float32)
    loss, g = M.loss_grad(X, V, T, w)
    lr_, gr = R.loss_grad(p, mean, istd, X, V, T, w, emulate_bf16=True)
    assert abs(loss - lr_) <= 1e-4 * abs(lr_)
    for lo, hi in [(R.OFF_W1, R.OFF_B1), (R.OFF_B1, R.OFF_W2), (R.OFF_W2, R.OFF_B2), (R.OFF_B2, R.OFF_W3),
                   (R.OFF_W3, R.OFF_B3), (R.OFF_B3, R.NPARAMS)]:
        a, b = g[lo:hi], gr[lo:hi]
        assert np.abs(a - b).max() <= 5e-3 * np.abs(b).max() + 1e-7, (lo, hi)
    _, g64 = R.loss_grad(p, mean, istd, X, V, T, w, emulate_bf16=False)
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)


def test_alpha_inv_objective_samples():
    """Every site gets the same anisotropic polarizability, so the sample map is one known L^T."""
    b = synthetic.water_box(300, 3100, anisotropic=True)
    a6 = np.array([5.0, 0.7, -0.3, 4.0, 0.4, 6.0])
    pol = np.tile(a6, (b.n, 1))
    iso = np.full(b.n, (a6[0] + a6[3] + a6[5]) / 3.0)
    S = capi.System(b.pos, pol, iso, b.excl_csr(), damping=2.1304)
    S.upload_rhs(rhs_of(b))
    S.solve_resident(0, 1e-10, 200)
    M0, M1 = capi.Mlp(), capi.Mlp()
    M1.set_objective("alpha_inv")
    for M in (M0, M1):
        M.reserve(b.n)
        M.collect(S)
    V0, T0, w0 = M0.samples(0, b.n)
    V1, T1, w1 = M1.samples(0, b.n)
    Lt = R.alpha_inv_factor(a6)[0]
    Ainv = np.linalg.inv(np.array([[a6[0], a6[1], a6[2]], [a6[1], a6[3], a6[4]], [a6[2], a6[4], a6[5]]]))
    e = np.random.default_rng(5).normal(size=(50, 3))
    assert np.allclose(np.sum((e @ Lt.T) ** 2, axis=1), np.einsum("si,ij,sj->s", e, Ainv, e), rtol=1e-12)
    exp_V = (V0.reshape(-1, 3, 3).astype(np.float64) @ Lt.T).reshape(-1, 9)
    exp_T = T0.astype(np.float64) @ Lt.T
    assert np.abs(V1 - exp_V).max() < 1e-6 * np.abs(exp_V).max()
    assert np.abs(T1 - exp_T).max() < 1e-6 * np.abs(exp_T).max()
    assert np.allclose(w1, 1.0 / np.sum(exp_V[:, :3] ** 2, axis=1), rtol=1e-5)
    with pytest.raises(capi.CppeHipError):  # the objective is fixed once samples are collected
        M1.set_objective("dipole")


@pytest.mark.parametrize("objective", ["dipole", "alpha_inv"])
def test_training_and_guess_gain(objective):
    """Train on a few small boxes; on a held-out box the learned guess must be closer to the solution
    than alpha F (error and initial residual), and CG from it must not need more iterations."""
    M = capi.Mlp()
    M.set_objective(objective)
    boxes = [synthetic.water_box(int(nw), 3000 + k, anisotropic=True) for k, nw in enumerate([500, 700, 600, 800])]
    M.reserve(sum(b.n for b in boxes))
    for b in boxes:
        S = capi.System(b.pos, b.pol, b.alpha_iso(), b.excl_csr(), damping=2.1304)
        S.upload_rhs(rhs_of(b))
        S.solve_resident(0, 1e-10, 200)
        M.collect(S)
    M.prepare()
    losses = M.train(epochs=30, batch=1024, lr=2e-3, seed=0)
    assert losses[-1] < 0.5 * losses[0]
    test = synthetic.water_box(700, 3999, anisotropic=True)
    F = rhs_of(test)
    S = capi.System(test.pos, test.pol, test.alpha_iso(), test.excl_csr(), damping=2.1304)
    S.upload_rhs(F)
    it0, _, _ = S.solve_resident(0, 1e-8, 200)
    mu_ref = S.download_mu()
    capi.ml_guess(S, M)
    g = S.download_mu()
    it1, _, _ = S.solve_resident(1, 1e-8, 200)
    mu = S.download_mu()
    aF = S.apply_alpha(F)
    err_alphaF = np.linalg.norm(aF - mu_ref)
    err_ml = np.linalg.norm(g - mu_ref)
    res_alphaF = np.linalg.norm(F - S.apply_b(aF))
    res_ml = np.linalg.norm(F - S.apply_b(g))
    print("guess error alpha F %.3e  learned %.3e; initial residual %.3e -> %.3e; iterations %d -> %d"
          % (err_alphaF, err_ml, res_alphaF, res_ml, it0, it1))
    assert err_ml < 0.5 * err_alphaF
    assert res_ml < res_alphaF
    assert it1 <= it0
    B = O.BSystem(test.pos, test.pol, test.excl_csr(), damping=2.1304)
    assert np.linalg.norm(F - B.apply(mu)) < 1e-8


def test_untimed_guess_equals_timed():
    """cppe_hip_ml_guess with device_ms NULL only enqueues the guess on the system's stream: the
    moments it leaves equal the timed call's bit for bit, and a resident solve right after it (no
    host synchronisation in between) starts from them and converges."""
    p, mean, istd = _random_model(7)
    M = capi.Mlp(p, mean, istd)
    w = synthetic.water_box(400, 3101, anisotropic=True)
    F = rhs_of(w)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), damping=2.1304)
    S.upload_rhs(F)
    ms = capi.ml_guess(S, M)
    assert ms is not None and ms > 0.0
    g_timed = S.download_mu()
    assert capi.ml_guess(S, M, timed=False) is None
    g_untimed = S.download_mu()
    np.testing.assert_array_equal(g_untimed, g_timed)
    capi.ml_guess(S, M, timed=False)
    it, norms, _ = S.solve_resident(1, 1e-8, 200)
    mu = S.download_mu()
    B = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=2.1304)
    print("untimed guess: iterations", it, "last norm", norms[-1])
    assert norms[-1] < 1e-8 and np.linalg.norm(F - B.apply(mu)) < 1e-8
