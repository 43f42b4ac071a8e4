"""CPU pins of the learned-warm-start reference (oracle/mlp_ref.py): backpropagation against central
finite differences, bfloat16 rounding against its definition, equivariance of the bases."""
import numpy as np

from oracle import mlp_ref as R


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, -3.14159, 65504.0, 1e-3])
    b = R.bf16(x)
    # representable: 8 significant bits
    m, e = np.frexp(b)
    assert np.all(np.abs(m * 2 ** 8 - np.round(m * 2 ** 8)) < 1e-12)
    assert np.all(np.abs(b - x) <= np.abs(x) * 2 ** -8)
    assert b[1] == 1.0  # tie rounds to even
    assert b[2] == 1.0 + 2 ** -7


def test_backprop_vs_finite_differences():
    rng = np.random.default_rng(0)
    ns = 7
    p = rng.normal(size=R.NPARAMS) * 0.1
    mean, istd = rng.normal(size=64), rng.uniform(0.5, 2, size=64)
    X = rng.normal(size=(ns, 64))
    V = rng.normal(size=(ns, 9))
    T = rng.normal(size=(ns, 3))
    w = rng.uniform(0.5, 2, size=ns)
    loss, g = R.loss_grad(p, mean, istd, X, V, T, w, emulate_bf16=False)
    for k in list(rng.integers(0, R.NPARAMS, 40)) + [R.OFF_B3, R.OFF_B3 + 2, R.OFF_W3 + 5, R.OFF_B1 + 3]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_grad(pp, mean, istd, X, V, T, w, False)[0] - R.loss_grad(pm, mean, istd, X, V, T, w, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k


def test_forward_matches_loss_path():
    rng = np.random.default_rng(1)
    p = rng.normal(size=R.NPARAMS) * 0.1
    mean, istd = np.zeros(64), np.ones(64)
    X = rng.normal(size=(5, 64))
    y = R.forward(p, mean, istd, X, emulate_bf16=False)
    V = np.tile(np.eye(3).reshape(1, 9), (5, 1))  # guess = y itself
    T = y.copy()
    loss, _ = R.loss_grad(p, mean, istd, X, V, T, np.ones(5), emulate_bf16=False)
    assert loss < 1e-20


def test_bases_rotate_with_the_system():
    """v1, v2, v3 are equivariant and the descriptors invariant under a rotation of everything."""
    from cppe_amd import synthetic
    w = synthetic.water_box(40, 3, anisotropic=True)
    rng = np.random.default_rng(2)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    F = rng.normal(size=(w.n, 3)) * 0.05
    X, V = R.features(w.pos, w.pol, F, w.excl_csr(), 2.1304)
    A = np.zeros((w.n, 3, 3))
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k, (i, j) in enumerate(idx):
        A[:, i, j] = A[:, j, i] = w.pol[:, k]
    Ar = q @ A @ q.T
    pol_r = np.stack([Ar[:, i, j] for (i, j) in idx], axis=1)
    Xr, Vr = R.features(w.pos @ q.T, pol_r, F @ q.T, w.excl_csr(), 2.1304)
    np.testing.assert_allclose(Xr, X, rtol=1e-9, atol=1e-9)
    for a in range(3):
        np.testing.assert_allclose(Vr[:, 3 * a:3 * a + 3], V[:, 3 * a:3 * a + 3] @ q.T, rtol=1e-9, atol=1e-12)
