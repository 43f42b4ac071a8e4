"""CPU reference of the learned warm start (TEST INFRASTRUCTURE; pins the GPU kernels).

* ``features``: brute-force O(N^2) restatement of the 64 descriptors and the three basis vectors
  (v1 = alpha F, v2/v3 = local damped Neumann terms within 6 A) documented in
  cppe_amd/csrc/device/ml_features.hip.
* ``forward`` / ``loss_grad``: FP64 MLP 64 -> 128 -> 128 -> 3 (SiLU), the loss
  sum_s w_s |sum_a y_sa v_sa - t_s|^2 / N and its exact gradient by backpropagation. With
  ``emulate_bf16`` the operands that the GPU rounds to bfloat16 (normalised inputs, W1, W2,
  s(Z1), dZ2, dZ1) are rounded the same way, so the GPU gradient can be compared tightly.
"""
import numpy as np

IN, HID, OUT = 64, 128, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * OUT
NPARAMS = OFF_B3 + OUT

RC = 6.0 * 1.8897261246
RK0, RKSTEP, ETA = 1.5, 1.4, 0.5


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(IN, HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(HID, HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(HID, OUT)
    b3 = p[OFF_B3:]
    return W1, b1, W2, b2, W3, b3


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sig(z)


def dsilu(z):
    s = sig(z)
    return s * (1.0 + z * (1.0 - s))


def normalise(X, mean, inv_std):
    return (np.asarray(X, dtype=np.float64) - mean) * inv_std


def forward(p, mean, inv_std, X, emulate_bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Xn = normalise(X, mean, inv_std)
    if emulate_bf16:
        Xn, W1, W2 = bf16(Xn), bf16(W1), bf16(W2)
    Z1 = Xn @ W1 + b1
    H1 = silu(Z1)
    if emulate_bf16:
        H1 = bf16(H1)
    Z2 = H1 @ W2 + b2
    return silu(Z2) @ W3 + b3


def loss_grad(p, mean, inv_std, X, V, T, w, emulate_bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Xn = normalise(X, mean, inv_std)
    if emulate_bf16:
        Xn, W1, W2 = bf16(Xn), bf16(W1), bf16(W2)
    ns = X.shape[0]
    Z1 = Xn @ W1 + b1
    H1 = silu(Z1)
    if emulate_bf16:
        H1 = bf16(H1)
    Z2 = H1 @ W2 + b2
    H2 = silu(Z2)
    y = H2 @ W3 + b3
    V = np.asarray(V, dtype=np.float64).reshape(ns, 3, 3)  # [s][a][xyz]
    g = np.einsum("sa,sax->sx", y, V)
    e = g - np.asarray(T, dtype=np.float64)
    wn = np.asarray(w, dtype=np.float64) / ns
    loss = float(np.sum(wn * np.sum(e * e, axis=1)))
    dy = 2.0 * wn[:, None] * np.einsum("sx,sax->sa", e, V)
    gW3 = H2.T @ dy
    gb3 = dy.sum(0)
    dZ2 = (dy @ W3.T) * dsilu(Z2)
    gb2 = dZ2.sum(0)
    dZ2m = bf16(dZ2) if emulate_bf16 else dZ2
    gW2 = H1.T @ dZ2m
    dZ1 = (dZ2m @ W2.T) * dsilu(Z1)
    gb1 = dZ1.sum(0)
    dZ1m = bf16(dZ1) if emulate_bf16 else dZ1
    gW1 = Xn.T @ dZ1m
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad


# ---------------------------------------------------------------------------------------------
def alpha_inv_factor(alpha6):
    """L^T with L L^T = alpha^-1 (Cholesky) per site, alpha6 packed (xx, xy, xz, yy, yz, zz): the map
    under which the squared dipole error becomes the alpha^-1-norm error, |L^T e|^2 = e^T alpha^-1 e
    (the training samples of the 'alpha_inv' objective are V and T mapped by it)."""
    a = np.asarray(alpha6, dtype=np.float64).reshape(-1, 6)
    A = np.empty((a.shape[0], 3, 3))
    for k, (r, c) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        A[:, r, c] = a[:, k]
        A[:, c, r] = a[:, k]
    L = np.linalg.cholesky(np.linalg.inv(A))
    return np.transpose(L, (0, 2, 1))


def _thole(u):
    e = np.exp(-u)
    fE = 1.0 - (1.0 + u + 0.5 * u * u) * e
    fT = fE - u ** 3 / 6.0 * e
    return fE, fT


def features(pos, alpha6, F, excl, damping):
    """Descriptors X (n x 64) and bases V (n x 9) by brute force (caller order)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    a6 = np.asarray(alpha6, dtype=np.float64).reshape(n, 6)
    A = np.zeros((n, 3, 3))
    A[:, 0, 0], A[:, 0, 1], A[:, 0, 2] = a6[:, 0], a6[:, 1], a6[:, 2]
    A[:, 1, 0], A[:, 1, 1], A[:, 1, 2] = a6[:, 1], a6[:, 3], a6[:, 4]
    A[:, 2, 0], A[:, 2, 1], A[:, 2, 2] = a6[:, 2], a6[:, 4], a6[:, 5]
    iso = np.trace(A, axis1=1, axis2=2) / 3.0
    s = np.where(iso > 0, iso ** (-1.0 / 6.0), 0.0)
    F = np.asarray(F, dtype=np.float64).reshape(n, 3)
    Fn = np.linalg.norm(F, axis=1) + 1e-30
    Fh = F / Fn[:, None]
    off, idx = excl
    X = np.zeros((n, 64))
    v1 = np.einsum("sij,sj->si", A, F)
    w = v1  # alpha_j F_j
    rk = RK0 + RKSTEP * np.arange(8)

    def local_T_sum(wvec):
        out = np.zeros((n, 3))
        for i in range(n):
            d = pos - pos[i]
            r = np.linalg.norm(d, axis=1)
            m = (r < RC) & (np.arange(n) != i)
            m[idx[off[i]:off[i + 1]]] = False
            j = np.nonzero(m)[0]
            dj, rj = d[j], r[j]
            fc = 0.5 * (np.cos(np.pi * rj / RC) + 1.0)
            a = damping * s[i] * s[j]
            fE, fT = np.ones_like(rj), np.ones_like(rj)
            dm = (a > 0) & (a * rj < 50.0)
            fE[dm], fT[dm] = _thole(a[dm] * rj[dm])
            dw = np.sum(dj * wvec[j], axis=1)
            t = (3.0 * fT * dw / rj ** 5)[:, None] * dj - (fE / rj ** 3)[:, None] * wvec[j]
            out[i] = np.sum(fc[:, None] * t, axis=0)
        return np.einsum("sij,sj->si", A, out)

    v2 = local_T_sum(w)
    v3 = local_T_sum(v2)
    for i in range(n):
        d = pos - pos[i]
        r = np.linalg.norm(d, axis=1)
        m = (r < RC) & (np.arange(n) != i)
        j = np.nonzero(m)[0]
        dj, rj = d[j], r[j]
        fc = 0.5 * (np.cos(np.pi * rj / RC) + 1.0)
        g = np.exp(-ETA * (rj[:, None] - rk[None, :]) ** 2) * fc[:, None]  # (nj, 8)
        cth = (dj @ Fh[i]) / rj
        p2 = 1.5 * cth ** 2 - 0.5
        al = Fh[j] @ Fh[i]
        bj = np.sum(dj * Fh[j], axis=1) / rj
        X[i, 0:8] = g.sum(0)
        X[i, 8:16] = (g * (iso[j] * 0.2)[:, None]).sum(0)
        X[i, 16:24] = (g * cth[:, None]).sum(0)
        X[i, 24:32] = (g * p2[:, None]).sum(0)
        X[i, 32:40] = (g * al[:, None]).sum(0)
        X[i, 40:48] = (g * bj[:, None]).sum(0)
        X[i, 54] = fc.sum()
        X[i, 55] = (fc / rj ** 3).sum() * 100.0
        e1 = Fh[i].copy()
        e2 = np.zeros(3)
        if off[i + 1] - off[i] >= 2:
            p, q = idx[off[i]], idx[off[i] + 1]
            u = pos[p] - pos[i]
            e1 = u / (np.linalg.norm(u) + 1e-30)
            ww = pos[q] - pos[i]
            ww = ww - (ww @ e1) * e1
            e2 = ww / (np.linalg.norm(ww) + 1e-30)
        e3 = np.cross(e1, e2)
        X[i, 48:51] = [Fh[i] @ e1, Fh[i] @ e2, Fh[i] @ e3]
        X[i, 51] = np.log(Fn[i])
        X[i, 52] = iso[i] * 0.2
        X[i, 53] = Fh[i] @ A[i] @ Fh[i] / iso[i]
    n1 = np.linalg.norm(v1, axis=1) + 1e-30
    n2 = np.linalg.norm(v2, axis=1) + 1e-30
    n3 = np.linalg.norm(v3, axis=1) + 1e-30
    X[:, 56] = np.sum(v2 * v1, 1) / n1 ** 2
    X[:, 57] = n2 / n1
    X[:, 58] = np.sum(v3 * v1, 1) / n1 ** 2
    X[:, 59] = n3 / n1
    X[:, 60] = np.sum(v3 * v2, 1) / (n2 * n3)
    X[:, 61] = np.log(n1)
    X[:, 62] = np.sum(v1 * F, 1) / (n1 * Fn)
    X[:, 63] = 1.0
    V = np.concatenate([v1, v2, v3], axis=1)
    return X, V
