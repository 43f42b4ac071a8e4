"""Learned CG warm start: training-set generation, device training and the shipped model.

The network (see include/cppe_hip.h, ``cppe_hip_mlp_*``) is trained on induced dipoles that the
FP64 solver itself produces: seeded synthetic water boxes of 3k-30k sites (the benchmark
recipe with per-snapshot seeds), static fields as right-hand side, FMM summation (theta 0.5,
order 5, ncrit 64), Thole damping 2.1304, CG to 1e-8. Every site of every snapshot is a sample.

    python -m cppe_amd.warmstart --snapshots 2000 --epochs 8 --objective alpha_inv --out cppe_amd/data/warmstart_mlp.npz

Objectives: 'dipole' (squared dipole error relative to alpha F) or 'alpha_inv' (the error in the
alpha^-1 norm, e^T alpha^-1 e per site: the diagonal blocks of the B-norm e^T B e that CG minimises).
The model file holds the FP32 parameters, the feature scaling and the training metadata.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi, synthetic

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(PKG, "data", "warmstart_mlp.npz")
TRAIN_SEED0 = 20000
DAMPING = 2.1304


def static_rhs(w, summation="fmm"):
    """Static field at the sites (nuclei + multipoles), computed on the device."""
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr(), summation=summation)
    Fm = Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)
    Sm.close()
    return Fn + Fm


def snapshot_sizes(n, seed=TRAIN_SEED0, lo=1000, hi=10000):
    """Water counts of the training snapshots: log-uniform in [lo, hi] waters (3k-30k sites)."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(int)


def solved_system(w, thresh=1e-8, summation="fmm"):
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation=summation, damping=DAMPING)
    S.upload_rhs(static_rhs(w, summation))
    it, _, _ = S.solve_resident(0, thresh, 500)
    return S, it


def train(n_snapshots=2000, epochs=8, batch=4096, lr=1e-3, objective="dipole", log=print):
    sizes = snapshot_sizes(n_snapshots)
    M = capi.Mlp()
    M.set_objective(objective)
    M.reserve(int(3 * sizes.sum()))
    t0 = time.time()
    iters = []
    for k, nw in enumerate(sizes):
        w = synthetic.water_box(int(nw), TRAIN_SEED0 + 1 + k, anisotropic=True)
        S, it = solved_system(w)
        iters.append(it)
        M.collect(S)
        S.close()
        if (k + 1) % 100 == 0:
            log("collected %d/%d snapshots (%.0f s)" % (k + 1, n_snapshots, time.time() - t0))
    M.prepare()
    log("training on %d sites" % int(3 * sizes.sum()))
    losses = []
    for e in range(epochs):  # cosine learning-rate schedule, one epoch per call (Adam state persists)
        lr_e = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * e / max(epochs, 1))))
        losses += list(M.train(1, batch, float(lr_e), seed=1 + e))
        log("epoch %d lr %.2e loss %.5e (%.0f s)" % (e, lr_e, losses[-1], time.time() - t0))
    meta = dict(n_snapshots=int(n_snapshots), sites=int(3 * sizes.sum()), epochs=int(epochs), batch=int(batch),
                lr=float(lr), lr_schedule="cosine to 5%", objective=objective, losses=[float(x) for x in losses],
                mean_cg_iterations=float(np.mean(iters)), seconds=time.time() - t0)
    meta["validation"] = validate(M, log=log)
    return M, meta


VALID_SEED0 = 40000


def validate(M, n=16, log=print):
    """Held-out snapshots (seeds outside the training range): error of the learned guess and of alpha F
    against the FP64 solution (dipole norm, and B-norm sqrt(e^T B e) with the system's FP64 operator),
    and CG iterations from each guess to 1e-8 with the alpha r preconditioner and with the local one."""
    rows = []
    for k, nw in enumerate(snapshot_sizes(n, seed=VALID_SEED0)):
        w = synthetic.water_box(int(nw), VALID_SEED0 + 1 + k, anisotropic=True)
        S, it_aF = solved_system(w)
        mu = S.download_mu()
        F = static_rhs(w)
        aF = S.apply_alpha(F)
        capi.ml_guess(S, M)
        g = S.download_mu()
        it_ml, _, _ = S.solve_resident(1, 1e-8, 500)
        S.set_preconditioner("local")
        it_aF_loc, _, _ = S.solve_resident(0, 1e-8, 500)
        capi.ml_guess(S, M)
        it_ml_loc, _, _ = S.solve_resident(1, 1e-8, 500)
        bnorm = lambda e: float(np.sqrt(max(e @ S.apply_b(e), 0.0)))
        rows.append(dict(sites=int(w.n), err_ml=float(np.linalg.norm(g - mu) / np.linalg.norm(mu)),
                         err_alphaF=float(np.linalg.norm(aF - mu) / np.linalg.norm(mu)),
                         berr_ml=bnorm(g - mu) / bnorm(mu), berr_alphaF=bnorm(aF - mu) / bnorm(mu),
                         iters_alphaF=int(it_aF), iters_ml=int(it_ml),
                         iters_alphaF_local=int(it_aF_loc), iters_ml_local=int(it_ml_loc)))
        S.close()
    out = dict(snapshots=rows)
    for key in ("err_ml", "err_alphaF", "berr_ml", "berr_alphaF", "iters_alphaF", "iters_ml", "iters_alphaF_local",
                "iters_ml_local"):
        out["mean_" + key] = float(np.mean([r[key] for r in rows]))
    log("validation: relative dipole error %.3e (alpha F %.3e), B-norm %.3e (alpha F %.3e), CG iterations "
        "%.2f (alpha F %.2f), with the local preconditioner %.2f (alpha F %.2f)"
        % (out["mean_err_ml"], out["mean_err_alphaF"], out["mean_berr_ml"], out["mean_berr_alphaF"],
           out["mean_iters_ml"], out["mean_iters_alphaF"], out["mean_iters_ml_local"], out["mean_iters_alphaF_local"]))
    return out


def save(M, meta, path=DEFAULT_MODEL):
    p, mean, istd = M.get()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=p, mean=mean, inv_std=istd, meta=json.dumps(meta))


def load(path=DEFAULT_MODEL):
    """The shipped model as a device network, or None when no model file exists."""
    if not os.path.exists(path):
        return None
    d = np.load(path, allow_pickle=False)
    return capi.Mlp(d["params"], d["mean"], d["inv_std"])


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--objective", default="alpha_inv", choices=sorted(capi.Mlp.OBJECTIVES))
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--validate-only", metavar="MODEL", default=None,
                    help="validate an existing model file and print its metrics")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.validate_only:
        print(json.dumps(validate(load(args.validate_only), log=log)))
        return 0
    M, meta = train(args.snapshots, args.epochs, args.batch, args.lr, args.objective, log=log)
    save(M, meta, args.out)
    print(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
