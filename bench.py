#!/usr/bin/env python
"""Induced-dipole solve benchmark on one MI355X.

Workload (default): 100,000 rigid waters = 300,000 polarizable sites (seeded synthetic box,
anisotropic polarizabilities, intramolecular exclusions, 20-nucleus QM core), Thole damping
2.1304 on induced fields, FMM summation (theta 0.5, expansion order 5, ncrit 64), Jacobi-
preconditioned CG to ||F - B mu||_2 < 1e-8. The right-hand side is the static field (nuclei +
order-0..2 multipoles), computed once on the device before timing.

One step = one full solve with the inputs resident in HBM (cppe_hip_solve_resident), for
  * the headline form: FP32 field applies inside FP64 iterative refinement (reliable
    updates), warm-started by the learned per-site guess when a trained model is available;
  * the plain form: FP64 throughout, the reference's first guess alpha * F.
Both stop on the same FP64 residual norm. The JSON line reports the headline form's ms per
solve as `value` (lower is better) and the plain form beside it.

roofline: the dominant kernel of the headline apply, timed live with HIP events recorded on the
system's stream right before and after its launch (cppe_hip_profile_apply, phases run one
after another); algorithmic flops per launch = 47 per damped
near-field site pair (self and excluded pairs not counted) + nothing else for the near-field
kernel, 950 per dipole M2L cell pair for the M2L kernel.
cpu_baseline: the reference's algorithm on the host cores (oracle/cpu_baseline.cc, built on the
running machine with the reference's -O3 -march=native -fopenmp): full solves (median of 3) where
they fit ~25 s, else the median of 3 B applies with their per-apply tree rebuild times the
solve's apply count, labelled extrapolated.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "induced-dipole solve ms @300k sites to 1e-8 resid: FP32+ML vs plain FP64"
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
FLOP_PAIR_DAMPED = 47
FLOP_M2L_M1_P5 = 950
# HBM bytes per launch of the staged near-field kernel (fused L2P) on the 300k workload, per
# precision and launch shape (near variant, block) the autotuner picks there, from rocprofv3
# FETCH_SIZE and WRITE_SIZE runs calibrated on a copy kernel (profiles/near_field_300k_hbm_traffic.md);
# other shapes: not measured (null)
MEASURED_TRAFFIC_300K = {("fp32", 8, 256): 871.1e6, ("fp64", 8, 256): 1472.4e6,
                         ("fp32", 8, 128): 1095.0e6, ("fp64", 8, 128): 1345.0e6,
                         ("fp32", 7, 128): 1215.5e6, ("fp64", 7, 128): 1417.4e6,
                         ("fp32", 7, 256): 1027.2e6, ("fp64", 7, 256): 1648.3e6}

WORKLOADS = {
    "300k": dict(n_waters=100000, seed=1003, anisotropic=True),
    "60k": dict(n_waters=20000, seed=1002, anisotropic=True),
    "3k": dict(n_waters=1000, seed=1001, anisotropic=False),
    "batch256": dict(n_snapshots=256, seed=1004),
}
BATCH_DIRECT_BELOW = 5000  # sites: such snapshots are summed directly (one leaf), the others by the tree


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_rhs(capi, w, summation):
    """Static field at the polarizable sites: nuclei + multipoles (device, FMM for the box)."""
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr(), summation=summation)
    Fm = Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)
    Sm.close()
    return Fn + Fm


def time_solves(S, steps, warmup, guess_mode, thresh=1e-8, maxiter=200):
    for _ in range(warmup):
        S.solve_resident(guess_mode, thresh, maxiter)
    wall, dev, its = [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        it, norms, ms = S.solve_resident(guess_mode, thresh, maxiter)
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(ms)
        its.append(it)
    return float(np.mean(wall)), float(np.mean(dev)), its


def run_batch(args, capi, synthetic):
    """256 snapshot environments (lognormal water counts), one handle, one CG recurrence each."""
    spec = WORKLOADS["batch256"]
    sizes = synthetic.batch_sizes(spec["n_snapshots"], spec["seed"])
    t0 = time.perf_counter()
    snaps = [synthetic.water_box(int(nw), spec["seed"] * 1000 + k, anisotropic=True) for k, nw in enumerate(sizes)]
    direct = [s.n < BATCH_DIRECT_BELOW for s in snaps]
    n = sum(s.n for s in snaps)
    log("batch: %d snapshots, %d sites, %d direct (%.1f s to generate)" % (len(snaps), n, sum(direct),
                                                                           time.perf_counter() - t0))
    common = dict(ncrit=64, order=5, theta=0.5)
    Bs = capi.Batch(snaps, direct, **common)
    Fn = np.concatenate([capi.nuclear_fields(s.pos, s.atoms_z.astype(float), s.atoms_pos) for s in snaps])
    rhs = Fn + Bs.multipole_fields(np.concatenate([s.charges for s in snaps]),
                                   np.concatenate([s.dipoles for s in snaps]),
                                   np.concatenate([s.quadrupoles for s in snaps]))
    Bs.close()
    out = {}
    for prec in ("fp64", "mixed"):
        t0 = time.perf_counter()
        B = capi.Batch(snaps, direct, damping=2.1304, precision=prec, **common)
        setup = (time.perf_counter() - t0) * 1e3
        B.upload_rhs(rhs)
        if prec == "mixed":  # the refined form with the headline's preconditioner (the FP64 form keeps alpha r)
            B.set_preconditioner(args.precond, weight=args.precond_weight)
        tune = None if args.no_tune else B.autotune(reps=2)
        res = {}
        # graphs of up to 8 steps per launch (the default), graphs of one step, kernel by kernel
        for name, graphs, gsteps in (("one_step_graphs", True, 1), ("no_graphs", False, 8), ("graphs", True, 8)):
            B.set_graphs(graphs)
            B.set_graph_steps(gsteps)
            for _ in range(args.warmup):
                B.solve_resident(0, 1e-8, 200)
            wall, dev = [], []
            for _ in range(args.steps):
                t1 = time.perf_counter()
                it, _, ms = B.solve_resident(0, 1e-8, 200)
                wall.append((time.perf_counter() - t1) * 1e3)
                dev.append(ms)
            iters, rn, conv = B.segment_results()
            res[name] = dict(ms=round(float(np.mean(wall)), 3), device_ms=round(float(np.mean(dev)), 3),
                             **B.last_solve_launches())
        assert conv.all()
        mu = B.download_mu()
        epol = [-0.5 * float(m @ f) for m, f in zip(B.split(mu), B.split(rhs))]
        out[prec] = dict(res, setup_ms=round(setup, 1), loop_iterations=int(it),
                         iterations={"min": int(iters.min()), "median": float(np.median(iters)), "max": int(iters.max())},
                         max_final_residual=float(rn.max()), info=B.last_solve_info(),
                         tune={p: {k: v for k, v in t.items() if k in ("near_variant", "near_block", "m2l_group",
                                                                        "overlap")} for p, t in tune.items()}
                         if tune else None, epol=epol)
        log(prec, {k: v for k, v in out[prec].items() if k != "epol"})
        B.close()
    e64, e32 = np.array(out["fp64"].pop("epol")), np.array(out["mixed"].pop("epol"))
    value = out["mixed"]["graphs"]["ms"]
    line = {
        "metric": "batched induced-dipole solve ms, 256 snapshots to 1e-8 resid each: FP32 refined vs plain FP64",
        "value": value, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": value,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 applies in f64 refinement",
        "data": "synthetic (seeded water boxes, seeds %d000 + k)" % spec["seed"],
        "config": {"workload": "256 snapshots, waters lognormal(ln 2000, 0.8) clipped to 500..8000, direct below "
                               "%d sites else tree (theta 0.5, order 5, ncrit 64), Thole 2.1304" % BATCH_DIRECT_BELOW,
                   "sites": int(n), "snapshots": len(snaps), "direct_snapshots": int(sum(direct)),
                   "parallelism": "single-gpu"},
        "plain_fp64_ms": out["fp64"]["graphs"]["ms"],
        "speedup_vs_fp64": round(out["fp64"]["graphs"]["ms"] / value, 3),
        "preconditioner": {"mixed": args.precond, "fp64": "jacobi"},
        "sites_per_s": round(n / (value * 1e-3), 1),
        "epol_max_rel_diff": float(np.max(np.abs(e32 - e64) / np.abs(e64))),
        "forms": out,
    }
    print(json.dumps(line))
    return 0


REBUILD_NOTE = {"fmm": "each B apply with its tree rebuild, bmatrix.cc:28-29",
                "direct": "the tree the reference also rebuilds per apply, unused by the direct sum, is not timed"}


def cpu_baseline(w, rhs, iters, summation, ncrit, order, theta, damping, budget_s=25.0):
    """The reference's algorithm on this machine's host cores (oracle/cpu_baseline.cc: per-apply tree
    rebuild, straight-line FMM operators or the direct sum, Jacobi-preconditioned CG from alpha F),
    compiled here with the reference's flags (-O3 -march=native, OpenMP). Full solves, median of 3
    after a warm-up, when three fit the time budget; otherwise the median of 3 B applies (each with
    its tree rebuild) times the solve's apply count (iterations + 1), labelled extrapolated."""
    import tempfile

    for k, v in (("OMP_PROC_BIND", "close"), ("OMP_PLACES", "cores"), ("OMP_SCHEDULE", "dynamic,16")):
        os.environ.setdefault(k, v)
    import oracle as O
    path = os.path.join(tempfile.mkdtemp(prefix="cppe_cpu_baseline_"), "libcpu_baseline.so")
    t0 = time.perf_counter()
    cb = O.CpuBaseline(w.pos, w.pol, w.excl_csr(), damping=damping, scheme=summation, ncrit=ncrit, order=order,
                       theta=theta, path=path)
    t_build = time.perf_counter() - t0
    x = np.random.default_rng(0).normal(size=3 * w.n)
    t0 = time.perf_counter()
    cb.apply(x)  # warm-up
    t_one = time.perf_counter() - t0
    common = dict(unit="ms", cores=cb.threads, kind="port", flags=cb.flags,
                  omp={k: os.environ[k] for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_SCHEDULE")},
                  build_s=round(t_build, 1))
    if 4 * (iters + 1) * t_one < budget_s:  # warm-up solve + 3 timed solves fit
        cb.solve(rhs, 1e-8, 200)
        runs = [cb.solve(rhs, 1e-8, 200) for _ in range(3)]
        ms = float(np.median([r[2] for r in runs]))
        return dict(common, value=round(ms, 1), measured="full solves", iterations=runs[0][1],
                    tree_ms=round(float(np.median([r[3] for r in runs])), 1),
                    sample="median of 3 full %s solves of the same %d-site system to 1e-8 (after one warm-up): "
                           "InducedMoments::compute, %s" % (summation, w.n, REBUILD_NOTE[summation]))
    applies = [cb.apply(x) for _ in range(3)]
    tot = sorted(a[1] + a[2] for a in applies)[1]
    tree = float(np.median([a[1] for a in applies]))
    return dict(common, value=round((iters + 1) * tot, 1), measured="extrapolated", per_apply_ms=round(tot, 1),
                tree_ms_per_apply=round(tree, 1), applies_per_solve=iters + 1,
                sample="median of 3 %s B applies of the same %d-site system (%s) times the solve's %d applies "
                       "(iterations + 1, the FP64 device solve's count); extrapolated, not a full CPU solve"
                       % (summation, w.n, REBUILD_NOTE[summation], iters + 1))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="300k", choices=sorted(WORKLOADS))
    ap.add_argument("--summation", default="fmm", choices=["fmm", "direct"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--no-ml", action="store_true", help="headline form without the learned warm start")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes instead of the autotuned ones")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the 2,000-site direct-summation check")
    ap.add_argument("--no-eef", action="store_true", help="skip the three-unit-field (induced_moments_eef) timing")
    ap.add_argument("--precond", default="local", choices=["jacobi", "local"],
                    help="CG preconditioner of the headline form (the plain FP64 form always uses the "
                         "reference's alpha r): local = alpha r + w alpha T_loc alpha r over pairs within 3 A")
    ap.add_argument("--precond-weight", type=float, default=0.8, help="weight w of the local preconditioner")
    ap.add_argument("--reliable-drop", type=float, default=None,
                    help="mixed form: residual drop that triggers an FP64 true residual (library default 1e-5)")
    ap.add_argument("--apply-damping", type=float, default=2.1304, help="Thole factor of --apply-only runs")
    ap.add_argument("--apply-ncrit", type=int, default=64, help="tree_ncrit of --apply-only runs")
    ap.add_argument("--apply-only", metavar="SHAPE", default=None,
                    help="profiling aid: only run FMM B applies of the workload with the launch shape "
                         "'variant,block,m2l[,overlap]' for both precisions (e.g. 8,256,16), print phase times, exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the induced-dipole path is built and timed on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build); rerun with --gpus 1", file=sys.stderr)
        return 2

    from cppe_amd import capi, synthetic
    if capi.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    ncrit, order, theta, damping = 64, 5, 0.5, 2.1304
    if args.workload == "batch256":
        return run_batch(args, capi, synthetic)
    spec = WORKLOADS[args.workload]
    log("generating workload", args.workload)
    w = synthetic.water_box(**spec)
    log("sites", w.n)
    if args.apply_only:
        shape = [int(t) for t in args.apply_only.split(",")]
        v, b, g = shape[:3]
        o = shape[3] if len(shape) > 3 else 0
        for prec in ("mixed", "fp64"):
            S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), precision=prec, summation="fmm",
                            ncrit=args.apply_ncrit, order=5, theta=0.5, damping=args.apply_damping)
            S.set_tune([v, b, g, o, v, b, g, o])
            ph = S.profile_apply(fp32=prec == "mixed", reps=args.profile_reps)
            st = S.tree_stats()
            log(prec, "shape", args.apply_only, "damping", args.apply_damping, "ncrit", args.apply_ncrit,
                "near ms %.4f m2l ms %.4f total ms %.4f" % (ph["near"], ph["m2l"], ph["total"]),
                "pairs", st["p2p_site_pairs"], "m2l", st["n_m2l"], "leaves", st["nleaves"])
            S.close()
        return 0
    rhs = build_rhs(capi, w, args.summation)
    common = dict(summation=args.summation, ncrit=ncrit, order=order, theta=theta, damping=damping)
    t0 = time.perf_counter()
    S64 = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), precision="fp64", **common)
    t_setup = (time.perf_counter() - t0) * 1e3
    S32 = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), precision="mixed", **common)
    S64.upload_rhs(rhs)
    S32.upload_rhs(rhs)
    if args.reliable_drop:
        S32.set_reliable_drop(args.reliable_drop)
    S32.set_preconditioner(args.precond, weight=args.precond_weight)
    log("setup (octree + lists) ms", round(t_setup, 1))
    tune = None
    if not args.no_tune:
        log("autotuning launch shapes ...")
        t0 = time.perf_counter()
        tune = {"fp64": S64.autotune(reps=3)["fp64"], "fp32": S32.autotune(reps=3)["fp32"]}
        tune["seconds"] = round(time.perf_counter() - t0, 2)
        log("autotune", tune)

    from cppe_amd import warmstart
    # the shipped network was trained on tree-summation solutions of anisotropic boxes; direct
    # summation (every pair damped) is outside that domain, where alpha F is the better guess
    model = None if (args.no_ml or args.summation != "fmm") else warmstart.load()
    ml_used = model is not None
    wall64, dev64, its64 = time_solves(S64, args.steps, args.warmup, 0)
    launch64 = S64.last_solve_launches()
    log("fp64 solve ms", wall64, "iters", its64)
    guess_ms = []
    if ml_used:
        # headline step = learned guess (descriptors + network + combination) + mixed solve; the
        # guess is enqueued ahead of the solve (its device time measured apart, below)
        for _ in range(max(1, args.warmup)):
            capi.ml_guess(S32, model)
        g_dev = float(np.mean([capi.ml_guess(S32, model) for _ in range(3)]))

        def head_solve():
            capi.ml_guess(S32, model, timed=False)
            it, norms, ms = S32.solve_resident(1, 1e-8, 200)
            return it, ms + g_dev, g_dev
        for _ in range(args.warmup):
            head_solve()
        wall32, dev32, its32 = [], [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            it, ms, g = head_solve()
            wall32.append((time.perf_counter() - t0) * 1e3)
            dev32.append(ms)
            its32.append(it)
            guess_ms.append(g)
        wall32, dev32 = float(np.mean(wall32)), float(np.mean(dev32))
    else:
        wall32, dev32, its32 = time_solves(S32, args.steps, args.warmup, 0)
    info32 = S32.last_solve_info()
    launch32 = S32.last_solve_launches()
    # the headline moments' residual under the FP64 operator (outside the timed region): the solve
    # confirms convergence from an FP64 true residual plus an FP32 apply of the correction since
    head_res = float(np.linalg.norm(rhs - S64.apply_b(S32.download_mu())))
    log("headline moments: FP64 residual", head_res)
    log("headline solve ms", wall32, "iters", its32, "ml", ml_used, "fp64 true-residual applies",
        info32["reliable_updates"], launch32)
    # the same mixed solve from alpha F and from zero, for the ML gain; then once more launched
    # kernel by kernel
    wallm, devm, itsm = time_solves(S32, max(1, args.steps // 2), 1, 0)
    wallz, devz, itsz = time_solves(S32, max(1, args.steps // 2), 1, 2)
    S32.set_graph_steps(1)  # one step per launch, a host status read after each
    wall1, _, _ = time_solves(S32, max(1, args.steps // 2), 1, 0)
    launch1 = S32.last_solve_launches()
    S32.set_graph_steps(8)
    S32.set_graphs(False)
    walln, _, _ = time_solves(S32, max(1, args.steps // 2), 1, 0)
    S32.set_graphs(True)
    # the preconditioner's share: the headline form with the reference's alpha r, and the plain FP64
    # form with the local preconditioner
    other = "jacobi" if args.precond == "local" else "local"
    S32.set_preconditioner(other, weight=args.precond_weight)
    if ml_used:
        for _ in range(max(1, args.warmup // 2)):
            head_solve()
        legs = []
        for _ in range(max(1, args.steps // 2)):
            t0 = time.perf_counter()
            it_o = head_solve()[0]
            legs.append(((time.perf_counter() - t0) * 1e3, it_o))
        wallo, itso = float(np.mean([l[0] for l in legs])), [l[1] for l in legs]
    else:
        wallo, _, itso = time_solves(S32, max(1, args.steps // 2), 1, 0)
    S32.set_preconditioner(args.precond, weight=args.precond_weight)
    S64.set_preconditioner("local", weight=args.precond_weight)
    wall64l, _, its64l = time_solves(S64, max(1, args.steps // 2), 1, 0)
    S64.set_preconditioner("jacobi")
    S64.solve_resident(0, 1e-8, 200)  # the plain form's moments again
    log("headline with %s preconditioner" % other, wallo, itso, "fp64 local", wall64l, its64l)
    mu64 = S64.download_mu()
    S32.solve_resident(0, 1e-8, 200)  # the headline form's moments, from alpha F
    mu32 = S32.download_mu()
    e64 = -0.5 * float(mu64 @ rhs)
    e32 = -0.5 * float(mu32 @ rhs)
    # the guesses' own dipole error against the FP64 solution
    nrm = float(np.linalg.norm(mu64))
    guess_err = {"alphaF": float(np.linalg.norm(S32.apply_alpha(rhs) - mu64)) / nrm, "zero": 1.0}
    if ml_used:
        capi.ml_guess(S32, model)
        guess_err["ml"] = float(np.linalg.norm(S32.download_mu() - mu64)) / nrm
    log("guess dipole error vs FP64 solution", guess_err)

    # first-call cost of a new system (what CppeState construction and its first
    # update_induced_moments pay): octree + lists, the lazily built near list and M2L records on the
    # first apply, the learned guess's 6 A neighbour list on its first use; default launch shapes
    first = None
    if args.summation == "fmm":
        t0 = time.perf_counter()
        Sf = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), precision="mixed", **common)
        t_create = (time.perf_counter() - t0) * 1e3
        Sf.upload_rhs(rhs)
        first = {"system_create_ms": round(t_create, 1)}
        if ml_used:
            t0 = time.perf_counter()
            capi.ml_guess(Sf, model)
            first["first_ml_guess_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        t0 = time.perf_counter()
        itf, _, _ = Sf.solve_resident(1 if ml_used else 0, 1e-8, 200)
        first["first_solve_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        first["first_call_total_ms"] = round(sum(first.values()), 1)
        first["autotune_s"] = tune["seconds"] if tune else None
        Sf.close()
        log("first call", first)

    acc = None
    if args.summation == "fmm" and not args.no_accuracy:
        # tree path against FP64 direct summation at 2,000 seeded sites, for the headline moments
        tg = np.sort(np.random.default_rng(2000).choice(w.n, min(2000, w.n), replace=False))
        Fd = S64.field_targets(tg, mu=mu32) + S64.apply_alpha_inv(mu32).reshape(-1, 3)[tg]
        B64 = S64.apply_b(mu32).reshape(-1, 3)[tg]
        B32 = S32.apply_b(mu32).reshape(-1, 3)[tg]
        nd = np.linalg.norm(Fd, axis=1)
        acc = {"targets": int(tg.size), "seed": 2000,
               "bmu_rel_err_fp64_tree_vs_direct": {"max": float((np.linalg.norm(B64 - Fd, axis=1) / nd).max()),
                                                   "rms": float(np.linalg.norm(B64 - Fd) / np.linalg.norm(Fd))},
               "bmu_rel_err_fp32_tree_vs_direct": {"max": float((np.linalg.norm(B32 - Fd, axis=1) / nd).max()),
                                                   "rms": float(np.linalg.norm(B32 - Fd) / np.linalg.norm(Fd))},
               "bmu_rel_diff_fp32_vs_fp64_tree": float(np.linalg.norm(B32 - B64) / np.linalg.norm(B64)),
               "residual_direct_operator_at_targets": float(np.linalg.norm(rhs.reshape(-1, 3)[tg] - Fd)),
               "residual_tree_operator_at_targets": float(np.linalg.norm(rhs.reshape(-1, 3)[tg] - B64))}
        log("accuracy", acc)

    st = S64.tree_stats()
    n_excl = int(w.excl_csr()[0][-1])
    pairs = st["p2p_site_pairs"] - w.n - n_excl if args.summation == "fmm" else w.n * (w.n - 1) - n_excl
    prof32 = S32.profile_apply(fp32=True, reps=args.profile_reps)
    prof64 = S64.profile_apply(fp32=False, reps=args.profile_reps)
    log("phases fp32", prof32)
    log("phases fp64", prof64)

    near_names = {0: "k_fmm_near", 1: "k_fmm_near_lw", 2: "k_fmm_p2p", 3: "k_fmm_near_sl", 4: "k_fmm_near_cl",
                  5: "k_fmm_near_st", 6: "k_fmm_near_st", 7: "k_fmm_near_st", 8: "k_fmm_near_st",
                  9: "k_fmm_near_sym", 10: "k_fmm_near_sym", 11: "k_fmm_near_sym"}

    def roofline(prof, peak, prec):
        near_flop = FLOP_PAIR_DAMPED * pairs
        m2l_flop = FLOP_M2L_M1_P5 * st["n_m2l"]
        if args.summation == "fmm" and prof["m2l"] > prof["near"]:
            name, flop, ms = "k_fmm_m2l (dipole M2L, order 5)", m2l_flop, prof["m2l"]
        else:
            v = tune[prec]["near_variant"] if tune else 0
            name = "%s (%s P2P + alpha^-1 mu)" % (near_names[v], "near field" if args.summation == "fmm"
                                                   else "all pairs, one-leaf tree")
            flop, ms = near_flop, (prof["near_kernel"] if prof["near_kernel"] > 0 else prof["near"])
            if v in (9, 10, 11):
                # symmetric form: each unordered pair evaluated once; the work is counted as the two
                # ordered pairs it replaces, over the pair kernel plus the gather of the reverse
                # blocks (with the L2P term) and the exclusion pass (the whole near pass)
                name = ("k_fmm_near_sym + k_fmm_sym_gather (symmetric near field P2P + alpha^-1 mu; "
                        "work counted as ordered pairs)")
                ms = prof["near"]
        ach = flop / (ms * 1e-3)
        traffic = None
        if args.workload == "300k" and "k_fmm_near_st" in name and tune:
            key = ("fp32" if prec == "fp32" else "fp64", tune[prec]["near_variant"], tune[prec]["near_block"])
            traffic = MEASURED_TRAFFIC_300K.get(key)
        return {"bound": "valu", "kernel": name, "achieved": round(ach / 1e12, 3),
                "peak": peak / 1e12, "unit": "TFLOP/s", "frac": round(ach / peak, 4), "traffic": traffic,
                "flop_per_launch": flop, "ms_per_launch": round(ms, 4)}

    roof = roofline(prof32, PEAK_FP32, "fp32")
    roof["bound"] = "fp32-valu"
    roof64 = roofline(prof64, PEAK_FP64, "fp64")
    roof64["bound"] = "fp64-valu"

    eef = None
    if args.summation == "fmm" and not args.no_eef:
        # CppeState::induced_moments_eef: three unit-field solves, one multi-right-hand-side device
        # solve against three single solves (host buffers in and out in both, FP64 and mixed form)
        R = np.zeros((3, 3 * w.n))
        for a in range(3):
            R[a, a::3] = 1.0
        eef = {}
        for name, S in (("fp64", S64), ("mixed", S32)):
            S.solve_multi(R, thresh=1e-8, maxiter=200)
            t0 = time.perf_counter()
            _, itm, _ = S.solve_multi(R, thresh=1e-8, maxiter=200)
            tm = (time.perf_counter() - t0) * 1e3
            t0 = time.perf_counter()
            its = [S.solve(R[a], thresh=1e-8, maxiter=200)[1] for a in range(3)]
            ts = (time.perf_counter() - t0) * 1e3
            eef[name] = {"multi_rhs_ms": round(tm, 3), "three_solves_ms": round(ts, 3),
                         "iterations": [int(v) for v in itm], "three_solves_iterations": its}
        log("eef", eef)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        try:
            cpu = cpu_baseline(w, rhs, int(np.median(its64)), args.summation, ncrit, order, theta, damping)
        except Exception as e:  # report, never hide
            cpu = {"value": None, "error": str(e)}
        log("cpu baseline", cpu)

    out = {
        "metric": METRIC,
        "value": round(wall32, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall32, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 applies in f64 refinement",
        "data": "synthetic (seeded water box)",
        "config": {"workload": "%d-site water box (%d waters), %s summation theta %.1f order %d ncrit %d, Thole %.4f, "
                               "CG to 1e-8" % (w.n, spec["n_waters"], args.summation, theta, order, ncrit, damping),
                   "sites": w.n, "parallelism": "single-gpu"},
        "plain_fp64_ms": round(wall64, 3),
        "speedup_vs_fp64": round(wall64 / wall32, 3),
        "device_ms": {"headline": round(dev32, 3), "fp64": round(dev64, 3)},
        "iterations": {"headline": its32, "fp64": its64},
        "ml_warm_start": ml_used,
        "fp64_true_residual_applies": info32["reliable_updates"],
        "fp32_corrected_confirmations": info32["confirmations"],
        "headline_fp64_residual": head_res,
        "ml_guess_ms": round(float(np.mean(guess_ms)), 3) if guess_ms else None,
        "mixed_alphaF_ms": round(wallm, 3),
        "mixed_alphaF_iterations": itsm,
        "mixed_alphaF_no_graphs_ms": round(walln, 3),
        "launches": {"headline": launch32, "fp64": launch64, "alphaF_one_step_per_launch_ms": round(wall1, 3),
                     "alphaF_one_step_per_launch": launch1},
        "preconditioner": {"headline": args.precond, "weight": args.precond_weight,
                           "headline_with_%s_ms" % other: round(wallo, 3),
                           "headline_with_%s_iterations" % other: itso,
                           "fp64_local_ms": round(wall64l, 3), "fp64_local_iterations": its64l},
        "mixed_zero_guess_ms": round(wallz, 3),
        "mixed_zero_guess_iterations": itsz,
        "guess_dipole_rel_err": guess_err,
        "first_call": first,
        "accuracy": acc,
        "epol_fp64": e64,
        "epol_rel_diff": abs(e32 - e64) / abs(e64),
        "sites_per_s": round(w.n / (wall32 * 1e-3), 1),
        "setup_ms": round(t_setup, 2),
        "tree": st,
        "autotune": tune,
        "phases_ms": {"headline": {k: round(v, 4) for k, v in prof32.items()},
                      "fp64": {k: round(v, 4) for k, v in prof64.items()}},
        "roofline": roof,
        "roofline_fp64": roof64,
        "eef_unit_fields": eef,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
