"""GPU checks of the launch-shape autotuner and of every launch shape it can pick.

Each near-field variant x block size, each M2L group width and each far/near overlap form must give the
same B apply as the default shape (FP64: 1e-13 relative, summation order only; FP32: 1e-6 relative), and a solve with
the tuned shapes must stop on the FP64 residual with dipoles equal to the default's to 1e-6 (both stop at 1e-8).
"""
import itertools
import sys

import numpy as np
import pytest

from cppe_amd import capi, synthetic

pytestmark = pytest.mark.gpu
DAMP = 2.1304


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def system(precision):
    w = synthetic.water_box(1500, 7200, anisotropic=True)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", ncrit=64, damping=DAMP,
                    precision=precision)
    return w, S


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_launch_shape_same_apply(precision):
    w, S = system(precision)
    x = np.random.default_rng(11).normal(size=3 * w.n)
    ref = S.apply_b(x)
    tol = 1e-13 if precision == "fp64" else 1e-6
    prec = 0 if precision == "fp64" else 1
    for (v, b), g, o in itertools.product(capi.System.NEAR_CANDIDATES, capi.System.M2L_CANDIDATES, (0, 1, 2)):
        choice = np.array([0, 256, 64, 0, 0, 256, 64, 0], dtype=np.int32)
        choice[4 * prec:4 * prec + 4] = (v, b, g, o)
        say("shape", precision, v, b, g, o)
        S.set_tune(choice)
        y = S.apply_b(x)
        assert np.abs(y - ref).max() < tol * np.abs(ref).max(), (v, b, g, o)


def test_autotune_then_solve():
    w, S = system("mixed")
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr(), summation="fmm")
    F = Fn + Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)
    S.upload_rhs(F)
    S.solve_resident(0, 1e-8, 100)
    mu0 = S.download_mu()
    rep = S.autotune(reps=2)
    say("autotune", rep)
    for prec in ("fp64", "fp32"):
        r = rep[prec]
        assert (r["near_variant"], r["near_block"]) in capi.System.NEAR_CANDIDATES
        assert r["m2l_group"] in capi.System.M2L_CANDIDATES and r["overlap"] in (0, 1, 2, 3)
        assert r["near_launches"] in (1, 2, 4) and (r["overlap"] > 0 or r["near_launches"] == 1)
        assert all(t > 0 for t in r["near_ms"].values()) and all(t > 0 for t in r["m2l_ms"].values())
    it, norms, _ = S.solve_resident(0, 1e-8, 100)
    assert norms[-1] < 1e-8
    mu = S.download_mu()
    assert np.abs(mu - mu0).max() < 1e-6 * np.abs(mu0).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_with_each_overlap_form(precision):
    """Graph-captured solves with the far field beside the near field (overlap 1), with the
    upward pass first (overlap 2) and with the far field enqueued ahead of the near field (overlap 3),
    each with the near field in 1, 2 and 4 launches (codes + 4, + 8), stop on the FP64 residual with
    the series form's dipoles."""
    w, S = system(precision)
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    F = Fn + capi.System(w.pos, excl=w.excl_csr(), summation="fmm").multipole_fields(w.charges, w.dipoles,
                                                                                        w.quadrupoles)
    S.upload_rhs(F)
    S.set_tune(np.array([8, 256, 16, 0, 8, 256, 16, 0], dtype=np.int32))
    S.solve_resident(0, 1e-8, 100)
    mu0 = S.download_mu()
    for o in (1, 2, 3, 5, 6, 7, 9, 10, 11):
        say("overlap", o, precision)
        S.set_tune(np.array([8, 256, 16, o, 8, 256, 16, o], dtype=np.int32))
        it, norms, _ = S.solve_resident(0, 1e-8, 100)
        assert norms[-1] < 1e-8
        mu = S.download_mu()
        assert np.abs(mu - mu0).max() < 1e-6 * np.abs(mu0).max(), o


def test_set_tune_rejects_unknown_shapes():
    _, S = system("fp64")
    with pytest.raises(capi.CppeHipError):
        S.set_tune([12, 256, 64, 0, 0, 256, 64, 0])
    with pytest.raises(capi.CppeHipError):
        S.set_tune([0, 96, 64, 0, 0, 256, 64, 0])
    for bad_overlap in (4, 8, 12, 13, 16, -1):  # split series form, 8 launches, ...
        with pytest.raises(capi.CppeHipError):
            S.set_tune([5, 256, 64, bad_overlap, 0, 256, 64, 0])


def test_fp32_staged_with_undamped_sites():
    """Sites with alpha_iso = 0 are undamped: the FP32 near field then takes the branching pair
    (the branch-free form assumes every site damped). Both forms, on the same system, against
    the FP64 apply of the same shape."""
    w = synthetic.water_box(1500, 7201, anisotropic=True)
    iso = w.alpha_iso().copy()
    iso[::7] = 0.0
    x = np.random.default_rng(12).normal(size=3 * w.n)
    out = {}
    for precision in ("fp64", "mixed"):
        S = capi.System(w.pos, w.pol, iso, w.excl_csr(), summation="fmm", ncrit=64, damping=DAMP,
                        precision=precision)
        S.set_tune(np.array([8, 256, 16, 0, 8, 256, 16, 0], dtype=np.int32))
        out[precision] = S.apply_b(x)
    ref = out["fp64"]
    assert np.abs(out["mixed"] - ref).max() < 1e-5 * np.abs(ref).max()
    # and the undamped sites really are undamped: the all-damped system differs
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", ncrit=64, damping=DAMP)
    assert np.abs(S.apply_b(x) - ref).max() > 1e-3 * np.abs(ref).max()
