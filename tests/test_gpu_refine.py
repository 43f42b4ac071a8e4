"""GPU checks of the FP32-refined solve's safety nets (reliable updates on a stalled residual, and
the switch to the plain FP64 form when the refinement stagnates), through the C ABI.

The conditions are forced from the host with ordinary settings (cppe_hip_set_refinement): a
replacement drop too small to ever trigger, and a progress bar that every replacement misses.
Each solve must still stop on the FP64 residual, checked with the oracle's operator. The
reference's solve is FP64 only and raises after maxiter (electric_fields.cc:235-237).
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
def case():
    w = synthetic.water_box(1000, 1001, anisotropic=True)
    rhs = O.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Bo = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=DAMP, scheme="fmm", ncrit=32)
    mu64, it64, _ = Bo.solve(rhs, thresh=1e-8)
    return w, rhs, Bo, mu64


def mixed_system(w):
    return capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", ncrit=32, damping=DAMP,
                       precision="mixed")


@pytest.mark.parametrize("stall", [10, 2])
def test_tiny_drop_still_converges(stall):
    """A drop that never triggers: the recursive residual is replaced when it stalls or falls
    below the threshold, and the solve converges on the FP64 residual."""
    w, rhs, Bo, mu64 = case()
    say("tiny drop, stall", stall)
    S = mixed_system(w)
    S.set_refinement(drop=1e-12, stall=stall, min_progress=0.5)
    mu, it, norms = S.solve(rhs, thresh=1e-8, maxiter=50)
    info = S.last_solve_info()
    say("iterations", it, info)
    assert np.linalg.norm(rhs - Bo.apply(mu)) < 1e-8
    e, e64 = -0.5 * mu @ rhs, -0.5 * mu64 @ rhs
    assert abs(e - e64) <= 1e-8 * abs(e64)
    assert info["reliable_updates"] >= 2
    S.close()


def test_stagnation_continues_in_fp64():
    """Every replacement after a step counts as stagnation (min_progress 0): the solve continues from
    its current moments in the plain FP64 form and converges there."""
    w, rhs, Bo, mu64 = case()
    S = mixed_system(w)
    S.set_refinement(drop=1e-5, stall=10, min_progress=0.0)
    mu, it, norms = S.solve(rhs, thresh=1e-8, maxiter=50)
    info = S.last_solve_info()
    say("forced fallback: iterations", it, info)
    assert info["fp64_fallback"] > 0
    assert np.linalg.norm(rhs - Bo.apply(mu)) < 1e-8
    e, e64 = -0.5 * mu @ rhs, -0.5 * mu64 @ rhs
    assert abs(e - e64) <= 1e-8 * abs(e64)
    assert len(norms) == it and norms[-1] < 1e-8
    # the default settings do not switch on this well-conditioned system
    S.set_refinement()
    S.solve(rhs, thresh=1e-8, maxiter=50)
    assert S.last_solve_info()["fp64_fallback"] == -1
    S.close()


def test_stagnation_in_a_batch():
    """Snapshots that stagnate restart in FP64 while finished ones stay finished."""
    sizes = [120, 400, 260]
    snaps = [synthetic.water_box(nw, 7100 + k, anisotropic=True) for k, nw in enumerate(sizes)]
    direct = [False, False, True]
    rhs = np.concatenate([O.nuclear_fields(s.pos, s.atoms_z.astype(float), s.atoms_pos) for s in snaps])
    B = capi.Batch(snaps, direct, ncrit=32, damping=DAMP, precision="mixed")
    B.set_refinement(drop=1e-5, stall=10, min_progress=0.0)
    B.upload_rhs(rhs)
    B.solve_resident(0, 1e-8, 60)
    iters, rn, conv = B.segment_results()
    info = B.last_solve_info()
    say("batch forced fallback", iters, rn, info)
    assert conv.all() and info["fp64_fallback"] > 0
    mu = B.split(B.download_mu())
    for g, s in enumerate(snaps):
        Bo = O.BSystem(s.pos, s.pol, s.excl_csr(), damping=DAMP, scheme="direct" if direct[g] else "fmm", ncrit=32)
        assert np.linalg.norm(B.split(rhs)[g] - Bo.apply(mu[g])) < 1e-8, g
    B.close()


def test_nonconvergence_after_fallback_raises():
    """Too few iterations for either form: the reference's error, unchanged."""
    w, rhs, Bo, mu64 = case()
    S = mixed_system(w)
    S.set_refinement(drop=1e-5, stall=10, min_progress=0.0)
    with pytest.raises(capi.CppeHipError, match="Failed to converge induced dipole moments."):
        S.solve(rhs, thresh=1e-8, maxiter=8)
    S.close()


def test_refinement_arguments_rejected():
    w, rhs, Bo, mu64 = case()
    S = mixed_system(w)
    for bad in ((0.0, 10, 0.5), (1.0, 10, 0.5), (1e-5, 0, 0.5), (1e-5, 10, -1.0)):
        with pytest.raises(capi.CppeHipError):
            S.set_refinement(*bad)
    S.close()
