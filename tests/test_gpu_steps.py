"""GPU checks of several CG steps per launch (cppe_hip_set_graph_steps), through the C ABI.

The device publishes after every decision how many steps the snapshots can take before one of them
needs the host; the host launches up to that many steps (1 to 8) as one sequence, and a step
that starts after an event changes nothing. So a solve must give bit-identical moments, iteration
counts and residual histories whatever the limit, with fewer launches and host status reads at 8
than at 1 -- for one system, a snapshot batch whose snapshots converge at different iterations,
the FP64 fallback of a stagnating refinement, and kernel-by-kernel launches. The reference's loop
is a host loop (electric_fields.cc:213-234); this has no counterpart there.
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
def water(nw, seed):
    return synthetic.water_box(nw, seed, anisotropic=True)


def rhs_of(w):
    Fn = capi.nuclear_fields(w.pos, w.atoms_z.astype(float), w.atoms_pos)
    Sm = capi.System(w.pos, excl=w.excl_csr(), summation="fmm")
    return Fn + Sm.multipole_fields(w.charges, w.dipoles, w.quadrupoles)


def run(S, rhs, steps, guess=0):
    S.set_graph_steps(steps)
    S.upload_rhs(rhs)
    it, norms, _ = S.solve_resident(guess, 1e-8, 100)
    return it, np.asarray(norms), S.download_mu(), S.last_solve_launches(), S.last_solve_info()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("precond", ["jacobi", "local"])
def test_single_system_identical_for_any_limit(precision, precond):
    w = water(2000, 41)
    rhs = rhs_of(w)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP, precision=precision)
    S.set_preconditioner(precond)
    out = {st: run(S, rhs, st) for st in (1, 2, 8)}
    it1, n1, mu1, l1, info1 = out[1]
    for st in (2, 8):
        it, n, mu, l, info = out[st]
        say(precision, precond, "steps", st, "iterations", it, "launches", l, "vs", l1)
        assert it == it1 and info == info1
        np.testing.assert_array_equal(n, n1)
        np.testing.assert_array_equal(mu, mu1)
    assert out[8][3]["step_launches"] < l1["step_launches"] == it1
    assert out[8][3]["host_reads"] < l1["host_reads"]
    Bo = O.BSystem(w.pos, w.pol, w.excl_csr(), damping=DAMP, scheme="fmm")
    assert np.linalg.norm(rhs - Bo.apply(out[8][2])) < 1e-8


def test_batch_identical_for_any_limit():
    sizes = [150, 900, 420, 2500, 300, 1200, 60]
    snaps = [synthetic.water_box(nw, 8100 + k, anisotropic=True) for k, nw in enumerate(sizes)]
    direct = [True, False, True, False, True, False, True]
    rhs = np.concatenate([O.nuclear_fields(s.pos, s.atoms_z.astype(float), s.atoms_pos) for s in snaps])
    for precision in ("mixed", "fp64"):
        B = capi.Batch(snaps, direct, ncrit=32, damping=DAMP, precision=precision)
        res = {}
        for st in (1, 8):
            B.set_graph_steps(st)
            B.upload_rhs(rhs)
            B.solve_resident(0, 1e-8, 100)
            res[st] = (B.segment_results(), B.download_mu(), B.last_solve_launches())
        (it1, rn1, cv1), mu1, l1 = res[1]
        (it8, rn8, cv8), mu8, l8 = res[8]
        say(precision, "batch iterations", it1.tolist(), "launches", l1, "->", l8)
        assert cv1.all() and cv8.all()
        assert len(set(it1.tolist())) > 1  # snapshots converge at different iterations
        np.testing.assert_array_equal(it8, it1)
        np.testing.assert_array_equal(rn8, rn1)
        np.testing.assert_array_equal(mu8, mu1)
        assert l8["step_launches"] < l1["step_launches"]
        mus = B.split(mu8)
        for g in (1, 3, 6):
            s = snaps[g]
            Bo = O.BSystem(s.pos, s.pol, s.excl_csr(), damping=DAMP, scheme="direct" if direct[g] else "fmm",
                           ncrit=32)
            assert np.linalg.norm(B.split(rhs)[g] - Bo.apply(mus[g])) < 1e-8, g
        B.close()


def test_fallback_and_eager_launches_identical():
    w = water(1500, 42)
    rhs = rhs_of(w)
    S = capi.System(w.pos, w.pol, w.alpha_iso(), w.excl_csr(), summation="fmm", damping=DAMP, precision="mixed")
    S.set_refinement(1e-5, 10, 0.0)  # every true residual after a step stagnates: FP64 fallback
    a = run(S, rhs, 1)
    b = run(S, rhs, 8)
    assert a[4]["fp64_fallback"] >= 0 and a[4] == b[4]
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[2], b[2])
    S.set_refinement()
    S.set_graphs(False)  # the same sequences launched kernel by kernel
    c = run(S, rhs, 1)
    d = run(S, rhs, 8)
    S.set_graphs(True)
    e = run(S, rhs, 8)
    assert c[0] == d[0] == e[0]
    np.testing.assert_array_equal(c[2], d[2])
    np.testing.assert_array_equal(d[2], e[2])
    for bad in (0, 9, 16, -1):
        with pytest.raises(capi.CppeHipError, match="1 to 8"):
            S.set_graph_steps(bad)
