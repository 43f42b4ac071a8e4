"""CPU tests of the device CG's per-segment decisions (cppe_amd/csrc/device/cg_policy.hh), built for
the host by tests/cg_policy_host_driver.cc.

Plain form: the reference recurrence's stopping rule (electric_fields.cc:213-229: one step before
the first test, stop when ||r|| < thresh). Mixed form: reliable updates on a residual drop, on a
stalled recursive residual, confirmation of a residual below the threshold by an FP64 true
residual or by its FP32-corrected form (accepted only with the error bound below the threshold),
and the stagnation test that sends the solve on in the plain FP64 form.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEV = os.path.join(ROOT, "cppe_amd", "csrc", "device")
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)

# indices of the per-segment tables (cg_policy.hh SegD / SegI)
D_RZ, D_BETA, D_RNORM, D_RLAST, D_RZN, D_RBEST, D_LOGQ, D_GAP = range(8)
I_ACTIVE, I_NEEDREL, I_ITERS, I_BAD, I_PENDING, I_ELIGIBLE, I_SINCE, I_STEPS, I_STAG, I_HASREL = range(10)
THRESH = 1e-8


@pytest.fixture(scope="module")
def pol(tmp_path_factory):
    out = tmp_path_factory.mktemp("cgpolicy") / "libcgpolicy.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", DEV,
                    os.path.join(ROOT, "tests", "cg_policy_host_driver.cc"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    dd, ii = ctypes.c_double, ctypes.c_int
    L.pol_sizes.argtypes = [_i, _i]
    L.pol_init.argtypes = [_d, _i, dd, dd, ii, _d]
    L.pol_restart.argtypes = [_d, _i, dd, dd, ii, _d]
    L.pol_step.argtypes = [_d, _i, dd, ii, ii, _d, _d, ii, dd, dd]
    L.pol_true.argtypes = [_d, _i, dd, ii, ii, _d, _d, ii]
    L.pol_beta.argtypes = [_d, _i, dd]
    L.pol_confirm.argtypes = [_d, _i, dd, dd, _d, _d, ii]
    L.pol_horizon.argtypes = [_d, _i, ii, _d, _i, _i]
    nd, ni = ctypes.c_int(), ctypes.c_int()
    L.pol_sizes(ctypes.byref(nd), ctypes.byref(ni))
    assert (nd.value, ni.value) == (8, 10)
    return L


class Seg:
    def __init__(self, L, drop=1e-5, stall=10, min_progress=0.5, eta=0.0, cap=64):
        self.L = L
        self.sd = np.zeros(8)
        self.si = np.zeros(10, dtype=np.int32)
        self.prm = np.array([THRESH, drop, min_progress, stall, eta, eta], dtype=np.float64)
        self.hist = np.full(cap, np.nan)

    def _a(self):
        return self.sd.ctypes.data_as(_d), self.si.ctypes.data_as(_i)

    def init(self, rn, mixed):
        self.L.pol_init(*self._a(), 1.0, rn * rn, int(mixed), self.prm.ctypes.data_as(_d))

    def restart(self, rn, it):
        self.L.pol_restart(*self._a(), 1.0, rn * rn, it, self.prm.ctypes.data_as(_d))

    def step(self, rn, it, mixed, pap=0.0, apap=0.0):
        self.L.pol_step(*self._a(), rn * rn, it, int(mixed), self.prm.ctypes.data_as(_d),
                        self.hist.ctypes.data_as(_d), self.hist.size, pap, apap)

    def horizon(self, mixed):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.L.pol_horizon(*self._a(), int(mixed), self.prm.ctypes.data_as(_d), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def true(self, rn, it, stepped):
        self.L.pol_true(*self._a(), rn * rn, it, stepped, self.prm.ctypes.data_as(_d),
                        self.hist.ctypes.data_as(_d), self.hist.size)

    def confirm(self, rc, y, eta):
        self.prm[4] = eta  # the corrected confirmation's eta (the gap bound keeps its own)
        return self.L.pol_confirm(*self._a(), rc * rc, y * y, self.prm.ctypes.data_as(_d),
                                  self.hist.ctypes.data_as(_d), self.hist.size)


def test_plain_form_reference_stopping_rule(pol):
    s = Seg(pol)
    s.init(1e-9, mixed=False)  # below the threshold: still one step (electric_fields.cc:213-229)
    assert s.si[I_ACTIVE] == 1 and s.si[I_NEEDREL] == 0
    s.step(5e-10, 0, mixed=False)
    assert s.si[I_ACTIVE] == 0 and s.si[I_ITERS] == 1 and s.hist[0] == 5e-10
    z = Seg(pol)
    z.init(0.0, mixed=False)  # an exactly zero residual stops at 0 iterations
    assert z.si[I_ACTIVE] == 0
    # a plain solve never asks for replacements, however flat its residual
    p = Seg(pol, stall=2)
    p.init(1.0, mixed=False)
    for it in range(8):
        p.step(0.9, it, mixed=False)
        assert p.si[I_NEEDREL] == 0 and p.si[I_ACTIVE] == 1 and p.si[I_STAG] == 0
    assert np.all(p.hist[:8] == 0.9)


def test_mixed_form_drop_pending_confirmation(pol):
    s = Seg(pol)
    s.init(1.0, mixed=True)
    assert s.si[I_ACTIVE] == 1 and s.si[I_NEEDREL] == 0
    rn, it = 1.0, 0
    while True:  # a healthy recursive residual, 2.5x per step
        rn /= 2.5
        s.step(rn, it, mixed=True)
        if s.si[I_NEEDREL]:
            break
        it += 1
    assert rn < 1e-5 <= rn * 2.5  # asked exactly when it fell by the drop
    s.true(1.3e-5, it, 1)  # FP64 true residual: a little drift, real progress
    assert s.si[I_STAG] == 0 and s.si[I_ACTIVE] == 1 and s.sd[D_RLAST] == 1.3e-5 and s.hist[it] == 1.3e-5
    it += 1
    rn = 1.3e-5
    while s.si[I_ACTIVE]:
        rn /= 2.5
        s.step(rn, it, mixed=True)
        it += 1
    assert s.si[I_PENDING] == 1 and s.si[I_ITERS] == it and rn < THRESH
    s.true(2e-8, it, 0)  # not confirmed: resumes from the true residual
    assert s.si[I_ACTIVE] == 1 and s.si[I_PENDING] == 0 and s.si[I_STAG] == 0
    assert s.hist[it - 1] == 2e-8
    s.step(4e-9, it, mixed=True)
    assert s.si[I_PENDING] == 1
    s.true(5e-9, it + 1, 0)  # confirmed
    assert s.si[I_ACTIVE] == 0 and s.si[I_PENDING] == 0 and s.si[I_ITERS] == it + 1


def test_mixed_form_fp32_corrected_confirmation(pol):
    s = Seg(pol)
    s.init(1.0, mixed=True)
    assert s.si[I_HASREL] == 0
    s.step(0.5e-8, 0, mixed=True)  # below the threshold before any FP64 true residual
    assert s.si[I_PENDING] == 1
    assert s.confirm(1e-9, 1.0, 1e-5) == 0 and s.si[I_PENDING] == 1  # nothing stored: not tried
    s.true(3e-6, 1, 0)  # FP64: not confirmed, resumes; its moments and residual are now stored
    assert s.si[I_ACTIVE] == 1 and s.si[I_HASREL] == 1
    rn, it = 3e-6, 1
    while s.si[I_ACTIVE]:
        rn /= 3.0
        s.step(rn, it, mixed=True)
        it += 1
    assert s.si[I_PENDING] == 1 and s.si[I_ITERS] == it
    # the corrected residual is below the threshold, but not with the error bound: stays pending
    assert s.confirm(9.5e-9, 3e-6, 1e-3) == 1
    assert s.si[I_PENDING] == 1 and s.si[I_ACTIVE] == 0 and np.isnan(s.hist[it - 1])
    # corrected residual and bound below the threshold: converged at the iteration where it froze
    assert s.confirm(9.5e-9, 3e-6, 1e-5) == 1
    assert s.si[I_PENDING] == 0 and s.si[I_ACTIVE] == 0 and s.si[I_ITERS] == it
    assert s.hist[it - 1] == 9.5e-9 and s.sd[D_RNORM] == 9.5e-9
    # a non-finite correction is never accepted
    t = Seg(pol)
    t.init(1.0, mixed=True)
    t.step(5e-6, 0, mixed=True)  # fell by the drop: FP64 true residual
    assert t.si[I_NEEDREL] == 1
    t.true(4e-6, 0, 1)
    assert t.si[I_HASREL] == 1 and t.si[I_ACTIVE] == 1
    t.step(1e-9, 1, mixed=True)
    assert t.si[I_PENDING] == 1
    assert t.confirm(float("nan"), 1.0, 1e-5) == 1
    assert t.si[I_PENDING] == 1


def test_mixed_form_gap_bound(pol):
    """After an FP64 true residual, ||r|| + eta sum |alpha| ||B32 p|| below the threshold ends the
    solve without a confirmation; anything short of that waits for one."""
    def replaced(eta):
        s = Seg(pol, eta=eta)
        s.init(1.0, mixed=True)  # D_RZ = 1: alpha = 1 / p.Ap below
        s.step(5e-6, 0, mixed=True, pap=1.0, apap=1e-2)
        assert s.si[I_NEEDREL] == 1 and s.sd[D_GAP] == (pytest.approx(0.1) if eta > 0 else 0.0)
        s.true(4e-6, 0, 1)
        assert s.si[I_HASREL] == 1 and s.sd[D_GAP] == 0.0  # the gap restarts at a true residual
        for it, rn in ((1, 4e-7), (2, 4e-8), (3, 2e-8)):  # ||alpha B32 p|| = 1e-4 per step
            s.step(rn, it, mixed=True, pap=2.0, apap=4e-8)
        assert s.si[I_ACTIVE] == 1
        assert s.sd[D_GAP] == (pytest.approx(3e-4) if eta > 0 else 0.0)  # tracked only with eta > 0
        return s

    s = replaced(1e-5)  # bound 5e-9 + 3e-9 < 1e-8: converged, never pending
    s.step(5e-9, 4, mixed=True, pap=1e9, apap=0.0)
    assert s.si[I_ACTIVE] == 0 and s.si[I_PENDING] == 0 and s.si[I_ITERS] == 5 and s.hist[4] == 5e-9
    s = replaced(1e-5)  # 8e-9 + 3e-9 > 1e-8: waits for a confirmation
    s.step(8e-9, 4, mixed=True, pap=1e9, apap=0.0)
    assert s.si[I_PENDING] == 1 and s.si[I_ITERS] == 5
    s = replaced(0.0)  # eta 0: every confirmation in FP64
    s.step(1e-12, 4, mixed=True, pap=1e9, apap=0.0)
    assert s.si[I_PENDING] == 1
    s = Seg(pol, eta=1e-5)  # no FP64 true residual yet: the FP32 initial residual bounds nothing
    s.init(1e-2, mixed=True)
    s.step(1e-12, 0, mixed=True, pap=1.0, apap=0.0)
    assert s.si[I_PENDING] == 1


def test_mixed_form_stalled_residual_asks_for_true_residual(pol):
    s = Seg(pol, drop=1e-12, stall=10)
    s.init(1.0, mixed=True)
    # with so small a drop, the first FP32 residual lies within one drop of the threshold and is
    # replaced before any step (no stagnation: no step taken)
    assert s.si[I_NEEDREL] == 1
    s.true(1.0, 0, 0)
    assert s.si[I_NEEDREL] == 0 and s.si[I_STAG] == 0 and s.si[I_ACTIVE] == 1
    for it, rn in enumerate([0.4, 0.15, 0.06, 0.02]):
        s.step(rn, it, mixed=True)
    assert s.si[I_NEEDREL] == 0 and s.si[I_SINCE] == 0
    # levels off: no halving for `stall` steps
    for k in range(10):
        s.step(0.015 * (1 - 0.01 * k), 4 + k, mixed=True)
        assert s.si[I_NEEDREL] == (1 if k == 9 else 0), k
    s.true(0.016, 13, 1)  # the true residual made progress since the start: no stagnation
    assert s.si[I_STAG] == 0 and s.si[I_SINCE] == 0 and s.si[I_STEPS] == 0


def test_mixed_form_stagnation_then_plain_restart(pol):
    s = Seg(pol, stall=3)
    s.init(1.0, mixed=True)
    s.step(0.4, 0, mixed=True)
    s.step(0.35, 1, mixed=True)
    s.step(0.3, 2, mixed=True)
    s.step(0.3, 3, mixed=True)
    assert s.si[I_NEEDREL] == 1
    s.true(0.9, 3, 1)  # true residual barely below the last one: the refinement stagnates
    assert s.si[I_STAG] == 1 and s.si[I_ACTIVE] == 1
    s.restart(0.9, 4)
    assert s.si[I_STAG] == 0 and s.si[I_ACTIVE] == 1 and s.sd[D_BETA] == 0.0 and s.sd[D_RLAST] == 0.9
    # a finished segment is left alone by the restart
    f = Seg(pol)
    f.init(1.0, mixed=True)
    f.si[I_ACTIVE] = 0
    f.si[I_ITERS] = 7
    f.restart(0.5, 12)
    assert f.si[I_ACTIVE] == 0 and f.si[I_ITERS] == 7 and f.sd[D_RLAST] == 1.0
    # a restart whose true residual already meets the threshold finishes at the current count
    c = Seg(pol)
    c.init(1.0, mixed=True)
    c.restart(5e-9, 12)
    assert c.si[I_ACTIVE] == 0 and c.si[I_ITERS] == 12


def test_mixed_form_no_false_stagnation_before_a_step(pol):
    # a guess whose FP32 residual is near the threshold is replaced before any step; the true
    # residual equals the FP32 one, which is not stagnation
    s = Seg(pol)
    s.init(5e-5, mixed=True)
    assert s.si[I_NEEDREL] == 1
    s.true(5.0001e-5, 0, 0)
    assert s.si[I_STAG] == 0 and s.si[I_ACTIVE] == 1
    c = Seg(pol)
    c.init(5e-9, mixed=True)
    c.true(6e-9, 0, 0)  # confirmed before any step: 0 iterations
    assert c.si[I_ACTIVE] == 0 and c.si[I_ITERS] == 0


def test_beta(pol):
    s = Seg(pol)
    s.init(1.0, mixed=False)
    s.sd[D_RZ] = 4.0
    pol.pol_beta(s.sd.ctypes.data_as(_d), s.si.ctypes.data_as(_i), 1.0)
    assert s.sd[D_BETA] == 0.25 and s.sd[D_RZ] == 1.0


def test_launch_horizon_predicts_events(pol):
    """The horizon (steps before a segment needs the host) follows the mean log reduction per step,
    rounded down, so a launch of that many steps ends at or before the event; 1 while the rate is
    unknown; 0 while a replacement is asked for; inactive segments never limit it."""
    far = pol.pol_far()
    s = Seg(pol)
    s.init(1.0, mixed=False)
    assert s.horizon(False) == (far, 1)  # no rate yet: one step at a time
    s.step(0.25, 0, mixed=False)
    assert s.sd[D_LOGQ] == pytest.approx(np.log(0.25))
    # log(1e-8 / 0.25) / log(0.25) = 12.29 -> 12 steps to the threshold
    assert s.horizon(False) == (far, 12)
    s.step(0.0625 * 0.5, 1, mixed=False)  # log q = mean(log 0.25, log 0.125)
    q = 0.5 * (np.log(0.25) + np.log(0.125))
    assert s.horizon(False) == (far, int(np.log(1e-8 / 0.03125) / q))
    # replay a geometric sequence: the predicted step count never overshoots the event
    for rate in (0.2, 0.3, 0.45):
        g = Seg(pol)
        g.init(1.0, mixed=False)
        rn, it, stops = 1.0, 0, []
        while g.si[I_ACTIVE]:
            rn *= rate
            g.step(rn, it, mixed=False)
            it += 1
            if g.si[I_ACTIVE]:
                stops.append((it, g.horizon(False)[1]))
        for at, h in stops:
            assert at + h <= it  # the segment leaves the active set at step `it`
    # mixed form: the replacement trigger (drop x last true residual) comes first
    m = Seg(pol, drop=1e-5)
    m.init(1.0, mixed=True)
    m.step(0.25, 0, mixed=True)
    to_rel, to_done = m.horizon(True)
    assert to_rel == int(np.log(1e-5 / 0.25) / np.log(0.25)) == 7
    assert to_done == 12
    # stall limit caps it
    st = Seg(pol, drop=1e-5, stall=3)
    st.init(1.0, mixed=True)
    st.step(0.9, 0, mixed=True)  # not halved: one step towards the stall limit
    assert st.si[I_SINCE] == 1
    assert st.horizon(True)[0] <= 2
    # a residual that does not fall: one step at a time
    u = Seg(pol)
    u.init(1.0, mixed=True)
    u.step(1.2, 0, mixed=True)
    assert u.horizon(True) == (1, 1)
    # a pending replacement request: the host is needed now
    r = Seg(pol, drop=1e-2)
    r.init(1.0, mixed=True)
    r.step(1e-3, 0, mixed=True)
    assert r.si[I_NEEDREL] == 1 and r.horizon(True) == (0, 0)
    # an inactive (converged) segment never limits the launch
    d = Seg(pol)
    d.init(1.0, mixed=False)
    d.step(1e-9, 0, mixed=False)
    assert d.si[I_ACTIVE] == 0 and d.horizon(False) == (far, 0)
