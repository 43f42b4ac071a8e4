"""ctypes binding of the C ABI in include/cppe_hip.h (libcppe_hip.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcppe_hip.so")

OK, ERR_ARG, ERR_DEVICE, ERR_NOCONV, ERR_TREE, ERR_UNSUPPORTED = 0, 1, 2, 3, 4, 5
SUM_DIRECT, SUM_FMM = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

EXPORTS = [
    "cppe_hip_last_error", "cppe_hip_device_count", "cppe_hip_system_create", "cppe_hip_system_destroy",
    "cppe_hip_field", "cppe_hip_multipole_fields", "cppe_hip_apply_b", "cppe_hip_apply_alpha",
    "cppe_hip_apply_alpha_inv", "cppe_hip_dense_b", "cppe_hip_dense_b_inverse", "cppe_hip_solve", "cppe_hip_upload_rhs",
    "cppe_hip_solve_resident", "cppe_hip_download_mu", "cppe_hip_nuclear_fields",
    "cppe_hip_nuclear_multipole_energy", "cppe_hip_nuclear_multipole_energy_noalloc", "cppe_hip_dot", "cppe_hip_tree_stats", "cppe_hip_tree_export",
    "cppe_hip_profile_apply", "cppe_hip_batch_create", "cppe_hip_segment_results", "cppe_hip_set_graphs",
    "cppe_hip_autotune", "cppe_hip_set_tune", "cppe_hip_set_reliable_drop", "cppe_hip_set_refinement", "cppe_hip_set_preconditioner",
    "cppe_hip_last_solve_info", "cppe_hip_set_graph_steps", "cppe_hip_last_solve_launches",
    "cppe_hip_field_targets", "cppe_hip_nuclear_field_gradient", "cppe_hip_nuclear_multipole_gradient",
    "cppe_hip_solve_multi", "cppe_hip_apply_b_multi",
    "cppe_hip_mlp_create", "cppe_hip_mlp_destroy", "cppe_hip_mlp_set", "cppe_hip_mlp_get", "cppe_hip_ml_features",
    "cppe_hip_mlp_forward", "cppe_hip_mlp_loss_grad", "cppe_hip_mlp_reserve", "cppe_hip_mlp_collect",
    "cppe_hip_mlp_prepare", "cppe_hip_mlp_train", "cppe_hip_ml_guess", "cppe_hip_mlp_set_objective",
    "cppe_hip_mlp_samples", "cppe_hip_set_confirmation", "cppe_hip_last_solve_confirmations",
]
MLP_IN = 64
MLP_PARAMS = 25219


class Options(ctypes.Structure):
    _fields_ = [("summation", ctypes.c_int), ("ncrit", ctypes.c_int), ("order", ctypes.c_int),
                ("theta", ctypes.c_double), ("damping", ctypes.c_double), ("precision", ctypes.c_int)]


class CppeHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_ll = ctypes.POINTER(ctypes.c_longlong)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m cppe_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.cppe_hip_last_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        L.cppe_hip_device_count.argtypes = [_i]
        L.cppe_hip_system_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _d, _d, _d, _i, _i,
                                             ctypes.POINTER(Options)]
        L.cppe_hip_system_destroy.argtypes = [vp]
        L.cppe_hip_field.argtypes = [vp, ctypes.c_int, _d, _d]
        L.cppe_hip_multipole_fields.argtypes = [vp, _d, _d, _d, _d]
        for f in ("cppe_hip_apply_b", "cppe_hip_apply_alpha", "cppe_hip_apply_alpha_inv"):
            getattr(L, f).argtypes = [vp, _d, _d]
        L.cppe_hip_dense_b.argtypes = [vp, _i, _d]
        L.cppe_hip_dense_b_inverse.argtypes = [vp, _i, _d]
        L.cppe_hip_solve.argtypes = [vp, _d, _d, ctypes.c_int, ctypes.c_double, ctypes.c_int, _d, _i, _d]
        L.cppe_hip_upload_rhs.argtypes = [vp, _d]
        L.cppe_hip_solve_resident.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _i, _d, _d]
        L.cppe_hip_download_mu.argtypes = [vp, _d]
        L.cppe_hip_nuclear_fields.argtypes = [ctypes.c_int, _d, ctypes.c_int, _d, _d, _d]
        L.cppe_hip_nuclear_multipole_energy.argtypes = [ctypes.c_int, _d, _i, _d, ctypes.c_int, _d, _d, _d]
        # device pointers (integers) and a stream handle
        L.cppe_hip_nuclear_multipole_energy_noalloc.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, vp, vp, vp,
                                                                vp]
        L.cppe_hip_nuclear_field_gradient.argtypes = [ctypes.c_int, _d, ctypes.c_int, _d, _d, _d]
        L.cppe_hip_apply_b_multi.argtypes = [vp, ctypes.c_int, _d, _d]
        L.cppe_hip_solve_multi.argtypes = [vp, ctypes.c_int, _d, _d, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           _d, _i, _d]
        L.cppe_hip_nuclear_multipole_gradient.argtypes = [ctypes.c_int, _d, _i, _d, ctypes.c_int, _d, _d, _d]
        L.cppe_hip_dot.argtypes = [ctypes.c_longlong, _d, _d, _d]
        L.cppe_hip_tree_stats.argtypes = [vp, _ll]
        L.cppe_hip_tree_export.argtypes = [vp, _d, _ll, _ll, _i, _ll, _ll]
        L.cppe_hip_profile_apply.argtypes = [vp, ctypes.c_int, ctypes.c_int, _d]
        L.cppe_hip_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _i, _i, _d, _d, _d, _i, _i,
                                            ctypes.POINTER(Options)]
        L.cppe_hip_segment_results.argtypes = [vp, _i, _d, _i]
        L.cppe_hip_set_graphs.argtypes = [vp, ctypes.c_int]
        L.cppe_hip_autotune.argtypes = [vp, ctypes.c_int, _i, _d]
        L.cppe_hip_set_tune.argtypes = [vp, _i]
        L.cppe_hip_set_reliable_drop.argtypes = [vp, ctypes.c_double]
        L.cppe_hip_set_refinement.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.cppe_hip_set_preconditioner.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.cppe_hip_last_solve_info.argtypes = [vp, _i]
        L.cppe_hip_set_confirmation.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.cppe_hip_last_solve_confirmations.argtypes = [vp, _i]
        L.cppe_hip_set_graph_steps.argtypes = [vp, ctypes.c_int]
        L.cppe_hip_last_solve_launches.argtypes = [vp, _i]
        L.cppe_hip_field_targets.argtypes = [vp, _d, ctypes.c_int, _i, _d]
        _f = ctypes.POINTER(ctypes.c_float)
        L.cppe_hip_mlp_create.argtypes = [ctypes.POINTER(vp)]
        L.cppe_hip_mlp_destroy.argtypes = [vp]
        L.cppe_hip_mlp_set.argtypes = [vp, _f, _f, _f]
        L.cppe_hip_mlp_get.argtypes = [vp, _f, _f, _f]
        L.cppe_hip_ml_features.argtypes = [vp, _f, _d]
        L.cppe_hip_mlp_forward.argtypes = [vp, _f, ctypes.c_longlong, _f]
        L.cppe_hip_mlp_loss_grad.argtypes = [vp, _f, _f, _f, _f, ctypes.c_longlong, _f, _d]
        L.cppe_hip_mlp_reserve.argtypes = [vp, ctypes.c_longlong]
        L.cppe_hip_mlp_collect.argtypes = [vp, vp]
        L.cppe_hip_mlp_prepare.argtypes = [vp]
        L.cppe_hip_mlp_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint, _d]
        L.cppe_hip_ml_guess.argtypes = [vp, vp, _d]
        L.cppe_hip_mlp_set_objective.argtypes = [vp, ctypes.c_int]
        L.cppe_hip_mlp_samples.argtypes = [vp, ctypes.c_longlong, ctypes.c_longlong, _f, _f, _f]
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise CppeHipError(rc, lib().cppe_hip_last_error().decode())


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def device_count():
    c = ctypes.c_int(0)
    lib().cppe_hip_device_count(ctypes.byref(c))
    return c.value


class System:
    """One device-resident set of sites (cppe_hip_system)."""

    def __init__(self, pos, alpha6=None, alpha_iso=None, excl=None, summation="direct", ncrit=64, order=5,
                 theta=0.5, damping=0.0, precision="fp64"):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        self.n = self.pos.size // 3
        a6 = None if alpha6 is None else np.ascontiguousarray(alpha6, dtype=np.float64).reshape(-1)
        ai = None if alpha_iso is None else np.ascontiguousarray(alpha_iso, dtype=np.float64)
        off = idx = None
        if excl is not None:
            off = np.ascontiguousarray(excl[0], dtype=np.int32)
            idx = np.ascontiguousarray(excl[1], dtype=np.int32)
            if idx.size == 0:
                idx = np.zeros(1, dtype=np.int32)
        self.opts = Options(SUM_FMM if summation == "fmm" else SUM_DIRECT, int(ncrit), int(order), float(theta),
                            float(damping), PREC_MIXED if precision == "mixed" else PREC_FP64)
        h = ctypes.c_void_p()
        _check(lib().cppe_hip_system_create(ctypes.byref(h), self.n, _p(self.pos), _p(a6), _p(ai), _p(off, _i),
                                            _p(idx, _i), ctypes.byref(self.opts)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().cppe_hip_system_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _vec_op(self, fn, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(3 * self.n)
        _check(getattr(lib(), fn)(self.h, _p(x), _p(out)))
        return out

    def apply_b(self, x):
        return self._vec_op("cppe_hip_apply_b", x)

    def apply_b_multi(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3 * self.n)
        out = np.zeros_like(X)
        _check(lib().cppe_hip_apply_b_multi(self.h, X.shape[0], _p(X), _p(out)))
        return out

    def apply_alpha(self, x):
        return self._vec_op("cppe_hip_apply_alpha", x)

    def apply_alpha_inv(self, x):
        return self._vec_op("cppe_hip_apply_alpha_inv", x)

    def field(self, m, S):
        S = np.ascontiguousarray(S, dtype=np.float64)
        out = np.zeros(3 * self.n)
        _check(lib().cppe_hip_field(self.h, int(m), _p(S), _p(out)))
        return out

    def multipole_fields(self, q, mu=None, theta6=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        th = None if theta6 is None else np.ascontiguousarray(theta6, dtype=np.float64)
        out = np.zeros(3 * self.n)
        _check(lib().cppe_hip_multipole_fields(self.h, _p(q), _p(mu), _p(th), _p(out)))
        return out

    def dense_b(self, index=None):
        index = np.ascontiguousarray(np.arange(self.n) if index is None else index, dtype=np.int32)
        out = np.zeros((3 * self.n, 3 * self.n))
        _check(lib().cppe_hip_dense_b(self.h, _p(index, _i), _p(out)))
        return out

    def dense_b_inverse(self, index=None):
        index = np.ascontiguousarray(np.arange(self.n) if index is None else index, dtype=np.int32)
        out = np.zeros((3 * self.n, 3 * self.n))
        _check(lib().cppe_hip_dense_b_inverse(self.h, _p(index, _i), _p(out)))
        return out

    def solve(self, rhs, guess=None, make_guess=True, thresh=1e-8, maxiter=50, raise_on_noconv=True):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
        mu = np.zeros(3 * self.n)
        it = ctypes.c_int(0)
        norms = np.zeros(max(maxiter, 1))
        rc = lib().cppe_hip_solve(self.h, _p(rhs), _p(g), int(make_guess), float(thresh), int(maxiter), _p(mu),
                                  ctypes.byref(it), _p(norms))
        if rc != OK and (raise_on_noconv or rc != ERR_NOCONV):
            _check(rc)
        return mu, it.value, norms[:it.value]

    def solve_multi(self, rhs, guess=None, make_guess=True, thresh=1e-8, maxiter=50):
        """Solves B mu_k = rhs_k for the rows of rhs (nrhs x 3n) together; returns (mu, iters, norms)."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1, 3 * self.n)
        k = rhs.shape[0]
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64).reshape(k, 3 * self.n)
        mu = np.zeros((k, 3 * self.n))
        it = np.zeros(k, dtype=np.int32)
        norms = np.zeros((k, max(maxiter, 1)))
        _check(lib().cppe_hip_solve_multi(self.h, k, _p(rhs), _p(g), int(make_guess), float(thresh), int(maxiter),
                                          _p(mu), _p(it, _i), _p(norms)))
        return mu, it, [norms[j, :it[j]] for j in range(k)]

    def upload_rhs(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        _check(lib().cppe_hip_upload_rhs(self.h, _p(rhs)))

    def solve_resident(self, guess_mode=0, thresh=1e-8, maxiter=50):
        it = ctypes.c_int(0)
        norms = np.zeros(max(maxiter, 1))
        ms = ctypes.c_double(0.0)
        _check(lib().cppe_hip_solve_resident(self.h, int(guess_mode), float(thresh), int(maxiter),
                                             ctypes.byref(it), _p(norms), ctypes.byref(ms)))
        return it.value, norms[:it.value], ms.value

    NEAR_CANDIDATES = [(0, 64), (0, 128), (0, 256), (1, 64), (1, 128), (1, 256), (2, 64), (3, 64), (3, 128),
                       (3, 256), (4, 64), (4, 128), (4, 256), (5, 64), (5, 128), (5, 256), (6, 64), (6, 128),
                       (6, 256), (7, 64), (7, 128), (7, 256), (8, 64), (8, 128), (8, 256), (9, 64), (9, 128),
                       (9, 256), (10, 64), (10, 128), (10, 256), (11, 64), (11, 128), (11, 256)]
    NEAR_NAMES = {0: "thread", 1: "leafwave", 2: "ldsleaf", 3: "slotted", 4: "cluster", 5: "staged64", 6: "staged32",
                  7: "staged16", 8: "staged8", 9: "sym-large", 10: "sym-mid", 11: "sym-small"}
    M2L_CANDIDATES = [16, 32, 64]

    def autotune(self, reps=3):
        """Times the B-apply launch candidates on this system and keeps the fastest."""
        choice = np.zeros(8, dtype=np.int32)
        nn, nm = len(self.NEAR_CANDIDATES), len(self.M2L_CANDIDATES)
        w = nn + nm + 2
        ms = np.zeros(2 * w)
        _check(lib().cppe_hip_autotune(self.h, int(reps), _p(choice, _i), _p(ms)))
        out = {}
        for k, prec in enumerate(["fp64", "fp32"]):
            m = ms[w * k:w * (k + 1)]
            out[prec] = dict(near_variant=int(choice[4 * k]), near_block=int(choice[4 * k + 1]),
                             m2l_group=int(choice[4 * k + 2]), overlap=int(choice[4 * k + 3]) & 3,
                             near_launches=1 << (int(choice[4 * k + 3]) >> 2),
                             apply_ms={"series": round(float(m[nn + nm]), 4), "overlap": round(float(m[nn + nm + 1]), 4)},
                             near_ms={"%s/%d" % (self.NEAR_NAMES[v], b): round(float(t), 4)
                                      for (v, b), t in zip(self.NEAR_CANDIDATES, m[:nn])},
                             m2l_ms={str(g): round(float(t), 4) for g, t in zip(self.M2L_CANDIDATES, m[nn:nn + nm])})
        self.tune_choice = choice
        return out

    def set_tune(self, choice):
        c = np.ascontiguousarray(choice, dtype=np.int32)
        _check(lib().cppe_hip_set_tune(self.h, _p(c, _i)))

    def field_targets(self, targets, mu=None):
        """FP64 direct dipole field (P2P<1,3> convention) at the given sites, for mu or the resident moments."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        m = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        out = np.zeros(3 * t.size)
        _check(lib().cppe_hip_field_targets(self.h, _p(m), t.size, _p(t, _i), _p(out)))
        return out.reshape(-1, 3)

    def set_reliable_drop(self, drop):
        _check(lib().cppe_hip_set_reliable_drop(self.h, float(drop)))

    def set_refinement(self, drop=1e-5, stall=10, min_progress=0.5):
        """Mixed form: reliable-update drop, stall steps and the stagnation test (cppe_hip_set_refinement)."""
        _check(lib().cppe_hip_set_refinement(self.h, float(drop), int(stall), float(min_progress)))

    PRECOND_RADIUS = 5.6691783738  # 3 A in bohr

    def set_preconditioner(self, kind="jacobi", radius=PRECOND_RADIUS, weight=0.8):
        """CG preconditioner: "jacobi" (alpha r, the reference's) or "local" (alpha r + weight alpha T_loc
        alpha r over pairs closer than radius bohr) (cppe_hip_set_preconditioner)."""
        k = {"jacobi": 0, "local": 1}[kind]
        _check(lib().cppe_hip_set_preconditioner(self.h, k, float(radius), float(weight)))

    def set_confirmation(self, eta=1e-5, gap=True, corrected=True):
        """Mixed form: how a converged recursive residual is confirmed, eta bounding an FP32 apply's
        relative error: by the gap bound since the last FP64 true residual (gap), by that residual
        and an FP32 apply of the correction since (corrected), else by an FP64 true residual
        (cppe_hip_set_confirmation)."""
        _check(lib().cppe_hip_set_confirmation(self.h, float(eta), int(bool(gap)) | 2 * int(bool(corrected))))

    def last_solve_info(self):
        info = np.zeros(3, dtype=np.int32)
        _check(lib().cppe_hip_last_solve_info(self.h, _p(info, _i)))
        conf = np.zeros(1, dtype=np.int32)
        _check(lib().cppe_hip_last_solve_confirmations(self.h, _p(conf, _i)))
        return dict(iterations=int(info[0]), reliable_updates=int(info[1]), fp64_fallback=int(info[2]),
                    confirmations=int(conf[0]))

    def set_graphs(self, enable=True):
        _check(lib().cppe_hip_set_graphs(self.h, int(bool(enable))))

    def set_graph_steps(self, steps=8):
        """Most CG steps launched as one sequence between host status reads: 1 to 8
        (cppe_hip_set_graph_steps)."""
        _check(lib().cppe_hip_set_graph_steps(self.h, int(steps)))

    def last_solve_launches(self):
        """Step-sequence launches, steps launched in them and host status reads of the last solve."""
        a = np.zeros(3, dtype=np.int32)
        _check(lib().cppe_hip_last_solve_launches(self.h, _p(a, _i)))
        return dict(step_launches=int(a[0]), steps_launched=int(a[1]), host_reads=int(a[2]))

    def download_mu(self):
        mu = np.zeros(3 * self.n)
        _check(lib().cppe_hip_download_mu(self.h, _p(mu)))
        return mu

    def profile_apply(self, fp32=False, reps=5):
        """Mean ms per apply: upward, m2l, l2l, near (near kernel + exclusion pass), total, and
        near_kernel (the staged near-field kernel alone; -1 for the other variants)."""
        out = np.zeros(6)
        _check(lib().cppe_hip_profile_apply(self.h, int(fp32), int(reps), _p(out)))
        return dict(upward=out[0], m2l=out[1], l2l=out[2], near=out[3], total=out[4], near_kernel=out[5])

    def tree_stats(self):
        s = np.zeros(6, dtype=np.int64)
        _check(lib().cppe_hip_tree_stats(self.h, _p(s, _ll)))
        return dict(ncells=int(s[0]), nleaves=int(s[1]), n_m2l=int(s[2]), n_p2p=int(s[3]), max_level=int(s[4]),
                    p2p_site_pairs=int(s[5]))

    def tree_export(self):
        st = self.tree_stats()
        nc = st["ncells"]
        geo = np.zeros((nc, 5))
        meta = np.zeros((nc, 4), dtype=np.int64)
        start = np.zeros(nc, dtype=np.int64)
        perm = np.zeros(self.n, dtype=np.int32)
        m2l = np.zeros((max(st["n_m2l"], 1), 2), dtype=np.int64)
        p2p = np.zeros((max(st["n_p2p"], 1), 2), dtype=np.int64)
        _check(lib().cppe_hip_tree_export(self.h, _p(geo), _p(meta, _ll), _p(start, _ll), _p(perm, _i),
                                          _p(m2l, _ll), _p(p2p, _ll)))
        return dict(geo=geo, meta=meta, start=start, perm=perm, m2l=m2l[:st["n_m2l"]], p2p=p2p[:st["n_p2p"]])


class Batch(System):
    """Independent snapshots held by one handle (cppe_hip_batch_create): one tree per snapshot, shared
    field applies, a CG recurrence per snapshot. ``direct[g]`` puts snapshot g in a single leaf (exact
    direct sum); the others use the tree with ``ncrit``."""

    def __init__(self, snapshots, direct, ncrit=64, order=5, theta=0.5, damping=0.0, precision="fp64"):
        sizes = [int(s.n) for s in snapshots]
        self.seg_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        self.nseg = len(sizes)
        self.pos = np.ascontiguousarray(np.concatenate([s.pos for s in snapshots]), dtype=np.float64).reshape(-1)
        self.n = self.pos.size // 3
        a6 = np.ascontiguousarray(np.concatenate([s.pol for s in snapshots]), dtype=np.float64).reshape(-1)
        ai = np.ascontiguousarray(np.concatenate([s.alpha_iso() for s in snapshots]), dtype=np.float64)
        offs, idxs, base_e = [np.zeros(1, dtype=np.int64)], [], 0
        for s, o in zip(snapshots, self.seg_off[:-1]):
            eo, ei = s.excl_csr()
            offs.append(eo[1:].astype(np.int64) + base_e)
            idxs.append(np.where(ei >= 0, ei + o, ei))
            base_e += int(eo[-1])
        off = np.ascontiguousarray(np.concatenate(offs), dtype=np.int32)
        idx = np.ascontiguousarray(np.concatenate(idxs) if base_e else np.zeros(1), dtype=np.int32)
        nc = np.ascontiguousarray([0 if d else int(ncrit) for d in direct], dtype=np.int32)
        self.opts = Options(SUM_FMM, int(ncrit), int(order), float(theta), float(damping),
                            PREC_MIXED if precision == "mixed" else PREC_FP64)
        h = ctypes.c_void_p()
        _check(lib().cppe_hip_batch_create(ctypes.byref(h), self.nseg, _p(self.seg_off, _i), _p(nc, _i),
                                           _p(self.pos), _p(a6), _p(ai), _p(off, _i), _p(idx, _i),
                                           ctypes.byref(self.opts)))
        self.h = h

    def segment_results(self):
        it = np.zeros(self.nseg, dtype=np.int32)
        rn = np.zeros(self.nseg)
        cv = np.zeros(self.nseg, dtype=np.int32)
        _check(lib().cppe_hip_segment_results(self.h, _p(it, _i), _p(rn), _p(cv, _i)))
        return it, rn, cv.astype(bool)

    def split(self, v):
        """Per-snapshot views of a concatenated AoS3 vector."""
        return [v[3 * a:3 * b] for a, b in zip(self.seg_off[:-1], self.seg_off[1:])]


def nuclear_fields(pos, charges, atom_pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    ap = np.ascontiguousarray(atom_pos, dtype=np.float64).reshape(-1)
    F = np.zeros(pos.size)
    _check(lib().cppe_hip_nuclear_fields(pos.size // 3, _p(pos), ch.size, _p(ch), _p(ap), _p(F)))
    return F


def nuclear_multipole_energy(pos, kmax, mul, charges, atom_pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    km = np.ascontiguousarray(kmax, dtype=np.int32)
    mul = np.ascontiguousarray(mul, dtype=np.float64).reshape(-1)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    ap = np.ascontiguousarray(atom_pos, dtype=np.float64).reshape(-1)
    e = ctypes.c_double(0.0)
    _check(lib().cppe_hip_nuclear_multipole_energy(pos.size // 3, _p(pos), _p(km, _i), _p(mul), ch.size, _p(ch),
                                                   _p(ap), ctypes.byref(e)))
    return e.value


def nuclear_multipole_energy_noalloc(n, pos, kmax, mul, natoms, charges, atom_pos, energy, stream=0):
    """Allocation-free energy: every argument but n / natoms is a device address (int), energy a
    device double written on `stream` without synchronisation (cppe_hip_nuclear_multipole_energy_noalloc)."""
    _check(lib().cppe_hip_nuclear_multipole_energy_noalloc(int(n), pos, kmax, mul, int(natoms), charges, atom_pos,
                                                           energy, stream))


def nuclear_field_gradient(pos, charges, atom_pos):
    """(3 natoms) x (3 n) derivative of the nuclear fields with respect to the atom positions."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    ap = np.ascontiguousarray(atom_pos, dtype=np.float64).reshape(-1)
    G = np.zeros((3 * ch.size, pos.size))
    _check(lib().cppe_hip_nuclear_field_gradient(pos.size // 3, _p(pos), ch.size, _p(ch), _p(ap), _p(G)))
    return G


def nuclear_multipole_gradient(pos, kmax, mul, charges, atom_pos):
    """natoms x 3 derivative of the nuclear-multipole energy with respect to the atom positions."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    km = np.ascontiguousarray(kmax, dtype=np.int32)
    mul = np.ascontiguousarray(mul, dtype=np.float64).reshape(-1)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    ap = np.ascontiguousarray(atom_pos, dtype=np.float64).reshape(-1)
    g = np.zeros((ch.size, 3))
    _check(lib().cppe_hip_nuclear_multipole_gradient(pos.size // 3, _p(pos), _p(km, _i), _p(mul), ch.size, _p(ch),
                                                     _p(ap), _p(g)))
    return g


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = ctypes.c_double(0.0)
    _check(lib().cppe_hip_dot(a.size, _p(a), _p(b), ctypes.byref(out)))
    return out.value


_f = ctypes.POINTER(ctypes.c_float)


def _pf(a):
    return None if a is None else a.ctypes.data_as(_f)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """Device-resident warm-start network (cppe_hip_mlp): parameters, scaling, training set."""

    def __init__(self, params=None, mean=None, inv_std=None):
        h = ctypes.c_void_p()
        _check(lib().cppe_hip_mlp_create(ctypes.byref(h)))
        self.h = h
        if params is not None:
            self.set(params, mean, inv_std)

    def close(self):
        if getattr(self, "h", None):
            lib().cppe_hip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set(self, params, mean=None, inv_std=None):
        p = _f32(params)
        assert p.size == MLP_PARAMS
        m = None if mean is None else _f32(mean)
        s = None if inv_std is None else _f32(inv_std)
        _check(lib().cppe_hip_mlp_set(self.h, _pf(p), _pf(m), _pf(s)))

    def get(self):
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        m = np.zeros(MLP_IN, dtype=np.float32)
        s = np.zeros(MLP_IN, dtype=np.float32)
        _check(lib().cppe_hip_mlp_get(self.h, _pf(p), _pf(m), _pf(s)))
        return p, m, s

    def forward(self, X):
        X = _f32(X)
        ns = X.shape[0]
        y = np.zeros((ns, 3), dtype=np.float32)
        _check(lib().cppe_hip_mlp_forward(self.h, _pf(X), ns, _pf(y)))
        return y

    def loss_grad(self, X, V, T, w):
        X, V, T, w = _f32(X), _f32(V), _f32(T), _f32(w)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double(0.0)
        _check(lib().cppe_hip_mlp_loss_grad(self.h, _pf(X), _pf(V), _pf(T), _pf(w), X.shape[0], _pf(g),
                                            ctypes.byref(loss)))
        return loss.value, g

    def reserve(self, ns):
        _check(lib().cppe_hip_mlp_reserve(self.h, int(ns)))

    def collect(self, system):
        _check(lib().cppe_hip_mlp_collect(self.h, system.h))

    def prepare(self):
        _check(lib().cppe_hip_mlp_prepare(self.h))

    OBJECTIVES = {"dipole": 0, "alpha_inv": 1}

    def set_objective(self, objective):
        """Training objective (before collecting): 'dipole' error, or its 'alpha_inv' norm (cppe_hip_mlp_set_objective)."""
        _check(lib().cppe_hip_mlp_set_objective(self.h, self.OBJECTIVES[objective]))

    def samples(self, s0, count):
        """Collected samples as the loss sees them: V (count, 9), T (count, 3), weights (count,)."""
        V = np.zeros((count, 9), dtype=np.float32)
        T = np.zeros((count, 3), dtype=np.float32)
        w = np.zeros(count, dtype=np.float32)
        _check(lib().cppe_hip_mlp_samples(self.h, int(s0), int(count), _pf(V), _pf(T), _pf(w)))
        return V, T, w

    def train(self, epochs, batch=4096, lr=1e-3, seed=0):
        losses = np.zeros(max(epochs, 1))
        _check(lib().cppe_hip_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed), _p(losses)))
        return losses[:epochs]


def ml_features(system):
    X = np.zeros((system.n, MLP_IN), dtype=np.float32)
    V = np.zeros((system.n, 9))
    _check(lib().cppe_hip_ml_features(system.h, _pf(X), _p(V)))
    return X, V


def ml_guess(system, mlp, timed=True):
    """Learned guess into the system's resident moments; timed: wait and return its device ms,
    otherwise only enqueue it ahead of the next solve (returns None)."""
    if not timed:
        _check(lib().cppe_hip_ml_guess(system.h, mlp.h, None))
        return None
    ms = ctypes.c_double(0.0)
    _check(lib().cppe_hip_ml_guess(system.h, mlp.h, ctypes.byref(ms)))
    return ms.value
