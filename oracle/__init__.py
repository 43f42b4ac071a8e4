"""CPU oracle for the polarizable-embedding field / induced-dipole path (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
this package, and only as the checker. The product (``cppe_amd``) never imports it.

``liboracle.so`` (built from ``oracle.cc`` by ``oracle/Makefile``) restates the reference's
algorithms in FP64; this module wraps it with ctypes and adds a restatement of the reference's
potfile reader (``src/core/potfile_reader.cc:31-179``) that returns plain numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_fmm_kernels.so")
ANG2BOHR = 1.8897261246  # potfile_reader.cc:10

_lib = None
_ref = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_ll = ctypes.POINTER(ctypes.c_longlong)


def build(force=False):
    """Compile liboracle.so (and oracle/_ref when the reference checkout is present)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/src/core/fmm") and (force or not os.path.exists(REF_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_direct_field.argtypes = [ctypes.c_int, ctypes.c_int, _d, _d, _d, _i, _i, ctypes.c_double, _d]
        L.orc_direct_field_targets.argtypes = [ctypes.c_int, ctypes.c_int, _d, _d, _d, _i, _i,
                                               ctypes.c_double, ctypes.c_int, _i, _d]
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_tree_build.argtypes = [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_sizes.argtypes = [ctypes.c_void_p, _ll]
        L.orc_tree_cells.argtypes = [ctypes.c_void_p, _d, _ll, _ll]
        L.orc_tree_lists.argtypes = [ctypes.c_void_p, _ll, _ll]
        L.orc_fmm_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _d, _d, _d, _i, _i,
                                    ctypes.c_double, _d]
        L.orc_coulomb_derivs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _d]
        L.orc_pair_field.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _d,
                                     ctypes.c_double, ctypes.c_int, _d]
        L.orc_thole_factors.argtypes = [ctypes.c_double, _d]
        L.orc_nuclear_fields.argtypes = [ctypes.c_int, _d, ctypes.c_int, _d, _d, _d]
        L.orc_nuclear_multipole_energy.restype = ctypes.c_double
        L.orc_nuclear_multipole_energy.argtypes = [ctypes.c_int, _d, _i, _d, ctypes.c_int, _d, _d]
        L.orc_bsys_create.restype = ctypes.c_void_p
        L.orc_bsys_create.argtypes = [ctypes.c_int, _d, _d, _i, _i, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_bsys_free.argtypes = [ctypes.c_void_p]
        for f in ("orc_bsys_apply", "orc_bsys_apply_alpha", "orc_bsys_apply_alpha_inv"):
            getattr(L, f).argtypes = [ctypes.c_void_p, _d, _d]
        L.orc_bsys_dense.argtypes = [ctypes.c_void_p, _i, _d]
        L.orc_bsys_solve.argtypes = [ctypes.c_void_p, _d, _d, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     _d, _i, _d]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def ref_kernels():
    """The reference's own generated FMM kernels (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        R = ctypes.CDLL(REF_LIB_PATH)
        dd = ctypes.c_double
        R.ref_p2p.argtypes = [ctypes.c_int, dd, dd, dd, _d, _d]
        R.ref_p2p_damped.argtypes = [ctypes.c_int, dd, dd, dd, dd, _d, _d]
        for f in ("ref_m2m", "ref_m2l", "ref_l2l", "ref_l2p"):
            getattr(R, f).argtypes = [ctypes.c_int, ctypes.c_int, dd, dd, dd, _d, _d]
        _ref = R
    return _ref


class CpuBaseline:
    """The reference's solve on the host cores (oracle/cpu_baseline.cc, -O3 -march=native, OpenMP):
    per-apply tree rebuild, straight-line FMM operators or the direct sum, Jacobi-preconditioned
    CG from alpha F. ``path`` selects a copy built for the running machine."""

    _libs = {}

    def __init__(self, pos, alpha6, excl, damping=0.0, scheme="direct", ncrit=64, order=5, theta=0.5,
                 path=None):
        self.L = self.load(path or os.path.join(HERE, "libcpu_baseline.so"))
        self.pos = _f64(pos).reshape(-1)
        self.n = self.pos.size // 3
        self.alpha6 = _f64(alpha6).reshape(-1)
        self.off = _i32(excl[0])
        self.idx = _i32(excl[1]) if len(excl[1]) else np.zeros(1, dtype=np.int32)
        self.h = self.L.cpb_create(self.n, _p(self.pos), _p(self.alpha6), _p(self.off, _i), _p(self.idx, _i),
                                   float(damping), int(scheme == "fmm"), int(ncrit), int(order), float(theta))
        if not self.h:
            raise RuntimeError(self.L.orc_last_error().decode())

    @classmethod
    def load(cls, path):
        if path not in cls._libs:
            if not os.path.exists(path):
                subprocess.run(["make", "-s", "-C", HERE, "baseline", "CPB_OUT=" + path], check=True)
            L = ctypes.CDLL(path)
            L.orc_last_error.restype = ctypes.c_char_p
            L.cpb_flags.restype = ctypes.c_char_p
            L.orc_num_threads.restype = ctypes.c_int
            L.cpb_create.restype = ctypes.c_void_p
            L.cpb_create.argtypes = [ctypes.c_int, _d, _d, _i, _i, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double]
            L.cpb_free.argtypes = [ctypes.c_void_p]
            L.cpb_apply.argtypes = [ctypes.c_void_p, _d, _d, _d]
            L.cpb_solve.argtypes = [ctypes.c_void_p, _d, ctypes.c_double, ctypes.c_int, _d, _i, _d]
            cls._libs[path] = L
        return cls._libs[path]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.cpb_free(self.h)
            self.h = None

    @property
    def flags(self):
        return self.L.cpb_flags().decode()

    @property
    def threads(self):
        return int(self.L.orc_num_threads())

    def apply(self, x):
        """B x (tree rebuilt for FMM); returns (Bx, tree ms, field ms)."""
        x = _f64(x)
        out = np.zeros(3 * self.n)
        t = np.zeros(2)
        self.L.cpb_apply(self.h, _p(x), _p(out), _p(t))
        return out, float(t[0]), float(t[1])

    def solve(self, rhs, thresh=1e-8, maxiter=50):
        """CG from alpha rhs; returns (mu, iterations, total ms, tree ms, field ms)."""
        rhs = _f64(rhs)
        x = np.zeros_like(rhs)
        it = ctypes.c_int(0)
        t = np.zeros(3)
        rc = self.L.cpb_solve(self.h, _p(rhs), float(thresh), int(maxiter), _p(x), ctypes.byref(it), _p(t))
        if rc:
            raise RuntimeError("Failed to converge induced dipole moments.")
        return x, it.value, float(t[0]), float(t[1]), float(t[2])


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def ncomp(k):
    return (k + 1) * (k + 2) // 2


# ----------------------------------------------------------------------------------------------
# Potfile reader restatement (potfile_reader.cc:31-179) returning plain arrays.
# ----------------------------------------------------------------------------------------------
class PotData:
    """Sites of a potential file: positions (bohr), multipoles, polarizabilities, exclusions."""

    def __init__(self, n):
        self.n = n
        self.elements = [""] * n
        self.pos = np.zeros((n, 3))
        self.mul = np.zeros((n, 10))          # q, mu(3), theta(6) (trace removed)
        self.has_order = np.zeros((n, 3), dtype=bool)
        self.pol = np.zeros((n, 6))
        self.polarizable = np.zeros(n, dtype=bool)
        self.excl = [[] for _ in range(n)]

    @property
    def kmax(self):
        k = np.full(self.n, -1, dtype=np.int32)
        for o in range(3):
            k[self.has_order[:, o]] = o
        return k

    def excl_csr(self, sites=None):
        sites = range(self.n) if sites is None else sites
        off = [0]
        idx = []
        for s in sites:
            idx.extend(self.excl[s])
            off.append(len(idx))
        return np.array(off, dtype=np.int32), np.array(idx, dtype=np.int32)

    def alpha_iso(self):
        return (self.pol[:, 0] + self.pol[:, 3] + self.pol[:, 5]) / 3.0


def read_potfile(path):
    opener = open
    if path.endswith(".gz"):
        import gzip
        opener = gzip.open
    with opener(path, "rt") as fh:
        lines = fh.read().split("\n")
    it = iter(lines)
    data = None
    num_sites = 0
    for line in it:
        if "@COORDINATES" in line:
            num_sites = int(next(it).split()[0])
            unit = " ".join(next(it).split())
            if unit == "AA":
                conv = ANG2BOHR
            elif unit == "AU":
                conv = 1.0
            else:
                raise RuntimeError("Invalid unit for potential file.")
            data = PotData(num_sites)
            for i in range(num_sites):
                t = next(it).split()
                data.elements[i] = t[0]
                data.pos[i] = [float(t[1]) * conv, float(t[2]) * conv, float(t[3]) * conv]
        if "ORDER" in line:
            t = line.split()
            if len(t) == 2:
                order = int(t[1])
                nm = int(next(it))
                o0 = [0, 1, 4][order]
                for _ in range(nm):
                    t = next(it).split()
                    s = int(t[0]) - 1
                    vals = [float(v) for v in t[1:1 + ncomp(order)]]
                    if order == 2:
                        tr = (vals[0] + vals[3] + vals[5]) / 3.0
                        vals[0] -= tr
                        vals[3] -= tr
                        vals[5] -= tr
                    data.mul[s, o0:o0 + ncomp(order)] = vals
                # the reader fills absent sites of an ORDER block with zeros
                data.has_order[:, order] = True
            elif len(t) == 3:
                if int(t[1]) != 1 or int(t[2]) != 1:
                    raise RuntimeError("Only dipole-dipole polarizabilities are currently supported.")
                npol = int(next(it))
                for _ in range(npol):
                    t = next(it).split()
                    s = int(t[0]) - 1
                    if data.polarizable[s]:
                        raise RuntimeError("Potential can only have one polarizability.")
                    data.pol[s] = [float(v) for v in t[1:7]]
                    data.polarizable[s] = True
            else:
                raise RuntimeError("Invalid number in potfile ORDER.")
        if "EXCLISTS" in line:
            nex = int(next(it).split()[0])
            for _ in range(nex):
                t = next(it).split()
                s = int(t[0]) - 1
                for v in t[1:]:
                    e = int(v) - 1
                    if e >= 0:
                        data.excl[s].append(e)
    return data


# ----------------------------------------------------------------------------------------------
# Fields, trees, energies
# ----------------------------------------------------------------------------------------------
def direct_field(m, pos, S, alpha_iso=None, excl=None, damping=0.0):
    pos = _f64(pos).reshape(-1)
    n = pos.size // 3
    S = _f64(S).reshape(-1)
    F = np.zeros(3 * n)
    off, idx = (None, None) if excl is None else (_i32(excl[0]), _i32(excl[1]))
    al = None if alpha_iso is None else _f64(alpha_iso)
    lib().orc_direct_field(m, n, _p(pos), _p(S), _p(al), _p(off, _i), _p(idx, _i), damping, _p(F))
    return F


def direct_field_targets(m, pos, S, targets, alpha_iso=None, excl=None, damping=0.0):
    pos = _f64(pos).reshape(-1)
    n = pos.size // 3
    S = _f64(S).reshape(-1)
    t = _i32(targets)
    F = np.zeros(3 * t.size)
    off, idx = (None, None) if excl is None else (_i32(excl[0]), _i32(excl[1]))
    al = None if alpha_iso is None else _f64(alpha_iso)
    lib().orc_direct_field_targets(m, n, _p(pos), _p(S), _p(al), _p(off, _i), _p(idx, _i), damping,
                                   t.size, _p(t, _i), _p(F))
    return F


class Tree:
    """Octree + interaction lists (fmm/tree.cc:134-237, fmm/calculate.cc:74-127)."""

    def __init__(self, pos, ncrit=64, order=5, theta=0.5):
        self.pos = _f64(pos).reshape(-1)
        self.n = self.pos.size // 3
        self.ncrit = ncrit
        self.order = order
        self.theta = theta
        self.h = lib().orc_tree_build(self.n, _p(self.pos), ncrit, order, theta)
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        sz = np.zeros(3, dtype=np.int64)
        lib().orc_tree_sizes(self.h, _p(sz, _ll))
        self.ncells, self.n_m2l, self.n_p2p = (int(v) for v in sz)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def cells(self):
        geo = np.zeros((self.ncells, 5))
        meta = np.zeros((self.ncells, 4), dtype=np.int64)
        leafs = np.zeros((self.ncells, self.ncrit), dtype=np.int64)
        lib().orc_tree_cells(self.h, _p(geo), _p(meta, _ll), _p(leafs, _ll))
        return geo, meta, leafs

    def lists(self):
        m2l = np.zeros((self.n_m2l, 2), dtype=np.int64)
        p2p = np.zeros((self.n_p2p, 2), dtype=np.int64)
        lib().orc_tree_lists(self.h, _p(m2l, _ll), _p(p2p, _ll))
        return m2l, p2p

    def field(self, m, S, alpha_iso=None, excl=None, damping=0.0):
        S = _f64(S).reshape(-1)
        F = np.zeros(3 * self.n)
        off, idx = (None, None) if excl is None else (_i32(excl[0]), _i32(excl[1]))
        al = None if alpha_iso is None else _f64(alpha_iso)
        rc = lib().orc_fmm_field(self.h, m, self.n, _p(self.pos), _p(S), _p(al), _p(off, _i), _p(idx, _i),
                                 damping, _p(F))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return F


def coulomb_derivs(p, r):
    nt = (p + 1) * (p + 2) * (p + 3) // 6
    D = np.zeros(nt)
    lib().orc_coulomb_derivs(p, float(r[0]), float(r[1]), float(r[2]), _p(D))
    return D


def pair_field(m, d, S, a=0.0, damped=False):
    F = np.zeros(3)
    S = _f64(S)
    lib().orc_pair_field(m, float(d[0]), float(d[1]), float(d[2]), _p(S), float(a), int(damped), _p(F))
    return F


def thole_factors(v):
    out = np.zeros(4)
    lib().orc_thole_factors(float(v), _p(out))
    return out  # fV, fE, fT, fD


def nuclear_fields(pos, charges, atom_pos):
    pos = _f64(pos).reshape(-1)
    n = pos.size // 3
    ch = _f64(charges)
    ap = _f64(atom_pos).reshape(-1)
    F = np.zeros(3 * n)
    lib().orc_nuclear_fields(n, _p(pos), ch.size, _p(ch), _p(ap), _p(F))
    return F


def nuclear_multipole_energy(pot, charges, atom_pos):
    pos = _f64(pot.pos).reshape(-1)
    ch = _f64(charges)
    ap = _f64(atom_pos).reshape(-1)
    mul = _f64(pot.mul).reshape(-1)
    kmax = _i32(pot.kmax)
    return lib().orc_nuclear_multipole_energy(pot.n, _p(pos), _p(kmax, _i), _p(mul), ch.size, _p(ch), _p(ap))


def multipole_fields(pot, scheme="direct", damp_multipole=False, damping_factor=2.1304,
                     ncrit=64, order=5, theta=0.5):
    """MultipoleFields::compute (electric_fields.cc:106-189): F(q) - F(mu) + F(theta/2 diag)."""
    n = pot.n
    kmax = int(pot.kmax.max())
    damping = damping_factor if damp_multipole else 0.0
    al = np.where(pot.polarizable, pot.alpha_iso(), 0.0) if damping > 0 else None
    excl = pot.excl_csr()
    q = pot.mul[:, 0].copy()
    mu = pot.mul[:, 1:4].copy()
    th = pot.mul[:, 4:10].copy()
    th[:, [0, 3, 5]] *= 0.5
    tree = Tree(pot.pos, ncrit, order, theta) if scheme == "fmm" else None

    def fld(m, S):
        if tree is not None:
            return tree.field(m, S, al, excl, damping)
        return direct_field(m, pot.pos, S, al, excl, damping)

    F = fld(0, q)
    if kmax > 0:
        F = F - fld(1, mu)
    if kmax > 1:
        F = F + fld(2, th)
    return F


class BSystem:
    """Response operator B = alpha^-1 - T2 over polarizable sites (bmatrix.{hh,cc})."""

    def __init__(self, pos, alpha6, excl, damping=0.0, scheme="direct", ncrit=64, order=5,
                 theta=0.5, rebuild_per_apply=False):
        self.pos = _f64(pos).reshape(-1)
        self.n = self.pos.size // 3
        self.alpha6 = _f64(alpha6).reshape(-1)
        self.off = _i32(excl[0])
        self.idx = _i32(excl[1])
        self.h = lib().orc_bsys_create(self.n, _p(self.pos), _p(self.alpha6), _p(self.off, _i),
                                       _p(self.idx, _i), float(damping), int(scheme == "fmm"), ncrit,
                                       order, float(theta), int(rebuild_per_apply))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bsys_free(self.h)
            self.h = None

    def _call(self, fn, x):
        x = _f64(x)
        out = np.zeros(3 * self.n)
        getattr(lib(), fn)(self.h, _p(x), _p(out))
        return out

    def apply(self, x):
        return self._call("orc_bsys_apply", x)

    def apply_diagonal_inverse(self, x):
        return self._call("orc_bsys_apply_alpha", x)

    def apply_diagonal(self, x):
        return self._call("orc_bsys_apply_alpha_inv", x)

    def dense(self, index=None):
        index = _i32(np.arange(self.n) if index is None else index)
        out = np.zeros((3 * self.n, 3 * self.n))
        lib().orc_bsys_dense(self.h, _p(index, _i), _p(out))
        return out

    def solve(self, rhs, guess=None, make_guess=True, thresh=1e-8, maxiter=50):
        rhs = _f64(rhs)
        g = _f64(np.zeros_like(rhs) if guess is None else guess)
        x = np.zeros_like(rhs)
        it = ctypes.c_int(0)
        norms = np.zeros(max(maxiter, 1))
        rc = lib().orc_bsys_solve(self.h, _p(rhs), _p(g), int(make_guess), float(thresh), int(maxiter),
                                  _p(x), ctypes.byref(it), _p(norms))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return x, it.value, norms[:it.value]
