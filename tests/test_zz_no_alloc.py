"""GPU check of the allocation-free entry point (cppe_hip_nuclear_multipole_energy_noalloc) against
cppe_hip_nuclear_multipole_energy, through the C ABI: device inputs placed by the test with the
HIP runtime the library itself uses, two calls in a row on one stream (the second with fewer
sites, so stale partials of the first would show), the reference's golden energy, and the size
limit."""
import ctypes
import os
import sys

import numpy as np
import pytest

from cppe_amd import capi, synthetic

from .conftest import GOLDEN

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """The few HIP runtime calls the test needs, from the runtime libcppe_hip.so is linked against."""

    def __init__(self):
        capi.lib()  # the library (and its HIP runtime) first
        path = next(p for p in ("/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so")
                    if os.path.exists(p))
        self.L = ctypes.CDLL(path)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        self.L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        self.L.hipFree.argtypes = [vp]
        self.L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.L.hipStreamDestroy.argtypes = [vp]
        self.L.hipStreamSynchronize.argtypes = [vp]
        self.held = []

    def ok(self, rc):
        assert rc == 0, "HIP runtime error %d" % rc

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.ok(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)))
        self.held.append(p)
        self.ok(self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D))
        return p.value

    def read_double(self, ptr):
        v = np.zeros(1)
        self.ok(self.L.hipMemcpy(v.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), 8, D2H))
        return float(v[0])

    def stream(self):
        s = ctypes.c_void_p()
        self.ok(self.L.hipStreamCreate(ctypes.byref(s)))
        return s

    def sync(self, s):
        self.ok(self.L.hipStreamSynchronize(s))

    def close(self):
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


def site_arrays(w):
    """kmax and the 10 packed multipole values per site (q, mu, traceless quadrupole)."""
    mul = np.concatenate([w.charges[:, None], w.dipoles, w.quadrupoles], axis=1)
    return np.full(w.n, 2, dtype=np.int32), np.ascontiguousarray(mul)


def test_noalloc_energy_matches_allocating_entry_point():
    hip = Hip()
    s = hip.stream()
    out = hip.upload(np.zeros(2))
    for k, nw in enumerate([30000, 400]):  # the second call is smaller
        w = synthetic.water_box(nw, 500 + k, anisotropic=True)
        kmax, mul = site_arrays(w)
        z = w.atoms_z.astype(float)
        ref = capi.nuclear_multipole_energy(w.pos, kmax, mul, z, w.atoms_pos)
        d = [hip.upload(a) for a in (w.pos.reshape(-1), kmax, mul.reshape(-1), z, w.atoms_pos.reshape(-1))]
        say("noalloc call", k, "sites", w.n)
        capi.nuclear_multipole_energy_noalloc(w.n, d[0], d[1], d[2], len(z), d[3], d[4], out + 8 * k, s.value)
        hip.sync(s)
        e = hip.read_double(out + 8 * k)
        say("energy", e, "allocating entry point", ref)
        assert e == ref  # same kernels, same summation order
    hip.L.hipStreamDestroy(s)
    hip.close()


def test_noalloc_energy_golden_and_edges():
    import oracle as O
    hip = Hip()
    m = np.load(os.path.join(GOLDEN, "molecules.npz"))
    pot = O.read_potfile(os.path.join(GOLDEN, "pna_6w.pot"))
    z = m["pna_z"].astype(float)
    d = [hip.upload(a) for a in (pot.pos.reshape(-1), pot.kmax, pot.mul.reshape(-1), z, m["pna_xyz"].reshape(-1))]
    out = hip.upload(np.array([7.0]))
    capi.nuclear_multipole_energy_noalloc(pot.n, d[0], d[1], d[2], len(z), d[3], d[4], out, 0)
    hip.sync(None)
    np.testing.assert_almost_equal(hip.read_double(out), -0.321349401430, decimal=9)  # test_functionality.py:45
    # no atoms: the energy is written as 0
    capi.nuclear_multipole_energy_noalloc(pot.n, d[0], d[1], d[2], 0, d[3], d[4], out, 0)
    hip.sync(None)
    assert hip.read_double(out) == 0.0
    # beyond the fixed scratch: rejected before any device work
    with pytest.raises(capi.CppeHipError, match="maximum of 4194304"):
        capi.nuclear_multipole_energy_noalloc(4194305, 0, 0, 0, 20, 0, 0, out, 0)
    hip.close()
