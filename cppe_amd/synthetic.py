"""Seeded synthetic water boxes with a point-nucleus QM core (the benchmark and test inputs).

Recipe (per water): O on a cubic lattice of 3.1 A spacing (x fastest) with +-0.3 A uniform
jitter; O-H 0.9572 A, HOH 104.52 deg, uniformly random orientation (uniform quaternion);
charges O -0.67 / H +0.335 plus N(0, 0.02); dipoles N(0, 0.1) per component; symmetric
quadrupoles N(0, 0.2) per packed component with the trace removed; polarizabilities either
isotropic (O 5.7, H 2.7 a.u.) or anisotropic R diag(lambda) R^T with lambda ~ U(2, 8) a.u.;
exclusions = the other two sites of the molecule. QM core: 20 point nuclei, Z uniform in 1..8,
uniform in a 4 A sphere at the lattice centre; every water with an atom within 4 A of any
nucleus is removed. Positions are returned in bohr.
"""
import math

import numpy as np

ANG2BOHR = 1.8897261246
SPACING = 3.1
JITTER = 0.3
R_OH = 0.9572
ANGLE_HOH = 104.52


class Sites:
    """Plain arrays describing a potential: one entry per site (all sites polarizable)."""

    def __init__(self, pos, charges, dipoles, quadrupoles, pol, excl, elements):
        self.pos = pos                  # (n, 3) bohr
        self.charges = charges          # (n,)
        self.dipoles = dipoles          # (n, 3)
        self.quadrupoles = quadrupoles  # (n, 6) xx xy xz yy yz zz, traceless
        self.pol = pol                  # (n, 6) packed symmetric polarizability
        self.excl = excl                # list of lists (0-based site indices)
        self.elements = elements
        self.atoms_z = None             # QM nuclei
        self.atoms_pos = None           # (natoms, 3) bohr

    @property
    def n(self):
        return len(self.pos)

    def excl_csr(self):
        off = np.zeros(self.n + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(e) for e in self.excl])
        idx = np.array([v for e in self.excl for v in e], dtype=np.int32)
        return off, idx

    def alpha_iso(self):
        return (self.pol[:, 0] + self.pol[:, 3] + self.pol[:, 5]) / 3.0

    def write_potfile(self, path):
        """Write a .pot file in AU with %.17g coordinates (no unit conversion on read)."""
        n = self.n
        with open(path, "w") as fh:
            fh.write("! synthetic water box\n@COORDINATES\n%d\nAU\n" % n)
            for i in range(n):
                fh.write("%s %.17g %.17g %.17g %d\n" % (self.elements[i], *self.pos[i], i + 1))
            fh.write("@MULTIPOLES\nORDER 0\n%d\n" % n)
            for i in range(n):
                fh.write("%d %.17g\n" % (i + 1, self.charges[i]))
            fh.write("ORDER 1\n%d\n" % n)
            for i in range(n):
                fh.write("%d %.17g %.17g %.17g\n" % (i + 1, *self.dipoles[i]))
            fh.write("ORDER 2\n%d\n" % n)
            for i in range(n):
                fh.write("%d %s\n" % (i + 1, " ".join("%.17g" % v for v in self.quadrupoles[i])))
            fh.write("@POLARIZABILITIES\nORDER 1 1\n%d\n" % n)
            for i in range(n):
                fh.write("%d %s\n" % (i + 1, " ".join("%.17g" % v for v in self.pol[i])))
            fh.write("EXCLISTS\n%d 3\n" % n)
            for i in range(n):
                fh.write("%d %s\n" % (i + 1, " ".join(str(e + 1) for e in self.excl[i])))


def _random_rotations(rng, k):
    q = rng.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((k, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box(n_waters, seed, anisotropic=False, n_nuclei=20, qm_radius=4.0, cutoff=4.0):
    """Build the site arrays of ``n_waters`` waters around a QM core (see module docstring)."""
    rng = np.random.default_rng(seed)
    side = int(math.ceil(n_waters ** (1.0 / 3.0)))
    half = math.radians(ANGLE_HOH) / 2.0
    h_local = np.array([[R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                        [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]])
    while True:
        g = np.arange(side, dtype=float) * SPACING
        zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
        O = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
        O = O + rng.uniform(-JITTER, JITTER, size=O.shape)
        rot = _random_rotations(rng, len(O))
        H = O[:, None, :] + np.einsum("kij,hj->khi", rot, h_local)
        centre = np.full(3, (side - 1) * SPACING / 2.0)
        u = rng.normal(size=(n_nuclei, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        rad = qm_radius * rng.uniform(0, 1, size=n_nuclei) ** (1.0 / 3.0)
        nuc = centre + u * rad[:, None]
        zs = rng.integers(1, 9, size=n_nuclei)
        atoms = np.concatenate([O[:, None, :], H], axis=1)  # (k, 3, 3)
        d = np.linalg.norm(atoms[:, :, None, :] - nuc[None, None, :, :], axis=-1)
        keep = np.nonzero((d.min(axis=(1, 2)) >= cutoff))[0]
        if len(keep) >= n_waters:
            keep = keep[:n_waters]
            break
        side += 1
    nw = n_waters
    atoms = atoms[keep]
    pos = atoms.reshape(-1, 3) * ANG2BOHR
    n = 3 * nw
    charges = np.tile([-0.67, 0.335, 0.335], nw) + rng.normal(0, 0.02, size=n)
    dipoles = rng.normal(0, 0.1, size=(n, 3))
    quad = rng.normal(0, 0.2, size=(n, 6))
    tr = (quad[:, 0] + quad[:, 3] + quad[:, 5]) / 3.0
    quad[:, 0] -= tr
    quad[:, 3] -= tr
    quad[:, 5] -= tr
    if anisotropic:
        lam = rng.uniform(2.0, 8.0, size=(n, 3))
        Rr = _random_rotations(rng, n)
        A = np.einsum("kij,kj,klj->kil", Rr, lam, Rr)
        pol = np.stack([A[:, 0, 0], A[:, 0, 1], A[:, 0, 2], A[:, 1, 1], A[:, 1, 2], A[:, 2, 2]], axis=1)
    else:
        iso = np.tile([5.7, 2.7, 2.7], nw)
        pol = np.zeros((n, 6))
        pol[:, 0] = pol[:, 3] = pol[:, 5] = iso
    excl = []
    for w in range(nw):
        b = 3 * w
        excl += [[b + 1, b + 2], [b, b + 2], [b, b + 1]]
    elements = ["O", "H", "H"] * nw
    s = Sites(pos, charges, dipoles, quad, pol, excl, elements)
    s.atoms_z = zs.astype(np.int32)
    s.atoms_pos = nuc * ANG2BOHR
    return s


# The five workloads (seeds 1001..1005); sizes are site counts / 3 waters.
CONFIGS = {
    "c1_3k_direct": dict(n_waters=1000, seed=1001, anisotropic=False),
    "c2_60k": dict(n_waters=20000, seed=1002, anisotropic=True),
    "c3_300k": dict(n_waters=100000, seed=1003, anisotropic=True),
}


def batch_sizes(n_snapshots=256, seed=1004):
    """Water counts of the batched workload: lognormal(ln 2000, 0.8) clipped to 500..8000."""
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(rng.lognormal(math.log(2000), 0.8, size=n_snapshots)), 500, 8000).astype(int)
