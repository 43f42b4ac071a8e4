"""The potential-file reader (src/core/potfile_reader.cc:31-179 of the reference) against the
oracle's restatement of it, on the reference's fixtures and on generated files that exercise the
section walk: AA units, absent multipole sites (zero fill), partial polarizability lists, exclusion
lists with non-positive entries, CRLF line ends, explicit '+' signs, and the error messages.
Host code only (no device)."""
import gzip
import os
import shutil

import numpy as np
import pytest

import oracle as O
from cppe_amd import synthetic

from .conftest import GOLDEN

cppe = pytest.importorskip("cppe_amd.pycppe")


def check_same(pots, ref):
    assert len(pots) == ref.n
    np.testing.assert_array_equal(np.array([p.position for p in pots]).reshape(-1, 3), ref.pos)
    for i, p in enumerate(pots):
        assert p.element == ref.elements[i]
        got = np.zeros(10)
        for m in p.multipoles:
            o0 = [0, 1, 4][m.k]
            got[o0:o0 + len(m.values)] = m.values
        np.testing.assert_array_equal(got, ref.mul[i])
        assert p.is_polarizable == bool(ref.polarizable[i])
        if ref.polarizable[i]:
            np.testing.assert_array_equal(p.polarizability.values, ref.pol[i])
        assert list(p.exclusions) == ref.excl[i]


def test_reader_fixtures_match_oracle(tmp_path):
    for name in ("pna_6w.pot", "pna_6w_isopol.pot"):
        path = os.path.join(GOLDEN, name)
        check_same(cppe.PotfileReader(path).read(), O.read_potfile(path))
    path = str(tmp_path / "loprop_solvated_20.pot")
    with gzip.open(os.path.join(GOLDEN, "loprop_solvated_20.pot.gz"), "rb") as fi, open(path, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    check_same(cppe.PotfileReader(path).read(), O.read_potfile(path))


def write_irregular(path, n_waters=3000, seed=5, crlf=False):
    """AA coordinates, charges on every site, dipoles on every third site only (zero fill of the
    others), no quadrupoles for the first sites, polarizabilities on even sites, exclusion lists
    padded with 0 and negative entries, some values written with a leading '+'."""
    w = synthetic.water_box(n_waters, seed, anisotropic=True)
    n = w.n
    nl = "\r\n" if crlf else "\n"
    lines = ["! irregular file", "@COORDINATES", "%d" % n, "AA"]
    for i in range(n):
        x, y, z = w.pos[i] / 1.8897261246
        lines.append("%s   %.12f\t%+.12f %.12e" % (w.elements[i], x, y, z))
    lines += ["@MULTIPOLES", "ORDER 0", "%d" % n]
    lines += ["%d %+.15g" % (i + 1, w.charges[i]) for i in range(n)]
    sel = list(range(1, n, 3))
    lines += ["ORDER 1", "%d" % len(sel)]
    lines += ["%d %.15g %.15g %.15g" % (i + 1, *w.dipoles[i]) for i in sel]
    sel = list(range(n // 2, n))
    lines += ["ORDER 2", "%d" % len(sel)]
    lines += ["%d %s" % (i + 1, " ".join("%.15g" % v for v in w.quadrupoles[i] + 0.01)) for i in sel]
    sel = list(range(0, n, 2))
    lines += ["@POLARIZABILITIES", "ORDER 1 1", "%d" % len(sel)]
    lines += ["%d %s" % (i + 1, " ".join("%.15g" % v for v in w.pol[i])) for i in sel]
    lines += ["EXCLISTS", "%d 5" % n]
    lines += ["%d %s 0 -1" % (i + 1, " ".join(str(e + 1) for e in w.excl[i])) for i in range(n)]
    with open(path, "w", newline="") as fh:
        fh.write(nl.join(lines) + nl)


@pytest.mark.parametrize("crlf", [False, True])
def test_reader_irregular_file_matches_oracle(tmp_path, crlf):
    path = str(tmp_path / "irregular.pot")
    write_irregular(path, crlf=crlf)
    pots = cppe.PotfileReader(path).read()
    check_same(pots, O.read_potfile(path))
    # the zero-filled dipoles exist, and quadrupole traces were removed on read
    assert all(len(p.multipoles) >= 2 for p in pots)
    q = pots[-1].multipoles[2].values
    assert abs(q[0] + q[3] + q[5]) < 1e-15


def test_reader_large_file_matches_oracle(tmp_path):
    """A file large enough to be parsed by several threads (30,000 sites)."""
    w = synthetic.water_box(10000, 9, anisotropic=True)
    path = str(tmp_path / "w30k.pot")
    w.write_potfile(path)
    pots = cppe.PotfileReader(path).read()
    check_same(pots, O.read_potfile(path))


def test_reader_errors(tmp_path):
    def bad(text, msg):
        p = tmp_path / "bad.pot"
        p.write_text(text)
        with pytest.raises(RuntimeError, match=msg):
            cppe.PotfileReader(str(p)).read()

    head = "@COORDINATES\n2\n%s\nO 0 0 0\nH 0 0 1\n"
    bad(head % "XX", "Invalid unit for potential file.")
    bad(head % "AU" + "@POLARIZABILITIES\nORDER 1 2\n0\n", "Only dipole-dipole polarizabilities")
    bad(head % "AU" + "ORDER 1 1\n2\n1 1 0 0 1 0 1\n1 1 0 0 1 0 1\n", "Potential can only have one polarizability.")
    bad(head % "AU" + "ORDER 1 1 1\n", "Invalid number in potfile ORDER.")
    with pytest.raises(RuntimeError, match="Potential file does not exist."):
        cppe.PotfileReader(str(tmp_path / "missing.pot"))
