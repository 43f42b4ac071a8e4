"""The reference's own Python test-suite, restated against ``cppe_amd`` (the drop-in API).

Mirrors tests/test_options.py, tests/test_functionality.py, tests/test_fields.py,
tests/test_solver.py and tests/test_fmm.py of the reference. Tests that touch the device are
marked ``gpu``; option handling and potfile reading run on the CPU.
"""
import os

import numpy as np
import pytest

from .conftest import GOLDEN

cppe = pytest.importorskip("cppe_amd.pycppe")

POT = os.path.join(GOLDEN, "pna_6w.pot")
POT_ISO = os.path.join(GOLDEN, "pna_6w_isopol.pot")


def pna_molecule():
    m = np.load(os.path.join(GOLDEN, "molecules.npz"))
    mol = cppe.Molecule()
    for z, c in zip(m["pna_z"], m["pna_xyz"]):
        mol.append(cppe.Atom(int(z), *c))
    return mol


def print_callback(output):
    print("cb: ", output)


DEFAULTS = {
    "potfile": "potential.pot", "iso_pol": False, "induced_thresh": 1e-8, "maxiter": 50,
    "damp_induced": False, "damp_multipole": False, "damping_factor_induced": 2.1304,
    "damping_factor_multipole": 2.1304, "summation_induced_fields": "direct", "tree_expansion_order": 5,
    "tree_ncrit": 64, "theta": 0.5, "pe_border": False, "border_type": "remove", "border_rmin": 2.2,
    "border_nredist": -1, "border_redist_order": 1, "border_redist_pol": False,
}


# ---- test_options.py ---------------------------------------------------------------------------
def test_options_defaults():
    s = cppe.CppeState({"potfile": POT})
    d = DEFAULTS.copy()
    d.pop("potfile")
    o = s.options.copy()
    o.pop("potfile")
    assert len(d) == len(o)
    for k in d:
        assert d[k] == s.options[k]


def test_options_set_all():
    custom = {"potfile": POT, "iso_pol": True, "induced_thresh": 1e-4, "maxiter": 500, "damp_induced": True,
              "damp_multipole": True, "damping_factor_induced": 213.04, "damping_factor_multipole": 213.04,
              "summation_induced_fields": "fmm", "tree_expansion_order": 4, "tree_ncrit": 12, "theta": 0.2,
              "pe_border": True, "border_type": "redist", "border_rmin": 2.6, "border_nredist": 2,
              "border_redist_order": 2, "border_redist_pol": True}
    s = cppe.CppeState(custom)
    for k in custom:
        assert custom[k] == s.options[k]


def test_options_invalid_key():
    d = DEFAULTS.copy()
    d["invalid_key"] = 42
    with pytest.raises(ValueError):
        cppe.CppeState(d)


def test_valid_option_keys_match_reference():
    assert list(cppe.valid_option_keys) == list(DEFAULTS.keys())[:3] + [
        "maxiter", "damp_induced", "damping_factor_induced", "damp_multipole", "damping_factor_multipole",
        "summation_induced_fields", "tree_expansion_order", "tree_ncrit", "theta", "pe_border", "border_type",
        "border_rmin", "border_nredist", "border_redist_order", "border_redist_pol"]


# ---- test_functionality.py (host part) ----------------------------------------------------------
def test_read_potfile():
    assert len(cppe.PotfileReader(POT).read()) == 18


def test_missing_potfile():
    with pytest.raises(RuntimeError, match="Potential file does not exist."):
        cppe.PotfileReader("/nonexistent.pot")


def test_potfile_reader_matches_oracle_reader():
    import oracle as O
    pots = cppe.PotfileReader(POT).read()
    ref = O.read_potfile(POT)
    for i, p in enumerate(pots):
        np.testing.assert_array_equal(p.position, ref.pos[i])
        np.testing.assert_array_equal(p.multipoles[2].values, ref.mul[i, 4:10])
        np.testing.assert_array_equal(p.polarizability.values, ref.pol[i])
        assert p.exclusions == ref.excl[i]


# ---- device tests ------------------------------------------------------------------------------
def solve_induced_moments(options, mol):
    s = cppe.CppeState(options, mol, print_callback)
    s.calculate_static_energies_and_fields()
    np.testing.assert_allclose(s.static_fields, s.nuclear_fields + s.multipole_fields)
    s.update_induced_moments(np.zeros_like(s.static_fields), False)
    return s


@pytest.mark.gpu
def test_nuclear_interaction_energy():
    s = cppe.CppeState({"potfile": POT}, pna_molecule(), print_callback)
    assert s.get_polarizable_site_number() == 18
    s.calculate_static_energies_and_fields()
    np.testing.assert_almost_equal(s.energies["Electrostatic"]["Nuclear"], -0.321349401430, decimal=9)
    bla = s.energies["Electrostatic"]
    bla["Nuclear"] = -10.0
    assert bla["Nuclear"] == -10.0
    s.energies["Electrostatic"]["Nuclear"] = -20.0
    assert s.energies["Electrostatic"]["Nuclear"] == -20.0
    s.energies["Electrostatic"]["Electronic"] = 1.0
    s.energies["Electrostatic"]["Nuclear"] = 2.0
    s.energies["Electrostatic"]["Multipoles"] = 3.0
    s.energies["Polarization"]["Electronic"] = 4.0
    s.energies["Polarization"]["Nuclear"] = 5.0
    s.energies["Polarization"]["Multipole"] = 6.0
    assert s.total_energy == 1.0 + 2.0 + 3.0 + 4.0 + 5.0 + 6.0


@pytest.mark.gpu
def test_iso_pol():
    mol = pna_molecule()
    t = solve_induced_moments({"potfile": POT, "iso_pol": True}, mol)
    r = solve_induced_moments({"potfile": POT_ISO}, mol)
    for pt, pr in zip(t.potentials, r.potentials):
        np.testing.assert_almost_equal(pt.polarizability.values, pr.polarizability.values)
    np.testing.assert_almost_equal(t.get_induced_moments(), r.get_induced_moments(), decimal=8)


@pytest.mark.gpu
def test_nuclear_fields_closed_form():
    mol = pna_molecule()
    s = cppe.CppeState({"potfile": POT}, mol, lambda x: None)
    s.calculate_static_energies_and_fields()
    ref = np.zeros((len(s.potentials), 3))
    for i, p in enumerate(s.potentials):
        for a in mol:
            d = p.position - a.position
            ref[i] += a.charge * d / np.linalg.norm(d) ** 3
    np.testing.assert_allclose(ref.flatten(), s.nuclear_fields, atol=1e-14)


def _dense_bmat(polsites, damped):
    from .test_oracle import sym_T2
    n = len(polsites)
    B = np.zeros((3 * n, 3 * n))
    for s1, p1 in enumerate(polsites):
        v = p1.polarizability.values
        a = np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])
        B[3 * s1:3 * s1 + 3, 3 * s1:3 * s1 + 3] = np.linalg.inv(a)
        for s2, p2 in enumerate(polsites):
            if p1.excludes_site(p2.index) or s1 == s2:
                continue
            d = p2.position - p1.position
            if damped:
                av = 2.1304 / (p1.polarizability.isotropic_value * p2.polarizability.isotropic_value) ** (1 / 6)
                B[3 * s1:3 * s1 + 3, 3 * s2:3 * s2 + 3] = -sym_T2(d, av, True)
            else:
                B[3 * s1:3 * s1 + 3, 3 * s2:3 * s2 + 3] = -sym_T2(d)
    return B


@pytest.mark.gpu
@pytest.mark.parametrize("damped", [False, True])
def test_solver_by_inversion(damped):
    """tests/test_solver.py:36-146 of the reference."""
    options = {"potfile": POT, "induced_thresh": 1e-12}
    if damped:
        options["damp_induced"] = True
    s = cppe.CppeState(options, pna_molecule(), print_callback)
    s.calculate_static_energies_and_fields()
    F = s.static_fields
    s.update_induced_moments(np.zeros_like(F), False)
    polsites = cppe.get_polarizable_sites(s.potentials)
    bmat = _dense_bmat(polsites, damped)
    b = cppe.BMatrix(polsites, options)
    np.testing.assert_allclose(b.apply(F), bmat @ F, atol=1e-12)
    Afull = np.stack([b.apply(e) for e in np.eye(F.size)], axis=1)
    np.testing.assert_allclose(Afull, bmat, atol=1e-12)
    np.testing.assert_allclose(b.to_dense_matrix(), bmat, atol=1e-13)
    ref = np.linalg.inv(bmat) @ F
    np.testing.assert_almost_equal(s.get_induced_moments(), ref, decimal=9)
    ind = cppe.InducedMoments(s.potentials, options)
    np.testing.assert_allclose(ind.compute(F, True), ref, atol=1e-10)
    np.testing.assert_allclose(b.direct_inverse() @ F, ref, atol=1e-12)


@pytest.mark.gpu
def test_solver_raises_when_not_converged():
    options = {"potfile": POT, "induced_thresh": 1e-14, "maxiter": 1}
    s = cppe.CppeState(options, pna_molecule(), lambda x: None)
    s.calculate_static_energies_and_fields()
    with pytest.raises(RuntimeError, match="Failed to converge induced dipole moments."):
        s.update_induced_moments(np.zeros_like(s.static_fields), False)


@pytest.mark.gpu
def test_printer_lines():
    lines = []
    s = cppe.CppeState({"potfile": POT}, pna_molecule(), lines.append)
    s.calculate_static_energies_and_fields()
    s.update_induced_moments(np.zeros_like(s.static_fields), False)
    assert len(lines) == s.last_iterations > 0
    for k, line in enumerate(lines):
        head, val = line.split(" --- Norm: ")
        assert int(head) == k and len(val.split(".")[1]) == 12


@pytest.mark.gpu
def test_set_potentials_keeps_guess_mode():
    """The reference's set_potentials (cppe_state.cc:34-55) zeroes the moments and leaves the
    first-guess flag alone: after one SCF update, a new set of potentials is solved from zero."""
    lines = []
    opts = {"potfile": POT, "damp_induced": True}
    s = cppe.CppeState(opts, pna_molecule(), lines.append)
    s.calculate_static_energies_and_fields()
    zeros = np.zeros_like(s.static_fields)
    s.update_induced_moments(zeros, False)
    pots = s.potentials
    s.set_potentials(pots)
    np.testing.assert_array_equal(s.get_induced_moments(), 0.0)
    del lines[:]
    s.update_induced_moments(zeros, False)
    from_zero, from_alpha_f = [], []
    for make_guess, out in ((False, from_zero), (True, from_alpha_f)):
        im = cppe.InducedMoments(pots, opts)
        im.set_print_callback(out.append)
        im.compute(s.static_fields, make_guess)
    assert lines == from_zero
    assert lines != from_alpha_f


@pytest.mark.gpu
def test_polarization_energies():
    s = solve_induced_moments({"potfile": POT, "damp_induced": True}, pna_molecule())
    mu = s.get_induced_moments()
    np.testing.assert_allclose(s.energies["Polarization"]["Nuclear"], -0.5 * mu @ s.nuclear_fields, rtol=1e-14)
    np.testing.assert_allclose(s.energies["Polarization"]["Multipoles"], -0.5 * mu @ s.multipole_fields,
                               rtol=1e-14)
    assert "Polarizable Embedding Summary:" in s.summary_string


@pytest.mark.gpu
@pytest.mark.parametrize("summation", ["direct", "fmm"])
def test_induced_moments_eef(summation):
    lines = []
    opts = {"potfile": POT, "induced_thresh": 1e-12, "summation_induced_fields": summation}
    s = cppe.CppeState(opts, pna_molecule(), lines.append)
    eef = s.induced_moments_eef()
    b = cppe.BMatrix(cppe.get_polarizable_sites(s.potentials), opts)
    n = s.get_polarizable_site_number()
    binv = np.linalg.inv(b.to_dense_matrix())
    for a in range(3):
        F = np.zeros(3 * n)
        F[a::3] = 1.0
        if summation == "direct":
            np.testing.assert_allclose(eef[:, a], binv @ F, atol=1e-10)
        else:  # the tree operator's own residual
            assert np.linalg.norm(F - b.apply(eef[:, a])) < 1e-12
        # the same dipoles as one InducedMoments solve of that unit field
        im = cppe.InducedMoments(s.potentials, opts)
        np.testing.assert_allclose(eef[:, a], im.compute(F, True), rtol=0, atol=1e-13)
    # three blocks of printer lines, one per unit field, each starting at iteration 0
    assert sum(1 for x in lines if x.startswith("0 --- Norm:")) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("theta,order", [(0.5, 3), (0.5, 5), (0.7, 3), (0.7, 5), (0.99, 3), (0.99, 5)])
def test_fmm_fields_error_columns(theta, order):
    """tests/test_fmm.py:14-59 of the reference."""
    import gzip
    import shutil
    import tempfile
    ref = np.load(os.path.join(GOLDEN, "loprop_fmm_ref.npz"))
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "loprop_solvated_20.pot")
        with gzip.open(os.path.join(GOLDEN, "loprop_solvated_20.pot.gz"), "rb") as fi, open(path, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        potentials = cppe.PotfileReader(path).read()
    fs = ref["fmuls"].reshape(-1, 3)
    F = cppe.MultipoleFields(potentials, {"summation_induced_fields": "fmm", "theta": theta,
                                          "tree_expansion_order": order}).compute().reshape(fs.shape)
    err = np.linalg.norm(fs - F, axis=1) / np.linalg.norm(fs, axis=1)
    col = list(ref["err_columns"]).index(f"err_F_{theta}_{order}")
    np.testing.assert_allclose(ref["errors"][:, col], err, atol=1e-8)


# ---- test_gradients.py ---------------------------------------------------------------------------
STENCIL_5P = ([-2, -1, 1, 2], np.array([1.0, -8.0, 8.0, -1.0]) / 12.0)


def _molecule(zs, coords):
    mol = cppe.Molecule()
    for z, c in zip(zs, coords):
        mol.append(cppe.Atom(int(z), *c))
    return mol


def _fdiff(zs, coords, fn, step=1e-3):
    """Five-point central differences of fn(molecule) in every atom coordinate (test_gradients.py:48-97)."""
    out = None
    for i in range(len(zs)):
        for c in range(3):
            acc = 0.0
            for f, w in zip(*STENCIL_5P):
                cp = coords.copy()
                cp[i, c] += f * step
                acc = acc + w * np.asarray(fn(_molecule(zs, cp))) / step
            if out is None:
                out = np.zeros((len(zs), 3) + np.shape(acc))
            out[i, c] = acc
    return out


@pytest.mark.gpu
def test_nuclear_gradients_vs_finite_differences():
    """tests/test_gradients.py:32-46: energy and field gradients against 5-point differences, atol 1e-10."""
    mol = pna_molecule()
    zs = [a.atomic_number for a in mol]
    coords = np.array([a.position for a in mol])
    potentials = cppe.PotfileReader(POT).read()
    grad_nuc = cppe.MultipoleExpansion(mol, potentials).nuclear_gradient()
    fd = _fdiff(zs, coords, lambda m: cppe.MultipoleExpansion(m, potentials).interaction_energy())
    assert grad_nuc.shape == (len(zs), 3)
    np.testing.assert_allclose(fd, grad_nuc, atol=1e-10)
    npol = len(cppe.get_polarizable_sites(potentials))
    grad_field = cppe.NuclearFields(mol, potentials).nuclear_gradient()
    assert grad_field.shape == (3 * len(zs), 3 * npol)
    fdf = _fdiff(zs, coords, lambda m: cppe.NuclearFields(m, potentials).compute().reshape(npol, 3))
    np.testing.assert_allclose(fdf, grad_field.reshape(len(zs), 3, npol, 3), atol=1e-10)
    # the state's entry points return the same matrices (cppe_state.cc:73-81)
    s = cppe.CppeState({"potfile": POT}, mol, lambda x: None)
    np.testing.assert_array_equal(s.nuclear_interaction_energy_gradient(), grad_nuc)
    np.testing.assert_array_equal(s.nuclear_field_gradient(), grad_field)


@pytest.mark.gpu
def test_nuclear_gradients_no_atoms_and_symmetry():
    """An empty molecule gives an empty gradient; each 3x3 block Z_a T2(r_i - R_a) of the field
    gradient is symmetric and traceless (the potential of a point charge is harmonic)."""
    potentials = cppe.PotfileReader(POT).read()
    assert cppe.MultipoleExpansion(cppe.Molecule(), potentials).nuclear_gradient().shape == (0, 3)
    mol = pna_molecule()
    npol = len(cppe.get_polarizable_sites(potentials))
    G = cppe.NuclearFields(mol, potentials).nuclear_gradient().reshape(len(mol), 3, npol, 3)
    # T2 is symmetric and traceless at every pair
    for a in range(len(mol)):
        blk = G[a].transpose(1, 0, 2)  # (site, p, q)
        np.testing.assert_allclose(blk, blk.transpose(0, 2, 1), atol=1e-14)
        np.testing.assert_allclose(np.trace(blk, axis1=1, axis2=2), 0.0, atol=1e-12)


@pytest.mark.gpu
def test_no_polarizable_sites(tmp_path):
    """A potential without polarizabilities: InducedMoments works on empty vectors and, as the
    reference's loop does (electric_fields.cc:213-229), prints one iteration with residual norm 0;
    B is the empty operator."""
    text = "@COORDINATES\n2\nAU\nO 0 0 0 1\nH 0 0 1.8 2\n@MULTIPOLES\nORDER 0\n2\n1 -0.4\n2 0.4\n"
    pot = tmp_path / "nopol.pot"
    pot.write_text(text)
    potentials = cppe.PotfileReader(str(pot)).read()
    assert len(cppe.get_polarizable_sites(potentials)) == 0
    lines = []
    im = cppe.InducedMoments(potentials, {"potfile": str(pot)})
    im.set_print_callback(lines.append)
    assert im.compute(np.zeros(0), True).size == 0
    assert lines == ["0 --- Norm: 0.000000000000"]
    s = cppe.CppeState({"potfile": str(pot)}, pna_molecule(), lambda x: None)
    assert s.induced_moments_eef().shape == (0, 3)
    b = cppe.BMatrix(cppe.get_polarizable_sites(potentials), {"potfile": str(pot)})
    assert b.apply(np.zeros(0)).size == 0 and b.to_dense_matrix().shape == (0, 0)
