"""Copy the reference's own test fixtures (data only) into compact files under tests/golden/.

Reads, in this container only, the data files that the reference's test suite holds:

* ``tests/pna.hdf5`` / ``tests/nilered.hdf5`` -- QM molecules (Z, bohr coordinates), read raw
  (h5py is not installed): int32 Z at byte 2048, float64 coordinates at 2112 / 2216.
  Used by reference tests ``tests/cache.py:33-43``.
* ``tests/potfiles/pna_6w.pot``, ``pna_6w_isopol.pot`` -- 18-site potentials
  (``tests/test_solver.py:33``, ``tests/test_functionality.py:29-30``).
* ``tests/potfiles/loprop_solvated_20.pot`` -- 8,805-site potential (``tests/test_fmm.py:12``).
* ``tests/potfiles/fmuls_loprop_solvated_20.txt`` -- direct multipole fields at all 8,805 sites
  (golden vector, ``tests/test_fmm.py:19-24``).
* ``tests/ref_fmm_errors.csv`` -- per-site relative FMM field errors (``tests/test_fmm.py:26,57-59``).

Outputs (data, no code): ``molecules.npz``, ``pna_6w.pot``, ``pna_6w_isopol.pot``,
``loprop_solvated_20.pot.gz``, ``loprop_fmm_ref.npz`` (direct fields + error columns).
"""
import gzip
import os
import shutil
import sys

import numpy as np

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def _raw_molecule(path, natoms, coord_offset):
    raw = open(path, "rb").read()
    z = np.frombuffer(raw[2048:2048 + 4 * natoms], dtype="<i4").astype(np.int32)
    xyz = np.frombuffer(raw[coord_offset:coord_offset + 24 * natoms], dtype="<f8").reshape(natoms, 3)
    return z, xyz.copy()


def main():
    if not os.path.isdir(REF):
        sys.exit("reference checkout not present; fixtures are already committed")
    pna_z, pna_xyz = _raw_molecule(os.path.join(REF, "pna.hdf5"), 16, 2112)
    nr_z, nr_xyz = _raw_molecule(os.path.join(REF, "nilered.hdf5"), 42, 2216)
    np.savez(os.path.join(HERE, "molecules.npz"), pna_z=pna_z, pna_xyz=pna_xyz,
             nilered_z=nr_z, nilered_xyz=nr_xyz)
    for fn in ("pna_6w.pot", "pna_6w_isopol.pot"):
        shutil.copyfile(os.path.join(REF, "potfiles", fn), os.path.join(HERE, fn))
    with open(os.path.join(REF, "potfiles", "loprop_solvated_20.pot"), "rb") as fi, \
            gzip.open(os.path.join(HERE, "loprop_solvated_20.pot.gz"), "wb", compresslevel=9) as fo:
        shutil.copyfileobj(fi, fo)
    fmuls = np.loadtxt(os.path.join(REF, "potfiles", "fmuls_loprop_solvated_20.txt"))
    import pandas as pd
    errs = pd.read_csv(os.path.join(REF, "ref_fmm_errors.csv"))
    cols = [c for c in errs.columns if c.startswith("err_F_")]
    np.savez_compressed(os.path.join(HERE, "loprop_fmm_ref.npz"), fmuls=fmuls,
                        err_columns=np.array(cols), errors=errs[cols].to_numpy())
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
