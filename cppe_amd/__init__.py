"""MI355X-native polarizable-embedding field / induced-dipole path with CPPE's Python API.

``import cppe_amd as cppe`` gives the reference's names (CppeState, PotfileReader,
MultipoleFields, InducedMoments, BMatrix, NuclearFields, Potential, Atom, Molecule, ...),
served by the pybind11 module ``cppe_amd.pycppe`` over the HIP library ``libcppe_hip.so``.
Submodules ``cppe_amd.build``, ``cppe_amd.capi`` and ``cppe_amd.synthetic`` import without it.
There is no CPU fallback: using the API without the built libraries raises ImportError.
"""
import importlib

__all__ = ["__build_type__", "__parallel__"]


def _pycppe():
    try:
        return importlib.import_module(".pycppe", __name__)
    except ImportError as e:  # pragma: no cover - exercised only when unbuilt
        raise ImportError("cppe_amd.pycppe is not built; run `python -m cppe_amd.build` "
                          f"(gfx950 HIP library required): {e}") from e


_SUBMODULES = ("build", "capi", "synthetic", "mlp", "pycppe")


def __getattr__(name):
    if name.startswith("__") and name not in ("__build_type__", "__parallel__"):
        raise AttributeError(name)
    if name in _SUBMODULES:
        return importlib.import_module("." + name, __name__)
    mod = _pycppe()
    try:
        return getattr(mod, name)
    except AttributeError:
        raise AttributeError(f"module 'cppe_amd' has no attribute '{name}'") from None


def __dir__():
    return sorted(set(globals()) | set(dir(_pycppe())))
