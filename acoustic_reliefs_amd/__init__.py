"""MI355X (gfx950) Helmholtz BEM: drop-in for acoustics3d (DiffBEM, sphere).

The native pieces are libarh.so (HIP kernels + the C-ABI of include/arh.h) and the pybind11
extension acoustics3d_impl built on that C-ABI. Importing DiffBEM/sphere fails loudly when the
extension has not been built (there is no CPU fallback).
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))


def _load():
    import importlib
    try:
        impl = importlib.import_module(__name__ + ".acoustics3d_impl")
    except ImportError as e:  # pragma: no cover - exercised when not built
        raise ImportError("acoustic_reliefs_amd: the native extension is not built (run __graft_entry__.build()): %s"
                          % e) from e
    return impl


def __getattr__(name):
    if name in ("DiffBEM", "sphere", "acoustics3d_impl"):
        impl = _load()
        return impl if name == "acoustics3d_impl" else getattr(impl, name)
    raise AttributeError(name)
