"""Drop-in for the reference package `acoustics3d` (acoustics3d_cpp/src/acoustics3d/__init__.py:1):
`import acoustics3d as ac3d; ac3d.DiffBEM(...)` runs the gfx950 implementation."""
from acoustic_reliefs_amd.acoustics3d_impl import DiffBEM, sphere  # noqa: F401
