"""MI355X-native (gfx950) matrix-free symmetry-sector exact diagonalization.

Host side of the sector solver: lattice/model/group builders (``lattice``), the
ctypes binding of ``libsed.so`` (``_lib``) and the reference-named API (``ed``).
"""
from . import lattice  # noqa: F401

__all__ = ["lattice"]
