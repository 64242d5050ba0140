"""hartree_fock_cpp_amd -- MI355X-native ERI and Fock-build engine.

Drop-in for the "eri_hardware":"GPU" path of KLYYCYNN/hartree_fock_cpp: the
native library (HIP for gfx950 + C++ host) computes the unique ERI array in the
reference's layout and the J/K/Fock matrices; this package is its Python mirror.
"""
from ._native import HfxError, lib  # noqa: F401  (fails loudly when the library is not built)
from .eri import ERI_GPU, Basis, FockBuilder, repulsion_matrix, schwarz_bounds  # noqa: F401

__all__ = ["ERI_GPU", "Basis", "FockBuilder", "repulsion_matrix", "schwarz_bounds", "HfxError"]
