"""MI355X-native QUMASUN LOBPCG H.psi hot path.

The compute path is the HIP C-ABI library built from csrc/ (libqumasun_hip.so);
this package holds its ctypes binding, the host-side mirrors of the reference
operator interface and the synthetic-system generator.
"""
from . import _lib, adjoint, hamiltonian, lobpcg, scf, system, upf  # noqa: F401

__all__ = ["_lib", "adjoint", "hamiltonian", "lobpcg", "scf", "system", "upf"]
