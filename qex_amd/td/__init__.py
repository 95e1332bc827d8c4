"""3D restricted Kohn-Sham / Hartree-Fock on the device (the reference's pyscf-based ``td`` path)."""
from . import dft, gto, linalg, numint, scf, xc  # noqa: F401
