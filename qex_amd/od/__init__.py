"""1D Kohn-Sham engine (jax_dft-style loop) on the device."""
from . import engine, eval, jit_scf, scf, utils, xc  # noqa: F401
