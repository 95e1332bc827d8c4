"""1D Kohn-Sham engine (jax_dft-style loop) on the device."""
from . import dataset, engine, eval, jit_scf, scf, train, utils, xc  # noqa: F401
