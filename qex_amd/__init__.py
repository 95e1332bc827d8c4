"""qex_amd -- MI355X-native Kohn-Sham SCF engine (gfx950 HIP kernels behind a C ABI).

Sub-packages mirror the reference's call surface:

* ``qex_amd.od``  -- the 1D ``jax_dft``-style loop (``scf``, ``jit_scf``, ``eval``, ``xc``, ``utils``);
* ``qex_amd.networks`` -- ``KohnShamNetwork`` / MLP builders / ``build_xc_functional``.

Numerics run in ``libqex.so`` (``qex_amd/csrc``); PyTorch only provides device memory and
streams.
"""

__all__ = ["od", "networks"]
