"""qex_amd -- MI355X-native Kohn-Sham SCF engine (gfx950 HIP kernels behind a C ABI).

Sub-packages mirror the reference's call surface:

* ``qex_amd.od``  -- the 1D ``jax_dft``-style loop (``scf``, ``jit_scf``, ``eval``, ``xc``, ``utils``,
  ``train``, ``trainer``, ``dataset``);
* ``qex_amd.td``  -- the 3D ``numint`` / RKS cycle (``gto``, ``dft``, ``numint``, ``scf``, ``linalg``, ``xc``);
* ``qex_amd.networks`` -- ``KohnShamNetwork`` / MLP builders / ``build_xc_functional``;
* ``qex_amd.serve`` -- the serving loop for mixed 1D / 3D request streams.

Numerics run in ``libqex.so`` (``qex_amd/csrc``); PyTorch only provides device memory and
streams.
"""

__all__ = ["od", "td", "networks", "serve"]
