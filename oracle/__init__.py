"""FP64 CPU oracle for the Kohn-Sham SCF hot path.

TEST INFRASTRUCTURE ONLY. This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
The product package ``qex_amd`` never imports it and never falls back to it.

* ``oracle.od`` restates the 1D ``jax_dft``-style loop (``qedft/train/od/scf.py``,
  ``qedft/train/od/jit_scf.py`` and the un-vendored ``jax_dft`` functions they call).
"""
