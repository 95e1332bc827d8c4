"""``lax.scan``-semantics Kohn-Sham SCF with the reference's call surface.

Mirrors jax_dft ``jit_scf.kohn_sham`` (called at ``od/train.py:117-121``) and
``qedft/train/od/jit_scf.py:247-330`` ``kohn_sham_amplitude_encoded``:

* there is no default initial density;
* convergence is MSE(mixed - input) < tol, taking effect from the next iteration, which
  then passes the previous state through with ``converged=True`` (``od/jit_scf.py:164-211``);
* reflection symmetry flips about the grid centre (``od/jit_scf.py:76-77``);
* ``stop_gradient_step`` affects only gradients and is accepted for signature parity.
"""

from __future__ import annotations

from .scf import KohnShamState, _global_functional, _local_functional, _single  # noqa: F401


def kohn_sham(locations, nuclear_charges, num_electrons, num_iterations, grids, xc_energy_density_fn,
              interaction_fn, initial_density, alpha=0.5, alpha_decay=0.9, enforce_reflection_symmetry=False,
              num_mixing_iterations=2, density_mse_converge_tolerance=-1.0, stop_gradient_step=-1):
    del stop_gradient_step
    if initial_density is None:
        raise ValueError("jit kohn_sham needs an initial_density")
    return _single(locations, nuclear_charges, num_electrons, num_iterations, grids,
                   _local_functional(xc_energy_density_fn), interaction_fn, initial_density, alpha, alpha_decay,
                   enforce_reflection_symmetry, num_mixing_iterations, density_mse_converge_tolerance, "scan")


def kohn_sham_amplitude_encoded(locations, nuclear_charges, num_electrons, num_iterations, grids,
                                xc_energy_density_fn, interaction_fn, initial_density, alpha=0.5, alpha_decay=0.9,
                                enforce_reflection_symmetry=False, num_mixing_iterations=2,
                                density_mse_converge_tolerance=-1.0, stop_gradient_step=-1):
    del stop_gradient_step
    if initial_density is None:
        raise ValueError("jit kohn_sham_amplitude_encoded needs an initial_density")
    return _single(locations, nuclear_charges, num_electrons, num_iterations, grids,
                   _global_functional(xc_energy_density_fn), interaction_fn, initial_density, alpha, alpha_decay,
                   enforce_reflection_symmetry, num_mixing_iterations, density_mse_converge_tolerance, "scan")
