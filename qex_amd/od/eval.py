"""``qedft/train/od/eval.py:23-89`` ``kohn_sham``: dispatch between the local (jax_dft) and the
amplitude-encoded loops, optional built-in LDA, exponential-Coulomb interaction."""

from __future__ import annotations

import functools

from . import scf
from . import xc as xcmod
from .utils import exponential_coulomb


def kohn_sham(params, locations, nuclear_charges, num_electrons, num_iterations, grids,
              neural_xc_energy_density_fn, use_amplitude_encoding, initial_density=None, use_lda=False,
              alpha=0.5, alpha_decay=0.9, num_mixing_iterations=2, density_mse_converge_tolerance=-1.0,
              enforce_reflection_symmetry=True):
    fn = xcmod.get_lda_xc_energy_density_fn() if use_lda else neural_xc_energy_density_fn
    kohn_sham_fn = scf.kohn_sham_amplitude_encoded if use_amplitude_encoding else scf.kohn_sham
    return kohn_sham_fn(
        locations=locations, nuclear_charges=nuclear_charges, num_electrons=num_electrons,
        num_iterations=num_iterations, grids=grids,
        xc_energy_density_fn=functools.partial(fn, params=params) if params is not None else fn,
        interaction_fn=exponential_coulomb, initial_density=initial_density, alpha=alpha,
        alpha_decay=alpha_decay, enforce_reflection_symmetry=enforce_reflection_symmetry,
        num_mixing_iterations=num_mixing_iterations, density_mse_converge_tolerance=density_mse_converge_tolerance)
