"""Python-loop Kohn-Sham SCF with the reference's call surface.

Mirrors ``jax_dft.scf`` as the reference calls it (``od/eval.py:67-89``) and
``qedft/train/od/scf.py``:

* ``kohn_sham``                   -- jax_dft ``scf.kohn_sham`` (local XC);
* ``kohn_sham_amplitude_encoded`` -- ``od/scf.py:434-560`` (global XC, E = dx sum f(n));
* ``solve_noninteracting_system``, ``get_initial_density``, ``get_hartree_potential``,
  ``get_hartree_energy``, ``get_external_potential_energy``, ``get_xc_energy``,
  ``get_xc_potential``, ``get_xc_energy_amplitude_encoded``, ``get_xc_potential_amplitude_encoded``,
  ``KohnShamState``.

Semantics of the loop (``od/scf.py:517-560``): convergence is tested on the unmixed
(reflected) output density of the same iteration; the stored density is the mixed one and
the energy is that of the unmixed output; once converged, later iterations repeat the
converged state; reflection symmetry flips about the centre of the nuclei.
All results are device tensors (torch, float64), stacked over iterations.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
import torch

from . import engine
from . import xc as xcmod
from .utils import get_dx


class KohnShamState(NamedTuple):
    density: torch.Tensor
    total_energy: torch.Tensor
    locations: torch.Tensor
    nuclear_charges: torch.Tensor
    external_potential: torch.Tensor
    grids: torch.Tensor
    num_electrons: torch.Tensor
    hartree_potential: torch.Tensor | None = None
    xc_potential: torch.Tensor | None = None
    xc_energy_density: torch.Tensor | None = None
    gap: torch.Tensor | None = None
    converged: torch.Tensor | bool | None = False

    def numpy(self) -> "KohnShamState":
        conv = lambda v: v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v
        return KohnShamState(*[conv(v) for v in self])


def _global_functional(fn):
    try:
        f = xcmod.resolve(fn)
    except TypeError:  # an opaque callable: evaluated between device passes
        return xcmod.ExternalXC(fn, "global")
    if f.encoding == "global":
        return f
    if isinstance(f, xcmod.PowerLaw):
        return xcmod.PowerLaw(f.c, f.p, f.eps, encoding="global")
    raise NotImplementedError(f"{type(f).__name__} has no amplitude-encoded (global) form on the device")


def _local_functional(fn):
    try:
        f = xcmod.resolve(fn)
    except TypeError:  # an opaque callable: evaluated between device passes
        return xcmod.ExternalXC(fn, "local")
    if f.encoding != "local":
        raise ValueError("kohn_sham needs a local xc_energy_density_fn; use kohn_sham_amplitude_encoded")
    return f


def stack_states(res: engine.BatchResult, prob: engine.Problem1D, b: int) -> KohnShamState:
    """State of system ``b`` of a batch, stacked over the T iterations."""
    T, dev = prob.T, prob.dev
    loc = torch.as_tensor(prob.locations[b], dtype=torch.float64, device=dev)
    z = torch.as_tensor(prob.nuclear_charges[b], device=dev)
    return KohnShamState(
        density=res.density[b],
        total_energy=res.total_energy[b],
        locations=loc.expand(T, -1),
        nuclear_charges=z.expand(T, -1),
        external_potential=res.external_potential[b].expand(T, -1),
        grids=prob.grids.expand(T, -1),
        num_electrons=torch.full((T,), int(prob.num_electrons[b]), device=dev),
        hartree_potential=res.hartree_potential[b],
        xc_potential=res.xc_potential[b],
        xc_energy_density=res.xc_energy_density[b],
        gap=res.gap[b],
        converged=res.converged[b].bool(),
    )


def run_batch(locations, nuclear_charges, num_electrons, num_iterations, grids, xc, interaction_fn,
              initial_density, alpha, alpha_decay, enforce_reflection_symmetry, num_mixing_iterations,
              density_mse_converge_tolerance, loop):
    prob = engine.Problem1D(grids, locations, nuclear_charges, num_electrons, interaction_fn, xc, num_iterations,
                            alpha=alpha, alpha_decay=alpha_decay, num_mixing_iterations=num_mixing_iterations,
                            density_mse_converge_tolerance=density_mse_converge_tolerance,
                            enforce_reflection_symmetry=enforce_reflection_symmetry, loop=loop)
    if initial_density is None:
        initial_density = engine.solve_noninteracting(prob.vext, prob.num_electrons, prob.grids).density
    res = prob.run(initial_density)
    return prob, res


def _single(locations, nuclear_charges, num_electrons, num_iterations, grids, xc, interaction_fn, initial_density,
            alpha, alpha_decay, enforce_reflection_symmetry, num_mixing_iterations, density_mse_converge_tolerance,
            loop):
    loc = np.asarray(locations.detach().cpu() if isinstance(locations, torch.Tensor) else locations,
                     dtype=np.float64).reshape(-1)
    z = np.asarray(nuclear_charges.detach().cpu() if isinstance(nuclear_charges, torch.Tensor) else nuclear_charges,
                   dtype=np.float64).reshape(-1)
    n0 = None if initial_density is None else engine.as_f64(initial_density, engine.device())[None]
    prob, res = run_batch([loc], [z], [int(num_electrons)], num_iterations, grids, xc, interaction_fn, n0, alpha,
                          alpha_decay, enforce_reflection_symmetry, num_mixing_iterations,
                          density_mse_converge_tolerance, loop)
    return stack_states(res, prob, 0)


def kohn_sham(locations, nuclear_charges, num_electrons, num_iterations, grids, xc_energy_density_fn,
              interaction_fn, initial_density=None, alpha=0.5, alpha_decay=0.9, enforce_reflection_symmetry=False,
              num_mixing_iterations=2, density_mse_converge_tolerance=-1.0) -> KohnShamState:
    """Runs Kohn-Sham (local XC) for ``num_iterations``; returns the states of every iteration."""
    return _single(locations, nuclear_charges, num_electrons, num_iterations, grids,
                   _local_functional(xc_energy_density_fn), interaction_fn, initial_density, alpha, alpha_decay,
                   enforce_reflection_symmetry, num_mixing_iterations, density_mse_converge_tolerance, "python")


def kohn_sham_amplitude_encoded(locations, nuclear_charges, num_electrons, num_iterations, grids,
                                xc_energy_density_fn, interaction_fn, initial_density=None, alpha=0.5,
                                alpha_decay=0.9, enforce_reflection_symmetry=False, num_mixing_iterations=2,
                                density_mse_converge_tolerance=-1.0) -> KohnShamState:
    """Runs Kohn-Sham with a global XC functional (``od/scf.py:434-560``)."""
    return _single(locations, nuclear_charges, num_electrons, num_iterations, grids,
                   _global_functional(xc_energy_density_fn), interaction_fn, initial_density, alpha, alpha_decay,
                   enforce_reflection_symmetry, num_mixing_iterations, density_mse_converge_tolerance, "python")


# ---- single-shot pieces (device) ------------------------------------------------------------
def solve_noninteracting_system(external_potential, num_electrons, grids):
    """(density, total_eigen_energies, gap) of T + diag(v) (``od/scf.py:233-268``)."""
    r = engine.solve_noninteracting(external_potential, num_electrons, grids)
    return r.density, r.total_eigen_energies, r.gap


def get_initial_density(states, method):
    """Initial densities of a dataset: ``'exact'`` returns ``states.density``;
    ``'noninteracting'`` solves T + v_ext for every system (``od/trainer.py:72-75``)."""
    if method == "exact":
        return states.density
    if method == "noninteracting":
        v = engine.as_f64(states.external_potential, engine.device())
        ne = np.asarray(states.num_electrons).reshape(-1)
        grids = states.grids
        if isinstance(grids, torch.Tensor) and grids.dim() > 1 or (not isinstance(grids, torch.Tensor)
                                                                    and np.asarray(grids).ndim > 1):
            grids = grids[0]
        return engine.solve_noninteracting(v, ne if ne.size > 1 else int(ne[0]), grids).density
    raise ValueError(f"Unknown initialization method {method}")


def get_hartree_potential(density, grids, interaction_fn):
    return engine.hartree_potential(density, grids, interaction_fn)


def get_hartree_energy(density, grids, interaction_fn):
    n = engine.as_f64(density, engine.device())
    vh = engine.hartree_potential(n, grids, interaction_fn)
    return 0.5 * (n * vh).sum(-1) * get_dx(grids)


def get_external_potential_energy(external_potential, density, grids):
    dev = engine.device()
    return (engine.as_f64(external_potential, dev) * engine.as_f64(density, dev)).sum(-1) * get_dx(grids)


def _xc(density, xc_energy_density_fn, grids, encoding):
    """(E_xc, v_xc) of a qex functional through ``qex1d_xc_evaluate``, None for other callables.
    A functional of the other encoding is summed as the reference's formula reads its values
    (global: dx sum_i e_i; local: dx f sum_i n_i), without a potential (v_xc None)."""
    try:
        f = xcmod.resolve(xc_energy_density_fn)
    except TypeError:
        return None
    dx = get_dx(grids)
    if f.encoding != encoding and encoding == "global" and isinstance(f, xcmod.PowerLaw):
        f = _global_functional(f)
    if f.encoding == encoding:
        _, exc, vxc = engine.xc_evaluate(f, density, dx)
        return exc, vxc
    eps, _, _ = engine.xc_evaluate(f, density, dx)
    if encoding == "global":
        return eps.sum(-1) * dx, None
    n = engine.as_f64(density, engine.device())
    return eps * n.sum(-1) * dx, None


def get_xc_energy(density, xc_energy_density_fn, grids):
    """dx * sum(e(n) n) (jax_dft ``scf.get_xc_energy``), on the device for qex functionals; any
    other callable is called on the device density."""
    r = _xc(density, xc_energy_density_fn, grids, "local")
    if r is not None:
        return r[0]
    n = engine.as_f64(density, engine.device())
    return (xc_energy_density_fn(n) * n).sum(-1) * get_dx(grids)


def get_xc_potential(density, xc_energy_density_fn, grids):
    """d E_xc / dn / dx (jax_dft ``scf.get_xc_potential``) for a qex functional, on the device."""
    r = _xc(density, xc_energy_density_fn, grids, "local")
    if r is None or r[1] is None:
        raise TypeError("get_xc_potential needs a local qex_amd functional")
    return r[1]


def get_xc_energy_amplitude_encoded(density, xc_energy_density_fn, grids):
    """dx * sum(f(n)) (``od/scf.py:32-50``), on the device for qex functionals."""
    r = _xc(density, xc_energy_density_fn, grids, "global")
    if r is not None:
        return r[0]
    n = engine.as_f64(density, engine.device())
    return torch.sum(xc_energy_density_fn(n)) * get_dx(grids)


def get_xc_potential_amplitude_encoded(density, xc_energy_density_fn, grids):
    """d/dn [dx * sum(f(n))] (``od/scf.py:53-76``) for a qex functional, on the device."""
    r = _xc(density, xc_energy_density_fn, grids, "global")
    if r is None or r[1] is None:
        raise TypeError("get_xc_potential_amplitude_encoded needs a global qex_amd functional or PowerLaw")
    return r[1]


def get_total_eigen_energies(num_electrons, eigen_energies):
    e = torch.repeat_interleave(torch.as_tensor(eigen_energies), 2)
    return e[:num_electrons].sum()


def get_gap(num_electrons, eigen_energies):
    e = torch.repeat_interleave(torch.as_tensor(eigen_energies), 2)
    return e[num_electrons] - e[num_electrons - 1]
