"""1D helpers with the names of ``jax_dft.utils`` as the reference calls them.

Interaction functions are callable objects: they evaluate u(r) on numpy arrays or torch
tensors, and carry a ``kind`` that the engine uses to build the Hartree table and the
external potential on the device.  Any other callable u(r) also works (evaluated once per
solve, on the device, through torch).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .. import _lib

EXPONENTIAL_COULOMB_AMPLITUDE = 1.071295
EXPONENTIAL_COULOMB_KAPPA = 1.0 / 2.385345


def _xp(x):
    return torch if isinstance(x, torch.Tensor) else np


class ExponentialCoulomb:
    """u(r) = A exp(-kappa |r|)  (jax_dft.utils.exponential_coulomb; od/eval.py:82)."""

    kind = _lib.QEX_INTERACTION_EXPONENTIAL

    def __init__(self, amplitude=EXPONENTIAL_COULOMB_AMPLITUDE, kappa=EXPONENTIAL_COULOMB_KAPPA):
        self.amplitude = float(amplitude)
        self.kappa = float(kappa)

    @property
    def params(self):
        return self.amplitude, self.kappa

    def __call__(self, displacements, amplitude=None, kappa=None):
        a = self.amplitude if amplitude is None else amplitude
        k = self.kappa if kappa is None else kappa
        xp = _xp(displacements)
        return a * xp.exp(-xp.abs(displacements) * k)


class SoftCoulomb:
    """u(r) = 1 / sqrt(r^2 + soften_factor)  (jax_dft.utils.soft_coulomb)."""

    kind = _lib.QEX_INTERACTION_SOFT

    def __init__(self, soften_factor=1.0):
        self.soften_factor = float(soften_factor)

    @property
    def params(self):
        return self.soften_factor, 0.0

    def __call__(self, displacements, soften_factor=None):
        f = self.soften_factor if soften_factor is None else soften_factor
        xp = _xp(displacements)
        return 1.0 / xp.sqrt(displacements ** 2 + f)


exponential_coulomb = ExponentialCoulomb()
soft_coulomb = SoftCoulomb()


def get_dx(grids) -> float:
    """(x_max - x_min) / (N - 1)."""
    if isinstance(grids, torch.Tensor):
        return float((grids[-1] - grids[0]).item()) / (grids.shape[0] - 1)
    grids = np.asarray(grids)
    return float((grids[-1] - grids[0]) / (len(grids) - 1))


def gaussian(grids, center, sigma=1.0):
    """Unit-norm Gaussian exp(-(x-c)^2 / 2 sigma^2) / (sigma sqrt(2 pi))."""
    xp = _xp(grids)
    return xp.exp(-0.5 * ((grids - center) / sigma) ** 2) / (math.sqrt(2 * math.pi) * sigma)


def check_uniform(grids) -> None:
    g = grids.detach().cpu().numpy() if isinstance(grids, torch.Tensor) else np.asarray(grids)
    if g.ndim != 1 or len(g) < 5:
        raise ValueError("grids must be a 1D array of at least 5 points")
    d = np.diff(g)
    if np.max(np.abs(d - d.mean())) > 1e-9 * max(1.0, abs(d.mean())):
        raise ValueError("grids must be uniform")


def reflection_window(locations, grids):
    """Inclusive index window [lo, hi] symmetric about the centre of ``locations``.

    The centre is the mean location; on a grid point the window is centred there, otherwise on
    the two grid points around it; the radius is the largest that stays on the grid (the
    behaviour of jax_dft.utils.flip_and_average used at od/scf.py:162-167).
    """
    g = grids.detach().cpu().numpy() if isinstance(grids, torch.Tensor) else np.asarray(grids, dtype=np.float64)
    loc = locations.detach().cpu().numpy() if isinstance(locations, torch.Tensor) else np.asarray(locations)
    center = float(np.mean(loc))
    on_grid = np.isclose(g, center, rtol=0.0, atol=1e-9)
    if np.any(on_grid):
        left = right = int(np.argmax(on_grid))
    else:
        dist = np.abs(g - center)
        left = int(np.argmin(np.where(g < center, dist, np.inf)))
        right = int(np.argmin(np.where(g > center, dist, np.inf)))
    radius = min(left, len(g) - right - 1)
    return left - radius, right + radius


def flip_and_average(locations, grids, array):
    """Flips and averages ``array`` inside the reflection window about the nuclei's centre."""
    lo, hi = reflection_window(locations, grids)
    if isinstance(array, torch.Tensor):
        out = array.clone()
        seg = array[lo:hi + 1]
        out[lo:hi + 1] = (seg + torch.flip(seg, dims=[0])) / 2
        return out
    out = np.array(array, copy=True)
    seg = out[lo:hi + 1].copy()
    out[lo:hi + 1] = (seg + seg[::-1]) / 2
    return out


def flip_and_average_on_center(array):
    """(a + reversed(a)) / 2 about the grid centre."""
    if isinstance(array, torch.Tensor):
        return (array + torch.flip(array, dims=[-1])) / 2
    return (array + np.asarray(array)[..., ::-1]) / 2


def get_atomic_chain_potential(grids, locations, nuclear_charges, interaction_fn):
    """v_ext(x) = -sum_a Z_a u(x - R_a), evaluated on the device."""
    from . import engine

    return engine.external_potential(grids, locations, nuclear_charges, interaction_fn)
