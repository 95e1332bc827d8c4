"""Losses and their gradients for training a LocalMLP XC functional through the batched 1D
Kohn-Sham scan loop (the reference's ``create_loss_fn``, ``qedft/train/od/train.py:250-323``, and
the simpler energy + density loss used for benchmarking), evaluated with the device forward and
reverse passes."""

from __future__ import annotations

import numpy as np
import torch

from . import engine


def c3_loss(result, target_energy, target_density, dx, density_weight=1e-2):
    """mean_b (E_b,T - E*_b)^2 + density_weight * mean_b dx sum_i (n_b,T - n*_b)^2 and its cotangents."""
    E = result.total_energy
    n = result.density
    B, T = E.shape
    tE = engine.as_f64(target_energy, E.device).reshape(B)
    tn = engine.as_f64(target_density, E.device).reshape(B, -1)
    e = E[:, -1] - tE
    d = n[:, -1] - tn
    loss = (e * e).mean() + density_weight * dx * (d * d).sum(dim=1).mean()
    dE = torch.zeros_like(E)
    dE[:, -1] = 2.0 * e / B
    dn = torch.zeros_like(n)
    dn[:, -1] = 2.0 * density_weight * dx * d / B
    return loss, dE, dn


def trajectory_loss(result, target_energy, target_density, num_electrons, dx, discount=0.9, start=10):
    """The reference trainer's loss (``od/train.py:286-319`` with jax_dft ``losses.trajectory_mse``):
    mean over the batch and the iterations t >= start of discount^(T-1-t) (E_t - E*)^2 / N_e, plus
    the mean square error of the final density times dx N / N_e."""
    E = result.total_energy
    n = result.density
    B, T = E.shape
    N = n.shape[-1]
    tE = engine.as_f64(target_energy, E.device).reshape(B, 1)
    tn = engine.as_f64(target_density, E.device).reshape(B, N)
    ne = float(num_electrons)
    S = T - start
    disc = torch.as_tensor(discount ** np.arange(S)[::-1].copy(), dtype=torch.float64, device=E.device)
    e = E[:, start:] - tE
    loss = (disc * e * e).mean() / ne
    dE = torch.zeros_like(E)
    dE[:, start:] = 2.0 * disc * e / (B * S) / ne
    d = n[:, -1] - tn
    loss = loss + (d * d).mean() * dx * N / ne
    dn = torch.zeros_like(n)
    dn[:, -1] = 2.0 * d / (B * N) * dx * N / ne
    return loss, dE, dn


def value_and_grad(problem, initial_density, loss_fn, *loss_args, **loss_kwargs):
    """(loss, flat parameter gradient) of ``loss_fn(result, ...)`` through the forward and reverse
    passes on the device.  ``problem`` needs ``want_orbitals=True``."""
    res = problem.run(initial_density)
    loss, dE, dn = loss_fn(res, *loss_args, **loss_kwargs)
    grad = engine.kohn_sham_vjp(problem, res, dE, dn)
    return loss, grad, res
