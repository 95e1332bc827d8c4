"""Molecular integration grid (``pyscf.dft.Grids`` as built at
``qedft/train/td/trainer_legacy_no_jit.py:248-251``): Treutler-Ahlrichs M4 radial points on a
Chebyshev (2nd kind) base, a Gauss-Legendre(cos theta) x uniform-phi angular product, and Becke
partitioning on the device (``qex3d_grid``)."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from ..od import engine


def radial_treutler(n, xi=0.8):
    i = np.arange(1, n + 1)
    th = i * np.pi / (n + 1)
    x = np.cos(th)
    wx = np.pi / (n + 1) * np.sin(th)
    ln2 = math.log(2.0)
    r = xi / ln2 * (1 + x) ** 0.6 * np.log(2 / (1 - x))
    drdx = xi / ln2 * (0.6 * (1 + x) ** (-0.4) * np.log(2 / (1 - x)) + (1 + x) ** 0.6 / (1 - x))
    return r, wx * drdx * r * r


def angular_product(n_theta=10, n_phi=11):
    ct, wt = np.polynomial.legendre.leggauss(n_theta)
    phi = 2 * np.pi * (np.arange(n_phi) + 0.5) / n_phi
    st = np.sqrt(1 - ct * ct)
    dirs = np.stack([np.outer(st, np.cos(phi)).ravel(), np.outer(st, np.sin(phi)).ravel(),
                     np.outer(ct, np.ones(n_phi)).ravel()], axis=1)
    w = np.outer(wt, np.full(n_phi, 2 * np.pi / n_phi)).ravel()
    return dirs, w


_TABLES = {}


def _device_tables(n_rad, n_theta, n_phi, dev):
    """Radial and angular tables on the device, made once per shape and device (a host-to-device
    copy from pageable memory waits for the stream, so it must not sit on the per-molecule path)."""
    key = (n_rad, n_theta, n_phi, str(dev))
    if key not in _TABLES:
        r, wr = radial_treutler(n_rad)
        dirs, wa = angular_product(n_theta, n_phi)
        _TABLES[key] = (len(r), len(wa), [engine.as_f64(x, dev) for x in (r, wr, dirs.reshape(-1), wa)])
    return _TABLES[key]


class Grids:
    def __init__(self, mol, n_rad=50, n_theta=10, n_phi=11):
        self.mol = mol
        self.n_rad, self.n_theta, self.n_phi = n_rad, n_theta, n_phi
        self.coords = None
        self.weights = None

    def build(self):
        dev = engine.device()
        nr, na, tabs = _device_tables(self.n_rad, self.n_theta, self.n_phi, dev)
        ng = self.mol.natm * nr * na
        self.coords = torch.empty((ng, 3), dtype=torch.float64, device=dev)
        self.weights = torch.empty(ng, dtype=torch.float64, device=dev)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        _lib.check3(_lib.load().qex3d_grid(ctypes.byref(self.mol.device_basis()), nr, p(tabs[0]), p(tabs[1]),
                                           na, p(tabs[2]), p(tabs[3]), p(self.coords), p(self.weights),
                                           engine._stream()), "qex3d_grid")
        self._keep = tabs
        return self

    @property
    def size(self):
        return 0 if self.weights is None else self.weights.shape[0]
