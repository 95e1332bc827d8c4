"""Molecular integration grid (``pyscf.dft.Grids`` as built at
``qedft/train/td/trainer_legacy_no_jit.py:248-251``): Treutler-Ahlrichs M4 radial points on a
Chebyshev (2nd kind) base, a Gauss-Legendre(cos theta) x uniform-phi angular product, and Becke or
Stratmann partitioning on the device (``qex3d_grid_partition``)."""

from __future__ import annotations

import ctypes
import itertools
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


# Lebedev-Laikov octahedral rules, n points -> [(orbit, parameter, weight)] with weights summing to 1
# over the points; orbits a1 (1,0,0) x6, a2 (0,1,1)/sqrt2 x12, a3 (1,1,1)/sqrt3 x8, b(l) (l,l,m) x24,
# c(p) (p,q,0) x24.  The rule with n points integrates every spherical polynomial up to degree 3, 5,
# 7, 9, 11 and 17 (n = 6 ... 110) exactly.
_LEBEDEV = {
    6: [("a1", None, 1 / 6)],
    14: [("a1", None, 1 / 15), ("a3", None, 3 / 40)],
    26: [("a1", None, 1 / 21), ("a2", None, 4 / 105), ("a3", None, 9 / 280)],
    38: [("a1", None, 1 / 105), ("a3", None, 9 / 280), ("c", 0.4597008433809831, 1 / 35)],
    50: [("a1", None, 0.012698412698412698), ("a2", None, 0.022574955908289243), ("a3", None, 0.02109375),
         ("b", 0.30151134457776363, 0.020173335537918871)],
    110: [("a1", None, 0.3828270494937162e-2), ("a3", None, 0.9793737512487512e-2),
          ("b", 0.1851156353447362, 0.8211737283191111e-2), ("b", 0.6904210483822922, 0.9942814891178103e-2),
          ("b", 0.3956894730559419, 0.9595471336070963e-2), ("c", 0.4783690288121502, 0.9694996361663028e-2)],
}


def angular_lebedev(n):
    """The Lebedev rule with n points (pyscf dft.Grids' angular grids; level 0 uses 50 for hydrogen):
    unit vectors (n, 3), weights summing to 4 pi, each orbit in sorted order."""
    if n not in _LEBEDEV:
        raise ValueError(f"Lebedev rules available: {sorted(_LEBEDEV)}")
    dirs, w = [], []
    for kind, par, wt in _LEBEDEV[n]:
        if kind == "a1":
            v = (1.0, 0.0, 0.0)
        elif kind == "a2":
            v = (0.0, 1 / math.sqrt(2), 1 / math.sqrt(2))
        elif kind == "a3":
            v = (1 / math.sqrt(3),) * 3
        elif kind == "b":
            v = (par, par, math.sqrt(1 - 2 * par * par))
        else:
            v = (par, math.sqrt(1 - par * par), 0.0)
        pts = {tuple(round(sg * c, 15) + 0.0 for sg, c in zip(sgn, perm))
               for perm in itertools.permutations(v) for sgn in itertools.product((1, -1), repeat=3)}
        o = np.array(sorted(pts))
        dirs.append(o)
        w.append(np.full(len(o), 4 * math.pi * wt))
    return np.concatenate(dirs), np.concatenate(w)


_TABLES = {}


def _device_tables(n_rad, n_theta, n_phi, dev, n_ang=None):
    """Radial and angular tables on the device, made once per shape and device (a host-to-device
    copy from pageable memory waits for the stream, so it must not sit on the per-molecule path).
    n_ang: a Lebedev rule with that many points instead of the product grid."""
    key = (n_rad, n_theta, n_phi, n_ang, str(dev))
    if key not in _TABLES:
        r, wr = radial_treutler(n_rad)
        dirs, wa = angular_product(n_theta, n_phi) if n_ang is None else angular_lebedev(n_ang)
        _TABLES[key] = (len(r), len(wa), [engine.as_f64(x, dev) for x in (r, wr, dirs.reshape(-1), wa)])
    return _TABLES[key]


def original_becke(g):
    """Becke's cell step, three iterations of 1.5 mu - 0.5 mu^3 (pyscf gen_grid.original_becke); as
    ``Grids.becke_scheme`` it selects that partition on the device."""
    for _ in range(3):
        g = 1.5 * g - 0.5 * g ** 3
    return g


def stratmann(g):
    """Stratmann-Scuseria-Frisch cell step (pyscf gen_grid.stratmann, chosen by the reference at
    td/trainer_legacy_no_jit.py:250); as ``Grids.becke_scheme`` it selects that partition on the device."""
    a = 0.64
    z = np.asarray(g, dtype=np.float64) / a
    out = (35 * z - 35 * z ** 3 + 21 * z ** 5 - 5 * z ** 7) / 16
    return np.where(z <= -1, -1.0, np.where(z >= 1, 1.0, out))


class Grids:
    """Atom-centred grid built on the device (``qex3d_grid_partition``); ``becke_scheme`` is
    ``original_becke`` (default) or ``stratmann``, as pyscf's attribute of that name."""

    def __init__(self, mol, n_rad=50, n_theta=10, n_phi=11, n_ang=None):
        self.mol = mol
        self.n_rad, self.n_theta, self.n_phi = n_rad, n_theta, n_phi
        self.n_ang = n_ang  # a Lebedev rule with this many points (6, 14, 26, 38, 50, 110) instead
        self.becke_scheme = original_becke
        self.coords = None
        self.weights = None

    def build(self):
        dev = engine.device()
        nr, na, tabs = _device_tables(self.n_rad, self.n_theta, self.n_phi, dev, self.n_ang)
        ng = self.mol.natm * nr * na
        self.coords = torch.empty((ng, 3), dtype=torch.float64, device=dev)
        self.weights = torch.empty(ng, dtype=torch.float64, device=dev)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        schemes = {original_becke: _lib.QEX3D_PARTITION_BECKE, stratmann: _lib.QEX3D_PARTITION_STRATMANN}
        if self.becke_scheme not in schemes:
            raise ValueError("becke_scheme must be qex_amd.td.dft.original_becke or qex_amd.td.dft.stratmann")
        _lib.check3(_lib.load().qex3d_grid_partition(ctypes.byref(self.mol.device_basis()), schemes[self.becke_scheme],
                                                     nr, p(tabs[0]), p(tabs[1]), na, p(tabs[2]), p(tabs[3]),
                                                     p(self.coords), p(self.weights), engine._stream()),
                    "qex3d_grid_partition")
        self._keep = tabs
        return self

    @property
    def size(self):
        return 0 if self.weights is None else self.weights.shape[0]
