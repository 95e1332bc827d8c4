"""Molecules and Cartesian s/p Gaussian basis sets, laid out for the device (``qex3d_basis``).

Mirrors the parts of ``pyscf.gto.Mole`` the reference reads (``atom``, ``basis``, ``nelectron``,
``nao``, ``energy_nuc``; ``qedft/train/td/hf_legacy.py``, ``td/numint.py``).  Contractions are
normalised to unit self-overlap; coefficients stored for the device include the primitive
normalisation (2a/pi)^(3/4) (2 sqrt(a))^l.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from ..od import engine

BOHR = 0.52917721092

ELEMENTS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


def _sto3g():
    """STO-3G (Hehre, Stewart, Pople, J. Chem. Phys. 51, 2657 (1969)): least-squares 3-Gaussian fits
    of 1s and 2sp Slater functions with exponent 1, scaled by the standard molecular zeta^2 of each
    element (H 1.24; He 1.69; Li-Ne 1s 2.69 ... 9.64, 2sp 0.80 ... 2.88).  Gives the tabulated
    exponents, e.g. H 3.42525091, 0.62391373, 0.16885540; O 130.70932, 23.808866, 6.4436083 and
    5.0331512, 1.1695960, 0.3803892."""
    e1 = np.array([2.227660584, 0.4057711562, 0.1098175104])
    c1 = [0.1543289673, 0.5353281423, 0.4446345422]
    e2 = np.array([0.9942027, 0.2310313, 0.0751386])
    c2s = [-0.09996722919, 0.3995128261, 0.7001154689]
    c2p = [0.1559162750, 0.6076837186, 0.3919573931]
    zeta = {"He": (1.69, None), "Li": (2.69, 0.80), "Be": (3.68, 1.15), "B": (4.68, 1.50), "C": (5.67, 1.72),
            "N": (6.67, 1.95), "O": (7.66, 2.25), "F": (8.65, 2.55), "Ne": (9.64, 2.88)}
    lib = {"H": [(0, [3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])]}
    for el, (z1, z2) in zeta.items():
        shells = [(0, list(z1 * z1 * e1), c1)]
        if z2 is not None:
            shells += [(0, list(z2 * z2 * e2), c2s), (1, list(z2 * z2 * e2), c2p)]
        lib[el] = shells
    return lib


BASIS_LIBRARY = {
    "sto-3g": _sto3g(),
    "6-31g": {"H": [(0, [18.7311370, 2.8253937, 0.6401217], [0.03349460, 0.23472695, 0.81375733]),
                    (0, [0.1612778], [1.0])]},
}


def prim_norm(alpha, l):
    return (2 * alpha / math.pi) ** 0.75 * (2 * math.sqrt(alpha)) ** l


class Mole:
    """atom: list of (symbol or Z, (x, y, z)); unit 'Angstrom' or 'Bohr'; basis: a name from
    BASIS_LIBRARY or {symbol: [(l, exps, coefs), ...]}; or per-atom shells via ``shells_per_atom``."""

    def __init__(self, atom, basis="6-31g", unit="Angstrom", shells_per_atom=None):
        zs, xyz = [], []
        for sym, pos in atom:
            zs.append(ELEMENTS[sym] if isinstance(sym, str) else int(sym))
            xyz.append(pos)
        self.charges = np.asarray(zs, dtype=np.float64)
        scale = 1.0 / BOHR if unit.lower().startswith("a") else 1.0
        self.coords = np.asarray(xyz, dtype=np.float64) * scale
        if shells_per_atom is None:
            lib = BASIS_LIBRARY[basis.lower()] if isinstance(basis, str) else basis
            inv = {v: k for k, v in ELEMENTS.items()}
            shells_per_atom = [lib[inv[int(z)]] for z in zs]
        self.shells = []  # (atom, l, exps, coefs incl. primitive norm)
        for a, shells in enumerate(shells_per_atom):
            for l, exps, coefs in shells:
                exps = np.asarray(exps, dtype=np.float64)
                coefs = np.asarray(coefs, dtype=np.float64)
                nrm = np.array([prim_norm(e, l) for e in exps])
                p = exps[:, None] + exps[None, :]
                sij = (math.pi / p) ** 1.5 * (1.0 / (2 * p)) ** l
                ss = float(np.sum((coefs * nrm)[:, None] * (coefs * nrm)[None, :] * sij))
                self.shells.append((a, int(l), exps, coefs * nrm / math.sqrt(ss)))
        self._dev = None

    @property
    def natm(self):
        return len(self.charges)

    @property
    def nelectron(self):
        return int(round(float(self.charges.sum())))

    @property
    def nao(self):
        return sum(1 if l == 0 else 3 for _, l, _, _ in self.shells)

    def energy_nuc(self):
        e = 0.0
        for a in range(self.natm):
            for b in range(a):
                e += self.charges[a] * self.charges[b] / float(np.linalg.norm(self.coords[a] - self.coords[b]))
        return e

    def device_basis(self):
        """qex3d_basis struct over device arrays (kept alive on the Mole)."""
        if self._dev is not None:
            return self._dev[0]
        dev = engine.device()
        sh_atom, sh_l, sh_n, sh_off, exps, coefs, ao_shell, ao_lmn = [], [], [], [], [], [], [], []
        for si, (a, l, e, c) in enumerate(self.shells):
            sh_atom.append(a)
            sh_l.append(l)
            sh_n.append(len(e))
            sh_off.append(len(exps))
            exps += list(e)
            coefs += list(c)
            if l == 0:
                ao_shell.append(si)
                ao_lmn.append(0)
            else:
                for lmn in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    ao_shell.append(si)
                    ao_lmn.append(lmn[0] | (lmn[1] << 2) | (lmn[2] << 4))
        f = lambda x: engine.upload(np.asarray(x, dtype=np.float64).reshape(-1), np.float64, dev)
        i = lambda x: engine.upload(np.asarray(x, dtype=np.int32), np.int32, dev)
        keep = [f(self.coords), f(self.charges), i(sh_atom), i(sh_l), i(sh_n), i(sh_off), f(exps), f(coefs),
                i(ao_shell), i(ao_lmn)]
        b = _lib.Qex3dBasis(self.natm, len(self.shells), self.nao, *[ctypes.c_void_p(t.data_ptr()) for t in keep],
                            max(sh_n))
        self._dev = (b, keep)
        return b

    # ---- integrals on the device ----
    def intor_1e(self):
        dev = engine.device()
        n = self.nao
        S, T, V = (torch.empty((n, n), dtype=torch.float64, device=dev) for _ in range(3))
        lib = _lib.load()
        _lib.check3(lib.qex3d_int1e(ctypes.byref(self.device_basis()), engine._ptr(S), engine._ptr(T), engine._ptr(V),
                                    engine._stream()), "qex3d_int1e")
        return S, T, V

    def intor_eri_pairs(self):
        dev = engine.device()
        npair = self.nao * (self.nao + 1) // 2
        g = torch.empty((npair, npair), dtype=torch.float64, device=dev)
        lib = _lib.load()
        b = self.device_basis()
        ws = engine.workspace(lib.qex3d_eri_workspace_bytes(ctypes.byref(b)), dev)
        _lib.check3(lib.qex3d_eri(ctypes.byref(b), engine._ptr(g), engine._ptr(ws), ws.numel(), engine._stream()),
                    "qex3d_eri")
        return g


def intor_eri_pairs_batch(mols):
    """The ERI pair matrices of several molecules in one launch sequence (``qex3d_eri_batch``), each
    equal to its ``mol.intor_eri_pairs()`` bit for bit."""
    mols = list(mols)
    if not mols:
        return []
    dev = engine.device()
    lib = _lib.load()
    B = len(mols)
    bases = (_lib.Qex3dBasis * B)(*[m.device_basis() for m in mols])
    outs = [torch.empty((m.nao * (m.nao + 1) // 2,) * 2, dtype=torch.float64, device=dev) for m in mols]
    ptrs = (ctypes.c_void_p * B)(*[o.data_ptr() for o in outs])
    nbytes = lib.qex3d_eri_batch_workspace_bytes(B, bases)
    if nbytes == 0:
        raise ValueError("invalid basis in the ERI batch")
    ws = engine.workspace(nbytes, dev)
    _lib.check3(lib.qex3d_eri_batch(B, bases, ptrs, engine._ptr(ws), ws.numel(), engine._stream()), "qex3d_eri_batch")
    return outs


def synthetic_chain(n_atoms, seed):
    """Hydrogen chain along z with bonds U[0.5, 1.5] Angstrom and, per atom, 2 contracted s shells and
    1 p shell of 3 primitives (exponents log-uniform in [0.1, 20], random positive coefficients),
    all from the seed: 5 AOs per atom."""
    rng = np.random.default_rng(seed)
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n_atoms - 1))]) / BOHR
    coords = np.stack([np.zeros(n_atoms), np.zeros(n_atoms), z - z.mean()], axis=1)
    shells = []
    for _ in range(n_atoms):
        per = []
        for l in (0, 0, 1):
            exps = np.sort(np.exp(rng.uniform(np.log(0.1), np.log(20.0), 3)))[::-1].copy()
            coefs = rng.uniform(0.2, 1.0, 3)
            per.append((l, exps, coefs))
        shells.append(per)
    return Mole([("H", tuple(c)) for c in coords], unit="Bohr", shells_per_atom=shells)
