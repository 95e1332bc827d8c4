"""Closed-shell SCF on the device with the reference's ``RKS(mol).kernel()`` surface
(``qedft/train/td/hf_legacy.py:119-272``, ``rks_legacy.py``): core-Hamiltonian guess, CDIIS from
cycle 1 (space 8), generalized eigh, aufbau occupation, pyscf convergence criterion."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from ..od import engine
from . import gto
from .dft import Grids
from .numint import eval_ao, eval_ao_f32, lda_functional


class _SCF:
    method = _lib.QEX3D_RHF

    def __init__(self, mol):
        self.mol = mol
        self.conv_tol = 1e-9
        self.max_cycle = 50
        self.diis_space = 8
        self.sync_every = 4  # host checks the device's convergence flag every 4 cycles (0: never)
        # grid work precision: "fp64", or "fp32_grid" (AO values stored in FP32, rho from them with FP64
        # sums, the Vxc contraction on FP32 matrix cores with FP64 sums across 16-point units;
        # everything else FP64)
        self.precision = "fp64"
        self.grids = None
        self.e_tot = None
        self.converged = False
        self.cycles = 0
        self.energies = None
        self.mo_energy = self.mo_coeff = self.dm = None
        self._ints = None

    def integrals(self):
        if self._ints is None:
            S, T, V = self.mol.intor_1e()
            self._ints = (S, (T + V).contiguous(), self.mol.intor_eri_pairs())
        return self._ints

    def kernel(self, dm0=None, **kwargs):
        """Runs the SCF on the device and returns e_tot (waits for the result)."""
        self.kernel_async(**kwargs)
        return self.collect()

    def kernel_async(self, **kwargs):
        """Enqueues the whole solve (integrals if not cached, grid, AO values, SCF cycles) on the
        current stream and returns at once; ``collect()`` reads the result.  No host
        synchronisation is made unless ``sync_every`` > 0."""
        p, ins, outs = self._prepare(**kwargs)
        lib = _lib.load()
        ws = engine.workspace(lib.qex3d_scf_workspace_bytes(ctypes.byref(p)), engine.device())
        _lib.check3(lib.qex3d_scf(ctypes.byref(p), ctypes.byref(ins), ctypes.byref(outs), engine._ptr(ws), ws.numel(),
                                  engine._stream()), "qex3d_scf")
        self._pending = (self._status, ws)
        return self

    def _prepare(self, **kwargs):
        """Integrals, grid, AO values and output tensors of this solve; the C structs that describe it."""
        mol = self.mol
        dev = engine.device()
        S, H, eri = self.integrals()
        ao = w = None
        ng = 0
        method = self.method
        net = getattr(self, "network", None)
        if method in (_lib.QEX3D_RKS_LDA, _lib.QEX3D_RKS_LDA_X) and net is not None:
            from .xc import LocalNNXC

            method = _lib.QEX3D_RKS_NN_LOCAL if isinstance(net, LocalNNXC) else _lib.QEX3D_RKS_NN_GLOBAL
            if kwargs.get("params") is not None:
                net = net.with_params(kwargs["params"])
        if self.precision not in ("fp64", "fp32_grid"):
            raise ValueError(f"precision must be 'fp64' or 'fp32_grid', got {self.precision!r}")
        f32 = self.precision == "fp32_grid" and method != _lib.QEX3D_RHF
        ao32 = None
        if method != _lib.QEX3D_RHF:
            if self.grids is None:
                self.grids = Grids(mol).build()
            if f32:
                if getattr(self, "_ao32", None) is None:
                    self._ao32 = eval_ao_f32(mol, self.grids.coords)
                ao32 = self._ao32
            else:
                if getattr(self, "_ao", None) is None:
                    self._ao = eval_ao(mol, self.grids.coords)
                ao = self._ao
            w, ng = self.grids.weights, self.grids.size
        n = mol.nao
        p = _lib.Qex3dScfProblem(n, mol.nelectron, method, self.max_cycle, self.diis_space, ng,
                                 float(self.conv_tol), float(mol.energy_nuc()), int(self.sync_every))
        p.precision = _lib.QEX3D_PRECISION_FP32_GRID if f32 else _lib.QEX3D_PRECISION_FP64
        if method == _lib.QEX3D_RKS_NN_GLOBAL:
            if net.n_grid != ng:
                raise ValueError(f"network input width {net.n_grid} != grid size {ng}")
            p.nn_h1, p.nn_h2, p.nn_scale = net.h1, net.h2, net.scale
        if method == _lib.QEX3D_RKS_NN_LOCAL:
            p.nn_h1, p.nn_h2, p.nn_scale = net.h1, net.h2, net.scale
        e_hist = torch.empty(self.max_cycle + 1, dtype=torch.float64, device=dev)
        mo = torch.empty(n, dtype=torch.float64, device=dev)
        C = torch.empty((n, n), dtype=torch.float64, device=dev)
        D = torch.empty((n, n), dtype=torch.float64, device=dev)
        status = torch.zeros(8, dtype=torch.int32, device=dev)
        ptr = engine._ptr
        packed = net.packed() if method in (_lib.QEX3D_RKS_NN_GLOBAL, _lib.QEX3D_RKS_NN_LOCAL) else None
        ins = _lib.Qex3dScfInputs(ptr(S), ptr(H), ptr(eri), ptr(ao), ptr(w), ptr(packed), ptr(ao32))
        outs = _lib.Qex3dScfOutputs(ptr(e_hist), ptr(mo), ptr(C), ptr(D), ptr(status))
        self._keep = packed
        self._status = status
        self.energies = e_hist
        self.mo_energy, self.mo_coeff, self.dm = mo, C, D
        return p, ins, outs

    def collect(self):
        status, _ = self._pending
        st = status.cpu().numpy()
        self.cycles = int(st[1])
        self.converged = bool(st[4])
        self.e_tot = float(self.energies[-1].item())
        self._pending = None
        return self.e_tot

    def make_rdm1(self, mo_coeff=None, mo_occ=None, **kwargs):
        """D = C_occ diag(occ) C_occ^T (``td/hf.py:303-308``); without arguments the converged density
        matrix.  Extra keywords (``a0_repr=True`` at ``trainer_legacy_no_jit.py:272``) are accepted."""
        if mo_coeff is None:
            return self.dm
        C = engine.as_f64(mo_coeff, engine.device())
        occ = engine.as_f64(mo_occ, engine.device())
        return (C * occ[None, :]) @ C.T


class RHF(_SCF):
    method = _lib.QEX3D_RHF


class RKS(_SCF):
    """Restricted Kohn-Sham.  ``xc`` is a pyscf xc code of an LDA (default "LDA,VWN", pyscf's and the
    reference's default, ``td/rks.py:172``; "lda" alone is Slater exchange), or a network installed
    with ``define_xc_``."""

    def __init__(self, mol, xc="LDA,VWN"):
        super().__init__(mol)
        self.network = None
        self.xc = xc

    @property
    def xc(self):
        return self._xc

    @xc.setter
    def xc(self, code):
        fn = lda_functional(code)
        self._xc = code
        self.method = _lib.QEX3D_RKS_LDA_X if fn == _lib.QEX3D_LDA_SLATER else _lib.QEX3D_RKS_LDA

    def define_xc_(self, description=None, xctype="NN-AmplitudeEncoding", network=None, **kwargs):
        """Install a network XC (``trainer_legacy_no_jit.py:256-261``: ``mf.define_xc_(description=
        eval_xc_func, xctype="NN-AmplitudeEncoding" | "NN")``); kernel(params=...) then evaluates it
        with those parameters.  ``description`` is a network (``GlobalNNXC`` / ``LocalNNXC``) or
        ``functools.partial(qex_amd.td.xc.eval_xc, network=..., is_global_xc=...)``."""
        from .xc import network_of

        self.network = network_of(description if description is not None else network, xctype)
        return self


def kernel_batch_async(mfs, **kwargs):
    """Enqueues the SCF of several molecules (any sizes and methods) as one batched launch sequence
    (the padded batch of ``scf_functions_masked.py:917-967``, without padding: each molecule keeps
    its own basis and grid, so each result equals its solo ``kernel()`` bit for bit).  ``collect()``
    on each object reads its result."""
    mfs = list(mfs)
    if not mfs:
        return mfs
    if len(mfs) > _lib.QEX3D_MAX_SCF_BATCH:
        raise ValueError(f"at most {_lib.QEX3D_MAX_SCF_BATCH} molecules per batch")
    # the two-electron integrals of every molecule without them, in one launch sequence
    todo = [mf for mf in mfs if mf._ints is None and hasattr(mf.mol, "device_basis")]
    for mf, eri in zip(todo, gto.intor_eri_pairs_batch([mf.mol for mf in todo])):
        S, T, V = mf.mol.intor_1e()
        mf._ints = (S, (T + V).contiguous(), eri)
    preps = [mf._prepare(**kwargs) for mf in mfs]
    B = len(mfs)
    P = (_lib.Qex3dScfProblem * B)(*[x[0] for x in preps])
    I = (_lib.Qex3dScfInputs * B)(*[x[1] for x in preps])
    O = (_lib.Qex3dScfOutputs * B)(*[x[2] for x in preps])
    lib = _lib.load()
    nbytes = lib.qex3d_scf_batch_workspace_bytes(B, P)
    if nbytes == 0:
        raise ValueError("invalid problem in the batch")
    ws = engine.workspace(nbytes, engine.device())
    _lib.check3(lib.qex3d_scf_batch(B, P, I, O, engine._ptr(ws), ws.numel(), engine._stream()), "qex3d_scf_batch")
    for mf in mfs:
        mf._pending = (mf._status, ws)
    return mfs


def kernel_batch(mfs, **kwargs):
    """Runs the batched SCF and returns the total energies (waits for the results)."""
    return [mf.collect() for mf in kernel_batch_async(mfs, **kwargs)]
