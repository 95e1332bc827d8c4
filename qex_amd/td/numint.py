"""``eval_ao`` / ``eval_rho`` / ``nr_rks`` with the reference's call surface
(``qedft/train/td/numint.py:16-273``, ``numint_legacy.py:351-397``), computed on the device."""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..od import engine


def eval_ao(mol, coords, deriv=0):
    if deriv != 0:
        raise NotImplementedError("only AO values (deriv=0) are needed by LDA")
    dev = engine.device()
    coords = engine.as_f64(coords, dev)
    ng = coords.shape[0]
    ao = torch.empty((ng, mol.nao), dtype=torch.float64, device=dev)
    _lib.check3(_lib.load().qex3d_eval_ao(ctypes.byref(mol.device_basis()), ng, engine._ptr(coords), engine._ptr(ao),
                                          engine._stream()), "qex3d_eval_ao")
    return ao


def eval_ao_f32(mol, coords):
    """The AO values rounded to FP32 in the layout of the FP32 grid SCF path: [rows, ld] with ld = nao
    rounded up to 16 and rows = ngrid rounded up to 64, padding zero."""
    dev = engine.device()
    coords = engine.as_f64(coords, dev)
    ng = coords.shape[0]
    ao = torch.empty((_lib.ao32_rows(ng), _lib.ao32_ld(mol.nao)), dtype=torch.float32, device=dev)
    _lib.check3(_lib.load().qex3d_eval_ao_f32(ctypes.byref(mol.device_basis()), ng, engine._ptr(coords),
                                              engine._ptr(ao), engine._stream()), "qex3d_eval_ao_f32")
    return ao


def eval_rho(mol, ao, dm, non0tab=None, xctype="LDA", hermi=0, verbose=None):
    """rho_g = sum_{mu nu} ao[g, mu] D[mu, nu] ao[g, nu]; D is symmetrised when hermi=0."""
    dev = engine.device()
    dm = engine.as_f64(dm, dev)
    if hermi == 0:
        dm = (0.5 * (dm + dm.T)).contiguous()
    ng, nao = ao.shape
    rho = torch.empty(ng, dtype=torch.float64, device=dev)
    _lib.check3(_lib.load().qex3d_eval_rho(ng, nao, engine._ptr(ao), engine._ptr(dm), engine._ptr(rho),
                                           engine._stream()), "qex3d_eval_rho")
    return rho


# pyscf xc codes of the local density approximations the device evaluates (libxc names): "lda"
# alone is Slater exchange (pyscf parses a bare exchange name as exchange only); "lda,vwn" and its
# aliases add VWN5 correlation.  pyscf's RKS default is "LDA,VWN" (td/rks.py:172).
_LDA_CODES = {
    "lda": _lib.QEX3D_LDA_SLATER, "lda,": _lib.QEX3D_LDA_SLATER, "slater": _lib.QEX3D_LDA_SLATER,
    "lda_x": _lib.QEX3D_LDA_SLATER, "lda,vwn": _lib.QEX3D_LDA_SVWN5, "lda,vwn5": _lib.QEX3D_LDA_SVWN5,
    "svwn": _lib.QEX3D_LDA_SVWN5, "svwn5": _lib.QEX3D_LDA_SVWN5, "slater,vwn": _lib.QEX3D_LDA_SVWN5,
    "slater,vwn5": _lib.QEX3D_LDA_SVWN5, "lda_x,lda_c_vwn": _lib.QEX3D_LDA_SVWN5,
}


def lda_functional(xc_code) -> int:
    """The device LDA functional (``QEX3D_LDA_*``) for a pyscf xc code; raises for any other code."""
    key = str(xc_code).lower().replace(" ", "")
    if key not in _LDA_CODES:
        raise NotImplementedError(f"xc_code {xc_code!r}: the device path implements {sorted(_LDA_CODES)} "
                                  "and network XC installed with define_xc_")
    return _LDA_CODES[key]


class NumInt:
    """The reference's ``ni`` dispatch handle; ``define_xc_`` installs a network XC
    (``trainer_legacy_no_jit.py:256-261``)."""

    def __init__(self):
        self.network = None

    def define_xc_(self, description=None, xctype="NN-AmplitudeEncoding", network=None, **kwargs):
        """``description``: a network (``GlobalNNXC`` / ``LocalNNXC``) or, as the reference passes it,
        ``functools.partial(qex_amd.td.xc.eval_xc, network=..., is_global_xc=...)``."""
        from .xc import network_of

        self.network = network_of(description if description is not None else network, xctype)
        return self


def nr_rks(ni, mol, grids, xc_code, dms, relativity=0, hermi=0, max_memory=2000, verbose=None, params=None,
           ao=None):
    """(nelec, excsum, vmat) on the device: the LDA named by ``xc_code`` (Slater exchange for "lda",
    Slater + VWN5 for "lda,vwn" and its aliases), or the network installed with
    ``ni.define_xc_`` (NN-AmplitudeEncoding branch, td/numint.py:206-231), whose parameters may be
    replaced by ``params``."""
    from .xc import LocalNNXC

    net = getattr(ni, "network", None)
    if net is not None:
        net = net if params is None else net.with_params(params)
        if isinstance(net, LocalNNXC):
            return _nr_rks_nn_local(net, mol, grids, dms, hermi, ao)
        return _nr_rks_nn(net, mol, grids, dms, hermi, ao)
    functional = lda_functional(xc_code)
    dev = engine.device()
    if ao is None:
        ao = eval_ao(mol, grids.coords)
    dm = engine.as_f64(dms, dev)
    if hermi == 0:
        dm = (0.5 * (dm + dm.T)).contiguous()
    ng, nao = ao.shape
    lib = _lib.load()
    ws = engine.workspace(lib.qex3d_nr_rks_workspace_bytes(ng, nao), dev)
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    vmat = torch.empty((nao, nao), dtype=torch.float64, device=dev)
    _lib.check3(lib.qex3d_nr_rks_lda(functional, ng, nao, engine._ptr(ao),
                                     engine._ptr(engine.as_f64(grids.weights, dev)), engine._ptr(dm),
                                     engine._ptr(out2), engine._ptr(vmat), engine._ptr(ws), ws.numel(),
                                     engine._stream()), "qex3d_nr_rks_lda")
    return out2[0], out2[1], vmat


def _nr_rks_nn_local(net, mol, grids, dms, hermi, ao):
    """The "NN" branch (td/numint.py:179-204): per-point exc and vrho = d exc / d rho on the device."""
    dev = engine.device()
    if ao is None:
        ao = eval_ao(mol, grids.coords)
    dm = engine.as_f64(dms, dev)
    if hermi == 0:
        dm = (0.5 * (dm + dm.T)).contiguous()
    ng, nao = ao.shape
    lib = _lib.load()
    ws = engine.workspace(lib.qex3d_nr_rks_workspace_bytes(ng, nao), dev)
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    vmat = torch.empty((nao, nao), dtype=torch.float64, device=dev)
    p = engine._ptr
    _lib.check3(lib.qex3d_nr_rks_nn_local(ng, nao, p(ao), p(engine.as_f64(grids.weights, dev)), p(dm), p(net.packed()),
                                          net.h1, net.h2, net.scale, p(out2), p(vmat), p(ws), ws.numel(),
                                          engine._stream()), "qex3d_nr_rks_nn_local")
    return out2[0], out2[1], vmat


def _nr_rks_nn(net, mol, grids, dms, hermi, ao):
    dev = engine.device()
    if ao is None:
        ao = eval_ao(mol, grids.coords)
    dm = engine.as_f64(dms, dev)
    if hermi == 0:
        dm = (0.5 * (dm + dm.T)).contiguous()
    ng, nao = ao.shape
    if net.n_grid != ng:
        raise ValueError(f"network input width {net.n_grid} != grid size {ng}")
    lib = _lib.load()
    ws = engine.workspace(lib.qex3d_nr_rks_nn_workspace_bytes(ng, nao, net.h1, net.h2), dev)
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    vmat = torch.empty((nao, nao), dtype=torch.float64, device=dev)
    p = engine._ptr
    _lib.check3(lib.qex3d_nr_rks_nn(ng, nao, p(ao), p(engine.as_f64(grids.weights, dev)), p(dm), p(net.packed()),
                                    net.h1, net.h2, net.scale, p(out2), p(vmat), p(ws), ws.numel(), engine._stream()),
                "qex3d_nr_rks_nn")
    return out2[0], out2[1], vmat
