"""Generalized symmetric eigensolver on the device (``generalized_eigh`` of
``qedft/train/td/generalized_eigensolver.py:264-330``): batched A v = w B v for n <= 96; and the
degenerate-safe eigendecomposition with its reverse pass (``degen_eigh``, ``:143-219``)."""

from __future__ import annotations

import torch

from .. import _lib
from ..od import engine


def _device_gen_eigh(Af, Bf):
    batch, n = Af.shape[0], Af.shape[-1]
    dev = Af.device
    lib = _lib.load()
    ws = engine.workspace(lib.qex3d_gen_eigh_workspace_bytes(batch, n), dev)
    w = torch.empty((batch, n), dtype=torch.float64, device=dev)
    v = torch.empty((batch, n, n), dtype=torch.float64, device=dev)
    p = engine._ptr
    _lib.check3(lib.qex3d_gen_eigh(batch, n, p(Af), p(Bf), p(w), p(v), p(ws), ws.numel(), engine._stream()),
                "qex3d_gen_eigh")
    return w, v


def generalized_eigh(A, B, *, eps=1.0e-12, scale=False, dtype=torch.float64):
    """A, B: [..., n, n] real symmetric, B (numerically) positive definite.  Returns (w [..., n]
    ascending, v [..., n, n] with eigenvectors in the columns, v^T B v = 1), as the reference does:

    * A and B are symmetrised;
    * scale=True applies the Jacobi scaling s = sqrt(diag B) to both (the eigenvectors are then those
      of the scaled problem, as in the reference);
    * B is shifted by (eps - lambda_min(B)) I when its smallest eigenvalue is below eps (lambda_min
      from the device solve of B v = lambda v);
    * Cholesky B = L L^T, the symmetric eigenproblem of L^-1 A L^-T, and v = L^-T U, in one workgroup
      per matrix on the device."""
    if dtype != torch.float64:
        raise ValueError("the device eigensolver works in float64")
    dev = engine.device()
    A = engine.as_f64(A, dev)
    B = engine.as_f64(B, dev)
    if A.shape != B.shape or A.dim() < 2 or A.shape[-1] != A.shape[-2]:
        raise ValueError("A and B must be [..., n, n] of the same shape")
    n = A.shape[-1]
    lead = A.shape[:-2]
    Af = A.reshape(-1, n, n)
    Bf = B.reshape(-1, n, n)
    Af = (0.5 * (Af + Af.transpose(1, 2))).contiguous()
    Bf = (0.5 * (Bf + Bf.transpose(1, 2))).contiguous()
    if scale:
        si = 1.0 / torch.sqrt(torch.diagonal(Bf, dim1=1, dim2=2))
        Af = (si[:, :, None] * Af * si[:, None, :]).contiguous()
        Bf = (si[:, :, None] * Bf * si[:, None, :]).contiguous()
    eye = torch.eye(n, dtype=torch.float64, device=dev).expand(Bf.shape[0], n, n).contiguous()
    lam_min = _device_gen_eigh(Bf, eye)[0][:, 0]
    shift = torch.where(lam_min < eps, eps - lam_min, torch.zeros_like(lam_min))
    Bf = (Bf + shift[:, None, None] * eye).contiguous()
    w, v = _device_gen_eigh(Af, Bf)
    return w.reshape(*lead, n), v.reshape(*lead, n, n)


def _device_eigh_vjp(w, v, gw, gv, exact):
    batch, n = v.shape[0], v.shape[-1]
    dev = v.device
    lib = _lib.load()
    ws = engine.workspace(lib.qex3d_eigh_vjp_workspace_bytes(batch, n), dev)
    gA = torch.empty((batch, n, n), dtype=torch.float64, device=dev)
    p = engine._ptr
    _lib.check3(lib.qex3d_eigh_vjp(batch, n, p(w), p(v), p(gw) if gw is not None else None,
                                   p(gv) if gv is not None else None, int(bool(exact)), p(gA), p(ws), ws.numel(),
                                   engine._stream()), "qex3d_eigh_vjp")
    return gA


def degen_eigh_fwd(A):
    """Forward of ``degen_eigh`` (``generalized_eigensolver.py:150-154``): ((w, v), residuals (w, v)),
    w ascending, eigenvectors in the columns of v, on the device."""
    dev = engine.device()
    A = engine.as_f64(A, dev)
    n = A.shape[-1]
    lead = A.shape[:-2]
    Af = A.reshape(-1, n, n).contiguous()
    eye = torch.eye(n, dtype=torch.float64, device=dev).expand(Af.shape[0], n, n).contiguous()
    w, v = _device_gen_eigh(Af, eye)
    w, v = w.reshape(*lead, n), v.reshape(*lead, n, n)
    return (w, v), (w, v)


def degen_eigh_bwd(res, grads, exact=False):
    """Reverse pass of ``degen_eigh`` (``generalized_eigensolver.py:158-219``): ``res`` = (eival,
    eivec), ``grads`` = (grad_eival, grad_eivec), either of them None; returns (grad_A,).  Degenerate
    pairs (|w_j - w_i| < eps^0.6) contribute nothing through the eigenvectors.  By default the
    reference's code as written: its ``eivect = jnp.transpose(eivec, axes=(-2, -1))`` leaves a 2-D
    array unchanged, so the formula uses V where the derivative of the decomposition has V^T;
    ``exact=True`` gives the latter.  Batched over leading dimensions."""
    w, v = res
    gw, gv = grads
    dev = engine.device()
    v = engine.as_f64(v, dev)
    n = v.shape[-1]
    lead = v.shape[:-2]
    f = lambda x, shp: None if x is None else engine.as_f64(x, dev).reshape(shp).contiguous()
    B = v.reshape(-1, n, n).shape[0]
    gA = _device_eigh_vjp(f(w, (B, n)), f(v, (B, n, n)), f(gw, (B, n)), f(gv, (B, n, n)), exact)
    return (gA.reshape(*lead, n, n),)


class _DegenEigh(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, exact):
        (w, v), res = degen_eigh_fwd(A)
        ctx.save_for_backward(*res)
        ctx.exact = exact
        return w, v

    @staticmethod
    def backward(ctx, gw, gv):
        w, v = ctx.saved_tensors
        return degen_eigh_bwd((w, v), (gw, gv), exact=ctx.exact)[0], None


def degen_eigh(A, exact=False):
    """Symmetric eigendecomposition with the degenerate-safe reverse pass of
    ``generalized_eigensolver.py:143-219`` (the JAX ``custom_vjp`` becomes a ``torch.autograd``
    function): returns (eival, eivec); gradients flow to A through ``degen_eigh_bwd`` on the device."""
    return _DegenEigh.apply(A, exact)
