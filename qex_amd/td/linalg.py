"""Generalized symmetric eigensolver on the device (``generalized_eigh`` of
``qedft/train/td/generalized_eigensolver.py:264-330``): batched A v = w B v for n <= 96."""

from __future__ import annotations

import torch

from .. import _lib
from ..od import engine


def generalized_eigh(A, B):
    """A, B: [..., n, n] (B symmetric positive definite).  Returns (w [..., n] ascending,
    v [..., n, n] with eigenvectors in the columns, v^T B v = 1).  A is symmetrised; the
    reference's diagonal shift for a numerically singular B is not applied (B must be SPD)."""
    dev = engine.device()
    A = engine.as_f64(A, dev)
    B = engine.as_f64(B, dev)
    if A.shape != B.shape or A.dim() < 2 or A.shape[-1] != A.shape[-2]:
        raise ValueError("A and B must be [..., n, n] of the same shape")
    n = A.shape[-1]
    lead = A.shape[:-2]
    Af = A.reshape(-1, n, n).contiguous()
    Bf = B.reshape(-1, n, n).contiguous()
    batch = Af.shape[0]
    lib = _lib.load()
    ws = engine.workspace(lib.qex3d_gen_eigh_workspace_bytes(batch, n), dev)
    w = torch.empty((batch, n), dtype=torch.float64, device=dev)
    v = torch.empty((batch, n, n), dtype=torch.float64, device=dev)
    p = engine._ptr
    _lib.check3(lib.qex3d_gen_eigh(batch, n, p(Af), p(Bf), p(w), p(v), p(ws), ws.numel(), engine._stream()),
                "qex3d_gen_eigh")
    return w.reshape(*lead, n), v.reshape(*lead, n, n)
