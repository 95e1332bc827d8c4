"""Neural XC for the 3D path: the MLP of ``qedft/train/td/trainer_legacy_no_jit.py:96-107``
(Dense -> gelu -> Dense -> gelu -> Dense(1) -> -scale * swish), applied either

* to the whole grid density (``GlobalNNXC``, xctype "NN-AmplitudeEncoding"), as
  ``exc_and_vrho_global`` (:46-53) does: the device evaluates it inside ``nr_rks`` / the SCF
  (``qex3d_nr_rks_nn``, method 2 of ``qex3d_scf``); or
* to every grid point's density (``LocalNNXC``, xctype "NN"), as ``exc_and_vrho_local`` (:56-63):
  ``qex3d_nr_rks_nn_local``, method 4 of ``qex3d_scf``.

``eval_xc`` has the reference's signature (:76-93); with ``functools.partial(eval_xc, network=...,
is_global_xc=...)`` as the ``description`` of ``define_xc_`` the network runs on the device."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from ..od import engine


class GlobalNNXC:
    """weights: [W1 (n_grid, h1), W2 (h1, h2), W3 (h2, 1)] (flax Dense kernels), biases alike."""

    xctype = "NN-AmplitudeEncoding"

    def __init__(self, weights, biases, scale=1e-2):
        if len(weights) != 3 or len(biases) != 3:
            raise ValueError("the device network has two hidden layers: weights [W1, W2, W3]")
        self.weights = [np.asarray(w, dtype=np.float64) for w in weights]
        self.biases = [np.asarray(b, dtype=np.float64).reshape(-1) for b in biases]
        self.scale = float(scale)
        self.h1 = self.weights[0].shape[1]
        self.h2 = self.weights[1].shape[1]
        if self.weights[1].shape[0] != self.h1 or self.weights[2].shape != (self.h2, 1):
            raise ValueError("inconsistent layer shapes")
        self._packed = None

    @property
    def n_grid(self):
        return self.weights[0].shape[0]

    def packed(self) -> torch.Tensor:
        """W1, b1, W2, b2, W3, b3 flattened into one device buffer (cached)."""
        if self._packed is None:
            parts = []
            for W, b in zip(self.weights, self.biases):
                parts += [W.reshape(-1), b]
            self._packed = engine.as_f64(np.concatenate(parts), engine.device())
        return self._packed

    def with_params(self, params):
        """A copy carrying new parameters: [(W, b), ...] or {'weights': [...], 'biases': [...]}."""
        if isinstance(params, dict):
            return GlobalNNXC(params["weights"], params["biases"], self.scale)
        W, b = zip(*params)
        return GlobalNNXC(list(W), list(b), self.scale)


class LocalNNXC(GlobalNNXC):
    """Per-point network: weights [W1 (1, h1), W2 (h1, h2), W3 (h2, 1)] (flax Dense kernels), biases
    alike; exc_g = net(rho_g), vrho_g = d exc_g / d rho_g.  Hidden widths <= 512."""

    xctype = "NN"

    def __init__(self, weights, biases, scale=1e-2):
        super().__init__(weights, biases, scale)
        if self.weights[0].shape[0] != 1:
            raise ValueError("the per-point network takes one input (the density at the point)")
        if max(self.h1, self.h2) > _lib.QEX3D_MAX_NN_LOCAL_HIDDEN:
            raise ValueError(f"hidden widths must be <= {_lib.QEX3D_MAX_NN_LOCAL_HIDDEN}")

    @property
    def n_grid(self):
        return None

    def with_params(self, params):
        if isinstance(params, dict):
            return LocalNNXC(params["weights"], params["biases"], self.scale)
        W, b = zip(*params)
        return LocalNNXC(list(W), list(b), self.scale)


def eval_xc(xc_code, rho, spin=0, relativity=0, deriv=2, verbose=None, params=None, network=None, is_global_xc=True):
    """``trainer_legacy_no_jit.py:76-93``: (exc, (vrho, None, None, None), None, None) of ``network`` on
    ``rho`` on the device; global: exc is the network's scalar, local: per point."""
    net = network if params is None else network.with_params(params)
    if not isinstance(net, GlobalNNXC) or bool(is_global_xc) == isinstance(net, LocalNNXC):
        raise ValueError("is_global_xc does not match the network kind")
    dev = engine.device()
    r = engine.as_f64(rho, dev).reshape(-1)
    ng = r.shape[0]
    lib = _lib.load()
    glob = 1 if is_global_xc else 0
    if glob and net.n_grid != ng:
        raise ValueError(f"network input width {net.n_grid} != grid size {ng}")
    ws = engine.workspace(lib.qex3d_eval_xc_nn_workspace_bytes(ng, net.h1, net.h2, glob), dev)
    exc = torch.empty(1 if glob else ng, dtype=torch.float64, device=dev)
    vrho = torch.empty(ng, dtype=torch.float64, device=dev)
    p = engine._ptr
    _lib.check3(lib.qex3d_eval_xc_nn(glob, ng, p(r), p(net.packed()), net.h1, net.h2, net.scale, p(exc), p(vrho), p(ws),
                                     ws.numel(), engine._stream()), "qex3d_eval_xc_nn")
    return (exc[0] if glob else exc), (vrho, None, None, None), None, None


def network_of(description, xctype):
    """The device network of a ``define_xc_`` description: a network object, or
    ``functools.partial(eval_xc, network=..., is_global_xc=...)`` as the reference builds it
    (``trainer_legacy_no_jit.py:153-165``)."""
    import functools

    net = description
    if isinstance(description, functools.partial):
        kw = description.keywords or {}
        net = kw.get("network")
        if "is_global_xc" in kw and bool(kw["is_global_xc"]) != (xctype == "NN-AmplitudeEncoding"):
            raise ValueError("is_global_xc does not match xctype")
    if xctype == "NN-AmplitudeEncoding" and isinstance(net, GlobalNNXC) and not isinstance(net, LocalNNXC):
        return net
    if xctype == "NN" and isinstance(net, LocalNNXC):
        return net
    raise NotImplementedError(f"xctype {xctype!r} with {type(net).__name__}: the device evaluates GlobalNNXC "
                              "(NN-AmplitudeEncoding) and LocalNNXC (NN) networks")


def init_local_mlp(seed, hidden=(512, 512)):
    """Seeded per-point network: kernels N(0, 1/fan_in), zero biases."""
    rng = np.random.default_rng(seed)
    sizes = [1, *hidden, 1]
    W = [rng.standard_normal((a, b)) / math.sqrt(a) for a, b in zip(sizes[:-1], sizes[1:])]
    return W, [np.zeros(b) for b in sizes[1:]]


def init_global_mlp(seed, n_in, hidden=(512, 512)):
    """Seeded weights: kernels N(0, 1/fan_in) (the variance of flax's lecun_normal), zero biases."""
    rng = np.random.default_rng(seed)
    sizes = [n_in, *hidden, 1]
    W = [rng.standard_normal((a, b)) / math.sqrt(a) for a, b in zip(sizes[:-1], sizes[1:])]
    return W, [np.zeros(b) for b in sizes[1:]]
