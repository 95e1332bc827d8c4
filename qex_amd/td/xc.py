"""Global (amplitude-encoded) neural XC for the 3D path: the MLP of
``qedft/train/td/trainer_legacy_no_jit.py:96-107`` (Dense -> gelu -> ... -> Dense(1) ->
-scale * swish) applied to the whole grid density, as ``exc_and_vrho_global`` (:46-53) does.
The device evaluates it inside ``nr_rks`` / the SCF (``qex3d_nr_rks_nn``, method 2 of ``qex3d_scf``)."""

from __future__ import annotations

import math

import numpy as np
import torch

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


def init_global_mlp(seed, n_in, hidden=(512, 512)):
    """Seeded weights: kernels N(0, 1/fan_in) (the variance of flax's lecun_normal), zero biases."""
    rng = np.random.default_rng(seed)
    sizes = [n_in, *hidden, 1]
    W = [rng.standard_normal((a, b)) / math.sqrt(a) for a, b in zip(sizes[:-1], sizes[1:])]
    return W, [np.zeros(b) for b in sizes[1:]]
