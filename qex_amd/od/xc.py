"""Exchange-correlation functionals the 1D engine evaluates on the device.

Each functional is a callable ``fn(density, params=None)`` (the ``xc_energy_density_fn``
convention of ``qedft/models/wrappers.py:85-113`` and ``jax_dft.xc``) and also describes
itself to the kernels (``xc_kind`` and device parameters).  Calling it evaluates the energy
density on the device (``qex1d_xc_evaluate``); ``energy_density`` is the same functional written
in torch, for host-side checks.  Inside the SCF loop only the HIP kernels run.  Any other
callable runs through ``ExternalXC``: the SCF then goes one device pass at a time and evaluates
the callable (torch, with autograd for the potential) between passes.

* ``LDAExp``: jax_dft ``xc.get_lda_xc_energy_density_fn()`` for the exponential
  interaction (``od/eval.py:78-81``).
* ``PowerLaw``: c (n + eps)^p, local (``tests/test_scf_od.py:127-129``) or summed to a scalar
  (global, ``tests/test_scf_od.py:46-48``).
* ``LocalMLPXC``: a per-point MLP on a (2h+1)-point density window (window 1 is
  ``build_local_mlp``, ``classical_models.py:123-174``).
"""

from __future__ import annotations

import functools
import math

import numpy as np
import torch

from .. import _lib

def _nd(x) -> int:
    """Array rank of a parameter leaf; -1 for containers (lists, tuples, dicts)."""
    return -1 if isinstance(x, (list, tuple, dict)) else np.ndim(x)


A_EXP = 1.071295
KAPPA_EXP = 1.0 / 2.385345


class XCFunctional:
    encoding = "local"
    xc_kind = _lib.QEX1D_XC_NONE

    def configure(self, problem) -> None:
        problem.xc_kind = self.xc_kind

    def device_params(self, device):
        return None

    def with_params(self, params):
        return self

    def energy_density(self, density: torch.Tensor) -> torch.Tensor:
        """The functional written in torch: its definition, for host-side checks of the
        parameters and wrappers.  Calling the functional evaluates it on the device instead."""
        return torch.zeros_like(density)

    def __call__(self, density, params=None):
        """xc_energy_density_fn(density[, params]) on the device (``qex1d_xc_evaluate``, the
        network kernels of the SCF): per point for local functionals, one value per system for
        global ones; density [N] or [B, N]."""
        from . import engine

        fn = self.with_params(params) if params is not None else self
        return engine.xc_evaluate(fn, density, 1.0)[0]


class NoXC(XCFunctional):
    xc_kind = _lib.QEX1D_XC_NONE


class LDAExp(XCFunctional):
    """1D LDA (exchange + correlation per electron) for u = A exp(-kappa|r|)."""

    xc_kind = _lib.QEX1D_XC_LDA_EXP

    def energy_density(self, n):
        y = math.pi * n / KAPPA_EXP
        big = y > 1e-15
        ys = torch.where(big, y, torch.ones_like(y))
        ex = torch.where(big, A_EXP / (2 * math.pi) * (torch.log(1 + ys ** 2) / ys - 2 * torch.atan(ys)),
                         A_EXP / (2 * math.pi) * (-y + y ** 3 / 6))
        al, be, ga, de, et, si, nu = 2.0, -1.00077, 6.26099, -11.9041, 9.62614, -1.48334, 1.0
        fy = torch.where(y == 0, torch.ones_like(y), y)
        s = torch.sqrt(fy)
        P = al + be * s + ga * fy + de * fy * s + et * fy ** 2 + si * fy ** 2 * s + nu * math.pi * KAPPA_EXP ** 2 / A_EXP * fy ** 3
        ec = torch.where(y == 0, -A_EXP * y / math.pi / al, -A_EXP * fy / math.pi / P)
        return ex + ec


class PowerLaw(XCFunctional):
    """c (n + eps)^p; ``encoding='global'`` sums it to a scalar (amplitude-encoded XC)."""

    def __init__(self, c=-0.73855, p=1.0 / 3, eps=0.0, encoding="local"):
        self.c, self.p, self.eps = float(c), float(p), float(eps)
        if encoding not in ("local", "global"):
            raise ValueError("encoding must be 'local' or 'global'")
        self.encoding = encoding
        self.xc_kind = _lib.QEX1D_XC_POWER_LOCAL if encoding == "local" else _lib.QEX1D_XC_POWER_GLOBAL

    def configure(self, problem):
        problem.xc_kind = self.xc_kind
        problem.xc_c, problem.xc_p, problem.xc_eps = self.c, self.p, self.eps

    def energy_density(self, n):
        e = self.c * (n + self.eps) ** self.p
        return e.sum() if self.encoding == "global" else e


class LocalMLPXC(XCFunctional):
    """Per-point MLP on a window of 2h+1 densities (zero padded), inputs divided by ``norm``.

    ``params``: list of (W, b) per Dense layer with W of shape (fan_in, fan_out) (stax
    convention, x @ W + b); activation after every layer but the last; output width 1.
    """

    xc_kind = _lib.QEX1D_XC_LOCAL_MLP

    def __init__(self, params=None, window=17, activation="silu", norm=2.0, output="linear"):
        if window < 1 or window % 2 == 0:
            raise ValueError("window must be a positive odd number")
        if activation not in _lib.ACTIVATIONS:
            raise ValueError(f"unsupported activation {activation!r}")
        if output not in _lib.OUTPUTS:
            raise ValueError(f"unsupported output transform {output!r}")
        self.window, self.activation, self.norm = int(window), activation, float(norm)
        self.output = output
        self.params = None if params is None else self._canon(params)

    @staticmethod
    def _canon(params):
        """(W, b) of every Dense layer, in order, from flat [(W, b), ...] or nested stax parameters
        (activation layers contribute ``()``; stax.serial nests a sub-network's list)."""
        layers = []

        def walk(p):
            if isinstance(p, dict):
                for k in sorted(p):
                    walk(p[k])
            elif isinstance(p, (list, tuple)):
                if len(p) == 2 and _nd(p[0]) == 2 and _nd(p[1]) in (1, 2):
                    layers.append((np.asarray(p[0], dtype=np.float64), np.asarray(p[1], dtype=np.float64).reshape(-1)))
                else:
                    for q in p:
                        walk(q)

        walk(params)
        if not layers:
            raise ValueError("no Dense layers in params")
        return layers

    def with_params(self, params):
        return LocalMLPXC(params, self.window, self.activation, self.norm, self.output)

    def widths(self):
        ws = [self.params[0][0].shape[0]] + [w.shape[1] for w, _ in self.params]
        return ws

    def configure(self, problem):
        if self.params is None:
            raise ValueError("LocalMLPXC has no parameters")
        ws = self.widths()
        if ws[0] != self.window:
            raise ValueError(f"first layer fan-in {ws[0]} != window {self.window}")
        if len(self.params) > _lib.QEX1D_MAX_MLP_LAYERS:
            raise ValueError("too many layers")
        problem.xc_kind = self.xc_kind
        problem.mlp_layers = len(self.params)
        for i, w in enumerate(ws):
            problem.mlp_widths[i] = int(w)
        problem.mlp_activation = _lib.ACTIVATIONS[self.activation]
        problem.mlp_norm = self.norm
        problem.mlp_output = _lib.OUTPUTS[self.output]

    def packed(self):
        return np.concatenate([np.concatenate([w.reshape(-1), b.reshape(-1)]) for w, b in self.params])

    def device_params(self, device):
        return torch.as_tensor(self.packed(), dtype=torch.float64, device=device)

    def energy_density(self, n):
        h = self.window // 2
        N = n.shape[-1]
        pad = torch.nn.functional.pad(n, (h, h))
        idx = torch.arange(N)[:, None] + torch.arange(self.window)[None, :]
        x = pad[..., idx] / self.norm
        acts = {"silu": torch.nn.functional.silu, "swish": torch.nn.functional.silu, "tanh": torch.tanh,
                "softplus": torch.nn.functional.softplus, "sigmoid": torch.sigmoid, "relu": torch.relu,
                "gelu": lambda z: torch.nn.functional.gelu(z, approximate="tanh")}
        act = acts[self.activation]
        for li, (w, b) in enumerate(self.params):
            x = x @ torch.as_tensor(w, dtype=x.dtype, device=x.device) + torch.as_tensor(b, dtype=x.dtype, device=x.device)
            if li < len(self.params) - 1:
                x = act(x)
        return _output(self.output, x[..., 0])


def _output(kind, y):
    """The network-output transform: identity, or -swish (jax_dft neural_xc.negativity_transform)."""
    return -torch.nn.functional.silu(y) if kind == "negativity" else y


class GlobalMLPXC(LocalMLPXC):
    """MLP on the whole density divided by ``norm`` -> one scalar f(n) (``build_global_mlp``,
    classical_models.py:177-224); the amplitude-encoded loop takes E_xc = f(n) dx and
    v_xc = dx df/dn (od/scf.py:32-76).

    ``params``: list of (W, b) per Dense layer, W of shape (fan_in, fan_out); the first fan-in is
    the number of grid points, hidden widths at most ``QEX1D_MAX_MLP_WIDTH`` (1024), output width 1.
    """

    encoding = "global"
    xc_kind = _lib.QEX1D_XC_GLOBAL_MLP

    def __init__(self, params=None, activation="tanh", norm=2.0, output="linear"):
        if activation not in _lib.ACTIVATIONS:
            raise ValueError(f"unsupported activation {activation!r}")
        if output not in _lib.OUTPUTS:
            raise ValueError(f"unsupported output transform {output!r}")
        self.window = None
        self.activation, self.norm = activation, float(norm)
        self.output = output
        self.params = None if params is None else self._canon(params)

    def with_params(self, params):
        return GlobalMLPXC(params, self.activation, self.norm, self.output)

    def configure(self, problem):
        if self.params is None:
            raise ValueError("GlobalMLPXC has no parameters")
        ws = self.widths()
        if ws[0] != problem.num_grids:
            raise ValueError(f"first layer fan-in {ws[0]} != number of grid points {problem.num_grids}")
        if len(self.params) > _lib.QEX1D_MAX_MLP_LAYERS:
            raise ValueError("too many layers")
        if ws[-1] != 1 or max(ws[1:-1], default=1) > _lib.QEX1D_MAX_MLP_WIDTH:
            raise ValueError(f"GlobalMLPXC needs output width 1 and hidden widths <= {_lib.QEX1D_MAX_MLP_WIDTH}")
        problem.xc_kind = self.xc_kind
        problem.mlp_layers = len(self.params)
        for i, w in enumerate(ws):
            problem.mlp_widths[i] = int(w)
        problem.mlp_activation = _lib.ACTIVATIONS[self.activation]
        problem.mlp_norm = self.norm
        problem.mlp_output = _lib.OUTPUTS[self.output]

    def energy_density(self, n):
        x = n / self.norm
        acts = {"silu": torch.nn.functional.silu, "swish": torch.nn.functional.silu, "tanh": torch.tanh,
                "softplus": torch.nn.functional.softplus, "sigmoid": torch.sigmoid, "relu": torch.relu,
                "gelu": lambda z: torch.nn.functional.gelu(z, approximate="tanh")}
        act = acts[self.activation]
        for li, (w, b) in enumerate(self.params):
            x = x @ torch.as_tensor(w, dtype=x.dtype, device=x.device) + torch.as_tensor(b, dtype=x.dtype, device=x.device)
            if li < len(self.params) - 1:
                x = act(x)
        return _output(self.output, x[..., 0])


class ExternalXC(XCFunctional):
    """An opaque ``xc_energy_density_fn(density)`` written in torch: the SCF runs one pass at a time
    (``qex1d_kohn_sham_step``) and evaluates it between passes on the densities the device hands
    over, with torch autograd for the potential.  Local: fn(n) is the energy density per point;
    global (amplitude encoded): the values of fn(n) are summed to f (``od/scf.py:32-76``)."""

    def __init__(self, fn, encoding="local"):
        if encoding not in ("local", "global"):
            raise ValueError("encoding must be 'local' or 'global'")
        self.fn, self.encoding = fn, encoding
        self.xc_kind = _lib.QEX1D_XC_EXTERNAL_LOCAL if encoding == "local" else _lib.QEX1D_XC_EXTERNAL_GLOBAL

    def energy_density(self, n):
        return self.fn(n)

    def __call__(self, density, params=None):
        if params is not None:
            return self.fn(density, params=params)
        return self.fn(density)

    def _values(self, n):
        """Per system: e [N] (local) or the summed scalar f (global)."""
        out = [torch.as_tensor(self.fn(n[b]), dtype=torch.float64, device=n.device) for b in range(n.shape[0])]
        if self.encoding == "local":
            return torch.stack([o.reshape(n.shape[1]) for o in out])
        return torch.stack([o.sum() for o in out])

    def exchange(self, n_in, n_out, dx):
        """The exchange buffer [3, B, N] of the EXTERNAL kinds: values at n_in, the potential at
        n_in (local: d(sum_j e_j n_j)/dn_i; global: d(dx f)/dn_i), values at n_out."""
        B, N = n_in.shape
        buf = torch.zeros((3, B, N), dtype=torch.float64, device=n_in.device)
        x = n_in.detach().clone().requires_grad_(True)
        with torch.enable_grad():
            e = self._values(x)
            total = (e * x).sum() if self.encoding == "local" else dx * e.sum()
            if not total.requires_grad:
                raise TypeError("xc_energy_density_fn must be a differentiable torch function of the density")
            (v,) = torch.autograd.grad(total, x)
        with torch.no_grad():
            eo = self._values(n_out.detach())
        if self.encoding == "local":
            buf[0] = e.detach()
            buf[2] = eo
        else:
            buf[0, :, 0] = e.detach()
            buf[2, :, 0] = eo
        buf[1] = v
        return buf


def is_external(f) -> bool:
    return isinstance(f, ExternalXC)


def get_lda_xc_energy_density_fn():
    """jax_dft ``xc.get_lda_xc_energy_density_fn()``: fn(density, params) -> e_xc per electron."""
    return LDAExp()


def resolve(fn) -> XCFunctional:
    """Turns an ``xc_energy_density_fn`` (a functional, or functools.partial of one carrying
    ``params=``) into the functional the kernels evaluate."""
    if isinstance(fn, XCFunctional):
        return fn
    if isinstance(fn, functools.partial):
        params = fn.keywords.get("params") if fn.keywords else None
        if params is None and fn.args:
            params = fn.args[0]
        inner = fn.func
        if isinstance(inner, XCFunctional):
            return inner.with_params(params) if params is not None else inner
        if isinstance(getattr(inner, "qex_functional", None), XCFunctional):
            return inner.with_params(params) if params is not None else inner.qex_functional
    inner = getattr(fn, "qex_functional", None)
    if isinstance(inner, XCFunctional):
        return inner
    raise TypeError(
        "xc_energy_density_fn must be a qex_amd functional (LDAExp, PowerLaw, LocalMLPXC, or one built by "
        f"qex_amd.networks.build_xc_functional), got {type(fn).__name__}")
