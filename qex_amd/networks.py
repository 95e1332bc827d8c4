"""The reference's network call surface for the classical XC networks the device evaluates.

Mirrors ``qedft/models/networks.py:43-138`` (the ``KohnShamNetwork`` protocol, ``LocalMLP``,
``GlobalMLP``), ``qedft/models/classical/classical_models.py:75-224`` (``build_mlp_layers``,
``build_local_mlp``, ``build_global_mlp``) and ``qedft/models/wrappers.py:25-299``
(``neural_xc_functional``, ``wrap_network``, ``wrap_network_from_config``,
``build_xc_functional``):

* ``network.build_network(grids)`` returns the stax pair ``(init_fn, apply_fn)``:
  ``init_fn(rng, input_shape) -> (output_shape, params)`` with ``params`` in stax's nesting (one
  ``(W, b)`` per Dense layer, ``()`` per activation layer), ``apply_fn(params, inputs)``;
* ``build_xc_functional(network, grids, config)`` returns ``(init_fn(rng) -> params,
  xc_energy_density_fn(density, params))`` with the flipped argument order of the reference.
  ``xc_energy_density_fn`` is a device functional (``qex_amd.od.xc``): passed to ``kohn_sham``
  (directly, or as ``functools.partial(fn, params=...)`` as the reference's trainers do), the SCF
  kernels evaluate it; calling it evaluates the same network with torch, for inspection.

``rng`` may be an int seed or a ``numpy.random.Generator`` (jax PRNG keys are not available here):
the parameter distributions are stax's (``Dense``: Glorot-normal weights, N(0, 1e-2) biases), the
values differ from jax's random streams.  Wrappers: ``wrap_with_negative_transform`` (the
``-swish`` output layer) runs on the device, forward and reverse pass;
``wrap_self_interaction`` is not implemented on the device and raises.
"""

from __future__ import annotations

import math
from typing import Callable, Protocol

import numpy as np
import torch

from .od import xc as xcmod

DEFAULT_N_NEURONS = 64
DEFAULT_N_LAYERS = 3
DEFAULT_ACTIVATION = "tanh"
DEFAULT_N_OUTPUTS = 1
DEFAULT_DENSITY_NORM = 2.0

# stax/jax names of the activations the kernels implement
ACTIVATION_MAP = {"tanh": "tanh", "relu": "relu", "sigmoid": "sigmoid", "softplus": "softplus", "silu": "silu",
                  "swish": "silu", "gelu": "gelu"}

# network_type values the reference treats as global (amplitude-encoded) models (wrappers.py:180-188)
GLOBAL_TYPES = ("gadqc", "gadqc_with_mlp", "conv_dqc", "conv_adqc", "mlp_ksr", "conv_dqc_tfm", "conv_adqc_tfm")


def _nd(x) -> int:
    """Array rank of a parameter leaf; -1 for containers (lists, tuples, dicts)."""
    return -1 if isinstance(x, (list, tuple, dict)) else np.ndim(x)


def _rng(rng) -> np.random.Generator:
    if isinstance(rng, np.random.Generator):
        return rng
    if rng is None:
        return np.random.default_rng(0)
    arr = np.asarray(rng).ravel()
    return np.random.default_rng([int(x) & 0xFFFFFFFF for x in arr] if arr.size > 1 else int(arr[0]) & 0xFFFFFFFF)


def _glorot_normal(rng: np.random.Generator, fan_in: int, fan_out: int) -> np.ndarray:
    """jax.nn.initializers.glorot_normal: truncated normal in [-2, 2] standard deviations, scaled so
    the variance is 2 / (fan_in + fan_out) (0.87962566 is the std of the unit normal truncated there)."""
    std = math.sqrt(2.0 / (fan_in + fan_out)) / 0.87962566103423978
    out = rng.standard_normal((fan_in, fan_out))
    bad = np.abs(out) > 2.0
    while bad.any():
        out[bad] = rng.standard_normal(int(bad.sum()))
        bad = np.abs(out) > 2.0
    return out * std


class Network(tuple):
    """``(init_fn, apply_fn)`` that also carries the description the device kernels need."""

    def __new__(cls, init_fn, apply_fn, spec):
        obj = super().__new__(cls, (init_fn, apply_fn))
        obj.spec = spec
        return obj


def build_mlp_layers(n_neurons: int, n_layers: int, activation: str, n_outputs: int = DEFAULT_N_OUTPUTS) -> list:
    """Layer list Input -> [Dense(n_neurons) -> act] x n_layers -> Dense(n_outputs)
    (classical_models.py:75-115), as ("dense", width) / ("act", name) entries."""
    if n_neurons <= 0:
        raise ValueError("n_neurons must be positive")
    if n_layers <= 0:
        raise ValueError("n_layers must be positive")
    if activation not in ACTIVATION_MAP:
        raise ValueError(f"Unknown activation '{activation}'. Valid options: {list(ACTIVATION_MAP)}")
    layers = []
    for _ in range(n_layers):
        layers += [("dense", n_neurons), ("act", ACTIVATION_MAP[activation])]
    layers.append(("dense", n_outputs))
    return layers


def _stax_init(layers, fan_in, rng):
    params = []
    for kind, arg in layers:
        if kind == "dense":
            W = _glorot_normal(rng, fan_in, arg)
            b = rng.standard_normal(arg) * 1e-2
            params.append((W, b))
            fan_in = arg
        else:
            params.append(())
    return params


def _dense_pairs(params):
    """The (W, b) pairs of nested stax parameters, in order (the leaves np_utils.flatten walks)."""
    out = []

    def walk(p):
        if isinstance(p, dict):
            for k in sorted(p):
                walk(p[k])
        elif isinstance(p, (list, tuple)):
            if len(p) == 2 and _nd(p[0]) == 2 and _nd(p[1]) == 1:
                out.append((np.asarray(p[0], dtype=np.float64), np.asarray(p[1], dtype=np.float64)))
            else:
                for q in p:
                    walk(q)

    walk(params)
    return out


def _torch_apply(pairs, activation, x):
    acts = {"silu": torch.nn.functional.silu, "tanh": torch.tanh, "softplus": torch.nn.functional.softplus,
            "sigmoid": torch.sigmoid, "relu": torch.relu,
            "gelu": lambda z: torch.nn.functional.gelu(z, approximate="tanh")}
    act = acts[activation]
    for li, (w, b) in enumerate(pairs):
        x = x @ torch.as_tensor(w, dtype=x.dtype, device=x.device) + torch.as_tensor(b, dtype=x.dtype, device=x.device)
        if li < len(pairs) - 1:
            x = act(x)
    return x


def build_local_mlp(n_neurons: int = DEFAULT_N_NEURONS, n_layers: int = DEFAULT_N_LAYERS,
                    activation: str = DEFAULT_ACTIVATION, n_outputs: int = DEFAULT_N_OUTPUTS,
                    density_normalization_factor: float = DEFAULT_DENSITY_NORM, grids=None, window: int = 1,
                    **kwargs) -> Network:
    """MLP applied to every grid point (``classical_models.py:123-174``): density_i / norm -> MLP ->
    output_i.  ``window`` > 1 (odd) feeds the 2h+1 neighbouring densities instead (zero padded at
    the edges), the windowed LocalMLP of the batched workload; 1 is the reference network."""
    if grids is None:
        raise ValueError("grids must be provided for local MLP")
    if n_outputs != 1:
        raise ValueError("the XC networks have one output per point")
    layers = build_mlp_layers(n_neurons, n_layers, activation, n_outputs)
    num_grids = int(np.shape(grids)[0])
    act = ACTIVATION_MAP[activation]

    def init_fn(rng, input_shape=None):
        del input_shape
        return (-1, num_grids, 1), _stax_init(layers, window, _rng(rng))

    def apply_fn(params, inputs, **kw):
        n = torch.as_tensor(np.asarray(inputs, dtype=np.float64)) if not isinstance(inputs, torch.Tensor) else inputs
        f = xcmod.LocalMLPXC(_dense_pairs(params), window=window, activation=act, norm=density_normalization_factor)
        return f.energy_density(n.to(torch.float64).reshape(-1))

    return Network(init_fn, apply_fn, {"kind": "local", "activation": act, "norm": float(density_normalization_factor),
                                       "window": int(window), "num_grids": num_grids})


def build_global_mlp(n_neurons: int = DEFAULT_N_NEURONS, n_layers: int = DEFAULT_N_LAYERS,
                     activation: str = DEFAULT_ACTIVATION, n_outputs: int = DEFAULT_N_OUTPUTS,
                     density_normalization_factor: float = DEFAULT_DENSITY_NORM, grids=None, **kwargs) -> Network:
    """MLP on the whole density (``classical_models.py:177-224``): density / norm -> MLP -> scalar."""
    if grids is None:
        raise ValueError("grids must be provided for global MLP")
    if n_outputs != 1:
        raise ValueError("the XC networks have one output")
    layers = build_mlp_layers(n_neurons, n_layers, activation, n_outputs)
    num_grids = int(np.shape(grids)[0])
    act = ACTIVATION_MAP[activation]

    def init_fn(rng, input_shape=None):
        del input_shape
        return (1,), _stax_init(layers, num_grids, _rng(rng))

    def apply_fn(params, inputs, **kw):
        n = torch.as_tensor(np.asarray(inputs, dtype=np.float64)) if not isinstance(inputs, torch.Tensor) else inputs
        return _torch_apply(_dense_pairs(params), act, n.to(torch.float64) / density_normalization_factor)

    return Network(init_fn, apply_fn, {"kind": "global", "activation": act, "norm": float(density_normalization_factor),
                                       "num_grids": num_grids})


class KohnShamNetwork(Protocol):
    """``qedft/models/networks.py:43-75``: built from a config dict; ``build_network(grids)`` returns
    the ``(init_fn, apply_fn)`` pair."""

    def __init__(self, config_dict: dict | None = None):
        ...

    def build_network(self, grids) -> tuple[Callable, Callable]:
        raise NotImplementedError


class LocalMLP:
    """``networks.py:83-110``: a LocalMLP with the reference's config defaults."""

    def __init__(self, config_dict: dict | None = None):
        self.config = {"network_type": "mlp", "wrap_self_interaction": False, "wrap_with_negative_transform": True,
                       "use_amplitude_encoding": False}
        if config_dict is not None:
            self.config.update(config_dict)
        if self.config.get("use_amplitude_encoding") is True:
            raise ValueError("Set use_amplitude_encoding to False.")

    def build_network(self, grids) -> Network:
        return build_local_mlp(n_neurons=self.config.get("n_neurons", 64), n_layers=self.config.get("n_layers", 3),
                               activation=self.config.get("activation", "tanh"),
                               density_normalization_factor=self.config.get("density_normalization_factor", 2.0),
                               grids=grids, window=self.config.get("window", 1))


class GlobalMLP:
    """``networks.py:113-138``: a GlobalMLP (amplitude encoding) with the reference's config defaults."""

    def __init__(self, config_dict: dict | None = None):
        self.config = {"network_type": "mlp_ksr", "wrap_self_interaction": True, "wrap_with_negative_transform": True,
                       "use_amplitude_encoding": True}
        if config_dict is not None:
            self.config.update(config_dict)
        if self.config.get("use_amplitude_encoding") is False:
            raise ValueError("Set use_amplitude_encoding to True.")

    def build_network(self, grids) -> Network:
        return build_global_mlp(n_neurons=self.config.get("n_neurons", 64), n_layers=self.config.get("n_layers", 3),
                                activation=self.config.get("activation", "tanh"),
                                density_normalization_factor=self.config.get("density_normalization_factor", 2.0),
                                grids=grids)


class _XCApply:
    """``xc_energy_density_fn(density, params)`` of a wrapped network: a device functional factory.
    ``resolve`` in qex_amd.od.xc turns it (or a functools.partial of it carrying ``params``) into the
    functional the kernels evaluate."""

    def __init__(self, functional, encoding):
        self.qex_functional = functional
        self.encoding = encoding

    def with_params(self, params):
        return self.qex_functional.with_params(_dense_pairs(params))

    def __call__(self, density, params=None, **kwargs):
        f = self.with_params(params) if params is not None else self.qex_functional
        return f(density)


def neural_xc_functional(network, grids, network_type: str, encoding: str = "local", output: str = "linear",
                         **kwargs):
    """``wrappers.py:25-139``: (init_fn(rng) -> params, xc_energy_density_fn(density, params)).
    Local encoding returns the energy density per point, amplitude encoding a scalar."""
    spec = getattr(network, "spec", None)
    if spec is None:
        raise TypeError("the device evaluates the networks built by qex_amd.networks (build_local_mlp / "
                        "build_global_mlp, LocalMLP / GlobalMLP); got an opaque (init_fn, apply_fn) pair")
    if kwargs.get("noise_std", 0.0) > 0.0:
        raise NotImplementedError("noise_std > 0 (Gaussian noise on the network output) is not implemented")
    init_net, _ = network
    if encoding == "local":
        if spec["kind"] != "local":
            raise ValueError("local encoding needs a per-point network")
        fn = xcmod.LocalMLPXC(None, window=spec["window"], activation=spec["activation"], norm=spec["norm"],
                              output=output)
    else:
        if spec["kind"] != "global":
            raise ValueError("amplitude encoding needs a whole-density network")
        fn = xcmod.GlobalMLPXC(None, activation=spec["activation"], norm=spec["norm"], output=output)

    def init_fn(rng, **kw):
        _, params = init_net(rng)
        if output == "negativity":
            params = [params, ()]  # stax.serial(network, negativity_transform()) nests the inner params
        return params

    return init_fn, _XCApply(fn, "local" if encoding == "local" else "global")


def wrap_network(network, grids, network_type: str, wrap_self_interaction: bool = False,
                 wrap_with_negative_transform: bool = False, **kwargs):
    """``wrappers.py:142-217``: global network types give amplitude encoding; optional layers."""
    if wrap_self_interaction:
        raise NotImplementedError("wrap_self_interaction (the jax_dft self-interaction layer) is not implemented "
                                  "on the device")
    encoding = "amplitude" if network_type in GLOBAL_TYPES else "local"
    return neural_xc_functional(network, grids, network_type, encoding,
                                output="negativity" if wrap_with_negative_transform else "linear", **kwargs)


def wrap_network_from_config(network, grids, config: dict):
    """``wrappers.py:220-264``."""
    return wrap_network(network=network, grids=grids, network_type=config["network_type"],
                        wrap_self_interaction=config.get("wrap_self_interaction", False),
                        wrap_with_negative_transform=config.get("wrap_with_negative_transform", False))


def build_xc_functional(network, grids, config: dict):
    """``wrappers.py:267-299``: the main entry point; ``(init_fn, xc_energy_density_fn)`` with
    ``xc_energy_density_fn(density, params)``."""
    return wrap_network_from_config(network, grids, config)


def flatten(params):
    """``jax_dft.np_utils.flatten``: (spec, flat vector) of nested stax parameters (Dense (W, b) leaves in
    order; the flat order is the packed layout the kernels read)."""
    pairs = _dense_pairs(params)
    spec = (params, [(w.shape, b.shape) for w, b in pairs])
    return spec, np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in pairs])


def unflatten(spec, flat):
    """``jax_dft.np_utils.unflatten``: nested parameters of ``spec``'s structure from a flat vector."""
    like, shapes = spec
    flat = np.asarray(flat, dtype=np.float64)
    it = iter(shapes)
    pos = [0]

    def take(shape):
        n = int(np.prod(shape))
        out = flat[pos[0]:pos[0] + n].reshape(shape)
        pos[0] += n
        return out

    def rebuild(p):
        if isinstance(p, dict):
            return {k: rebuild(p[k]) for k in p}
        if isinstance(p, (list, tuple)):
            if len(p) == 2 and _nd(p[0]) == 2 and _nd(p[1]) == 1:
                ws, bs = next(it)
                return (take(ws), take(bs))
            out = [rebuild(q) for q in p]
            return out if isinstance(p, list) else tuple(out)
        return p

    res = rebuild(like)
    if pos[0] != flat.size:
        raise ValueError(f"flat vector has {flat.size} entries, the parameters {pos[0]}")
    return res
