"""Host-side pieces of the 1D path that run without a GPU: the functional wrappers' own torch
evaluation and the problem description they write for the device."""

import numpy as np
import pytest
import torch

from oracle import od as O
from qex_amd import _lib
from qex_amd.od import xc


def test_global_mlp_wrapper_matches_oracle_value():
    """GlobalMLPXC(n) is build_global_mlp's apply_fn (classical_models.py:177-224) on n / norm."""
    W, b = O.glorot_normal_mlp(3, [41, 16, 8, 1])
    rng = np.random.default_rng(4)
    b = [rng.normal(0, 0.1, bb.shape) for bb in b]
    n = np.abs(rng.normal(0, 0.5, 41))
    for act in ("tanh", "silu", "softplus"):
        f = xc.GlobalMLPXC(list(zip(W, b)), activation=act, norm=2.0)
        got = float(f.energy_density(torch.as_tensor(n)))
        assert got == pytest.approx(O.GlobalMLP(W, b, act, 2.0).value(n), abs=1e-13)


def test_global_mlp_configure_checks_shapes():
    W, b = O.glorot_normal_mlp(3, [41, 16, 1])
    p = _lib.Qex1dProblem()
    p.num_grids = 41
    xc.GlobalMLPXC(list(zip(W, b))).configure(p)
    assert p.xc_kind == _lib.QEX1D_XC_GLOBAL_MLP and p.mlp_layers == 2
    assert list(p.mlp_widths[:3]) == [41, 16, 1]
    p.num_grids = 40
    with pytest.raises(ValueError, match="grid points"):
        xc.GlobalMLPXC(list(zip(W, b))).configure(p)
    W2, b2 = O.glorot_normal_mlp(3, [41, 513, 1])  # the reference notebook's width is accepted
    p.num_grids = 41
    xc.GlobalMLPXC(list(zip(W2, b2))).configure(p)
    assert list(p.mlp_widths[:3]) == [41, 513, 1]
    W3, b3 = O.glorot_normal_mlp(3, [41, _lib.QEX1D_MAX_MLP_WIDTH + 1, 1])
    with pytest.raises(ValueError, match="hidden"):
        xc.GlobalMLPXC(list(zip(W3, b3))).configure(p)


def test_global_mlp_is_amplitude_encoded():
    from qex_amd.od import scf

    W, b = O.glorot_normal_mlp(3, [41, 16, 1])
    f = xc.GlobalMLPXC(list(zip(W, b)))
    assert scf._global_functional(f) is f
    with pytest.raises(ValueError, match="local"):
        scf._local_functional(f)


def test_solve_noninteracting_rejects_electrons_beyond_max():
    """The non-interacting solve sizes its orbital slots from max_electrons; a system with more
    electrons than that (or beyond the library limit) is refused before any device work."""
    from qex_amd.od import engine

    v = np.zeros((2, 33))
    with pytest.raises(ValueError, match="exceeds max_electrons"):
        engine.solve_noninteracting(v, [2, 6], np.linspace(-4, 4, 33), max_electrons=4)
    with pytest.raises(ValueError, match="<= 62"):
        engine.solve_noninteracting(v, [2, 64], np.linspace(-4, 4, 33))
    with pytest.raises(ValueError, match=">= 1"):
        engine.solve_noninteracting(v, [0, 2], np.linspace(-4, 4, 33))
