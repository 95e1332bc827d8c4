"""Host side of the network and training surface (no GPU): KohnShamNetwork / build_xc_functional
(qedft/models/networks.py, wrappers.py), the jax_dft dataset layout loader
(qedft/data_io/dataset_loader.py) on the committed h2 rows, parameter flattening, and the functional
descriptions the kernels receive."""

import functools
import os

import numpy as np
import pytest
import torch

from oracle import od as O
from qex_amd import _lib
from qex_amd import networks as N
from qex_amd.od import dataset as D
from qex_amd.od import xc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_local_mlp_defaults_and_stax_params():
    grids = np.linspace(-10, 10, 101)
    net = N.LocalMLP().build_network(grids)
    out_shape, params = net[0](0, (-1, 101, 1))
    assert out_shape == (-1, 101, 1)
    # stax.serial(Dense, act, Dense, act, Dense, act, Dense): n_layers = 3, n_neurons = 64
    assert len(params) == 7 and params[1] == () and params[3] == ()
    assert params[0][0].shape == (1, 64) and params[6][0].shape == (64, 1)
    # Glorot-normal weights (variance 2 / (fan_in + fan_out)), N(0, 1e-2) biases
    _, big = N.build_local_mlp(n_neurons=512, n_layers=2, grids=grids)[0](1, None)
    W = big[2][0]
    assert W.std() == pytest.approx(np.sqrt(2.0 / 1024), rel=0.02)
    assert np.abs(big[0][1]).max() < 0.06
    with pytest.raises(ValueError):
        N.LocalMLP({"use_amplitude_encoding": True})
    with pytest.raises(ValueError):
        N.GlobalMLP({"use_amplitude_encoding": False})


def test_apply_fn_matches_oracle_network():
    grids = np.linspace(-10, 10, 101)
    net = N.build_local_mlp(n_neurons=12, n_layers=2, activation="tanh", grids=grids)
    _, params = net[0](3)
    pairs = [p for p in params if p != ()]
    n = np.exp(-grids ** 2) + 0.01
    ref = O.LocalMLP([w for w, _ in pairs], [b for _, b in pairs], "tanh", 1, 2.0).energy_density(n)
    np.testing.assert_allclose(net[1](params, n).numpy(), ref, atol=1e-14)
    gnet = N.build_global_mlp(n_neurons=8, n_layers=2, activation="tanh", grids=grids)
    _, gp = gnet[0](4)
    gpairs = [p for p in gp if p != ()]
    gref = O.GlobalMLP([w for w, _ in gpairs], [b for _, b in gpairs], "tanh", 2.0).value(n)
    assert float(gnet[1](gp, n)[0]) == pytest.approx(gref, abs=1e-13)


def test_build_xc_functional_wrappers_and_device_description():
    grids = np.linspace(-10, 10, 101)
    net = N.LocalMLP({"n_neurons": 16, "n_layers": 2}).build_network(grids)
    init_fn, fn = N.build_xc_functional(net, grids, {"network_type": "mlp", "wrap_with_negative_transform": True})
    params = init_fn(0)
    # stax.serial(network, negativity_transform()) nests the network's parameters
    assert isinstance(params, list) and len(params) == 2 and params[1] == ()
    f = xc.resolve(functools.partial(fn, params=params))
    assert isinstance(f, xc.LocalMLPXC) and f.output == "negativity" and f.widths() == [1, 16, 16, 1]
    p = _lib.Qex1dProblem()
    f.configure(p)
    assert p.mlp_output == _lib.QEX1D_MLP_OUTPUT_NEGATIVITY and p.xc_kind == _lib.QEX1D_XC_LOCAL_MLP
    n = np.exp(-grids ** 2) + 0.01
    pairs = [q for q in params[0] if q != ()]
    ref = O.LocalMLP([w for w, _ in pairs], [b for _, b in pairs], "tanh", 1, 2.0,
                     output="negativity").energy_density(n)
    # the torch definition of the wrapped functional (calling fn evaluates on the device, which
    # tests/test_xc_gpu.py covers)
    got = fn.with_params(params).energy_density(torch.as_tensor(n)).numpy()
    np.testing.assert_allclose(got, ref, atol=1e-14)
    assert float(got.max()) <= 0.2785
    # global network types are amplitude encoded (wrappers.py:180-193)
    gnet = N.GlobalMLP({"n_neurons": 8, "n_layers": 1, "wrap_self_interaction": False}).build_network(grids)
    ginit, gfn = N.build_xc_functional(gnet, grids, {"network_type": "mlp_ksr"})
    assert xc.resolve(functools.partial(gfn, params=ginit(1))).encoding == "global"
    with pytest.raises(NotImplementedError, match="self-interaction"):
        N.build_xc_functional(gnet, grids, {"network_type": "mlp_ksr", "wrap_self_interaction": True})
    with pytest.raises(TypeError, match="opaque"):
        N.build_xc_functional((lambda r: r, lambda p, x: x), grids, {"network_type": "mlp"})


def test_flatten_unflatten_round_trip_keeps_stax_nesting():
    grids = np.linspace(-10, 10, 51)
    init_fn, _ = N.build_xc_functional(N.LocalMLP({"n_neurons": 6, "n_layers": 2}).build_network(grids), grids,
                                       {"network_type": "mlp", "wrap_with_negative_transform": True})
    params = init_fn(2)
    spec, flat = N.flatten(params)
    assert flat.size == (1 * 6 + 6) + (6 * 6 + 6) + (6 + 1)
    back = N.unflatten(spec, flat * 2)
    assert isinstance(back, list) and back[1] == () and back[0][1] == ()
    np.testing.assert_array_equal(back[0][2][0], 2 * params[0][2][0])
    # the flat order is the packed order the kernels read
    f = xc.LocalMLPXC(params, window=1)
    np.testing.assert_array_equal(f.packed(), flat)
    with pytest.raises(ValueError):
        N.unflatten(spec, flat[:-1])


def test_dataset_loader_reads_the_jax_dft_layout():
    cfg = {"molecule_names": ["od_h2_train"], "num_grids": 513, "dataset1": [128, 384]}
    (ds, train), = D.load_molecular_datasets_from_config(cfg, GOLDEN)
    assert ds.num_electrons == 2 and ds.grids.shape == (513,)
    assert train.density.shape == (2, 513) and train.total_energy.shape == (2,)
    np.testing.assert_allclose(train.locations[:, 1] - train.locations[:, 0], [1.28, 3.84], atol=1e-12)
    np.testing.assert_allclose(train.density.sum(axis=1) * (ds.grids[1] - ds.grids[0]), 2.0, atol=1e-6)
    # the external potentials are the exponential-Coulomb chain potential of the stored nuclei
    for k in range(2):
        v = O.get_atomic_chain_potential(ds.grids, train.locations[k], train.nuclear_charges[k], O.exponential_coulomb)
        np.testing.assert_allclose(train.external_potential[k], v, atol=1e-13)
    assert D.location_center_at_grids_center_point(train.locations, ds.grids)
    # a smaller grid keeps the centre points
    small = D.Dataset(path=os.path.join(GOLDEN, "od_h2_train"), num_grids=401)
    assert small.grids.shape == (401,) and small.densities.shape == (2, 401)
    assert small.grids[200] == ds.grids[256]
    with pytest.raises(ValueError):
        D.Dataset(path=os.path.join(GOLDEN, "od_h2_train"), num_grids=400)
    with pytest.raises(ValueError, match="distances_x100"):
        ds.get_molecules([128, 999])
    with pytest.raises(ValueError, match="Number of molecule names"):
        D.load_molecular_datasets(["a", "b"], {"dataset1": [1]}, 513, base_path=GOLDEN)


def test_trainer_rejects_what_the_device_reverse_pass_lacks():
    from qex_amd.od.trainer import KSDFTTrainer
    tr = KSDFTTrainer({"density_mse_converge_tolerance": 1e-6, "num_iterations": 15}, N.LocalMLP(), GOLDEN)
    with pytest.raises(NotImplementedError):
        tr.value_and_grad_fn(None, None, None, None, None, None)
