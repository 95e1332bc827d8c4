"""KSDFTTrainer on the device (qedft/train/od/trainer.py:26-240) with the reference's call sites:
a LocalMLP network, build_xc_functional with the negativity transform, the committed h2 rows at
distances x100 = 128 and 384, reflection symmetry, 15 iterations.  The device loss and gradient are
checked against the oracle's reverse pass; three L-BFGS-B iterations must lower the loss and the
checkpoints must load back into the stax parameter nesting."""

import os
import pickle

import numpy as np
import pytest
import torch

from oracle import od as O
from oracle import od_adjoint as A

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _config(**kw):
    cfg = {"molecule_names": ["od_h2_train"], "num_grids": 513, "dataset1": [128, 384], "network_type": "mlp",
           "wrap_with_negative_transform": True, "n_neurons": 16, "n_layers": 2, "activation": "tanh",
           "num_iterations": 15, "alpha": 0.5, "alpha_decay": 0.9, "num_mixing_iterations": 1,
           "density_mse_converge_tolerance": -1.0, "discount_factor": 0.9, "use_amplitude_encoding": False,
           "save_every_n": 1, "maxiter": 3, "maxfun": 8, "factr": 1, "m": 20, "pgtol": 1e-14}
    cfg.update(kw)
    return cfg


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd import networks
    from qex_amd.od import trainer
    return networks, trainer


def test_trainer_loss_and_gradient_match_oracle(mods):
    networks, trainer = mods
    cfg = _config()
    tr = trainer.KSDFTTrainer(cfg, networks.LocalMLP(cfg), GOLDEN, seed=3, log=lambda *a: None)
    dataset, train_set, n0, grids = tr.prepare_dataset()
    init_fn, xc_fn = tr._functional(grids)
    params = init_fn(3)
    spec, flat = networks.flatten(params)
    vg = tr.value_and_grad_fn(dataset, train_set, n0, grids, xc_fn, spec)
    loss, grad = vg(flat)
    # oracle: the same network, the oracle's non-interacting initial densities, per-system reverse passes
    pairs = [p for p in params[0] if p != ()]
    oxc = O.LocalMLP([w for w, _ in pairs], [b for _, b in pairs], "tanh", 1, 2.0, output="negativity")
    dx = float(grids[1] - grids[0])
    B = len(train_set.total_energy)
    lo, go = 0.0, np.zeros_like(grad)
    for s in range(B):
        on0 = O.solve_noninteracting_system(train_set.external_potential[s], 2, grids)[0]
        np.testing.assert_allclose(n0[s].cpu().numpy(), on0, atol=1e-10)
        fw = A.forward_trajectory(train_set.locations[s], train_set.nuclear_charges[s], 2, 15, grids, oxc,
                                  O.exponential_coulomb, on0, 0.5, 0.9, 1, reflect=True)
        L, dE, dn = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], train_set.total_energy[s],
                                                     train_set.density[s], 2, dx, discount=0.9, start=10)
        lo += L / B
        go += A.backward(fw, dE, dn) / B
    assert abs(loss - lo) < 1e-9 * max(1.0, abs(lo))
    np.testing.assert_allclose(grad, go, rtol=0, atol=1e-6 * np.abs(go).max())


def test_trainer_lbfgs_lowers_the_loss_and_checkpoints(mods, tmp_path):
    networks, trainer = mods
    cfg = _config()
    tr = trainer.KSDFTTrainer(cfg, networks.LocalMLP(cfg), GOLDEN, seed=5, log=lambda *a: None)
    params, loss, info = tr.train(checkpoint_save_dir=tmp_path)
    losses = tr.loss_record
    print(f"L-BFGS-B: {len(losses)} evaluations, loss {losses[0]:.6e} -> {loss:.6e}, {info['nit']} iterations",
          flush=True)
    assert info["nit"] >= 1
    assert loss < losses[0]
    ckpts = sorted(p for p in os.listdir(tmp_path) if p.startswith("ckpt-"))
    assert ckpts and ckpts[0] == "ckpt-00000"
    with open(tmp_path / ckpts[-1], "rb") as fh:
        back = pickle.load(fh)
    assert isinstance(back, list) and back[1] == () and len(back[0]) == 5
    # resuming from a checkpoint starts from its parameters
    tr2 = trainer.KSDFTTrainer(_config(maxiter=1, maxfun=2), networks.LocalMLP(cfg), GOLDEN, log=lambda *a: None)
    _, loss2, _ = tr2.train(checkpoint_path=tmp_path / ckpts[-1], checkpoint_save_dir=tmp_path / "again")
    assert tr2.loss_record[0] < losses[0]


def test_trainer_evaluate_runs_the_python_loop(mods, tmp_path):
    networks, trainer = mods
    cfg = _config(num_iterations=12)
    tr = trainer.KSDFTTrainer(cfg, networks.LocalMLP(cfg), GOLDEN, log=lambda *a: None)
    _, _, _, grids = tr.prepare_dataset()
    init_fn, _ = tr._functional(grids)
    ck = tmp_path / "ck"
    with open(ck, "wb") as fh:
        pickle.dump(init_fn(1), fh)
    st = tr.evaluate(ck, [128, 384], output_path=str(tmp_path / "states.pkl"))
    assert st.total_energy.shape == (2, 12) and st.density.shape == (2, 12, 513)
    assert np.all(np.isfinite(st.total_energy))
    lda = tr.evaluate(ck, [128], use_lda=True)
    assert lda.total_energy.shape == (1, 12)


def test_trainer_amplitude_encoding_loss_and_gradient_match_oracle(mods):
    """use_amplitude_encoding: a GlobalMLP network (no self-interaction layer) trained through the
    amplitude-encoded scan loop (od/train.py:117-121 picks jit_scf_od.kohn_sham_amplitude_encoded);
    the device loss and gradient against the oracle's reverse pass."""
    import functools

    from qex_amd.od import xc as xcmod

    networks, trainer = mods
    cfg = _config(network_type="mlp_ksr", use_amplitude_encoding=True, wrap_self_interaction=False)
    tr = trainer.KSDFTTrainer(cfg, networks.GlobalMLP(cfg), GOLDEN, seed=3, log=lambda *a: None)
    dataset, train_set, n0, grids = tr.prepare_dataset()
    init_fn, xc_fn = tr._functional(grids)
    params = init_fn(4)
    spec, flat = networks.flatten(params)
    vg = tr.value_and_grad_fn(dataset, train_set, n0, grids, xc_fn, spec)
    loss, grad = vg(flat)
    f = xcmod.resolve(functools.partial(xc_fn, params=params))
    assert f.encoding == "global" and f.output == "negativity"
    oxc = O.GlobalMLP([w for w, _ in f.params], [b for _, b in f.params], "tanh", 2.0, output="negativity")
    dx = float(grids[1] - grids[0])
    B = len(train_set.total_energy)
    lo, go = 0.0, np.zeros_like(grad)
    for s in range(B):
        on0 = O.solve_noninteracting_system(train_set.external_potential[s], 2, grids)[0]
        fw = A.forward_trajectory(train_set.locations[s], train_set.nuclear_charges[s], 2, 15, grids, oxc,
                                  O.exponential_coulomb, on0, 0.5, 0.9, 1, reflect=True)
        L, dE, dn = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], train_set.total_energy[s],
                                                     train_set.density[s], 2, dx, discount=0.9, start=10)
        lo += L / B
        go += A.backward(fw, dE, dn) / B
    assert abs(loss - lo) < 1e-9 * max(1.0, abs(lo))
    np.testing.assert_allclose(grad, go, rtol=0, atol=1e-6 * np.abs(go).max())
