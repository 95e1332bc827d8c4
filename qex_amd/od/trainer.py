"""``KSDFTTrainer``: training a neural XC functional through the batched 1D Kohn-Sham scan loop with
the reference's interface (``qedft/train/od/trainer.py:26-240``, ``qedft/train/od/train.py:58-466``).

The forward solves run in ``qex1d_kohn_sham`` and the gradient in ``qex1d_kohn_sham_vjp`` (the device
reverse pass replaces ``jax.value_and_grad`` of the jitted loss); the optimiser is the reference's
``scipy.optimize.fmin_l_bfgs_b`` over the flat parameter vector, with its keyword settings taken
from the config as the reference does (``trainer.py:165-178``).  As in ``create_kohn_sham_fn``
(``train.py:84-122``) every system runs the scan loop with reflection symmetry about the grid
centre; the loss is ``create_loss_fn``'s (``train.py:250-323``): the discounted trajectory MSE of
the energies from iteration 10 on, divided by the electron count, plus the final-density MSE times
dx N / N_e.  Checkpoints are pickles of the unflattened parameters every ``save_every_n`` steps
(``train.py:456-461``).

Device reverse-pass limits: a LocalMLP of 1 or 2 hidden layers of width <= 64 (window <= 33) with an
optional negativity output transform; ``density_mse_converge_tolerance`` < 0.
"""

from __future__ import annotations

import functools
import os
import pickle
import time
from pathlib import Path

import numpy as np
import scipy.optimize
import torch

from .. import networks as nets
from . import engine, scf, train
from . import xc as xcmod
from .dataset import load_molecular_datasets_from_config
from .utils import exponential_coulomb, get_dx

KS_ITER_TO_IGNORE = 10  # the loss starts at iteration 10 (train.py:305)


def create_training_step(value_and_grad_fn, save_every_n, initial_checkpoint_index=0, checkpoint_dir=None, spec=None,
                         ckpt_prefix="ckpt", log=print):
    """``train.py:402-466``: a function flat -> (loss, gradient) for the optimiser that logs every step
    and writes a checkpoint (pickle of the unflattened parameters) every ``save_every_n`` steps."""
    loss_record = []

    def step_fn(flat):
        t = time.time()
        loss, grad = value_and_grad_fn(flat)
        step = initial_checkpoint_index + len(loss_record)
        log(f"step {step}, loss {loss} in {time.time() - t} sec")
        if checkpoint_dir and len(loss_record) % save_every_n == 0:
            path = f"{checkpoint_dir}/{ckpt_prefix}-{step:05d}"
            os.makedirs(checkpoint_dir, exist_ok=True)
            with open(path, "wb") as handle:
                pickle.dump(nets.unflatten(spec, flat), handle)
        loss_record.append(loss)
        return loss, grad

    step_fn.loss_record = loss_record
    return step_fn


def load_model_params(ckpt_path):
    """``eval.py:90-101``: parameters from a checkpoint this package wrote (pickle)."""
    with open(ckpt_path, "rb") as handle:
        return pickle.load(handle)


class KSDFTTrainer:
    """Trainer for Kohn-Sham DFT models (``trainer.py:26-240``)."""

    def __init__(self, config_dict: dict, network, data_path, seed: int = 0, log=print):
        self.network = network
        self.data_path = Path(data_path)
        self.seed = seed
        self.config = dict(config_dict)
        self.default_checkpoint_save_dir = Path("ckpts")
        self.log = log

    def prepare_dataset(self):
        """(dataset, train_set, initial_density, grids) of the config's first molecule
        (``trainer.py:60-77``); the initial density is the non-interacting solve on the device."""
        dataset, train_set = load_molecular_datasets_from_config(self.config, self.data_path,
                                                                 check_grid_centering=True)[0]
        initial_density = scf.get_initial_density(train_set, method="noninteracting")
        return dataset, train_set, initial_density, dataset.grids

    def _functional(self, grids):
        base = self.network.build_network(grids)
        return nets.build_xc_functional(network=base, grids=grids, config=self.config)

    def value_and_grad_fn(self, dataset, train_set, initial_density, grids, xc_fn, spec):
        """flat parameters -> (loss, gradient), forward and reverse pass on the device."""
        cfg = self.config
        if cfg.get("density_mse_converge_tolerance", -1.0) >= 0:
            raise NotImplementedError("the device reverse pass needs density_mse_converge_tolerance < 0")
        T = int(cfg["num_iterations"])
        if T <= KS_ITER_TO_IGNORE:
            raise ValueError(f"num_iterations must exceed {KS_ITER_TO_IGNORE} (the loss starts there)")
        params0 = nets.unflatten(spec, np.zeros(sum(int(np.prod(w)) + int(np.prod(b)) for w, b in spec[1])))
        f = xcmod.resolve(functools.partial(xc_fn, params=params0))
        prob = engine.Problem1D(grids, list(train_set.locations), list(train_set.nuclear_charges),
                                np.asarray(train_set.num_electrons), exponential_coulomb, f, T,
                                alpha=cfg["alpha"], alpha_decay=cfg["alpha_decay"],
                                num_mixing_iterations=cfg["num_mixing_iterations"],
                                density_mse_converge_tolerance=cfg.get("density_mse_converge_tolerance", -1.0),
                                enforce_reflection_symmetry=True, loop="scan", want_orbitals=True,
                                stop_gradient_step=cfg.get("stop_gradient_step", -1))
        n0 = engine.as_f64(initial_density, prob.dev)
        dx = get_dx(np.asarray(grids))
        discount = cfg["discount_factor"]

        def vg(flat):
            prob.xc_params.copy_(torch.as_tensor(np.asarray(flat, dtype=np.float64)))
            loss, grad, _ = train.value_and_grad(prob, n0, train.trajectory_loss, train_set.total_energy,
                                                 train_set.density, dataset.num_electrons, dx, discount=discount,
                                                 start=KS_ITER_TO_IGNORE)
            return float(loss), grad.cpu().numpy()

        vg.problem = prob
        return vg

    def train(self, checkpoint_path=None, checkpoint_save_dir=None):
        """L-BFGS-B over the flat parameters; returns (params, loss, info) of fmin_l_bfgs_b."""
        dataset, train_set, initial_density, grids = self.prepare_dataset()
        init_fn, xc_fn = self._functional(grids)
        if checkpoint_path and Path(checkpoint_path).exists():
            self.log(f"Loading parameters from checkpoint: {checkpoint_path}")
            init_params = load_model_params(checkpoint_path)
        else:
            init_params = init_fn(self.seed)
        spec, flat0 = nets.flatten(init_params)
        vg = self.value_and_grad_fn(dataset, train_set, initial_density, grids, xc_fn, spec)
        if checkpoint_save_dir is None:
            checkpoint_save_dir = self.default_checkpoint_save_dir
        else:
            Path(checkpoint_save_dir).mkdir(parents=True, exist_ok=True)
        step = create_training_step(vg, save_every_n=self.config.get("save_every_n", 10),
                                    initial_checkpoint_index=self.config.get("initial_checkpoint_index", 0),
                                    checkpoint_dir=str(checkpoint_save_dir), spec=spec,
                                    ckpt_prefix=self.config.get("ckpt_prefix", "ckpt"), log=self.log)
        kwargs = {"maxfun": self.config.get("maxfun", 1000), "factr": self.config.get("factr", 1),
                  "m": self.config.get("m", 20), "pgtol": self.config.get("pgtol", 1e-14),
                  "maxiter": self.config.get("maxiter", 1000)}
        self.log(f"Optimizing with L-BFGS-B: {kwargs}")
        params, loss, info = scipy.optimize.fmin_l_bfgs_b(step, x0=np.array(flat0), **kwargs)
        self.loss_record = step.loss_record
        return params, loss, info

    def evaluate(self, checkpoint_path, plot_distances, use_lda=False, output_path=None):
        """States of the python-loop solves at ``plot_distances`` with the checkpoint's parameters
        (``trainer.py:182-240``, ``eval.py:158-236``)."""
        from .eval import kohn_sham

        dataset, _, _, grids = self.prepare_dataset()
        plot_set = dataset.get_molecules(plot_distances)
        plot_n0 = scf.get_initial_density(plot_set, method="noninteracting")
        _, xc_fn = self._functional(grids)
        params = load_model_params(checkpoint_path)
        cfg = self.config
        states = []
        for i in range(len(plot_distances)):
            st = kohn_sham(params, plot_set.locations[i], plot_set.nuclear_charges[i], dataset.num_electrons,
                           cfg.get("num_iterations", 100), grids, xc_fn, cfg.get("use_amplitude_encoding", False),
                           initial_density=plot_n0[i], use_lda=use_lda, alpha=cfg.get("alpha", 0.5),
                           alpha_decay=cfg.get("alpha_decay", 0.9),
                           num_mixing_iterations=cfg.get("num_mixing_iterations", 2),
                           density_mse_converge_tolerance=cfg.get("density_mse_converge_tolerance", -1.0),
                           enforce_reflection_symmetry=cfg.get("enforce_reflection_symmetry", True))
            states.append(st.numpy())
        stacked = scf.KohnShamState(*[np.stack([getattr(s, f) for s in states]) if getattr(states[0], f) is not None
                                      else None for f in scf.KohnShamState._fields])
        if output_path is not None:
            os.makedirs(os.path.dirname(os.path.abspath(output_path)), exist_ok=True)
            with open(output_path, "wb") as handle:
                pickle.dump(stacked, handle)
        return stacked
