"""1D dataset I/O in the jax_dft ``Dataset`` layout that the reference's trainer consumes
(``qedft/data_io/dataset_loader.py:11-162``; ``data/od/<molecule>/``): one ``.npy`` file per field
(``num_electrons``, ``grids``, ``locations``, ``nuclear_charges``, ``distances_x100``, ``distances``,
``total_energies``, ``densities``, ``external_potentials``), read with ``numpy.load`` (no pickles).

``Dataset(path, num_grids)`` keeps the centre ``num_grids`` points of the stored grid (and of every
per-point field) when a smaller grid is asked for; ``get_molecules(distances_x100)`` returns the
selected systems as a batched ``KohnShamState`` (numpy arrays), the form the trainer passes to
``get_initial_density`` and the loss.
"""

from __future__ import annotations

import os
from pathlib import Path

import numpy as np

from .scf import KohnShamState

_PER_POINT = ("densities", "external_potentials", "exact_v_s")


def location_center_at_grids_center_point(locations, grids, atol=1e-8) -> bool:
    """True when the mean nuclear position of every system sits on the centre grid point (an odd
    grid has one): the condition the flip about the grid centre of the scan loop relies on."""
    grids = np.asarray(grids)
    n = grids.shape[0]
    if n % 2 == 0:
        return False
    loc = np.atleast_2d(np.asarray(locations, dtype=np.float64))
    return bool(np.all(np.abs(loc.mean(axis=-1) - grids[n // 2]) < atol))


class Dataset:
    """The stored systems of one molecule family."""

    def __init__(self, path=None, data=None, num_grids=None, name=None):
        if (path is None) == (data is None):
            raise ValueError("give exactly one of path and data")
        self.name = name
        self.data = self.load_data(path) if data is None else {k: np.asarray(v) for k, v in data.items()}
        self._set_num_grids(num_grids)
        self.num_electrons = int(np.asarray(self.data["num_electrons"]))
        self.total_num_samples = int(self.data["distances_x100"].shape[0])
        for key in ("total_energies", "densities", "external_potentials", "locations", "nuclear_charges",
                    "distances_x100", "distances"):
            setattr(self, key, self.data[key])

    @staticmethod
    def load_data(path):
        path = Path(path)
        if not path.is_dir():
            raise FileNotFoundError(f"dataset directory {path} does not exist")
        data = {}
        for f in ("num_electrons", "grids", "locations", "nuclear_charges", "distances_x100", "distances",
                  "total_energies", "densities", "external_potentials", "exact_v_s"):
            fp = path / f"{f}.npy"
            if fp.exists():
                data[f] = np.load(fp, allow_pickle=False)
            elif f != "exact_v_s":
                raise FileNotFoundError(f"{fp} is missing")
        data["distances_x100"] = data["distances_x100"].astype(np.int64)
        return data

    def _set_num_grids(self, num_grids):
        full = int(self.data["grids"].shape[0])
        if num_grids is None:
            num_grids = full
        if num_grids > full:
            raise ValueError(f"num_grids ({num_grids}) can not be greater than the stored grid ({full})")
        diff = full - num_grids
        if diff % 2:
            raise ValueError(f"the stored grid ({full}) can not be cut to {num_grids} points symmetrically")
        lo, hi = diff // 2, full - diff // 2
        self.num_grids = int(num_grids)
        self.grids = np.asarray(self.data["grids"][lo:hi], dtype=np.float64)
        for key in _PER_POINT:
            if key in self.data:
                self.data[key] = np.asarray(self.data[key][:, lo:hi], dtype=np.float64)

    def get_molecules(self, selected_distance_x100=None) -> KohnShamState:
        """The systems whose distance x 100 is listed (all of them for None), as a batched state."""
        if selected_distance_x100 is None:
            mask = np.ones(self.total_num_samples, dtype=bool)
        else:
            sel = set(int(d) for d in selected_distance_x100)
            mask = np.array([int(d) in sel for d in self.distances_x100])
            if len(sel) != int(mask.sum()):
                raise ValueError(f"not all distances_x100 {sorted(sel)} are in the dataset")
        count = int(mask.sum())
        return KohnShamState(
            density=self.densities[mask], total_energy=self.total_energies[mask], locations=self.locations[mask],
            nuclear_charges=self.nuclear_charges[mask], external_potential=self.external_potentials[mask],
            grids=np.tile(self.grids[None, :], (count, 1)), num_electrons=np.repeat(self.num_electrons, count),
            converged=np.repeat(True, count))


def load_molecular_datasets(molecule_names, datasets_params, num_grids, base_path=None, check_grid_centering=True):
    """``dataset_loader.py:11-119``: [(Dataset, train_set)] per molecule; distances in x100 units
    under the keys ``dataset1``, ``dataset2``, ... of ``datasets_params``."""
    keys = [k for k in datasets_params if k.startswith("dataset")]
    if len(molecule_names) != len(keys):
        raise ValueError("Number of molecule names must match number of datasets in params. "
                         f"Got {len(molecule_names)} names and {len(keys)} datasets.")
    out = []
    for idx, mol in enumerate(molecule_names, 1):
        ng = num_grids[mol] if isinstance(num_grids, dict) else num_grids
        mol_path = Path(base_path) / mol if base_path is not None else Path(mol)
        try:
            dataset = Dataset(path=mol_path, num_grids=ng, name=mol)
        except Exception as e:
            raise RuntimeError(f"Failed to load dataset for {mol} from path {mol_path}") from e
        train_set = dataset.get_molecules(datasets_params[f"dataset{idx}"])
        if check_grid_centering and not location_center_at_grids_center_point(train_set.locations, dataset.grids):
            raise ValueError(f"Training set for {mol} contains examples not centered at the center of the grids.")
        out.append((dataset, train_set))
    return out


def load_molecular_datasets_from_config(config, base_path, check_grid_centering=True):
    """``dataset_loader.py:122-162``: the ``dataset*`` keys, ``molecule_names`` and ``num_grids`` of a config."""
    params = {k: v for k, v in config.items() if "dataset" in k.lower() and k != "dataset"}
    return load_molecular_datasets(config["molecule_names"], params, config["num_grids"], base_path=base_path,
                                   check_grid_centering=check_grid_centering)


def default_data_path() -> Path:
    """``$QEX_DATA_PATH`` or ``./data/od`` (the reference keeps its datasets under ``data/od``)."""
    return Path(os.environ.get("QEX_DATA_PATH", "data/od"))
