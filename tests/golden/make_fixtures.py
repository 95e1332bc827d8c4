"""Writes the small golden fixtures under tests/golden/ from the reference's own data files.

Run once in a container where the reference tree is mounted:

    python tests/golden/make_fixtures.py /root/reference

Nothing from the reference is executed or unpickled:
* ``data/od/<mol>/*.npy`` are read with ``numpy.load(allow_pickle=False)``;
* ``qedft/models/initial_density.pkl`` holds one float64 array of shape (2, 513)
  (the non-interacting initial densities of H2 at distances x100 = 128 and 384,
  i.e. nuclei at +-0.64 and +-1.92 bohr, written by ``od/trainer.py:72-75``).
  Its raw little-endian buffer is located with ``pickletools.genops``, which only
  walks the opcode stream and never builds objects; the shape (2, 513) and dtype
  ``<f8`` are read from the same opcode stream and checked.

Outputs (all float64 unless noted):
* ``od_h2_rows.npz``: ``grids`` (513,), and for dataset rows ``ROWS`` of data/od/h2:
  ``locations`` (R, 2), ``nuclear_charges`` (R, 2), ``external_potentials`` (R, 513),
  ``distances_x100`` (R,) int64.
* ``od_h4_rows.npz``, ``od_h2_h2_rows.npz``: the same for the 4-nucleus sets.
* ``initial_density_h2_128_384.npy``: (2, 513).
* ``od_h2_train/``: the jax_dft dataset layout (one ``.npy`` per field, as ``data/od/h2/``) restricted to
  the rows at distances x100 = 128 and 384 (the reference trainer's ``dataset1``,
  ``qedft/config/train_config.yaml``): ``densities`` (2, 513), ``total_energies`` (2,),
  ``external_potentials``, ``locations``, ``nuclear_charges`` (int32), ``distances``,
  ``distances_x100`` (int64), ``num_electrons``, ``grids``.
"""

import os
import pickletools
import sys

import numpy as np

ROWS = {"h2": [16, 57], "h4": [0, 40], "h2_h2": [5, 30]}


def _extract_float_buffer(path, nbytes):
    data = open(path, "rb").read()
    blobs = []
    ints = []
    for op, arg, _pos in pickletools.genops(data):
        if op.name in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8") and len(arg) == nbytes:
            blobs.append(arg)
        if op.name in ("BININT1", "BININT2", "BININT"):
            ints.append(arg)
        if op.name == "SHORT_BINUNICODE" and arg not in (
            "jax._src.array", "_reconstruct_array", "numpy._core.multiarray", "numpy.core.multiarray",
            "_reconstruct", "numpy", "ndarray", "dtype", "f8", "<", "weak_type",
        ):
            raise ValueError(f"unexpected name in pickle stream: {arg!r}")
    if len(blobs) != 1:
        raise ValueError(f"expected one {nbytes}-byte buffer, found {len(blobs)}")
    if 2 not in ints or 513 not in ints:
        raise ValueError("shape (2, 513) not found in the opcode stream")
    return np.frombuffer(blobs[0], dtype="<f8").reshape(2, 513).copy()


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for mol, rows in ROWS.items():
        d = os.path.join(ref, "data", "od", mol)
        ld = lambda f: np.load(os.path.join(d, f), allow_pickle=False)
        out = dict(
            grids=ld("grids.npy"),
            locations=ld("locations.npy")[rows],
            nuclear_charges=ld("nuclear_charges.npy")[rows].astype(np.int64),
            external_potentials=ld("external_potentials.npy")[rows],
            distances_x100=ld("distances_x100.npy")[rows].astype(np.int64),
            num_electrons=np.int64(ld("num_electrons.npy")),
        )
        np.savez(os.path.join(here, f"od_{mol}_rows.npz"), **out)
    d = os.path.join(ref, "data", "od", "h2")
    ld = lambda f: np.load(os.path.join(d, f), allow_pickle=False)
    sel = np.isin(ld("distances_x100.npy"), [128, 384])
    out_dir = os.path.join(here, "od_h2_train")
    os.makedirs(out_dir, exist_ok=True)
    for f in ("densities", "total_energies", "external_potentials", "locations", "nuclear_charges", "distances",
              "distances_x100"):
        np.save(os.path.join(out_dir, f + ".npy"), ld(f + ".npy")[sel])
    np.save(os.path.join(out_dir, "num_electrons.npy"), ld("num_electrons.npy"))
    np.save(os.path.join(out_dir, "grids.npy"), ld("grids.npy"))
    dens = _extract_float_buffer(os.path.join(ref, "qedft", "models", "initial_density.pkl"), 2 * 513 * 8)
    np.save(os.path.join(here, "initial_density_h2_128_384.npy"), dens)
    print("fixtures written to", here)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
