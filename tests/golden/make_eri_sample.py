"""Sampled ERIs of the 8-atom hydrogen chain of the 3D chain workload, from the FP64 oracle.

oracle.td.synthetic_chain(8, seed=108) (40 AOs): the full (ij|kl) tensor is 20 MB, so the fixture
keeps 4096 seeded random index quadruples plus every Coulomb-type (ii|jj) entry, with values.
Writes eri_chain8_sample.npz beside this script.

    python tests/golden/make_eri_sample.py [--procs 8]
"""

import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import td  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=8)
    args = ap.parse_args()
    mol = td.synthetic_chain(8, 108)
    g = td.eri(mol, procs=args.procs)
    n = mol.nao
    rng = np.random.default_rng(8)
    idx = rng.integers(0, n, size=(4096, 4))
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    coul = np.stack([ii.ravel(), ii.ravel(), jj.ravel(), jj.ravel()], axis=1)
    idx = np.concatenate([idx, coul]).astype(np.int32)
    vals = g[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eri_chain8_sample.npz")
    np.savez(path, index=idx, value=vals, nao=np.int32(n), seed=np.int32(108), n_atoms=np.int32(8))
    print(path, idx.shape, float(np.abs(vals).max()))


if __name__ == "__main__":
    main()
