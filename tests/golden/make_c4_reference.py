"""Reference energies of the 3D LDA chain workload from the FP64 oracle (oracle/td.py).

Hydrogen chains H2, H4, H8, H16 (oracle.td.synthetic_chain, seed = 100 + atoms): bonds U[0.5, 1.5]
Angstrom, per atom 2 s + 1 p contracted shells of 3 primitives (5 AOs/atom), grid 50 radial x 110
angular points per atom with Becke partition, Slater + VWN5, core-Hamiltonian guess, CDIIS(8),
at most 30 cycles, |dE| < 1e-8.  Writes c4_chain_reference.json beside this script.

    python tests/golden/make_c4_reference.py [--procs 8]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import td  # noqa: E402

CHAINS = (2, 4, 8, 16)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=8)
    args = ap.parse_args()
    out = []
    for n in CHAINS:
        mol = td.synthetic_chain(n, 100 + n)
        t = time.perf_counter()
        S, T, V = td.one_electron(mol)
        g = td.eri(mol, procs=args.procs)
        grid = td.molecular_grid(mol)
        r = td.scf(mol, "rks", grid=grid, ints=(S, T, V, g), max_cycle=30, conv_tol=1e-8, diis_space=8)
        dt = time.perf_counter() - t
        out.append({"n_atoms": n, "seed": 100 + n, "nao": mol.nao, "ngrid": int(len(grid[1])),
                    "e_tot": r.e_tot, "converged": bool(r.converged), "cycles": int(r.cycles),
                    "nelec_grid": r.nelec_grid, "oracle_seconds": dt})
        print(out[-1], flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_chain_reference.json")
    with open(path, "w") as f:
        json.dump({"workload": "LDA hydrogen chains, 5 AOs/atom, 50x110 grid/atom, CDIIS(8), <=30 cycles, "
                               "conv_tol 1e-8", "chains": out}, f, indent=1)


if __name__ == "__main__":
    main()
