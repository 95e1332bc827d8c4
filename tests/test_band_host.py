"""The eigen routines of the gfx950 kernels (qex_amd/csrc/band.h), built for the host with hipcc and
checked against LAPACK on random 1D Kohn-Sham Hamiltonians (CPU only)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import od as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("band") / "band_harness")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "qex_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "band_harness.hip"), "-o", exe], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


def _run(exe, v, k, dx, mode="1"):
    inp = f"{len(v)} {k} {dx:.17g}\n" + " ".join(f"{x:.17g}" for x in v)
    out = subprocess.run([exe, mode], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    lam = np.array([float(out[2 * j]) for j in range(k)])
    vec = np.array([[float(t) for t in out[2 * j + 1].split()] for j in range(k)])
    if mode == "2":
        # shifts whose two-sided Sturm count differs from the one-sided count
        assert int(out[2 * k]) == 0
    return lam, vec


@pytest.mark.parametrize("mode", ["1", "2"])
def test_lowest_eigenpairs_match_lapack(harness, mode):
    """mode 1: one-sided LDL^T; mode 2: the two-sided elimination (top and bottom halves meeting at
    the middle pair) that the device eigen kernel uses, odd and even grid sizes."""
    rng = np.random.default_rng(0)
    for trial in range(24):
        grids = O.h2_config_grid() if trial % 2 else np.linspace(-10, 10, 201 + trial % 3)
        dx = O.get_dx(grids)
        A = int(rng.integers(1, 5))
        loc, Z = np.sort(rng.uniform(-4, 4, A)), rng.integers(1, 5, A)
        if trial % 6 == 0:
            loc, Z = np.array([-4.0, 4.0]), np.array([2, 2])  # nearly degenerate pairs
        v = O.get_atomic_chain_potential(grids, loc, Z, O.exponential_coulomb)
        if trial % 4 == 1:
            v = v + rng.normal(0, 0.2, len(v))
        lam, vec = _run(harness, v, 5, dx, mode)
        w, V = np.linalg.eigh(O.get_kinetic_matrix(grids) + np.diag(v))
        np.testing.assert_allclose(lam, w[:5], rtol=0, atol=1e-11)
        for j in range(5):
            gap = min(abs(w[j] - w[i]) for i in range(6) if i != j)
            if gap > 1e-6:
                err = min(np.linalg.norm(vec[j] - V[:, j]), np.linalg.norm(vec[j] + V[:, j]))
                assert err < 1e-8, (trial, j, err, gap)
