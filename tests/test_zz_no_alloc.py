"""The allocation-free non-interacting solve (qex1d_solve_noninteracting_noalloc) against the
workspace entry point (qex1d_solve_noninteracting) through the C ABI: the same kernels on the same
inputs, so every output must be bitwise equal; two calls in a row, the second smaller, so that
the fixed scratch array keeping the first call's contents cannot change the second's results."""

import ctypes

import numpy as np
import pytest

from oracle import od as O
from qex_amd import _lib

pytestmark = pytest.mark.gpu


def _case(batch, n, seed):
    grids = np.linspace(-n // 2 * 0.1, n // 2 * 0.1, n)
    mols = O.random_molecules(seed, batch, grids=grids)
    v = np.stack([O.get_atomic_chain_potential(grids, loc, z, O.exponential_coulomb) for loc, z, _ in mols])
    ne = np.array([m[2] for m in mols], dtype=np.int32)
    return grids, v, ne


def _solve(lib, torch, entry, grids, v, ne):
    dev = torch.device("cuda", 0)
    B, N = v.shape
    me = int(ne.max())
    K = me // 2 + 1
    vd = torch.as_tensor(v, dtype=torch.float64, device=dev)
    ned = torch.as_tensor(ne, dtype=torch.int32, device=dev)
    out = {"density": torch.full((B, N), np.nan, dtype=torch.float64, device=dev),
           "energy": torch.full((B,), np.nan, dtype=torch.float64, device=dev),
           "gap": torch.full((B,), np.nan, dtype=torch.float64, device=dev),
           "eig": torch.full((B, K), np.nan, dtype=torch.float64, device=dev),
           "wf": torch.full((B, K, N), np.nan, dtype=torch.float64, device=dev)}
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dx = float(grids[1] - grids[0])
    args = [B, N, dx, me, p(vd), p(ned), p(out["density"]), p(out["energy"]), p(out["gap"]), p(out["eig"]),
            p(out["wf"])]
    if entry == "workspace":
        ws = torch.empty(lib.qex1d_noninteracting_workspace_bytes(B, N, me), dtype=torch.uint8, device=dev)
        rc = lib.qex1d_solve_noninteracting(*args, p(ws), ws.numel(), stream)
        _lib.check(rc, "qex1d_solve_noninteracting")
    else:
        rc = lib.qex1d_solve_noninteracting_noalloc(*args, stream)
        _lib.check(rc, "qex1d_solve_noninteracting_noalloc")
    torch.cuda.synchronize()
    return {k: t.cpu().numpy() for k, t in out.items()}


def test_noalloc_matches_workspace_entry_point_on_two_calls():
    import torch

    lib = _lib.load()
    for batch, n, seed in ((48, 513, 21), (5, 257, 22)):  # second call smaller than the first
        grids, v, ne = _case(batch, n, seed)
        print(f"noalloc case: batch {batch}, N {n}, seed {seed}, N_e max {int(ne.max())}", flush=True)
        ref = _solve(lib, torch, "workspace", grids, v, ne)
        got = _solve(lib, torch, "noalloc", grids, v, ne)
        for key in ref:
            np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
        assert np.all(np.isfinite(got["density"]))
        dx = grids[1] - grids[0]
        np.testing.assert_allclose(got["density"].sum(axis=1) * dx, ne, atol=1e-10)


def test_noalloc_refuses_problems_beyond_its_scratch():
    lib = _lib.load()
    need = lib.qex1d_noninteracting_workspace_bytes(4096, 1024, 62)
    assert need > _lib.QEX1D_NOALLOC_SCRATCH_BYTES
    rc = lib.qex1d_solve_noninteracting_noalloc(4096, 1024, 0.1, 62, None, None, None, None, None, None, None, None)
    assert rc == _lib.QEX_E_WORKSPACE
    assert b"scratch" in lib.qex_last_error()
