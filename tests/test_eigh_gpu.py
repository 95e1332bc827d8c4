"""Device generalized eigensolver (qex3d_gen_eigh) against LAPACK (scipy.linalg.eigh), including
exact degeneracies and matrices whose tridiagonal form splits."""

import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def geigh():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd.td.linalg import generalized_eigh
    return generalized_eigh


def _spd(n, rng, cond=10.0):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return Q @ np.diag(np.geomspace(1.0, 1.0 / cond, n)) @ Q.T


def _check(w, v, A, B, tol=1e-10):
    w0 = scipy.linalg.eigh(A, B, eigvals_only=True)
    scale = max(1.0, np.abs(w0).max())
    assert np.abs(w - w0).max() < tol * scale
    n = len(w)
    np.testing.assert_allclose(v.T @ B @ v, np.eye(n), atol=tol * 10)
    np.testing.assert_allclose(A @ v, B @ v * w, atol=tol * 10 * scale)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 40, 80, 96])
def test_random_pairs(geigh, n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M + M.T
    B = _spd(n, rng)
    w, v = geigh(A, B)
    _check(w.cpu().numpy(), v.cpu().numpy(), A, B)


def test_batched_matches_single(geigh):
    rng = np.random.default_rng(7)
    n, b = 24, 5
    As = np.stack([(lambda M: M + M.T)(rng.standard_normal((n, n))) for _ in range(b)])
    Bs = np.stack([_spd(n, rng) for _ in range(b)])
    w, v = geigh(As, Bs)
    for k in range(b):
        w1, v1 = geigh(As[k], Bs[k])
        assert torch.equal(w[k], w1) and torch.equal(v[k], v1)  # bitwise: one workgroup per matrix
        _check(w[k].cpu().numpy(), v[k].cpu().numpy(), As[k], Bs[k])


def test_exact_degeneracies_and_splits(geigh):
    rng = np.random.default_rng(3)
    n = 12
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag([1, 1, 2, 3, 3, 3, 4, 5, 6, 6, 7, 8.0]) @ Q.T
    A = 0.5 * (A + A.T)
    w, v = geigh(A, np.eye(n))
    _check(w.cpu().numpy(), v.cpu().numpy(), A, np.eye(n))
    blk = rng.standard_normal((4, 4))
    blk = blk + blk.T
    A = scipy.linalg.block_diag(blk, blk, np.array([[3.0]]), blk)
    B = scipy.linalg.block_diag(*[_spd(4, rng)] * 2, np.eye(1), _spd(4, rng))
    w, v = geigh(A, B)
    _check(w.cpu().numpy(), v.cpu().numpy(), A, B)
    for A in (np.diag(rng.standard_normal(9)), np.eye(6), np.zeros((5, 5))):
        w, v = geigh(A, np.eye(len(A)))
        _check(w.cpu().numpy(), v.cpu().numpy(), A, np.eye(len(A)))


def test_linear_chain_fock_like(geigh):
    """p_x / p_y pairs of a linear chain are exactly degenerate (the 3D SCF case)."""
    from oracle import td as otd
    mol = otd.synthetic_chain(6, 2)
    S, T, V = otd.one_electron(mol)
    w, v = geigh(T + V, S)
    _check(w.cpu().numpy(), v.cpu().numpy(), T + V, S)
