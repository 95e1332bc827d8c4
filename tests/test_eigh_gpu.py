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


@pytest.mark.parametrize("exact", [False, True])
def test_degen_eigh_bwd_matches_oracle(exact):
    """The device reverse pass of degen_eigh (generalized_eigensolver.py:158-219) against the oracle on a
    batch of sizes up to the limit, with eigenvalue and eigenvector cotangents alone and together, and
    a matrix with an exactly degenerate pair."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import td as otd
    from qex_amd.td import linalg
    rng = np.random.default_rng(11)
    for n in (5, 40, 96):
        A = rng.standard_normal((3, n, n))
        A = 0.5 * (A + A.transpose(0, 2, 1))
        Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
        d = np.sort(rng.standard_normal(n))
        d[1] = d[2]  # exact pair
        A[2] = Q @ np.diag(d) @ Q.T
        (w, v), res = linalg.degen_eigh_fwd(torch.tensor(A))
        gw, gv = rng.standard_normal((3, n)), rng.standard_normal((3, n, n))
        for cw, cv in ((gw, None), (None, gv), (gw, gv)):
            got = linalg.degen_eigh_bwd(res, (None if cw is None else torch.tensor(cw),
                                              None if cv is None else torch.tensor(cv)), exact=exact)[0]
            got = got.cpu().numpy()
            for b in range(3):
                ref = otd.degen_eigh_bwd(w[b].cpu().numpy(), v[b].cpu().numpy(), None if cw is None else cw[b],
                                         None if cv is None else cv[b], exact=exact)
                scale = max(1.0, np.abs(ref).max())
                assert np.abs(got[b] - ref).max() <= 1e-11 * scale, (n, b, exact, np.abs(got[b] - ref).max())
                assert np.all(np.isfinite(got[b]))


def test_degen_eigh_autograd():
    """degen_eigh as a torch.autograd function: A.grad of c.w + <G, V> equals the device reverse pass and
    the oracle's exact form; eigenvalues agree with LAPACK."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import td as otd
    from qex_amd.td import linalg
    rng = np.random.default_rng(12)
    n = 24
    A0 = rng.standard_normal((n, n))
    A0 = 0.5 * (A0 + A0.T)
    A = torch.tensor(A0, device="cuda", requires_grad=True)
    w, v = linalg.degen_eigh(A, exact=True)
    np.testing.assert_allclose(w.detach().cpu().numpy(), np.linalg.eigvalsh(A0), rtol=0, atol=1e-12)
    c, G = rng.standard_normal(n), rng.standard_normal((n, n))
    loss = (w * torch.tensor(c, device="cuda")).sum() + (v * torch.tensor(G, device="cuda")).sum()
    loss.backward()
    ref = otd.degen_eigh_bwd(w.detach().cpu().numpy(), v.detach().cpu().numpy(), c, G, exact=True)
    np.testing.assert_allclose(A.grad.cpu().numpy(), ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))
