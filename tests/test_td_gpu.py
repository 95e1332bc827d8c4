"""3D RKS/RHF device path against the FP64 oracle (oracle/td.py), on small molecules.

Integrals, grid, AO values, density and LDA potential matrix are compared elementwise; SCF runs
are compared by total energy (1e-8 Ha), density matrix and orbital energies.  H2/6-31G RHF
energies are the values printed in notebooks/03_notebook_td_dataset.ipynb:40 and
notebooks/04_notebook_td_trainer.ipynb (cell 5).
"""

import numpy as np
import pytest
import torch

from oracle import td as otd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qtd():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd import td
    return td


def _mol_pair(qtd, kind, arg):
    if kind == "h2":
        from qex_amd.td.gto import Mole
        return (Mole([("H", (0, 0, 0)), ("H", (0, 0, arg))], basis="6-31g"), otd.h2_molecule(arg))
    return qtd.gto.synthetic_chain(arg, 7), otd.synthetic_chain(arg, 7)


def _pairs_to_full(gp, n):
    idx = np.array([[max(i, j) * (max(i, j) + 1) // 2 + min(i, j) for j in range(n)] for i in range(n)])
    return gp[idx[:, :, None, None], idx[None, None, :, :]]


@pytest.mark.parametrize("kind,arg", [("h2", 0.74), ("chain", 3), ("chain", 4)])
def test_one_electron_integrals(qtd, kind, arg):
    mol, omol = _mol_pair(qtd, kind, arg)
    S, T, V = (x.cpu().numpy() for x in mol.intor_1e())
    oS, oT, oV = otd.one_electron(omol)
    for a, b in ((S, oS), (T, oT), (V, oV)):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * max(1.0, np.abs(b).max()))


@pytest.mark.parametrize("kind,arg", [("h2", 0.74), ("chain", 2), ("chain", 3)])
def test_eri_pair_matrix(qtd, kind, arg):
    mol, omol = _mol_pair(qtd, kind, arg)
    g = mol.intor_eri_pairs().cpu().numpy()
    full = _pairs_to_full(g, mol.nao)
    ref = otd.eri(omol)
    np.testing.assert_allclose(full, ref, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(g, g.T)


def test_grid_matches_oracle(qtd):
    mol, omol = _mol_pair(qtd, "chain", 4)
    grids = qtd.dft.Grids(mol).build()
    pts, w = otd.molecular_grid(omol)
    np.testing.assert_allclose(grids.coords.cpu().numpy(), pts, rtol=0, atol=1e-12)
    np.testing.assert_allclose(grids.weights.cpu().numpy(), w, rtol=1e-11, atol=1e-13 * w.max())


def test_stratmann_grid_matches_oracle(qtd):
    """grids.becke_scheme = stratmann (trainer_legacy_no_jit.py:250) on the device against the
    oracle's Stratmann partition (a different partition from Becke's); other schemes are refused."""
    mol, omol = _mol_pair(qtd, "chain", 4)
    grids = qtd.dft.Grids(mol)
    grids.becke_scheme = qtd.dft.stratmann
    grids.build()
    pts, w = otd.molecular_grid(omol, scheme="stratmann")
    np.testing.assert_allclose(grids.coords.cpu().numpy(), pts, rtol=0, atol=1e-12)
    np.testing.assert_allclose(grids.weights.cpu().numpy(), w, rtol=1e-11, atol=1e-13 * w.max())
    wb = qtd.dft.Grids(mol).build().weights.cpu().numpy()
    assert np.max(np.abs(wb - w)) > 1e-6  # a different partition
    grids.becke_scheme = print
    with pytest.raises(ValueError, match="becke_scheme"):
        grids.build()


def test_lebedev_grid_and_scf_energy_match_oracle(qtd):
    """A pyscf-level-0-like grid (10 Treutler radial points x 50-point Lebedev, Stratmann cells)
    built on the device against the oracle's, and the RKS energy of the H4 chain on it."""
    mol, omol = _mol_pair(qtd, "chain", 4)
    grids = qtd.dft.Grids(mol, n_rad=10, n_ang=50)
    grids.becke_scheme = qtd.dft.stratmann
    grids.build()
    pts, w = otd.molecular_grid(omol, n_rad=10, n_ang=50, scheme="stratmann")
    assert grids.size == 4 * 10 * 50
    np.testing.assert_allclose(grids.coords.cpu().numpy(), pts, rtol=0, atol=1e-12)
    np.testing.assert_allclose(grids.weights.cpu().numpy(), w, rtol=1e-11, atol=1e-13 * w.max())
    mf = qtd.scf.RKS(mol)
    mf.grids = grids
    e = mf.kernel()
    ref = otd.scf(omol, "rks", grid=(pts, w), conv_tol=1e-10, max_cycle=60)
    assert abs(e - ref.e_tot) < 1e-8, (e, ref.e_tot)


def test_water_sto3g_rhf_and_rks_match_oracle(qtd):
    """A first-row molecule: H2O/STO-3G (O with s and p shells, Z = 8) at the experimental geometry.
    RHF on the device against the oracle and the standard -74.962928 Ha; RKS-LDA on the device
    grid against the oracle's SCF on the same grid."""
    import math

    a, r = math.radians(104.52), 0.9572
    h1 = (r * math.sin(a / 2), 0.0, r * math.cos(a / 2))
    h2 = (-r * math.sin(a / 2), 0.0, r * math.cos(a / 2))
    mol = qtd.gto.Mole([("O", (0.0, 0.0, 0.0)), ("H", h1), ("H", h2)], basis="sto-3g")
    omol = otd.water_molecule(otd.sto3g_shells(7.66, 2.25), otd.STO3G_H)
    assert mol.nao == 7
    e_rhf = qtd.scf.RHF(mol).kernel()
    ref = otd.scf(omol, method="rhf", conv_tol=1e-12, max_cycle=100)
    assert abs(e_rhf - ref.e_tot) < 1e-8 and abs(e_rhf - -74.962928) < 2e-6, (e_rhf, ref.e_tot)
    e_rks = qtd.scf.RKS(mol).kernel()
    ref = otd.scf(omol, "rks", grid=otd.molecular_grid(omol), conv_tol=1e-10, max_cycle=60)
    assert abs(e_rks - ref.e_tot) < 1e-8, (e_rks, ref.e_tot)


def _random_dm(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n, n)) * 0.3
    return c @ c.T


@pytest.mark.parametrize("n_atoms", [1, 4, 9])
def test_eval_ao_rho_nr_rks(qtd, n_atoms):
    mol = qtd.gto.synthetic_chain(n_atoms, 11) if n_atoms > 1 else None
    omol = otd.synthetic_chain(n_atoms, 11) if n_atoms > 1 else None
    if n_atoms == 1:
        from qex_amd.td.gto import Mole
        mol = Mole([("H", (0, 0, 0.1))], basis="6-31g")
        omol = otd.build_molecule(np.array([[0, 0, 0.1 / otd.BOHR]]), [1], [otd.B631G_H])
    grids = qtd.dft.Grids(mol).build()
    pts = grids.coords.cpu().numpy()
    w = grids.weights.cpu().numpy()
    ao = qtd.numint.eval_ao(mol, grids.coords)
    oao = otd.eval_ao(omol, pts)
    np.testing.assert_allclose(ao.cpu().numpy(), oao, rtol=0, atol=1e-13 * np.abs(oao).max())
    dm = _random_dm(mol.nao, n_atoms)
    rho = qtd.numint.eval_rho(mol, ao, dm).cpu().numpy()
    orho = otd.eval_rho(oao, dm)
    np.testing.assert_allclose(rho, orho, rtol=1e-11, atol=1e-13 * np.abs(orho).max())
    # pyscf's "lda,vwn" (Slater + VWN5) and bare "lda" (Slater exchange alone)
    for code, fn in (("lda,vwn", "svwn5"), ("lda", "slater")):
        nel, exc, vmat = qtd.numint.nr_rks(qtd.numint.NumInt(), mol, grids, code, dm)
        onel, oexc, ovmat = otd.nr_rks_lda(oao, w, dm, fn)
        assert abs(float(nel) - onel) < 1e-10 * max(1, abs(onel))
        assert abs(float(exc) - oexc) < 1e-10 * max(1, abs(oexc))
        np.testing.assert_allclose(vmat.cpu().numpy(), ovmat, rtol=0, atol=1e-10 * np.abs(ovmat).max())


def test_eval_rho_nonsymmetric_dm_is_symmetrised(qtd):
    mol, omol = _mol_pair(qtd, "chain", 2)
    grids = qtd.dft.Grids(mol).build()
    ao = qtd.numint.eval_ao(mol, grids.coords)
    rng = np.random.default_rng(3)
    dm = rng.standard_normal((mol.nao, mol.nao))
    rho = qtd.numint.eval_rho(mol, ao, dm).cpu().numpy()
    orho = otd.eval_rho(ao.cpu().numpy(), dm)
    np.testing.assert_allclose(rho, orho, rtol=1e-10, atol=1e-12 * np.abs(orho).max())


@pytest.mark.parametrize("r,energy", [(0.74, -1.12675531719693), (0.5, -1.05802481296927),
                                      (1.5, -0.997497294328357)])
def test_h2_631g_rhf_notebook_energies(qtd, r, energy):
    mol, _ = _mol_pair(qtd, "h2", r)
    mf = qtd.scf.RHF(mol)
    mf.conv_tol = 1e-11
    e = mf.kernel()
    assert mf.converged
    assert abs(e - energy) < 1e-9


@pytest.mark.parametrize("method", ["rks", "rhf"])
@pytest.mark.parametrize("n_atoms", [2, 4])
def test_scf_matches_oracle(qtd, method, n_atoms):
    mol, omol = _mol_pair(qtd, "chain", n_atoms)
    mf = qtd.scf.RKS(mol) if method == "rks" else qtd.scf.RHF(mol)
    mf.conv_tol = 1e-10
    e = mf.kernel()
    S, H, g = mf.integrals()
    ints = (S.cpu().numpy(), (H.cpu().numpy()), None)
    grid = None
    if method == "rks":
        grid = (mf.grids.coords.cpu().numpy(), mf.grids.weights.cpu().numpy())
    oS, oT, oV = otd.one_electron(omol)
    ref = otd.scf(omol, method, grid=grid, ints=(oS, oT, oV, _pairs_to_full(g.cpu().numpy(), mol.nao)),
                  conv_tol=1e-10, max_cycle=50)
    assert ref.converged and mf.converged
    assert abs(e - ref.e_tot) < 1e-8
    np.testing.assert_allclose(mf.dm.cpu().numpy(), ref.dm, rtol=0, atol=1e-6)
    nocc = mol.nelectron // 2
    np.testing.assert_allclose(mf.mo_energy.cpu().numpy()[:nocc + 1], ref.mo_energy[:nocc + 1], atol=1e-6)
    # per-cycle energies follow the same trajectory while both run
    ne = min(len(ref.energies), mf.cycles)
    np.testing.assert_allclose(mf.energies.cpu().numpy()[1:1 + ne], ref.energies[:ne], atol=1e-7)


def test_scf_larger_chain_against_oracle_with_device_integrals(qtd):
    """16 atoms, 80 AOs (near the single-workgroup eigensolver limit of 96)."""
    mol = qtd.gto.synthetic_chain(16, 5)
    omol = otd.synthetic_chain(16, 5)
    mf = qtd.scf.RKS(mol)
    mf.conv_tol = 1e-9
    mf.max_cycle = 60
    e = mf.kernel()
    S, H, g = mf.integrals()
    oS, oT, oV = otd.one_electron(omol)
    np.testing.assert_allclose(S.cpu().numpy(), oS, atol=1e-12)
    grid = (mf.grids.coords.cpu().numpy(), mf.grids.weights.cpu().numpy())
    ref = otd.scf(omol, "rks", grid=grid, ints=(oS, oT, oV, _pairs_to_full(g.cpu().numpy(), mol.nao)),
                  conv_tol=1e-9, max_cycle=60)
    assert ref.converged == mf.converged
    assert abs(e - ref.e_tot) < 1e-7


def test_scf_invariants(qtd):
    mol, _ = _mol_pair(qtd, "chain", 4)
    mf = qtd.scf.RKS(mol)
    mf.kernel()
    S, _, _ = mf.integrals()
    C = mf.mo_coeff
    eye = (C.T @ S @ C).cpu().numpy()
    np.testing.assert_allclose(eye, np.eye(mol.nao), atol=1e-10)
    D = mf.dm
    assert abs(float((D * S).sum()) - mol.nelectron) < 1e-10
    mo = mf.mo_energy.cpu().numpy()
    assert np.all(np.diff(mo) >= -1e-12)


def _nets(ng, hidden, seed):
    from qex_amd.td.xc import GlobalNNXC
    W, b = otd.init_global_mlp(seed, ng, hidden=hidden)
    rng = np.random.default_rng(seed + 1)
    b = [rng.standard_normal(x.shape) * 0.1 for x in b]
    return GlobalNNXC(W, b, scale=0.05), otd.GlobalMLP3D(W, b, scale=0.05)


@pytest.mark.parametrize("hidden", [(64, 32), (512, 512)])
def test_nr_rks_global_network_matches_oracle(qtd, hidden):
    mol, omol = _mol_pair(qtd, "chain", 4)
    grids = qtd.dft.Grids(mol).build()
    w = grids.weights.cpu().numpy()
    ao = qtd.numint.eval_ao(mol, grids.coords)
    oao = ao.cpu().numpy()
    net, onet = _nets(grids.size, hidden, 3)
    dm = _random_dm(mol.nao, 5)
    ni = qtd.numint.NumInt().define_xc_(net)
    nel, exc, vmat = qtd.numint.nr_rks(ni, mol, grids, "NN", dm)
    onel, oexc, ovmat = otd.nr_rks_nn_global(oao, w, dm, onet)
    assert abs(float(nel) - onel) < 1e-10 * abs(onel)
    assert abs(float(exc) - oexc) < 1e-11 * max(1.0, abs(oexc))
    np.testing.assert_allclose(vmat.cpu().numpy(), ovmat, rtol=0, atol=1e-10 * np.abs(ovmat).max())


def test_rks_global_network_scf_matches_oracle(qtd):
    mol, omol = _mol_pair(qtd, "chain", 4)
    mf = qtd.scf.RKS(mol)
    mf.grids = qtd.dft.Grids(mol).build()
    net, onet = _nets(mf.grids.size, (64, 32), 9)
    mf.define_xc_(net)
    mf.conv_tol = 1e-10
    e = mf.kernel()
    S, H, g = mf.integrals()
    oS, oT, oV = otd.one_electron(omol)
    grid = (mf.grids.coords.cpu().numpy(), mf.grids.weights.cpu().numpy())
    ref = otd.scf(omol, "rks_nn", grid=grid, ints=(oS, oT, oV, _pairs_to_full(g.cpu().numpy(), mol.nao)),
                  conv_tol=1e-10, max_cycle=50, net=onet)
    assert ref.converged and mf.converged
    assert abs(e - ref.e_tot) < 1e-8
    np.testing.assert_allclose(mf.dm.cpu().numpy(), ref.dm, atol=1e-6)
    # new parameters through kernel(params=...) (no rebuild of the molecule or grid)
    W2, b2 = otd.init_global_mlp(10, mf.grids.size, hidden=(64, 32))
    e2 = mf.kernel(params=list(zip(W2, b2)))
    ref2 = otd.scf(omol, "rks_nn", grid=grid, ints=(oS, oT, oV, _pairs_to_full(g.cpu().numpy(), mol.nao)),
                   conv_tol=1e-10, max_cycle=50, net=otd.GlobalMLP3D(W2, b2, scale=0.05))
    assert abs(e2 - ref2.e_tot) < 1e-8


def _local_nets(hidden, seed):
    from qex_amd.td.xc import LocalNNXC
    W, b = otd.init_global_mlp(seed, 1, hidden=hidden)
    rng = np.random.default_rng(seed + 2)
    b = [rng.standard_normal(x.shape) * 0.1 for x in b]
    return LocalNNXC(W, b, scale=0.2), otd.LocalMLP3D(W, b, scale=0.2)


@pytest.mark.parametrize("hidden", [(16, 8), (512, 512)])
def test_nr_rks_local_network_matches_oracle(qtd, hidden):
    """The "NN" branch (td/numint.py:179-204) with the per-point network on the device, installed the
    way the reference's trainer does (define_xc_(description=partial(eval_xc, network=...),
    xctype="NN"))."""
    import functools
    mol, omol = _mol_pair(qtd, "chain", 4)
    grids = qtd.dft.Grids(mol).build()
    w = grids.weights.cpu().numpy()
    ao = qtd.numint.eval_ao(mol, grids.coords)
    net, onet = _local_nets(hidden, 5)
    dm = _random_dm(mol.nao, 6)
    desc = functools.partial(qtd.xc.eval_xc, network=net, is_global_xc=False)
    ni = qtd.numint.NumInt().define_xc_(description=desc, xctype="NN")
    nel, exc, vmat = qtd.numint.nr_rks(ni, mol, grids, "NN", dm)
    onel, oexc, ovmat = otd.nr_rks_nn_local(ao.cpu().numpy(), w, dm, onet)
    assert abs(float(nel) - onel) < 1e-10 * abs(onel)
    assert abs(float(exc) - oexc) < 1e-11 * max(1.0, abs(oexc))
    np.testing.assert_allclose(vmat.cpu().numpy(), ovmat, rtol=0, atol=1e-10 * np.abs(ovmat).max())
    # eval_xc itself: per-point exc and vrho
    rho = qtd.numint.eval_rho(mol, ao, dm)
    e, (v, _, _, _), _, _ = qtd.xc.eval_xc("NN", rho, network=net, is_global_xc=False)
    oe, ov = onet.exc_vrho(rho.cpu().numpy())
    np.testing.assert_allclose(e.cpu().numpy(), oe, rtol=0, atol=1e-12 * np.abs(oe).max())
    np.testing.assert_allclose(v.cpu().numpy(), ov, rtol=0, atol=1e-12 * np.abs(ov).max())


def test_eval_xc_global_network_matches_oracle(qtd):
    mol, _ = _mol_pair(qtd, "chain", 2)
    grids = qtd.dft.Grids(mol).build()
    net, onet = _nets(grids.size, (64, 32), 4)
    rho = torch.rand(grids.size, dtype=torch.float64, device="cuda") * 0.3
    e, (v, _, _, _), _, _ = qtd.xc.eval_xc("NN-AmplitudeEncoding", rho, network=net, is_global_xc=True)
    oe, ov = onet.exc_vrho(rho.cpu().numpy())
    assert abs(float(e) - oe) < 1e-11 * max(1.0, abs(oe))
    np.testing.assert_allclose(v.cpu().numpy(), ov, rtol=0, atol=1e-12 * np.abs(ov).max())


def test_rks_local_network_scf_matches_oracle(qtd):
    import functools
    mol, omol = _mol_pair(qtd, "chain", 4)
    mf = qtd.scf.RKS(mol)
    mf.grids = qtd.dft.Grids(mol).build()
    net, onet = _local_nets((32, 16), 7)
    mf = mf.define_xc_(description=functools.partial(qtd.xc.eval_xc, network=net, is_global_xc=False), xctype="NN")
    mf.conv_tol, mf.max_cycle = 1e-10, 60
    e = mf.kernel()
    S, H, g = mf.integrals()
    oS, oT, oV = otd.one_electron(omol)
    grid = (mf.grids.coords.cpu().numpy(), mf.grids.weights.cpu().numpy())
    ref = otd.scf(omol, "rks_nn_local", grid=grid, ints=(oS, oT, oV, _pairs_to_full(g.cpu().numpy(), mol.nao)),
                  conv_tol=1e-10, max_cycle=60, net=onet)
    assert ref.converged and mf.converged
    assert abs(e - ref.e_tot) < 1e-8
    np.testing.assert_allclose(mf.make_rdm1(a0_repr=True).cpu().numpy(), ref.dm, atol=1e-6)
    # kernel(params=...) replaces the parameters, as the trainer's batch_loss_fn does
    W2, b2 = otd.init_global_mlp(8, 1, hidden=(32, 16))
    e2 = mf.kernel(params=list(zip(W2, b2)))
    ref2 = otd.scf(omol, "rks_nn_local", grid=grid, ints=(oS, oT, oV, _pairs_to_full(g.cpu().numpy(), mol.nao)),
                   conv_tol=1e-10, max_cycle=60, net=otd.LocalMLP3D(W2, b2, scale=0.2))
    assert abs(e2 - ref2.e_tot) < 1e-8


def test_generalized_eigh_shifts_a_singular_metric(qtd):
    """generalized_eigh (generalized_eigensolver.py:264-330): symmetrised inputs, B shifted by
    eps - lambda_min(B) when lambda_min(B) < eps, optional Jacobi scaling; against scipy on the
    same shifted problem."""
    import scipy.linalg
    rng = np.random.default_rng(5)
    n = 12
    X = rng.standard_normal((n, n - 2))
    B = X @ X.T  # rank n - 2: two zero eigenvalues
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, v = qtd.linalg.generalized_eigh(A, B, eps=1e-6)
    Bs = B + (1e-6 - np.linalg.eigvalsh(B)[0]) * np.eye(n)
    ref = scipy.linalg.eigh(A, Bs, eigvals_only=True)
    np.testing.assert_allclose(w.cpu().numpy(), ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
    V = v.cpu().numpy()
    np.testing.assert_allclose(V.T @ Bs @ V, np.eye(n), atol=1e-6)
    # a well-conditioned B is not shifted; Jacobi scaling leaves the eigenvalues unchanged
    S = X @ X.T + np.eye(n)
    w1, _ = qtd.linalg.generalized_eigh(A, S)
    w2, _ = qtd.linalg.generalized_eigh(A, S, scale=True)
    ref = scipy.linalg.eigh(A, S, eigvals_only=True)
    np.testing.assert_allclose(w1.cpu().numpy(), ref, atol=1e-10 * np.abs(ref).max())
    np.testing.assert_allclose(w2.cpu().numpy(), ref, atol=1e-9 * np.abs(ref).max())
