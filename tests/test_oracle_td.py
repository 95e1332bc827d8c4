"""Pins of the 3D FP64 oracle (CPU): textbook integrals, the reference notebook's PySCF energies,
finite differences and grid quadrature checks."""

import math

import numpy as np
import pytest
import scipy.linalg

from oracle import td as T


def _h2_sto3g():
    coords = np.array([[0, 0, 0], [0, 0, 1.4]])
    return T.build_molecule(coords, [1, 1], [T.STO3G_H, T.STO3G_H])


def test_szabo_ostlund_h2_sto3g_integrals():
    """Szabo & Ostlund, Modern Quantum Chemistry, Sec. 3.5.2 (R = 1.4 bohr, zeta = 1.24)."""
    mol = _h2_sto3g()
    S, Tk, V = T.one_electron(mol)
    g = T.eri(mol)
    assert S[0, 1] == pytest.approx(0.6593, abs=1e-4)
    assert Tk[0, 0] == pytest.approx(0.7600, abs=1e-4)
    assert Tk[0, 1] == pytest.approx(0.2365, abs=1e-4)
    assert V[0, 0] == pytest.approx(-1.8804, abs=1e-4)
    assert V[0, 1] == pytest.approx(-1.1948, abs=1e-4)
    assert g[0, 0, 0, 0] == pytest.approx(0.7746, abs=1e-4)
    assert g[0, 0, 1, 1] == pytest.approx(0.5697, abs=1e-4)
    assert g[1, 0, 0, 0] == pytest.approx(0.4441, abs=1e-4)
    assert g[1, 0, 1, 0] == pytest.approx(0.2970, abs=1e-4)
    res = T.scf(mol, method="rhf")
    assert res.e_tot == pytest.approx(-1.1167, abs=1e-4)


@pytest.mark.parametrize("r,energy", [(0.74, -1.12675531719693), (0.5, -1.05802481296927), (1.5, -0.997497294328357)])
def test_h2_631g_rhf_matches_reference_notebook(r, energy):
    """PySCF RHF energies captured in notebooks/03_notebook_td_dataset.ipynb:40 and
    notebooks/04_notebook_td_trainer.ipynb (cell 5)."""
    res = T.scf(T.h2_molecule(r), method="rhf", conv_tol=1e-12)
    assert res.converged
    assert res.e_tot == pytest.approx(energy, abs=1e-9)


def test_p_shell_integrals_symmetry_and_translation():
    """p shells: ERI 8-fold symmetry, S positive definite, and invariance under a rigid shift."""
    mol = T.synthetic_chain(2, seed=3)
    S, Tk, V = T.one_electron(mol)
    g = T.eri(mol)
    assert np.all(np.linalg.eigvalsh(S) > 0)
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-12)
    np.testing.assert_allclose(g, g.transpose(1, 0, 2, 3), atol=1e-14)
    np.testing.assert_allclose(g, g.transpose(2, 3, 0, 1), atol=1e-14)
    mol2 = T.synthetic_chain(2, seed=3)
    mol2.coords = mol2.coords + np.array([0.3, -0.7, 1.1])
    S2, T2, V2 = T.one_electron(mol2)
    np.testing.assert_allclose(S2, S, atol=1e-12)
    np.testing.assert_allclose(T2, Tk, atol=1e-12)
    np.testing.assert_allclose(V2, V, atol=1e-12)
    np.testing.assert_allclose(T.eri(mol2), g, atol=1e-12)


def test_kinetic_is_laplacian_expectation_on_grid():
    """T_ij = 1/2 <grad phi_i | grad phi_j>, checked by quadrature of finite-difference gradients."""
    mol = T.synthetic_chain(2, seed=5)
    S, Tk, _ = T.one_electron(mol)
    pts, w = T.molecular_grid(mol, n_rad=80, n_theta=20, n_phi=21)
    h = 1e-4
    grads = []
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        grads.append((T.eval_ao(mol, pts + e) - T.eval_ao(mol, pts - e)) / (2 * h))
    tq = 0.5 * sum(gd.T @ (gd * w[:, None]) for gd in grads)
    np.testing.assert_allclose(tq, Tk, atol=2e-4)
    ao = T.eval_ao(mol, pts)
    np.testing.assert_allclose(ao.T @ (ao * w[:, None]), S, atol=2e-5)


def test_grid_integrates_density_to_electron_count():
    mol = T.synthetic_chain(4, seed=1)
    res = T.scf(mol, method="rks", max_cycle=30)
    assert res.nelec_grid == pytest.approx(4.0, abs=1e-3)


def test_lda_potential_is_energy_derivative():
    rho = np.array([1e-4, 1e-3, 0.01, 0.1, 0.5, 1.0, 3.0])
    e, v = T.lda_exc_vxc(rho)
    h = 1e-7 * rho
    ep, _ = T.lda_exc_vxc(rho + h)
    em, _ = T.lda_exc_vxc(rho - h)
    np.testing.assert_allclose(v, ((rho + h) * ep - (rho - h) * em) / (2 * h), rtol=1e-6)
    assert np.all(e < 0)
    ex, vx = T.lda_exc_vxc(rho, "slater")
    hp, hm = T.lda_exc_vxc(rho + h, "slater")[0], T.lda_exc_vxc(rho - h, "slater")[0]
    np.testing.assert_allclose(vx, ((rho + h) * hp - (rho - h) * hm) / (2 * h), rtol=1e-6)


def test_slater_exchange_is_the_uniform_gas_exchange():
    """Exchange per electron of the uniform electron gas, -3 k_F / (4 pi) with k_F = (3 pi^2 rho)^(1/3)
    (the Hartree-Fock exchange of plane waves filled to k_F), at densities over ten decades."""
    rho = np.logspace(-6, 4, 23)
    kf = (3 * np.pi ** 2 * rho) ** (1 / 3)
    np.testing.assert_allclose(T.lda_exc_vxc(rho, "slater")[0], -3 * kf / (4 * np.pi), rtol=1e-13)


def _correlation(rs):
    rho = 3 / (4 * np.pi * np.asarray(rs, dtype=np.float64) ** 3)
    return T.lda_exc_vxc(rho)[0] - T.lda_exc_vxc(rho, "slater")[0]


def test_vwn5_correlation_high_density_coefficient():
    """As r_s -> 0 the correlation energy per electron of the uniform gas goes as A ln r_s + C with
    the random-phase coefficient A = (1 - ln 2) / pi^2 = 0.0310907 Ha (Gell-Mann and Brueckner); the
    slope d e_c / d ln r_s at r_s = 1e-6 must reproduce it."""
    rs = 1e-6
    slope = (_correlation(rs * 1.01) - _correlation(rs)) / math.log(1.01)
    assert slope == pytest.approx((1 - math.log(2)) / math.pi ** 2, rel=1e-4)


def test_vwn5_correlation_against_perdew_wang_fit():
    """VWN5 and the Perdew-Wang 1992 parametrisation are independent fits of the same Ceperley-Alder
    quantum Monte Carlo correlation energies; they agree to about 5e-4 Ha for 0.5 <= r_s <= 10.  The
    VWN-RPA parameters (x0 = -0.409286, b = 13.0720, c = 42.7198) miss by more than 1e-2 Ha here."""
    rs = np.array([0.5, 1.0, 2.0, 5.0, 10.0])
    A, a1, b1, b2, b3, b4 = 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294
    pw92 = -2 * A * (1 + a1 * rs) * np.log(1 + 1 / (2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** 2)))
    np.testing.assert_allclose(_correlation(rs), pw92, atol=5e-4)


def test_h2_631g_slater_exchange_rks_matches_reference_notebook():
    """pyscf RKS with xc="lda" (qedft/data_io/td/dataset_generation.py:331) for H2/6-31G at 0.74 A
    printed -1.03718794786902 Ha on its level-0 grid of 1240 points
    (notebooks/03_notebook_td_dataset.ipynb:64, 71).  pyscf's bare "lda" is Slater exchange alone;
    on a converged grid it gives -1.0385739 Ha here, and the 1.4e-3 Ha left is the error of that
    1240-point grid.  Adding VWN5 correlation would move the energy by about -9.5e-2 Ha, and a 1% error
    in the exchange constant by about 6e-3 Ha, so the 2e-3 tolerance pins the exchange."""
    mol = T.h2_molecule(0.74)
    grid = T.molecular_grid(mol, n_rad=75, n_theta=18, n_phi=19)
    res = T.scf(mol, method="rks", grid=grid, conv_tol=1e-12, max_cycle=60, functional="slater")
    assert res.converged
    assert res.e_tot == pytest.approx(-1.03718794786902, abs=2e-3)
    svwn = T.scf(mol, method="rks", grid=grid, conv_tol=1e-12, max_cycle=60)
    assert abs(svwn.e_tot - -1.03718794786902) > 5e-2


def test_generalized_eigh_and_rdm1():
    """eigh(F, S): F C = S C e, C^T S C = 1; make_rdm1 trace = N_e, idempotent in the S metric
    (tests/test_make_rdm1.py invariants)."""
    mol = T.synthetic_chain(4, seed=2)
    S, Tk, V = T.one_electron(mol)
    F = Tk + V
    e, C = T.gen_eigh(F, S)
    np.testing.assert_allclose(F @ C, S @ C * e[None, :], atol=1e-10)
    np.testing.assert_allclose(C.T @ S @ C, np.eye(len(e)), atol=1e-10)
    D = T.make_rdm1(C, 2)
    assert np.trace(D @ S) == pytest.approx(4.0, abs=1e-10)
    np.testing.assert_allclose(D @ S @ D, 2 * D, atol=1e-10)
    np.testing.assert_allclose(D, D.T, atol=1e-14)


def test_rks_scf_converges_and_is_variational():
    """The converged RKS energy is stationary: perturbing the converged density matrix along an
    occupied-virtual rotation raises the energy to second order."""
    mol = T.synthetic_chain(2, seed=4)
    res = T.scf(mol, method="rks", conv_tol=1e-11)
    assert res.converged
    assert np.all(np.diff(res.energies[-3:]) <= 1e-9)


def test_gelu_tanh_and_its_derivative():
    x = np.linspace(-6, 6, 101)
    # tanh-approximate GELU against the exact one within the approximation's known error (< 1e-3)
    from scipy.special import erf
    exact = 0.5 * x * (1 + erf(x / np.sqrt(2)))
    assert np.abs(T.gelu_tanh(x) - exact).max() < 1e-3
    h = 1e-6
    fd = (T.gelu_tanh(x + h) - T.gelu_tanh(x - h)) / (2 * h)
    np.testing.assert_allclose(T.gelu_tanh_grad(x), fd, atol=1e-8)


def test_global_mlp_vjp_matches_finite_differences():
    rng = np.random.default_rng(0)
    ng = 300
    W, b = T.init_global_mlp(5, ng, hidden=(32, 16))
    b = [rng.standard_normal(x.shape) * 0.1 for x in b]
    net = T.GlobalMLP3D(W, b, scale=0.5)
    rho = rng.uniform(0, 1, ng)
    e, v = net.exc_vrho(rho)
    for g in rng.choice(ng, 8, replace=False):
        d = np.zeros(ng)
        d[g] = 1e-5
        fd = (net.exc_vrho(rho + d)[0] - net.exc_vrho(rho - d)[0]) / 2e-5
        assert abs(fd - v[g]) < 1e-8 * max(1.0, abs(v[g]))


def test_nn_amplitude_vmat_is_density_matrix_derivative_with_unit_weights():
    """With unit grid weights the NN-Amplitude vmat is d excsum / d D; with real weights vrho is
    scaled by them once more (the reference's convention, td/numint.py:206-231)."""
    mol = T.synthetic_chain(2, 4)
    pts, w = T.molecular_grid(mol, n_rad=12, n_theta=4, n_phi=5)
    ao = T.eval_ao(mol, pts)
    W, b = T.init_global_mlp(1, len(w), hidden=(16, 8))
    net = T.GlobalMLP3D(W, b, scale=1.0)
    rng = np.random.default_rng(2)
    c = rng.standard_normal((mol.nao, mol.nao)) * 0.2
    D = c @ c.T
    ones = np.ones_like(w)
    _, exc, vmat = T.nr_rks_nn_global(ao, ones, D, net)
    for (i, j) in ((0, 0), (1, 3), (4, 2)):
        E = np.zeros_like(D)
        E[i, j] += 0.5e-6
        E[j, i] += 0.5e-6
        fd = (T.nr_rks_nn_global(ao, ones, D + E, net)[1] - T.nr_rks_nn_global(ao, ones, D - E, net)[1]) / 2e-6
        assert abs(fd - vmat[i, j]) < 1e-7 * max(1.0, abs(vmat[i, j]))
    _, _, vw = T.nr_rks_nn_global(ao, w, D, net)
    rho = T.eval_rho(ao, D)
    _, vrho = net.exc_vrho(rho)
    np.testing.assert_allclose(vw, ao.T @ (ao * (w * vrho)[:, None]), atol=1e-12)


def test_local_network_vrho_is_the_pointwise_derivative():
    """exc_and_vrho_local: per point exc(rho) and d exc / d rho (central differences, each point
    independent of the others), with the reference's nr_rks "NN" assembly."""
    rng = np.random.default_rng(3)
    W = [rng.standard_normal((1, 12)), rng.standard_normal((12, 7)) / 3, rng.standard_normal((7, 1)) / 2]
    b = [rng.standard_normal(12) * 0.1, rng.standard_normal(7) * 0.1, rng.standard_normal(1) * 0.1]
    net = T.LocalMLP3D(W, b, scale=0.3)
    rho = rng.uniform(0, 2, 50)
    exc, vrho = net.exc_vrho(rho)
    h = 1e-6
    np.testing.assert_allclose(vrho, (net.exc_vrho(rho + h)[0] - net.exc_vrho(rho - h)[0]) / (2 * h), atol=1e-8)
    # a point's value does not depend on the other points
    r2 = rho.copy()
    r2[7] += 0.3
    np.testing.assert_array_equal(net.exc_vrho(r2)[0][:7], exc[:7])
    # one point of the batch equals the same network on that point alone
    assert net.exc_vrho(rho[3:4])[0][0] == pytest.approx(exc[3], abs=1e-15)


def test_local_network_scf_runs_and_uses_the_pointwise_derivative():
    mol = T.synthetic_chain(2, 4)
    pts, w = T.molecular_grid(mol, n_rad=20, n_theta=6, n_phi=7)
    W, b = T.init_global_mlp(2, 1, hidden=(16, 8))
    net = T.LocalMLP3D(W, b, scale=0.5)
    res = T.scf(mol, "rks_nn_local", grid=(pts, w), net=net, conv_tol=1e-10, max_cycle=60)
    assert res.converged and np.isfinite(res.e_tot)


def _sym(rng, n):
    a = rng.standard_normal((n, n))
    return 0.5 * (a + a.T)


def test_degen_eigh_bwd_exact_matches_finite_differences():
    """generalized_eigensolver.py:158-219 in its exact form (U = V^T) is the derivative of the
    decomposition: <gA, E> equals central differences of L = c.w + <G, V> along symmetric E
    (eigenvector signs aligned with the unperturbed ones)."""
    rng = np.random.default_rng(3)
    n = 6
    A = _sym(rng, n)
    c, G = rng.standard_normal(n), rng.standard_normal((n, n))
    w0, V0 = np.linalg.eigh(A)

    def loss(M):
        w, V = np.linalg.eigh(M)
        V = V * np.sign(np.sum(V * V0, axis=0))[None, :]
        return c @ w + np.sum(G * V)

    gA = T.degen_eigh_bwd(w0, V0, c, G, exact=True)
    for k in range(4):
        E = _sym(rng, n)
        h = 1e-6
        fd = (loss(A + h * E) - loss(A - h * E)) / (2 * h)
        assert abs(np.sum(gA * E) - fd) < 1e-7 * max(1.0, abs(fd)), k


def test_degen_eigh_bwd_reference_form_equals_exact_for_symmetric_eigenvectors():
    """The reference's code uses U = V (its transpose with axes (-2, -1) is the identity on a 2-D
    array); that agrees with the exact form exactly when V = V^T (a reflector) and not otherwise."""
    rng = np.random.default_rng(4)
    n = 7
    u = rng.standard_normal(n)
    H = np.eye(n) - 2.0 * np.outer(u, u) / (u @ u)  # symmetric and orthogonal
    w = np.sort(rng.standard_normal(n))
    gw, gV = rng.standard_normal(n), rng.standard_normal((n, n))
    np.testing.assert_allclose(T.degen_eigh_bwd(w, H, gw, gV), T.degen_eigh_bwd(w, H, gw, gV, exact=True),
                               rtol=1e-12, atol=1e-12)
    w1, V1 = np.linalg.eigh(_sym(rng, n))
    assert not np.allclose(T.degen_eigh_bwd(w1, V1, gw, gV), T.degen_eigh_bwd(w1, V1, gw, gV, exact=True))


def test_degen_eigh_bwd_masks_degenerate_pairs():
    """A pair closer than eps^0.6 contributes nothing through its eigenvector coupling: the result stays
    finite, and an eigenvector cotangent inside the degenerate pair's span (antisymmetric in the pair)
    gives zero from that pair."""
    rng = np.random.default_rng(5)
    n = 5
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    w = np.array([-1.0, 0.5, 0.5 + 1e-11, 2.0, 3.0])
    gV = np.zeros((n, n))
    gV[:, 1], gV[:, 2] = Q[:, 2], -Q[:, 1]  # U gV is nonzero only on the (1, 2) pair
    g = T.degen_eigh_bwd(w, Q, None, gV, exact=True)
    assert np.all(np.isfinite(g)) and np.abs(g).max() < 1e-12


def test_stratmann_step_closed_form_properties():
    """The Stratmann cell step (CPL 257, 213 (1996), eq. 14): odd, zero at 0, +-1 with zero slope at
    mu = +-0.64 (C1 join), monotone in between; the Becke step is its three-fold iterated cubic."""
    mu = np.linspace(-1, 1, 2001)
    g = T.stratmann_step(mu)
    np.testing.assert_allclose(g, -T.stratmann_step(-mu), atol=0)
    assert T.stratmann_step(0.0) == 0.0
    assert T.stratmann_step(0.64) == 1.0 and T.stratmann_step(-0.64) == -1.0
    h = 1e-6
    assert abs((T.stratmann_step(0.64 - h) - T.stratmann_step(0.64 - 2 * h)) / h) < 1e-4
    assert np.all(np.diff(g) >= 0)
    # polynomial value at z = 1/2 worked by hand: (35/2 - 35/8 + 21/32 - 5/128) / 16
    assert T.stratmann_step(0.32) == pytest.approx((35 / 2 - 35 / 8 + 21 / 32 - 5 / 128) / 16, abs=1e-15)


@pytest.mark.parametrize("scheme", ["becke", "stratmann"])
def test_partition_of_unity_and_gaussian_integral(scheme):
    """Cell weights of every atom sum to one at any point; a normalised Gaussian on each atom of a
    chain integrates to the atom count on the partitioned grid."""
    mol = T.synthetic_chain(4, 104)
    pts = np.random.default_rng(0).normal(size=(200, 3)) * 2.0
    natm = len(mol.coords)
    tot = sum(T.becke_weights(pts, mol.coords, np.full(len(pts), a), scheme) for a in range(natm))
    np.testing.assert_allclose(tot, 1.0, atol=1e-13)
    p, w = T.molecular_grid(mol, scheme=scheme)
    dens = sum((2 / np.pi) ** 1.5 * np.exp(-2 * np.sum((p - c) ** 2, axis=1)) for c in mol.coords)
    assert abs(w @ dens - natm) < 5e-4  # quadrature error of the 50 x 110-point atomic grids


def _sphere_monomial(a, b, c):
    """Exact integral of x^a y^b z^c over the unit sphere (Gamma-function closed form)."""
    import math

    if a % 2 or b % 2 or c % 2:
        return 0.0
    g = math.gamma
    return 2 * g((a + 1) / 2) * g((b + 1) / 2) * g((c + 1) / 2) / g((a + b + c + 3) / 2)


@pytest.mark.parametrize("n", sorted(T.LEBEDEV))
def test_lebedev_rules_are_exact_to_their_degree(n):
    """Each Lebedev rule integrates every monomial up to its degree exactly (a wrong weight or orbit
    parameter breaks this) and misses at the next even degree, so the degree is its own."""
    deg = T.LEBEDEV[n][0]
    d, w = T.angular_lebedev(n)
    assert d.shape == (n, 3)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-15)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                v = w @ (d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c)
                assert abs(v - _sphere_monomial(a, b, c)) < 1e-13, (a, b, c)
    D = deg + 1
    miss = max(abs(w @ (d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** (D - a - b)) - _sphere_monomial(a, b, D - a - b))
               for a in range(0, D + 1, 2) for b in range(0, D + 1 - a, 2))
    assert miss > 1e-8


def test_device_lebedev_tables_equal_the_oracle():
    from qex_amd.td import dft

    for n in sorted(T.LEBEDEV):
        d, w = dft.angular_lebedev(n)
        do, wo = T.angular_lebedev(n)
        np.testing.assert_allclose(d, do, rtol=0, atol=2e-15)
        np.testing.assert_allclose(w, wo, rtol=1e-15, atol=0)


def test_sto3g_first_row_data_and_water_energy():
    """The zeta-scaled STO-3G fits reproduce the tabulated exponents (C, O), and the RHF energy of
    water at the experimental geometry (0.9572 A, 104.52 deg) is the standard -74.962928 Ha."""
    o = T.sto3g_shells(7.66, 2.25)
    np.testing.assert_allclose(o[0][1], [130.70932, 23.808861, 6.4436083], rtol=1e-6)
    np.testing.assert_allclose(o[1][1], [5.0331513, 1.1695961, 0.3803890], rtol=1e-6)
    c = T.sto3g_shells(5.67, 1.72)
    np.testing.assert_allclose(c[0][1], [71.616837, 13.045096, 3.5305122], rtol=1e-6)
    np.testing.assert_allclose(c[1][1], [2.9412494, 0.6834831, 0.2222899], rtol=1e-6)
    r = T.scf(T.water_molecule(o, T.STO3G_H), method="rhf", conv_tol=1e-12, max_cycle=100)
    assert r.e_tot == pytest.approx(-74.962928, abs=2e-6)


def test_device_basis_library_equals_the_oracle_sto3g():
    from qex_amd.td import gto

    lib = gto.BASIS_LIBRARY["sto-3g"]
    for el, (z1, z2) in {"O": (7.66, 2.25), "C": (5.67, 1.72), "He": (1.69, None), "Li": (2.69, 0.80)}.items():
        ref = T.sto3g_shells(z1, z2)
        assert len(lib[el]) == len(ref)
        for (l, e, c), (lo, eo, co) in zip(lib[el], ref):
            assert l == lo
            np.testing.assert_allclose(e, eo, rtol=1e-15)
            np.testing.assert_allclose(c, co, rtol=0)
