"""Pins of the FP64 oracle (CPU).  Each check is independent of the oracle's own formula:
reference fixtures, closed forms, textbook limits or finite differences."""

import math
import os

import numpy as np
import pytest
from scipy.special import erfc

from oracle import od as O


def test_external_potential_matches_reference_datasets(golden_dir):
    """data/od/{h2,h4,h2_h2}/external_potentials.npy rows (written by jax_dft in the reference)."""
    for mol in ("h2", "h4", "h2_h2"):
        z = np.load(os.path.join(golden_dir, f"od_{mol}_rows.npz"))
        for i in range(len(z["locations"])):
            v = O.get_atomic_chain_potential(z["grids"], z["locations"][i], z["nuclear_charges"][i],
                                             O.exponential_coulomb)
            np.testing.assert_allclose(v, z["external_potentials"][i], rtol=0, atol=1e-14)


def test_noninteracting_density_matches_reference_pickle(golden_dir):
    """qedft/models/initial_density.pkl: 5-point kinetic stencil + eigh + density, H2 at 1.28 / 3.84."""
    z = np.load(os.path.join(golden_dir, "od_h2_rows.npz"))
    gold = np.load(os.path.join(golden_dir, "initial_density_h2_128_384.npy"))
    for i, R in enumerate((1.28, 3.84)):
        v = O.get_atomic_chain_potential(z["grids"], [-R / 2, R / 2], [1, 1], O.exponential_coulomb)
        n = O.solve_noninteracting_system(v, 2, z["grids"])[0]
        np.testing.assert_allclose(n, gold[i], rtol=0, atol=1e-12)


def test_three_point_stencil_would_miss_the_golden(golden_dir):
    """The golden vector discriminates the stencil (a 3-point Laplacian misses it by ~2e-4)."""
    z = np.load(os.path.join(golden_dir, "od_h2_rows.npz"))
    gold = np.load(os.path.join(golden_dir, "initial_density_h2_128_384.npy"))
    g = z["grids"]
    dx = O.get_dx(g)
    n = len(g)
    L3 = np.diag(-2 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    v = O.get_atomic_chain_potential(g, [-0.64, 0.64], [1, 1], O.exponential_coulomb)
    w, V = np.linalg.eigh(-0.5 * L3 / dx ** 2 + np.diag(v))
    n3 = O.wavefunctions_to_density(2, V.T, g)
    assert np.max(np.abs(n3 - gold[0])) > 1e-5


def test_global_xc_energy_known_answer():
    """tests/test_scf_od.py:52-61."""
    grids = np.linspace(-5, 5, 10001)
    density = np.exp(-((grids - 1) ** 2))
    e = O.xc_energy(O.PowerGlobal(-0.73855, 1 / 3), density, O.get_dx(grids))
    assert e == pytest.approx(-2.2660978689082367, abs=1e-5)


def test_global_xc_potential_closed_form():
    """tests/test_scf_od.py:64-86 (note the factor dx)."""
    grids = np.linspace(-5, 5, 10001)
    density = np.exp(-((grids - 1) ** 2))
    dx = O.get_dx(grids)
    v = O.xc_potential(O.PowerGlobal(-0.73855, 1 / 3), density, dx)
    np.testing.assert_allclose(v, -0.73855 * (1 / 3) * density ** (-2 / 3) * dx, rtol=1e-7)


def _gauss_conv_exp(t, sigma, kappa):
    """E[exp(-kappa |Y - t|)], Y ~ N(0, sigma^2) (closed form with erfc)."""
    a = kappa * sigma ** 2
    return 0.5 * np.exp(0.5 * (kappa * sigma) ** 2) * (
        np.exp(-kappa * t) * erfc((a - t) / (sigma * math.sqrt(2))) + np.exp(kappa * t) * erfc((a + t) / (sigma * math.sqrt(2))))


def test_hartree_potential_and_energy_closed_forms():
    """v_H and E_H of a Gaussian density under A exp(-kappa|r|) against their erfc closed forms
    (pins the dx factor, the 1/2, A and kappa)."""
    A, kappa = O.EXPONENTIAL_COULOMB_AMPLITUDE, O.EXPONENTIAL_COULOMB_KAPPA
    grids = np.linspace(-30, 30, 6001)
    mu, sigma, ne = 0.4, 1.3, 2.0
    n = ne * O.gaussian(grids, mu, sigma)
    vh = O.get_hartree_potential(n, grids, O.exponential_coulomb)
    exact = A * ne * _gauss_conv_exp(grids - mu, sigma, kappa)
    assert np.max(np.abs(vh - exact)[np.abs(grids) < 10]) < 2e-5
    eh = O.get_hartree_energy(n, grids, O.exponential_coulomb)
    eh_exact = 0.5 * A * ne ** 2 * math.exp((kappa * sigma) ** 2) * erfc(kappa * sigma)
    assert eh == pytest.approx(eh_exact, rel=2e-5)


def test_harmonic_oscillator_spectrum():
    """v = x^2/2: eigenvalues k + 1/2 (pins the kinetic sign/scale and the repeat-x2 occupation)."""
    grids = np.linspace(-12, 12, 961)
    v = 0.5 * grids ** 2
    dens, e_eig, gap, eigs, _ = O.solve_noninteracting_system(v, 5, grids, return_eigs=True)
    np.testing.assert_allclose(eigs[:4], [0.5, 1.5, 2.5, 3.5], atol=1e-6)
    assert e_eig == pytest.approx(2 * 0.5 + 2 * 1.5 + 2.5, abs=1e-5)
    assert gap == pytest.approx(0.0, abs=1e-12)  # odd N_e: HOMO = LUMO in the repeat-x2 convention
    assert O.solve_noninteracting_system(v, 4, grids)[2] == pytest.approx(1.0, abs=1e-5)
    assert np.sum(dens) * O.get_dx(grids) == pytest.approx(5.0, abs=1e-10)


def test_energy_assembly_equals_direct_functional():
    """E = E_eig - dx v_KS.n + E_H + E_xc + dx v_ext.n equals T_s + E_ext + E_H + E_xc with T_s taken
    directly from the occupied orbitals (an independent evaluation of the same energy)."""
    grids = np.linspace(-10, 10, 201)
    dx = O.get_dx(grids)
    loc, Z, ne = np.array([-0.8, 0.9]), np.array([2, 1]), 3
    v_ext = O.get_atomic_chain_potential(grids, loc, Z, O.exponential_coulomb)
    n_in = O.gaussian_initial_density(grids, loc, Z, ne, weighted=True)
    xc = O.LDAExp()
    n_out, E, vh, vxc, exc, gap = O.kohn_sham_iteration(n_in, v_ext, grids, ne, xc, O.exponential_coulomb, None)
    v_ks = vh + vxc + v_ext
    w, V = np.linalg.eigh(O.get_kinetic_matrix(grids) + np.diag(v_ks))
    T = O.get_kinetic_matrix(grids)
    occ = [2, 1]
    ts = 0.0
    for j, f in enumerate(occ):
        psi = V[:, j] / math.sqrt(dx * np.sum(V[:, j] ** 2))
        ts += f * dx * psi @ T @ psi
    e_direct = (ts + dx * v_ext @ n_out + O.get_hartree_energy(n_out, grids, O.exponential_coulomb)
                + O.xc_energy(xc, n_out, dx))
    assert E == pytest.approx(e_direct, abs=1e-10)


def test_lda_exchange_matches_exchange_hole_integral():
    """e_x(n) = -(2A/(pi^2 n)) int_0^inf exp(-kappa r) sin^2(k_F r)/r^2 dr, k_F = pi n / 2
    (exchange of the spin-unpolarised 1D uniform gas), integrated numerically."""
    from scipy.integrate import quad

    A, kappa = O.EXPONENTIAL_COULOMB_AMPLITUDE, O.EXPONENTIAL_COULOMB_KAPPA
    for n in (0.01, 0.1, 0.5, 1.0, 3.0):
        kf = math.pi * n / 2
        f = lambda r: math.exp(-kappa * r) * (math.sin(kf * r) / r) ** 2 if r > 0 else kf ** 2
        val = quad(f, 0, np.inf, limit=800)[0]
        ex = -2 * A / (math.pi ** 2 * n) * val
        assert O.lda_exchange(np.array([n]))[0][0] == pytest.approx(ex, rel=1e-7)


def test_lda_small_density_branches_are_continuous():
    y_eps = 1e-15
    n_eps = y_eps * O.EXPONENTIAL_COULOMB_KAPPA / math.pi
    lo, hi = O.lda_exchange(np.array([n_eps * 0.999, n_eps * 1.001]))[0]
    assert abs(lo - hi) < 1e-15
    ec0 = O.lda_correlation(np.array([0.0, 1e-12]))[0]
    assert ec0[0] == 0.0 and abs(ec0[1]) < 1e-11


def _fd_check(xc, n, dx, idx, h=1e-6, rtol=1e-6):
    v = O.xc_potential(xc, n, dx)
    for i in idx:
        e1, e2 = n.copy(), n.copy()
        e1[i] += h
        e2[i] -= h
        d = (O.xc_energy(xc, e1, dx) - O.xc_energy(xc, e2, dx)) / (2 * h)
        expected = d / dx if xc.kind == "local" else d
        assert v[i] == pytest.approx(expected, rel=rtol, abs=1e-8), (type(xc).__name__, i)


@pytest.mark.parametrize("kind", ["lda", "power_local", "power_global", "local_mlp", "local_mlp_w1", "global_mlp",
                                  "sym_mlp"])
def test_xc_potentials_are_energy_derivatives(kind):
    rng = np.random.default_rng(3)
    grids = np.linspace(-6, 6, 61)
    dx = O.get_dx(grids)
    n = 1.5 * O.gaussian(grids, 0.3, 1.0) + 0.4 * O.gaussian(grids, -1.5, 0.7) + 1e-3
    if kind == "lda":
        xc = O.LDAExp()
    elif kind == "power_local":
        xc = O.PowerLocal(-0.73855, 1 / 3, 1e-10)
    elif kind == "power_global":
        xc = O.PowerGlobal(-0.73855, 1 / 3, 1e-10)
    elif kind in ("local_mlp", "sym_mlp"):
        W, b = O.glorot_normal_mlp(11, [17, 16, 16, 1])
        b = [rng.normal(0, 0.1, bb.shape) for bb in b]
        xc = O.LocalMLP(W, b, "silu", 17, 2.0)
        if kind == "sym_mlp":
            xc = O.SymmetrizedLocalXC(xc, O.flip_and_average_on_center)
    elif kind == "local_mlp_w1":
        W, b = O.glorot_normal_mlp(12, [1, 8, 8, 1])
        xc = O.LocalMLP(W, b, "tanh", 1, 2.0)
    else:
        W, b = O.glorot_normal_mlp(13, [61, 8, 1])
        xc = O.GlobalMLP(W, b, "tanh", 2.0)
    _fd_check(xc, n, dx, idx=[0, 5, 17, 30, 44, 60])


def test_local_mlp_window_one_is_pointwise():
    """Window 1 is build_local_mlp (classical_models.py:163-172): the same MLP at every point on n/2."""
    W, b = O.glorot_normal_mlp(5, [1, 7, 7, 1])
    xc = O.LocalMLP(W, b, "tanh", 1, 2.0)
    n = np.linspace(0, 2, 9)
    e = xc.energy_density(n)
    for i, x in enumerate(n):
        h = np.tanh(np.array([[x / 2]]) @ W[0] + b[0])
        h = np.tanh(h @ W[1] + b[1])
        assert e[i] == pytest.approx(float((h @ W[2] + b[2])[0, 0]), abs=1e-14)


def test_global_mlp_is_the_dense_stack_on_the_whole_density():
    """build_global_mlp (classical_models.py:177-224): one scalar from the whole density / 2."""
    W, b = O.glorot_normal_mlp(6, [9, 5, 1])
    xc = O.GlobalMLP(W, b, "tanh", 2.0)
    n = np.linspace(0.1, 1.7, 9)
    h = np.tanh(n / 2 @ W[0] + b[0])
    assert xc.value(n) == pytest.approx(float(h @ W[1][:, 0] + b[1][0]), abs=1e-14)
    # linear network: f = (n/2) . W0 W1 + b0 W1 + b1, df/dn = W0 W1 / 2
    lin = O.GlobalMLP([W[0] @ W[1]], [b[0] @ W[1] + b[1]], "tanh", 2.0)
    np.testing.assert_allclose(lin.grad(n), (W[0] @ W[1])[:, 0] / 2, rtol=0, atol=1e-15)
    assert O.xc_energy(xc, n, 0.1) == pytest.approx(0.1 * xc.value(n), abs=1e-15)
    np.testing.assert_allclose(O.xc_potential(xc, n, 0.1), 0.1 * xc.grad(n), rtol=0, atol=1e-16)


def test_reflection_window_and_flip():
    g = np.linspace(-2, 2, 41)  # dx = 0.1
    lo, hi = O.reflection_window([-0.5, 0.5], g)
    assert (lo, hi) == (0, 40)
    lo, hi = O.reflection_window([-0.5, 0.7], g)  # centre 0.1 (on grid, index 21)
    assert lo + hi == 42 and lo == 2 and hi == 40
    lo, hi = O.reflection_window([-0.5, 0.6], g)  # centre 0.05 (between 20 and 21)
    assert lo + hi == 41 and hi == 40
    a = np.arange(41.0)
    f = O.flip_and_average([-0.5, 0.6], g, a)
    np.testing.assert_allclose(f[lo:hi + 1], (a[lo:hi + 1] + a[lo:hi + 1][::-1]) / 2)
    np.testing.assert_allclose(f[:lo], a[:lo])


def test_mixing_limits():
    """alpha = 1, m = 1: the stored density is the iteration's own output; alpha = 0: it stays n0."""
    grids = np.linspace(-8, 8, 81)
    loc, Z = np.array([-0.6, 0.6]), np.array([1, 1])
    v = O.get_atomic_chain_potential(grids, loc, Z, O.exponential_coulomb)
    n0 = O.gaussian_initial_density(grids, loc, Z, 2)
    st = O.kohn_sham(loc, Z, 2, 2, grids, O.LDAExp(), O.exponential_coulomb, n0, alpha=1.0, alpha_decay=1.0,
                     num_mixing_iterations=1)
    n_out = O.kohn_sham_iteration(n0, v, grids, 2, O.LDAExp(), O.exponential_coulomb, None)[0]
    np.testing.assert_allclose(st.density[0], n_out, atol=1e-15)
    st0 = O.jit_kohn_sham(loc, Z, 2, 3, grids, O.LDAExp(), O.exponential_coulomb, n0, alpha=0.0)
    for t in range(3):
        np.testing.assert_allclose(st0.density[t], n0, atol=1e-15)


@pytest.mark.parametrize("tol,expected", [(-1.0, [False, False, False]), (np.inf, [True, True, True])])
def test_python_loop_convergence_flags(tol, expected):
    """tests/test_scf_od.py:184-215."""
    grids = np.linspace(-5, 5, 101)
    st = O.kohn_sham(np.array([-0.5, 0.5]), np.array([1, 1]), 2, 3, grids, O.PowerGlobal(), O.exponential_coulomb,
                     density_mse_converge_tolerance=tol)
    assert list(st.converged) == expected


@pytest.mark.parametrize("tol,expected", [(1e-2, [False, False, False]), (1e10, [False, True, True])])
def test_scan_loop_convergence_flags(tol, expected):
    """tests/test_jit_scf_od.py:186-229."""
    grids = np.linspace(-5, 5, 101)
    n0 = 2 * O.gaussian(grids, 0.0, 1.0)
    st = O.jit_kohn_sham(np.array([-0.5, 0.5]), np.array([1, 1]), 2, 3, grids, O.PowerGlobal(),
                         O.exponential_coulomb, n0, density_mse_converge_tolerance=tol)
    assert list(st.converged) == expected


def test_connection_weights_average_last_m():
    w = O.connection_weights(5, 2)
    np.testing.assert_allclose(w[0], [1, 0, 0, 0, 0])
    np.testing.assert_allclose(w[3], [0, 0, 0.5, 0.5, 0])


def test_random_molecules_constraints():
    g = O.h2_config_grid()
    for loc, z, ne in O.random_molecules(3, 200, grids=g):
        assert 1 <= len(loc) <= 4 and ne == z.sum() <= 8 and set(z) <= {1, 2, 3, 4}
        assert np.all(np.abs(loc) <= 4.0 + 1e-9)
        if len(loc) > 1:
            assert np.min(np.diff(loc)) >= 0.4 - 1e-9
        assert np.allclose(np.round((loc - g[0]) / 0.1), (loc - g[0]) / 0.1)
