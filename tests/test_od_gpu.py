"""GPU parity of the 1D engine (libqex.so kernels) against the FP64 oracle.

Bars: total energy within 1e-6 Ha and density L2 = sqrt(dx sum (n - n_ref)^2) within 1e-6 on
every stored iteration (both are usually met to ~1e-10); eigen-path golden vector to 1e-10.
"""

import os

import numpy as np
import pytest

from oracle import od as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

E_TOL = 1e-6
L2_TOL = 1e-6


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _l2(a, b, dx):
    return float(np.sqrt(dx * np.sum((np.asarray(a) - np.asarray(b)) ** 2)))


def _compare(state, ref, dx, e_tol=E_TOL, l2_tol=L2_TOL, fields=True):
    s = state.numpy()
    T = ref.total_energy.shape[0]
    for t in range(T):
        assert abs(s.total_energy[t] - ref.total_energy[t]) < e_tol, (t, s.total_energy[t], ref.total_energy[t])
        assert _l2(s.density[t], ref.density[t], dx) < l2_tol, t
        if fields:
            assert _l2(s.hartree_potential[t], ref.hartree_potential[t], dx) < 1e-6, t
            assert _l2(s.xc_potential[t], ref.xc_potential[t], dx) < 1e-6, t
            assert abs(s.gap[t] - ref.gap[t]) < 1e-6, t
    np.testing.assert_array_equal(s.converged.astype(bool), ref.converged.astype(bool))


def test_external_potential_matches_reference_dataset(golden_dir):
    from qex_amd.od import utils

    for mol in ("h2", "h4", "h2_h2"):
        z = np.load(os.path.join(golden_dir, f"od_{mol}_rows.npz"))
        v = utils.get_atomic_chain_potential(z["grids"], z["locations"], z["nuclear_charges"],
                                             utils.exponential_coulomb)
        np.testing.assert_allclose(_np(v), z["external_potentials"], rtol=0, atol=1e-13)


def test_noninteracting_density_matches_golden(golden_dir):
    """od/trainer.py:72-75 initial densities of H2 at 1.28 and 3.84 bohr (initial_density.pkl)."""
    from qex_amd.od import scf, utils

    z = np.load(os.path.join(golden_dir, "od_h2_rows.npz"))
    gold = np.load(os.path.join(golden_dir, "initial_density_h2_128_384.npy"))
    grids = z["grids"]
    for i, R in enumerate((1.28, 3.84)):
        v = utils.get_atomic_chain_potential(grids, [-R / 2, R / 2], [1, 1], utils.exponential_coulomb)
        dens, eeig, gap = scf.solve_noninteracting_system(v, 2, grids)
        np.testing.assert_allclose(_np(dens), gold[i], rtol=0, atol=1e-10)
        ref = O.solve_noninteracting_system(_np(v), 2, grids)
        assert abs(float(eeig) - ref[1]) < 1e-10 and abs(float(gap) - ref[2]) < 1e-10


@pytest.mark.parametrize("ne", [1, 2, 3, 5, 8, 14, 16])
def test_noninteracting_many_electrons(ne):
    """ne // 2 + 1 orbitals: up to 8 are orthonormalised by Cholesky QR (ne = 14 is the largest
    such case), more by Gram-Schmidt (ne = 16)."""
    from qex_amd.od import engine, utils

    grids = O.h2_config_grid()
    loc, Z = np.array([-2.0, 0.3, 2.1]), np.array([2, 3, 3])
    v = O.get_atomic_chain_potential(grids, loc, Z, O.exponential_coulomb)
    r = engine.solve_noninteracting(v, ne, grids, with_states=True)
    dens, ee, gap, eigs, _ = O.solve_noninteracting_system(v, ne, grids, return_eigs=True)
    assert _l2(_np(r.density), dens, O.get_dx(grids)) < 1e-9
    np.testing.assert_allclose(_np(r.eigenvalues)[: ne // 2 + 1], eigs[: ne // 2 + 1], rtol=0, atol=1e-10)
    assert abs(float(r.total_eigen_energies) - ee) < 1e-10 and abs(float(r.gap) - gap) < 1e-10


@pytest.mark.parametrize("n_grid", [64, 65, 200, 201])
def test_noninteracting_even_and_odd_grids(n_grid):
    """The device eigen kernel splits the grid into a top and a bottom half that meet at a middle
    pair of rows (halves of equal length for even N, the bottom one row longer for odd N)."""
    from qex_amd.od import engine

    grids = np.linspace(-10, 10, n_grid)
    loc, Z = np.array([-1.5, 1.2]), np.array([2, 1])
    v = O.get_atomic_chain_potential(grids, loc, Z, O.exponential_coulomb)
    ne = 3
    r = engine.solve_noninteracting(v, ne, grids, with_states=True)
    dens, ee, gap, eigs, _ = O.solve_noninteracting_system(v, ne, grids, return_eigs=True)
    assert _l2(_np(r.density), dens, O.get_dx(grids)) < 1e-9
    np.testing.assert_allclose(_np(r.eigenvalues)[: ne // 2 + 1], eigs[: ne // 2 + 1], rtol=0, atol=1e-10)
    assert abs(float(r.gap) - gap) < 1e-10


@pytest.mark.parametrize("interaction", ["exp", "soft"])
def test_hartree_potential(interaction):
    from qex_amd.od import scf, utils

    grids = np.linspace(-10, 10, 257)
    n = 2 * O.gaussian(grids, 0.3, 1.2)
    fo = O.exponential_coulomb if interaction == "exp" else O.soft_coulomb
    fd = utils.exponential_coulomb if interaction == "exp" else utils.soft_coulomb
    np.testing.assert_allclose(_np(scf.get_hartree_potential(n, grids, fd)), O.get_hartree_potential(n, grids, fo),
                               rtol=0, atol=1e-12)


def test_c1_plain_lda_scan_parity():
    """Config 1: H2-like, 513 points on [-25.6, 25.6], LDA, T=15, alpha=0.3, reflection on."""
    from qex_amd.od import jit_scf, utils, xc

    grids = O.h2_config_grid()
    loc, Z = np.array([-0.7, 0.7]), np.array([1, 1])
    n0 = O.gaussian_initial_density(grids, loc, Z, 2)
    ref = O.jit_kohn_sham(loc, Z, 2, 15, grids, O.LDAExp(), O.exponential_coulomb, n0, alpha=0.3, alpha_decay=1.0,
                          enforce_reflection_symmetry=True, num_mixing_iterations=1)
    st = jit_scf.kohn_sham(loc, Z, 2, 15, grids, xc.get_lda_xc_energy_density_fn(), utils.exponential_coulomb, n0,
                           alpha=0.3, alpha_decay=1.0, enforce_reflection_symmetry=True, num_mixing_iterations=1)
    _compare(st, ref, O.get_dx(grids))
    s = st.numpy()
    np.testing.assert_allclose(s.xc_energy_density, ref.xc_energy_density, rtol=0, atol=1e-10)


def test_python_loop_off_centre_reflection(golden_dir):
    """Python-loop semantics (jax_dft scf.kohn_sham): dataset grid, nuclei -2.4 / 2.48 (centre between
    grid points), reflection about the nuclei, default mixing (alpha 0.5, decay 0.9, m=2)."""
    from qex_amd.od import scf, utils, xc

    z = np.load(os.path.join(golden_dir, "od_h2_rows.npz"))
    grids, loc, Z = z["grids"], z["locations"][1], z["nuclear_charges"][1]
    ref = O.kohn_sham(loc, Z, 2, 10, grids, O.LDAExp(), O.exponential_coulomb, None,
                      enforce_reflection_symmetry=True)
    st = scf.kohn_sham(loc, Z, 2, 10, grids, xc.LDAExp(), utils.exponential_coulomb,
                       enforce_reflection_symmetry=True)
    _compare(st, ref, O.get_dx(grids))


@pytest.mark.parametrize("interaction", ["soft", "exp"])
def test_reference_single_iteration_invariants(interaction):
    """tests/test_scf_od.py:88-142: grid 101 on [-5,5], nuclei +-0.5, Z=1, N_e=2, n0 = 2 gaussian,
    f = -0.73855 (n + 1e-10)^(1/3) summed (amplitude encoding), one iteration (alpha = 1 keeps the
    unmixed output).  (n + 1e-10)^(-2/3) makes v_xc ~ -1e5 Ha where n ~ 0, so the lowest two states are
    degenerate edge states and the output density is not unique: like the reference test, this checks
    the state's invariants; parity is pinned on the well-posed variant below."""
    from qex_amd.od import scf, utils, xc

    grids = np.linspace(-5, 5, 101)
    loc, Z = np.array([-0.5, 0.5]), np.array([1, 1])
    n0 = 2 * O.gaussian(grids, 0.0, 1.0)
    fo = O.exponential_coulomb if interaction == "exp" else O.soft_coulomb
    fd = utils.exponential_coulomb if interaction == "exp" else utils.soft_coulomb
    st = scf.kohn_sham_amplitude_encoded(loc, Z, 2, 1, grids, xc.PowerLaw(eps=1e-10), fd, initial_density=n0,
                                         alpha=1.0)
    s = st.numpy()
    assert abs(np.sum(s.density[0]) * 0.1 - 2.0) < 1e-7
    assert np.isfinite(s.total_energy[0]) and s.gap[0] >= 0
    assert s.hartree_potential.shape[-1] == 101 and s.xc_potential.shape[-1] == 101
    np.testing.assert_allclose(s.locations[0], loc)
    np.testing.assert_allclose(s.nuclear_charges[0], Z)
    np.testing.assert_allclose(s.grids[0], grids)
    np.testing.assert_allclose(s.external_potential[0], O.get_atomic_chain_potential(grids, loc, Z, fo), atol=1e-13)
    assert int(s.num_electrons[0]) == 2
    # the input-density quantities are unique: compare them with the oracle
    ref = O.kohn_sham(loc, Z, 2, 1, grids, O.PowerGlobal(eps=1e-10), fo, n0, alpha=1.0)
    np.testing.assert_allclose(s.hartree_potential, ref.hartree_potential, rtol=0, atol=1e-12)
    np.testing.assert_allclose(s.xc_potential, ref.xc_potential, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(s.xc_energy_density, ref.xc_energy_density, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("interaction", ["soft", "exp"])
def test_single_iteration_parity_well_posed(interaction):
    """The same case with f = -0.73855 (n + 1e-2)^(1/3): a unique ground state; two iterations."""
    from qex_amd.od import scf, utils, xc

    grids = np.linspace(-5, 5, 101)
    loc, Z = np.array([-0.5, 0.5]), np.array([1, 1])
    n0 = 2 * O.gaussian(grids, 0.0, 1.0)
    fo = O.exponential_coulomb if interaction == "exp" else O.soft_coulomb
    fd = utils.exponential_coulomb if interaction == "exp" else utils.soft_coulomb
    ref = O.kohn_sham(loc, Z, 2, 2, grids, O.PowerGlobal(eps=1e-2), fo, n0)
    st = scf.kohn_sham_amplitude_encoded(loc, Z, 2, 2, grids, xc.PowerLaw(eps=1e-2), fd, initial_density=n0)
    _compare(st, ref, O.get_dx(grids))
    assert st.numpy().gap[0] > 0


def _global_mlp(widths, seed, scale=2.0, activation="tanh"):
    W, b = O.glorot_normal_mlp(seed, widths)
    rng = np.random.default_rng(seed + 100)
    W = [w * scale for w in W]
    b = [rng.normal(0, 0.05, bb.shape) for bb in b]
    return W, b


@pytest.mark.parametrize("widths,activation", [([101, 32, 1], "tanh"), ([101, 64, 32, 16, 1], "silu"),
                                               ([101, 128, 1], "softplus")])
def test_global_mlp_python_loop_parity(widths, activation):
    """Amplitude-encoded Kohn-Sham (od/scf.py:434-560) with a global MLP functional
    (classical_models.py:177-224): E_xc = f(n) dx, v_xc = dx df/dn, three iterations, H2."""
    from qex_amd.od import scf, utils, xc

    grids = np.linspace(-5, 5, 101)
    loc, Z = np.array([-0.7, 0.7]), np.array([1, 1])
    n0 = 2 * O.gaussian(grids, 0.0, 1.0)
    W, b = _global_mlp(widths, 21 + len(widths), activation=activation)
    ref = O.kohn_sham(loc, Z, 2, 3, grids, O.GlobalMLP(W, b, activation, 2.0), O.exponential_coulomb, n0)
    f = xc.GlobalMLPXC(list(zip(W, b)), activation=activation, norm=2.0)
    st = scf.kohn_sham_amplitude_encoded(loc, Z, 2, 3, grids, f, utils.exponential_coulomb, initial_density=n0)
    _compare(st, ref, O.get_dx(grids))
    s = st.numpy()
    np.testing.assert_allclose(s.xc_energy_density, ref.xc_energy_density, rtol=1e-10, atol=1e-10)


def test_global_mlp_scan_loop_parity_batched():
    """The scan loop (od/jit_scf.py) with a global MLP over a batch of chains of different sizes
    (H2, H3+ padded, H4): every system against its own oracle run."""
    from qex_amd.od import engine, utils, xc

    grids = np.linspace(-6, 6, 121)
    W, b = _global_mlp([121, 48, 24, 1], 5)
    f = xc.GlobalMLPXC(list(zip(W, b)), activation="tanh", norm=2.0)
    systems = [(np.array([-0.7, 0.7]), np.array([1.0, 1.0]), 2),
               (np.array([-1.4, 0.0, 1.4]), np.array([1.0, 1.0, 1.0]), 2),
               (np.array([-2.1, -0.7, 0.7, 2.1]), np.array([1.0, 1.0, 1.0, 1.0]), 4)]
    prob = engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, 4, loop="scan")
    n0 = engine.solve_noninteracting(prob.vext, prob.num_electrons, prob.grids).density
    res = prob.run(n0)
    for i, (loc, Z, ne) in enumerate(systems):
        ref = O.jit_kohn_sham(loc, Z, ne, 4, grids, O.GlobalMLP(W, b, "tanh", 2.0), O.exponential_coulomb,
                              n0[i].cpu().numpy())
        for t in range(4):
            assert abs(float(res.total_energy[i, t]) - ref.total_energy[t]) < E_TOL, (i, t)
            assert _l2(res.density[i, t].cpu().numpy(), ref.density[t], 0.1) < L2_TOL, (i, t)
            assert _l2(res.xc_potential[i, t].cpu().numpy(), ref.xc_potential[t], 0.1) < 1e-6, (i, t)


@pytest.mark.parametrize("tol,expected", [(-1.0, [False, False, False]), (np.inf, [True, True, True])])
def test_python_loop_convergence_flags(tol, expected):
    """tests/test_scf_od.py:184-215."""
    from qex_amd.od import scf, utils, xc

    grids = np.linspace(-5, 5, 101)
    st = scf.kohn_sham_amplitude_encoded(np.array([-0.5, 0.5]), np.array([1, 1]), 2, 3, grids, xc.PowerLaw(),
                                         utils.exponential_coulomb, density_mse_converge_tolerance=tol)
    assert list(st.numpy().converged) == expected


@pytest.mark.parametrize("tol,expected", [(1e-2, [False, False, False]), (1e10, [False, True, True])])
def test_scan_loop_convergence_flags(tol, expected):
    """tests/test_jit_scf_od.py:186-229."""
    from qex_amd.od import jit_scf, utils, xc

    grids = np.linspace(-5, 5, 101)
    n0 = 2 * O.gaussian(grids, 0.0, 1.0)
    st = jit_scf.kohn_sham_amplitude_encoded(np.array([-0.5, 0.5]), np.array([1, 1]), 2, 3, grids, xc.PowerLaw(),
                                             utils.exponential_coulomb, n0, density_mse_converge_tolerance=tol)
    s = st.numpy()
    assert list(s.converged) == expected
    # f' = c p n^(-2/3) diverges where n ~ 0, so the output density is not unique (degenerate edge
    # states); parity is checked on the regularised functional instead
    ref = O.jit_kohn_sham(np.array([-0.5, 0.5]), np.array([1, 1]), 2, 3, grids, O.PowerGlobal(eps=1e-2),
                          O.exponential_coulomb, n0, density_mse_converge_tolerance=tol)
    st = jit_scf.kohn_sham_amplitude_encoded(np.array([-0.5, 0.5]), np.array([1, 1]), 2, 3, grids,
                                             xc.PowerLaw(eps=1e-2), utils.exponential_coulomb, n0,
                                             density_mse_converge_tolerance=tol)
    _compare(st, ref, 0.1)


def _c2_batch(count, seed=1234):
    grids = O.h2_config_grid()
    mols = O.random_molecules(seed, count, grids=grids)
    W, b = O.glorot_normal_mlp(seed, [17, 64, 64, 1])
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in mols])
    return grids, mols, W, b, n0


def test_c2_local_mlp_batch_parity():
    """Config 2 shape: uneven 1-4 nucleus systems, LocalMLP-17 [64,64] SiLU, T=20, alpha=0.3, FP64 plain path."""
    from qex_amd.od import engine, xc

    grids, mols, W, b, n0 = _c2_batch(6)
    f = xc.LocalMLPXC(list(zip(W, b)), window=17, activation="silu", norm=2.0)
    from qex_amd.od import utils

    prob = engine.Problem1D(grids, [m[0] for m in mols], [m[1] for m in mols], [m[2] for m in mols],
                            utils.exponential_coulomb, f, 20, alpha=0.3, alpha_decay=1.0, num_mixing_iterations=1,
                            loop="scan")
    res = prob.run(n0)
    E = _np(res.total_energy)
    D = _np(res.density)
    oxc = O.LocalMLP(W, b, "silu", 17, 2.0)
    dx = O.get_dx(grids)
    for i, (l, z, ne) in enumerate(mols):
        ref = O.jit_kohn_sham(l, z, ne, 20, grids, oxc, O.exponential_coulomb, n0[i], alpha=0.3, alpha_decay=1.0,
                              num_mixing_iterations=1)
        # early iterations are well conditioned: tight; the full trajectory to the metric's bar
        assert np.max(np.abs(E[i, :5] - ref.total_energy[:5])) < 1e-9, i
        assert abs(E[i, -1] - ref.total_energy[-1]) < E_TOL, (i, E[i, -1], ref.total_energy[-1])
        assert _l2(D[i, -1], ref.density[-1], dx) < L2_TOL, i
        np.testing.assert_allclose(_np(res.xc_potential[i, 0]), ref.xc_potential[0], rtol=0, atol=1e-10)


def test_batch_independence_bitwise():
    """A system's result is bit-identical alone and inside a batch (fixed-order reductions)."""
    from qex_amd.od import engine, utils, xc

    grids, mols, W, b, n0 = _c2_batch(5, seed=7)
    f = xc.LDAExp()
    common = dict(alpha=0.3, alpha_decay=1.0, num_mixing_iterations=1, loop="scan")
    prob = engine.Problem1D(grids, [m[0] for m in mols], [m[1] for m in mols], [m[2] for m in mols],
                            utils.exponential_coulomb, f, 12, **common)
    res = prob.run(n0)
    for i in (0, 3):
        p1 = engine.Problem1D(grids, [mols[i][0]], [mols[i][1]], [mols[i][2]], utils.exponential_coulomb, f, 12,
                              **common)
        r1 = p1.run(n0[i:i + 1])
        assert torch.equal(r1.total_energy[0], res.total_energy[i])
        assert torch.equal(r1.density[0], res.density[i])


def test_eval_kohn_sham_lda_python_loop():
    """qedft od/eval.py:23-89 with use_lda=True (python loop, reflection about the nuclei)."""
    from qex_amd.od import eval as qeval

    grids = np.linspace(-20.48, 20.48, 513)
    loc = np.array([-0.64, 0.64])
    ref = O.kohn_sham(loc, np.array([1, 1]), 2, 8, grids, O.LDAExp(), O.exponential_coulomb, None,
                      enforce_reflection_symmetry=True)
    st = qeval.kohn_sham(None, loc, np.array([1, 1]), 2, 8, grids, None, use_amplitude_encoding=False, use_lda=True)
    _compare(st, ref, O.get_dx(grids))


def test_notebook_width_global_mlp_513():
    """The reference notebook's global network (notebooks/models/02b_notebook_od_mlp.ipynb:645:
    build_global_mlp(n_neurons=513, n_layers=2, activation="tanh")) on the 513-point grid: hidden
    layers wider than a workgroup, amplitude-encoded loop, three iterations against the oracle."""
    from qex_amd.od import scf, utils, xc

    grids = np.linspace(-25.6, 25.6, 513)
    loc, Z = np.array([-0.7, 0.7]), np.array([1, 1])
    n0 = 2 * O.gaussian(grids, 0.0, 1.0)
    W, b = _global_mlp([513, 513, 513, 1], 31, scale=0.5)
    ref = O.kohn_sham(loc, Z, 2, 3, grids, O.GlobalMLP(W, b, "tanh", 2.0), O.exponential_coulomb, n0)
    f = xc.GlobalMLPXC(list(zip(W, b)), activation="tanh", norm=2.0)
    st = scf.kohn_sham_amplitude_encoded(loc, Z, 2, 3, grids, f, utils.exponential_coulomb, initial_density=n0)
    _compare(st, ref, O.get_dx(grids))
    np.testing.assert_allclose(st.numpy().xc_energy_density, ref.xc_energy_density, rtol=1e-10, atol=1e-10)


def test_notebook_width_local_mlp_513():
    """The reference notebook's local network (02b_notebook_od_mlp.ipynb:370: build_local_mlp(
    n_neurons=num_grids, n_layers=2, activation="tanh"), one density value per point) with 513-wide
    hidden layers: the generic kernel with 64 lanes per point (32 points' activations do not fit in
    LDS), scan loop, four iterations against the oracle."""
    from qex_amd.od import engine, utils, xc

    grids = np.linspace(-25.6, 25.6, 513)
    W, b = _global_mlp([1, 513, 513, 1], 37, scale=0.5)
    f = xc.LocalMLPXC(list(zip(W, b)), window=1, activation="tanh", norm=2.0)
    loc, Z = np.array([-0.7, 0.7]), np.array([1.0, 1.0])
    prob = engine.Problem1D(grids, [loc], [Z], [2], utils.exponential_coulomb, f, 4, loop="scan")
    n0 = engine.solve_noninteracting(prob.vext, prob.num_electrons, prob.grids).density
    res = prob.run(n0)
    ref = O.jit_kohn_sham(loc, Z, 2, 4, grids, O.LocalMLP(W, b, "tanh", 1, 2.0), O.exponential_coulomb,
                          n0[0].cpu().numpy())
    for t in range(4):
        assert abs(float(res.total_energy[0, t]) - ref.total_energy[t]) < E_TOL, t
        assert _l2(res.density[0, t].cpu().numpy(), ref.density[t], O.get_dx(grids)) < L2_TOL, t
        np.testing.assert_allclose(_np(res.xc_potential[0, t]), ref.xc_potential[t], rtol=0, atol=1e-9)
