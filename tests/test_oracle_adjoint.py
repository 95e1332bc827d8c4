"""The hand-written reverse pass of the 1D scan trajectory (oracle/od_adjoint.py) against central
finite differences of the forward loss, for several activations, odd electron counts and mixing
over two iterations; and its pieces (MLP second-order adjoint, eigen adjoint) on their own."""

import numpy as np
import pytest

from oracle import od as O
from oracle import od_adjoint as A


def _net(seed, sizes, act):
    W, b = O.glorot_normal_mlp(seed, sizes)
    rng = np.random.default_rng(seed + 7)
    b = [rng.standard_normal(x.shape) * 0.1 for x in b]
    return list(zip(W, b))


@pytest.mark.parametrize("act", ["silu", "tanh", "gelu", "softplus", "sigmoid"])
def test_mlp_second_order_matches_finite_differences(act):
    rng = np.random.default_rng(1)
    params = _net(2, [5, 7, 6, 1], act)
    W, b = [w for w, _ in params], [x for _, x in params]
    X = rng.standard_normal((11, 5))
    Xd = rng.standard_normal((11, 5))
    fbar, fdbar = rng.standard_normal(11), rng.standard_normal(11)

    def phi(X_, W_):
        f, fd, _, _, _ = A.mlp_second_order(W_, b, act, X_, Xd, np.zeros(11), np.zeros(11))
        return float(fbar @ f + fdbar @ fd)

    _, _, Xbar, dW, _ = A.mlp_second_order(W, b, act, X, Xd, fbar, fdbar)
    h = 1e-6
    for (p, q) in [(0, 0), (3, 2), (10, 4)]:
        E = np.zeros_like(X)
        E[p, q] = h
        fd_ = (phi(X + E, W) - phi(X - E, W)) / (2 * h)
        assert abs(fd_ - Xbar[p, q]) < 1e-7 * max(1, abs(fd_))
    for li, (p, q) in [(0, (1, 2)), (1, (3, 4)), (2, (5, 0))]:
        Wp = [w.copy() for w in W]
        Wm = [w.copy() for w in W]
        Wp[li][p, q] += h
        Wm[li][p, q] -= h
        fd_ = (phi(X, Wp) - phi(X, Wm)) / (2 * h)
        assert abs(fd_ - dW[li][p, q]) < 1e-7 * max(1, abs(fd_))


def test_vxc_vjp_is_hessian_vector_product():
    grids = np.linspace(-5, 5, 41)
    p = _net(3, [5, 6, 6, 1], "silu")
    xc = O.LocalMLP([w for w, _ in p], [b for _, b in p], "silu", 5, 2.0)
    adj = A.LocalMLPAdjoint(xc)
    rng = np.random.default_rng(4)
    n = np.abs(rng.standard_normal(41)) * 0.3
    vbar = rng.standard_normal(41)
    nbar, _ = adj.vxc_vjp(n, vbar)
    h = 1e-6
    fd = (xc.potential(n + h * vbar) - xc.potential(n - h * vbar)) / (2 * h)  # H v (H symmetric)
    np.testing.assert_allclose(nbar, fd, atol=1e-7)
    # gradient of F matches the forward potential
    _, g, _ = adj.energy_and_grads(n)
    np.testing.assert_allclose(g, xc.potential(n), atol=1e-12)
    assert grids.size == 41


def test_eig_adjoint_matches_finite_differences():
    grids = np.linspace(-6, 6, 61)
    dx = O.get_dx(grids)
    T = O.get_kinetic_matrix(grids)
    rng = np.random.default_rng(5)
    v = -2.0 * np.exp(-grids ** 2) + 0.01 * rng.standard_normal(61)
    occ = A.occupations(3)

    def nout(v_):
        eps, U = np.linalg.eigh(T + np.diag(v_))
        return (occ[:, None] * U[:, :2].T ** 2).sum(0) / dx, eps, U

    nbar = rng.standard_normal(61)
    n0, eps, U = nout(v)
    vbar = A.eig_adjoint(eps, U, occ, nbar, dx)
    d = rng.standard_normal(61)
    h = 1e-6
    fd = (nbar @ nout(v + h * d)[0] - nbar @ nout(v - h * d)[0]) / (2 * h)
    assert abs(fd - vbar @ d) < 1e-7 * max(1, abs(fd))


@pytest.mark.parametrize("act,ne,m", [("silu", 2, 1), ("silu", 3, 2), ("tanh", 4, 1), ("gelu", 1, 2)])
def test_trajectory_gradient_matches_finite_differences(act, ne, m):
    grids = np.linspace(-8, 8, 129)
    dx = O.get_dx(grids)
    params = _net(3, [5, 8, 8, 1], act)
    loc = np.array([-0.7, 0.8]) if ne > 1 else np.array([0.1])
    z = np.array([1.0, float(ne - 1)]) if ne > 1 else np.array([1.0])
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True)
    rng = np.random.default_rng(11)

    def run(flat):
        p = A.unflatten(flat, params)
        xc = O.LocalMLP([w for w, _ in p], [b for _, b in p], act, 5, 2.0)
        fw = A.forward_trajectory(loc, z, ne, 6, grids, xc, O.exponential_coulomb, n0, 0.3, 0.9, m)
        L1, e1, d1 = A.c3_loss_and_cotangents(fw["energies"], fw["densities"], -1.5, 0.95 * n0, dx)
        L2, e2, d2 = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], -1.5, 0.95 * n0, ne, dx,
                                                      start=2)
        return L1 + L2, fw, e1 + e2, d1 + d2

    th = A.flatten(params)
    _, fw, dE, dn = run(th)
    g = A.backward(fw, dE, dn)
    h = 1e-5
    for _ in range(2):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (run(th + h * d)[0] - run(th - h * d)[0]) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * max(1.0, abs(fd))


def test_forward_trajectory_matches_scan_oracle():
    grids = np.linspace(-8, 8, 129)
    params = _net(5, [17, 8, 8, 1], "silu")
    xc = O.LocalMLP([w for w, _ in params], [b for _, b in params], "silu", 17, 2.0)
    loc, z, ne = np.array([-0.7, 0.7]), np.array([1.0, 1.0]), 2
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True)
    fw = A.forward_trajectory(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, 0.3, 1.0, 1)
    st = O.jit_kohn_sham(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, alpha=0.3, alpha_decay=1.0,
                         num_mixing_iterations=1)
    np.testing.assert_allclose(fw["energies"], st.total_energy, atol=1e-10)
    np.testing.assert_allclose(fw["densities"], st.density, atol=1e-10)


def test_negativity_transform_is_minus_swish():
    """-swish(y) = -y sigmoid(y): 0 at 0, maximum 0.2785 at y = -1.2785 (where sigmoid(y)(1 + y(1 -
    sigmoid(y))) = 0), -> -y for large y; the derivatives are checked by finite differences."""
    y = np.linspace(-8, 8, 161)
    g, d1, d2 = A.output3("negativity", y)
    np.testing.assert_allclose(g, -y / (1 + np.exp(-y)), rtol=1e-15)
    assert g.max() < 0.2785 and g.max() > 0.27
    h = 1e-6
    np.testing.assert_allclose(d1, (A.output3("negativity", y + h)[0] - A.output3("negativity", y - h)[0]) / (2 * h),
                               atol=1e-8)
    np.testing.assert_allclose(d2, (A.output3("negativity", y + h)[1] - A.output3("negativity", y - h)[1]) / (2 * h),
                               atol=1e-8)
    e, de = O.output_transform("negativity", y)
    np.testing.assert_array_equal(e, g)
    np.testing.assert_array_equal(de, d1)


def test_mlp_second_order_with_output_transform_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = _net(4, [5, 7, 1], "tanh")
    W, b = [w for w, _ in params], [x for _, x in params]
    X, Xd = rng.standard_normal((9, 5)), rng.standard_normal((9, 5))
    fbar, fdbar = rng.standard_normal(9), rng.standard_normal(9)

    def phi(X_, W_):
        f, fd, _, _, _ = A.mlp_second_order(W_, b, "tanh", X_, Xd, np.zeros(9), np.zeros(9), "negativity")
        return float(fbar @ f + fdbar @ fd)

    _, _, Xbar, dW, db = A.mlp_second_order(W, b, "tanh", X, Xd, fbar, fdbar, "negativity")
    h = 1e-6
    for (p, q) in [(0, 1), (5, 3)]:
        E = np.zeros_like(X)
        E[p, q] = h
        assert abs((phi(X + E, W) - phi(X - E, W)) / (2 * h) - Xbar[p, q]) < 1e-7
    for li, (p, q) in [(0, (2, 3)), (1, (4, 0))]:
        Wp, Wm = [w.copy() for w in W], [w.copy() for w in W]
        Wp[li][p, q] += h
        Wm[li][p, q] -= h
        assert abs((phi(X, Wp) - phi(X, Wm)) / (2 * h) - dW[li][p, q]) < 1e-7


@pytest.mark.parametrize("output,window", [("negativity", 1), ("negativity", 5), ("linear", 3)])
def test_reflected_trajectory_gradient_matches_finite_differences(output, window):
    """enforce_reflection_symmetry about the grid centre (what KSDFTTrainer always sets,
    qedft/train/od/train.py:110) with the output transform of wrap_with_negative_transform."""
    grids = np.linspace(-8, 8, 129)
    dx = O.get_dx(grids)
    params = _net(6, [window, 8, 8, 1], "tanh")
    loc, z, ne = np.array([-0.75, 0.75]), np.array([1.0, 1.0]), 2
    rng = np.random.default_rng(2)
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True) * (1 + 0.05 * rng.standard_normal(129) ** 2)

    def run(flat):
        p = A.unflatten(flat, params)
        xc = O.LocalMLP([w for w, _ in p], [b for _, b in p], "tanh", window, 2.0, output=output)
        fw = A.forward_trajectory(loc, z, ne, 6, grids, xc, O.exponential_coulomb, n0, 0.5, 0.9, 1, reflect=True)
        L, e1, d1 = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], -1.4, 0.95 * n0, ne, dx,
                                                     start=2)
        return L, fw, e1, d1

    th = A.flatten(params)
    _, fw, dE, dn = run(th)
    g = A.backward(fw, dE, dn)
    h = 1e-5
    for _ in range(2):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (run(th + h * d)[0] - run(th - h * d)[0]) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * max(1.0, abs(fd))
    # the reflected forward is jit_kohn_sham with enforce_reflection_symmetry (S e(S n) wrapper)
    xc = O.LocalMLP([w for w, _ in params], [b for _, b in params], "tanh", window, 2.0, output=output)
    st = O.jit_kohn_sham(loc, z, ne, 6, grids, xc, O.exponential_coulomb, n0, alpha=0.5, alpha_decay=0.9,
                         enforce_reflection_symmetry=True, num_mixing_iterations=1)
    np.testing.assert_allclose(fw["energies"], st.total_energy, atol=1e-10)
    np.testing.assert_allclose(fw["densities"], st.density, atol=1e-10)


def test_stop_gradient_step_is_the_gradient_of_the_restarted_loss():
    """With states of iterations <= s held constant (jax.lax.stop_gradient in jit_scf), and one mixing
    difference, the gradient is that of the loss of a run that restarts from the stored density of
    iteration s (mixing factor alpha decay^(s+1)) and only sees iterations s+1..T-1; checked by central
    differences of that restarted run."""
    grids = np.linspace(-8, 8, 129)
    dx = O.get_dx(grids)
    params = _net(8, [5, 8, 1], "tanh")
    loc, z, ne = np.array([-0.7, 0.8]), np.array([1.0, 1.0]), 2
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True)
    T, s, alpha, decay = 7, 2, 0.5, 0.9
    xc0 = O.LocalMLP([w for w, _ in params], [b for _, b in params], "tanh", 5, 2.0)
    fw = A.forward_trajectory(loc, z, ne, T, grids, xc0, O.exponential_coulomb, n0, alpha, decay, 1)
    _, dE, dn = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], -1.3, 0.9 * n0, ne, dx, start=0)
    g = A.backward(fw, dE, dn, stop_gradient_step=s)
    n_restart = fw["densities"][s]
    tail_E, tail_n = dE[s + 1:].copy(), dn[s + 1:].copy()

    def restarted(flat):
        p = A.unflatten(flat, params)
        xc = O.LocalMLP([w for w, _ in p], [b for _, b in p], "tanh", 5, 2.0)
        f2 = A.forward_trajectory(loc, z, ne, T - s - 1, grids, xc, O.exponential_coulomb, n_restart,
                                  alpha * decay ** (s + 1), decay, 1)
        return float(tail_E @ f2["energies"] + np.sum(tail_n * f2["densities"]))

    th = A.flatten(params)
    rng = np.random.default_rng(4)
    h = 1e-5
    for _ in range(2):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (restarted(th + h * d) - restarted(th - h * d)) / (2 * h)
        assert abs(fd - g @ d) < 1e-7 * max(1.0, abs(fd))
    assert np.abs(A.backward(fw, dE, dn, stop_gradient_step=T - 1)).max() == 0.0
    assert np.abs(g - A.backward(fw, dE, dn)).max() > 1e-6


def test_global_vxc_vjp_is_hessian_vector_product():
    """GlobalMLPAdjoint: nbar of vbar . v_xc is the derivative of v_xc = dx grad f along vbar, and the
    energy gradient is the forward grad f."""
    grids = np.linspace(-5, 5, 41)
    dx = O.get_dx(grids)
    p = _net(3, [41, 9, 7, 1], "tanh")
    xc = O.GlobalMLP([w for w, _ in p], [b for _, b in p], "tanh", 2.0, output="negativity")
    adj = A.GlobalMLPAdjoint(xc, dx)
    rng = np.random.default_rng(4)
    n = np.abs(rng.standard_normal(41)) * 0.3
    vbar = rng.standard_normal(41)
    nbar, dth = adj.vxc_vjp(n, vbar)
    h = 1e-6
    fd = (adj.potential(n + h * vbar) - adj.potential(n - h * vbar)) / (2 * h)
    np.testing.assert_allclose(nbar, fd, atol=1e-8)
    f, g, _ = adj.energy_and_grads(n)
    assert f == pytest.approx(xc.value(n), abs=1e-14)
    np.testing.assert_allclose(g, xc.grad(n), atol=1e-13)
    # theta-gradient of vbar . v_xc against differences in one weight of each layer
    W, b = [w for w, _ in p], [x for _, x in p]
    for li, (i, j) in [(0, (5, 2)), (1, (3, 4)), (2, (6, 0))]:
        Wp = [w.copy() for w in W]
        Wm = [w.copy() for w in W]
        Wp[li][i, j] += h
        Wm[li][i, j] -= h
        vp = A.GlobalMLPAdjoint(O.GlobalMLP(Wp, b, "tanh", 2.0, output="negativity"), dx).potential(n)
        vm = A.GlobalMLPAdjoint(O.GlobalMLP(Wm, b, "tanh", 2.0, output="negativity"), dx).potential(n)
        assert abs((vp - vm) @ vbar / (2 * h) - dth[li][0][i, j]) < 1e-7


@pytest.mark.parametrize("act,ne,m", [("tanh", 2, 1), ("silu", 3, 2)])
def test_global_trajectory_gradient_matches_finite_differences(act, ne, m):
    grids = np.linspace(-8, 8, 65)
    dx = O.get_dx(grids)
    params = _net(6, [65, 8, 6, 1], act)
    loc = np.array([-0.7, 0.8])
    z = np.array([1.0, float(ne - 1)])
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True)
    rng = np.random.default_rng(12)

    def run(flat):
        p = A.unflatten(flat, params)
        xc = O.GlobalMLP([w for w, _ in p], [b for _, b in p], act, 2.0)
        fw = A.forward_trajectory(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, 0.3, 0.9, m)
        L, e, d = A.c3_loss_and_cotangents(fw["energies"], fw["densities"], -1.5, 0.95 * n0, dx)
        return L, fw, e, d

    th = A.flatten(params)
    _, fw, dE, dn = run(th)
    g = A.backward(fw, dE, dn)
    h = 1e-5
    for _ in range(2):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (run(th + h * d)[0] - run(th - h * d)[0]) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * max(1.0, abs(fd))


def test_global_forward_trajectory_matches_scan_oracle():
    grids = np.linspace(-8, 8, 65)
    params = _net(5, [65, 8, 8, 1], "tanh")
    xc = O.GlobalMLP([w for w, _ in params], [b for _, b in params], "tanh", 2.0)
    loc, z, ne = np.array([-0.7, 0.7]), np.array([1.0, 1.0]), 2
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True)
    fw = A.forward_trajectory(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, 0.3, 1.0, 1)
    st = O.jit_kohn_sham(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, alpha=0.3, alpha_decay=1.0,
                         num_mixing_iterations=1)
    np.testing.assert_allclose(fw["energies"], st.total_energy, atol=1e-10)
    np.testing.assert_allclose(fw["densities"], st.density, atol=1e-10)


def test_global_reflected_trajectory_matches_scan_oracle_and_differences():
    """Global functional under reflection (od/jit_scf.py:15-78): only the density passed on is
    flipped.  The trajectory equals the scan oracle's, and its gradient central differences."""
    grids = np.linspace(-8, 8, 65)
    dx = O.get_dx(grids)
    params = _net(8, [65, 8, 1], "tanh")
    loc, z, ne = np.array([-0.9, 0.6]), np.array([1.0, 1.0]), 2
    n0 = O.gaussian_initial_density(grids, loc, z, ne, weighted=True)

    def run(flat):
        p = A.unflatten(flat, params)
        xc = O.GlobalMLP([w for w, _ in p], [b for _, b in p], "tanh", 2.0)
        fw = A.forward_trajectory(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, 0.3, 0.9, 2, reflect=True)
        L, e, d = A.c3_loss_and_cotangents(fw["energies"], fw["densities"], -1.5, 0.95 * n0, dx)
        return L, fw, e, d

    th = A.flatten(params)
    _, fw, dE, dn = run(th)
    xc = O.GlobalMLP([w for w, _ in params], [b for _, b in params], "tanh", 2.0)
    st = O.jit_kohn_sham(loc, z, ne, 5, grids, xc, O.exponential_coulomb, n0, alpha=0.3, alpha_decay=0.9,
                         num_mixing_iterations=2, enforce_reflection_symmetry=True)
    np.testing.assert_allclose(fw["energies"], st.total_energy, atol=1e-10)
    np.testing.assert_allclose(fw["densities"], st.density, atol=1e-10)
    g = A.backward(fw, dE, dn)
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(2):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (run(th + h * d)[0] - run(th - h * d)[0]) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * max(1.0, abs(fd))
