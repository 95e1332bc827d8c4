"""Device reverse pass (qex1d_kohn_sham_vjp) against the oracle's hand-written reverse pass
(oracle/od_adjoint.py, itself pinned by finite differences) and against central finite
differences of the device forward loss."""

import numpy as np
import pytest
import torch

from oracle import od as O
from oracle import od_adjoint as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd.od import engine, train, utils, xc
    return engine, train, utils, xc


def _params(seed, sizes):
    W, b = O.glorot_normal_mlp(seed, sizes)
    rng = np.random.default_rng(seed + 3)
    b = [rng.standard_normal(x.shape) * 0.05 for x in b]
    return list(zip(W, b))


def _problem(mods, grids, systems, params, act, window, T, alpha, decay, m):
    engine, _, utils, xc = mods
    f = xc.LocalMLPXC([(w, b) for w, b in params], window=window, activation=act, norm=2.0)
    return engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, T, alpha=alpha, alpha_decay=decay,
                            num_mixing_iterations=m, loop="scan", want_orbitals=True)


def _systems():
    return [(np.array([-0.7, 0.7]), np.array([1.0, 1.0]), 2),
            (np.array([-1.1, 0.4]), np.array([1.0, 2.0]), 3),
            (np.array([-0.3, 1.2, 2.0]), np.array([2.0, 1.0, 1.0]), 4)]


@pytest.mark.parametrize("act,sizes,m", [("silu", [17, 16, 16, 1], 1), ("tanh", [9, 12, 1], 2),
                                         ("gelu", [5, 8, 8, 1], 2)])
def test_device_gradient_matches_oracle(mods, act, sizes, m):
    engine, train, _, _ = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = _systems()
    params = _params(4, sizes)
    T, alpha, decay = 6, 0.3, 0.9
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    prob = _problem(mods, grids, systems, params, act, sizes[0], T, alpha, decay, m)
    tE = np.array([-1.2, -2.5, -4.0])
    tn = 0.9 * n0
    loss, g, res = train.value_and_grad(prob, n0, train.c3_loss, tE, tn, dx)
    g = g.cpu().numpy()
    # oracle: per-system reverse passes with the same cotangents
    xc = O.LocalMLP([w for w, _ in params], [b for _, b in params], act, sizes[0], 2.0)
    go = np.zeros_like(g)
    B = len(systems)
    for s, (loc, z, ne) in enumerate(systems):
        fw = A.forward_trajectory(loc, z, ne, T, grids, xc, O.exponential_coulomb, n0[s], alpha, decay, m)
        np.testing.assert_allclose(res.total_energy[s].cpu().numpy(), fw["energies"], atol=1e-8)
        dE = np.zeros(T)
        dn = np.zeros((T, len(grids)))
        dE[-1] = 2 * (fw["energies"][-1] - tE[s]) / B
        dn[-1] = 2 * 1e-2 * dx * (fw["densities"][-1] - tn[s]) / B
        go += A.backward(fw, dE, dn)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-6 * scale)


def test_device_gradient_matches_finite_differences(mods):
    engine, train, _, _ = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = _systems()
    params = _params(9, [17, 16, 16, 1])
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    prob = _problem(mods, grids, systems, params, "silu", 17, 8, 0.3, 1.0, 1)
    tE, tn = np.array([-1.0, -2.0, -3.5]), 0.95 * n0
    _, g, _ = train.value_and_grad(prob, n0, train.trajectory_loss, tE, tn, 3, dx, start=4)
    th0 = prob.xc_params.clone()
    rng = np.random.default_rng(0)
    h = 1e-4
    for _ in range(4):
        d = torch.as_tensor(rng.standard_normal(th0.numel()), dtype=torch.float64, device=th0.device)
        d /= d.norm()
        vals = []
        for sgn in (1.0, -1.0):
            prob.xc_params.copy_(th0 + sgn * h * d)
            r = prob.run(n0)
            vals.append(float(train.trajectory_loss(r, tE, tn, 3, dx, start=4)[0]))
        prob.xc_params.copy_(th0)
        fd = (vals[0] - vals[1]) / (2 * h)
        gd = float(g @ d)
        assert abs(fd - gd) < 1e-6 * max(1.0, abs(fd))


def test_gradient_is_batch_additive_and_deterministic(mods):
    engine, train, _, _ = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = _systems()
    params = _params(2, [17, 16, 16, 1])
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    tE, tn = np.array([-1.0, -2.0, -3.5]), 0.95 * n0

    def grad_of(idx):
        prob = _problem(mods, grids, [systems[i] for i in idx], params, "silu", 17, 5, 0.3, 1.0, 1)
        res = prob.run(n0[idx])
        E = res.total_energy
        dE = torch.zeros_like(E)
        dE[:, -1] = 2.0 * (E[:, -1] - torch.as_tensor(tE[idx], device=E.device))
        return engine.kohn_sham_vjp(prob, res, dE, None)

    g_all = grad_of([0, 1, 2])
    g_sum = grad_of([0]) + grad_of([1]) + grad_of([2])
    torch.testing.assert_close(g_all, g_sum, rtol=1e-11, atol=1e-11 * float(g_all.abs().max()))
    assert torch.equal(grad_of([0, 1, 2]), g_all)


@pytest.mark.parametrize("output,window,act", [("negativity", 1, "tanh"), ("negativity", 17, "silu"),
                                               ("linear", 5, "tanh")])
def test_reflected_gradient_matches_oracle(mods, output, window, act):
    """The trainer's setting (enforce_reflection_symmetry about the grid centre, od/train.py:110) with
    the negativity output transform (wrappers.py:201-203): device reverse pass against the oracle's."""
    engine, train, utils, xc = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = [(np.array([-0.7, 0.7]), np.array([1.0, 1.0]), 2), (np.array([-1.5, 1.5]), np.array([1.0, 1.0]), 2),
               (np.array([-0.4, 0.4]), np.array([2.0, 2.0]), 4)]
    params = _params(7, [window, 16, 16, 1])
    T, alpha, decay = 6, 0.5, 0.9
    rng = np.random.default_rng(1)
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    n0 = n0 * (1 + 0.05 * rng.standard_normal(n0.shape) ** 2)  # not symmetric: the flip matters
    f = xc.LocalMLPXC(params, window=window, activation=act, norm=2.0, output=output)
    prob = engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, T, alpha=alpha, alpha_decay=decay, num_mixing_iterations=1,
                            loop="scan", want_orbitals=True, enforce_reflection_symmetry=True)
    tE = np.array([-1.2, -0.9, -4.0])
    tn = 0.9 * n0
    _, g, res = train.value_and_grad(prob, n0, train.trajectory_loss, tE, tn, 2, dx, start=2)
    g = g.cpu().numpy()
    oxc = O.LocalMLP([w for w, _ in params], [b for _, b in params], act, window, 2.0, output=output)
    go = np.zeros_like(g)
    B = len(systems)
    for s, (loc, z, ne) in enumerate(systems):
        fw = A.forward_trajectory(loc, z, ne, T, grids, oxc, O.exponential_coulomb, n0[s], alpha, decay, 1,
                                  reflect=True)
        np.testing.assert_allclose(res.total_energy[s].cpu().numpy(), fw["energies"], atol=1e-8)
        np.testing.assert_allclose(res.density[s].cpu().numpy(), fw["densities"], atol=1e-8)
        _, dE, dn = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], tE[s], tn[s], 2, dx, start=2)
        go += A.backward(fw, dE / B, dn / B)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-6 * scale)


def test_python_loop_off_centre_reflection_gradient_matches_oracle(mods):
    """scf.kohn_sham's loop (od/scf.py:434-560; reflection about the centre of the nuclei) with a
    negative convergence tolerance: the device reverse pass against the oracle's with that flip."""
    engine, train, utils, xc = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = [(np.array([0.3, 1.7]), np.array([1.0, 1.0]), 2), (np.array([-2.2, -0.6]), np.array([1.0, 2.0]), 3)]
    params = _params(11, [5, 12, 12, 1])
    T, alpha, decay = 5, 0.5, 0.9
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    f = xc.LocalMLPXC(params, window=5, activation="tanh", norm=2.0, output="negativity")
    prob = engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, T, alpha=alpha, alpha_decay=decay, num_mixing_iterations=1,
                            loop="python", want_orbitals=True, enforce_reflection_symmetry=True)
    tE, tn = np.array([-1.1, -2.3]), 0.9 * n0
    _, g, res = train.value_and_grad(prob, n0, train.trajectory_loss, tE, tn, 2, dx, start=1)
    g = g.cpu().numpy()
    oxc = O.LocalMLP([w for w, _ in params], [b for _, b in params], "tanh", 5, 2.0, output="negativity")
    go = np.zeros_like(g)
    for s, (loc, z, ne) in enumerate(systems):
        flip = lambda a, loc=loc: O.flip_and_average(loc, grids, a)
        fw = A.forward_trajectory(loc, z, ne, T, grids, oxc, O.exponential_coulomb, n0[s], alpha, decay, 1,
                                  reflect=True, flip=flip)
        np.testing.assert_allclose(res.total_energy[s].cpu().numpy(), fw["energies"], atol=1e-8)
        _, dE, dn = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], tE[s], tn[s], 2, dx, start=1)
        go += A.backward(fw, dE / 2, dn / 2)
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-6 * np.abs(go).max())


def test_stop_gradient_step_matches_oracle(mods):
    """stop_gradient_step (jit_scf.py:199-204 via train.py:113): the device reverse pass with the
    first iterations held constant against the oracle's (itself pinned by differences of the
    restarted run)."""
    engine, train, utils, xc = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = _systems()
    params = _params(13, [9, 12, 12, 1])
    T, s_stop = 7, 3
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    f = xc.LocalMLPXC(params, window=9, activation="silu", norm=2.0)
    prob = engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, T, alpha=0.4, alpha_decay=0.9, num_mixing_iterations=1,
                            loop="scan", want_orbitals=True, stop_gradient_step=s_stop)
    tE, tn = np.array([-1.2, -2.5, -4.0]), 0.9 * n0
    _, g, _ = train.value_and_grad(prob, n0, train.trajectory_loss, tE, tn, 3, dx, start=0)
    oxc = O.LocalMLP([w for w, _ in params], [b for _, b in params], "silu", 9, 2.0)
    go = 0
    for s, (loc, z, ne) in enumerate(systems):
        fw = A.forward_trajectory(loc, z, ne, T, grids, oxc, O.exponential_coulomb, n0[s], 0.4, 0.9, 1)
        _, dE, dn = A.trajectory_loss_and_cotangents(fw["energies"], fw["densities"], tE[s], tn[s], 3, dx, start=0)
        go = go + A.backward(fw, dE / 3, dn / 3, stop_gradient_step=s_stop)
    g = g.cpu().numpy()
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-6 * np.abs(go).max())


@pytest.mark.parametrize("act,sizes,m,out", [("tanh", [257, 16, 8, 1], 1, "linear"),
                                             ("silu", [257, 24, 1], 2, "negativity")])
def test_global_mlp_device_gradient_matches_oracle(mods, act, sizes, m, out):
    """GLOBAL_MLP (amplitude encoded) reverse pass: the device gradient against the oracle's
    reverse pass (GlobalMLPAdjoint, pinned by finite differences in tests/test_oracle_adjoint.py)."""
    engine, train, utils, xc = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = _systems()
    params = _params(6, sizes)
    T, alpha, decay = 6, 0.3, 0.9
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    f = xc.GlobalMLPXC([(w, b) for w, b in params], activation=act, norm=2.0, output=out)
    prob = engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, T, alpha=alpha, alpha_decay=decay,
                            num_mixing_iterations=m, loop="scan", want_orbitals=True)
    tE = np.array([-1.2, -2.5, -4.0])
    tn = 0.9 * n0
    _, g, res = train.value_and_grad(prob, n0, train.c3_loss, tE, tn, dx)
    g = g.cpu().numpy()
    oxc = O.GlobalMLP([w for w, _ in params], [b for _, b in params], act, 2.0, output=out)
    go = np.zeros_like(g)
    B = len(systems)
    for s, (loc, z, ne) in enumerate(systems):
        fw = A.forward_trajectory(loc, z, ne, T, grids, oxc, O.exponential_coulomb, n0[s], alpha, decay, m)
        np.testing.assert_allclose(res.total_energy[s].cpu().numpy(), fw["energies"], atol=1e-8)
        dE = np.zeros(T)
        dn = np.zeros((T, len(grids)))
        dE[-1] = 2 * (fw["energies"][-1] - tE[s]) / B
        dn[-1] = 2 * 1e-2 * dx * (fw["densities"][-1] - tn[s]) / B
        go += A.backward(fw, dE, dn)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-6 * scale)


def test_global_mlp_reflected_device_gradient_matches_oracle(mods):
    """GLOBAL_MLP with reflection about the grid centre (the trainer's setting): only the density
    passed on is flipped, the functional sees unflipped densities (od/jit_scf.py:15-78)."""
    engine, train, utils, xc = mods
    grids = np.linspace(-12.8, 12.8, 257)
    dx = float(grids[1] - grids[0])
    systems = _systems()
    params = _params(7, [257, 16, 16, 1])
    T, alpha, decay, m = 6, 0.3, 0.9, 2
    n0 = np.stack([O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in systems])
    f = xc.GlobalMLPXC([(w, b) for w, b in params], activation="tanh", norm=2.0)
    prob = engine.Problem1D(grids, [s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                            utils.exponential_coulomb, f, T, alpha=alpha, alpha_decay=decay,
                            num_mixing_iterations=m, loop="scan", want_orbitals=True,
                            enforce_reflection_symmetry=True)
    tE = np.array([-1.2, -2.5, -4.0])
    tn = 0.9 * n0
    _, g, res = train.value_and_grad(prob, n0, train.c3_loss, tE, tn, dx)
    g = g.cpu().numpy()
    oxc = O.GlobalMLP([w for w, _ in params], [b for _, b in params], "tanh", 2.0)
    go = np.zeros_like(g)
    B = len(systems)
    for s, (loc, z, ne) in enumerate(systems):
        fw = A.forward_trajectory(loc, z, ne, T, grids, oxc, O.exponential_coulomb, n0[s], alpha, decay, m,
                                  reflect=True)
        np.testing.assert_allclose(res.total_energy[s].cpu().numpy(), fw["energies"], atol=1e-8)
        dE = np.zeros(T)
        dn = np.zeros((T, len(grids)))
        dE[-1] = 2 * (fw["energies"][-1] - tE[s]) / B
        dn[-1] = 2 * 1e-2 * dx * (fw["densities"][-1] - tn[s]) / B
        go += A.backward(fw, dE, dn)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-6 * scale)
