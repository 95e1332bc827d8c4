"""Functional evaluation on the device (``qex1d_xc_evaluate``) against the FP64 oracle.

Covers every XC kind the SCF runs (closed forms, windowed local MLP on both network kernels,
global MLP), the functional's call ``xc_energy_density_fn(density)`` and the scf helpers
get_xc_energy / get_xc_potential (jax_dft scf) and their amplitude-encoded forms
(od/scf.py:32-76).  Bar: 1e-12 relative to the largest value (FP64 throughout, only the
summation order differs).  The potential is also pinned by central differences of the device
energy itself (v_i = dE/dn_i, divided by dx for local functionals).
"""

import numpy as np
import pytest

from oracle import od as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from qex_amd.od import engine, scf, xc  # noqa: E402

N = 101
GRIDS = np.linspace(-10, 10, N)
DX = float(GRIDS[1] - GRIDS[0])


def _densities(B=3):
    rng = np.random.default_rng(7)
    out = []
    for _ in range(B):
        c = rng.uniform(-3, 3, 2)
        s = rng.uniform(0.7, 1.6, 2)
        n = sum(np.exp(-((GRIDS - ci) / si) ** 2) for ci, si in zip(c, s)) + 1e-3
        out.append(n / (n.sum() * DX) * 2.0)
    return np.stack(out)


def _mlp(widths, seed):
    W, b = O.glorot_normal_mlp(seed, widths)
    rng = np.random.default_rng(seed + 100)
    return W, [rng.normal(0, 0.1, bb.shape) for bb in b]


def _cases():
    W1, b1 = _mlp([17, 16, 16, 1], 1)    # matrix-core kernel (hidden <= 64, <= 4 hidden layers)
    W2, b2 = _mlp([9, 96, 1], 2)         # generic kernel, 8 lanes per point
    W3, b3 = _mlp([5, 513, 1], 3)        # generic kernel, 64 lanes per point (notebook width)
    Wg, bg = _mlp([N, 32, 16, 1], 4)
    return {
        "none": (xc.NoXC(), None),
        "lda": (xc.LDAExp(), O.LDAExp()),
        "power_local": (xc.PowerLaw(-0.7, 0.4, 1e-3), O.PowerLocal(-0.7, 0.4, 1e-3)),
        "power_global": (xc.PowerLaw(-0.7, 0.4, 1e-3, encoding="global"), O.PowerGlobal(-0.7, 0.4, 1e-3)),
        "mlp_mfma": (xc.LocalMLPXC(list(zip(W1, b1)), window=17, activation="silu", norm=2.0),
                     O.LocalMLP(W1, b1, "silu", 17, 2.0)),
        "mlp_lpp8": (xc.LocalMLPXC(list(zip(W2, b2)), window=9, activation="tanh", norm=1.5, output="negativity"),
                     O.LocalMLP(W2, b2, "tanh", 9, 1.5, output="negativity")),
        "mlp_lpp64": (xc.LocalMLPXC(list(zip(W3, b3)), window=5, activation="softplus", norm=2.0),
                      O.LocalMLP(W3, b3, "softplus", 5, 2.0)),
        "global_mlp": (xc.GlobalMLPXC(list(zip(Wg, bg)), activation="tanh", norm=2.0),
                       O.GlobalMLP(Wg, bg, "tanh", 2.0)),
    }


def _close(got, ref, rtol=1e-12):
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(ref))) if ref.size else 1.0)
    assert float(np.max(np.abs(got - ref))) <= rtol * scale, float(np.max(np.abs(got - ref)))


@pytest.mark.parametrize("name", list(_cases()))
def test_xc_evaluate_matches_oracle(name):
    f, ref = _cases()[name]
    n = _densities()
    eps, exc, vxc = (t.cpu().numpy() for t in engine.xc_evaluate(f, n, DX))
    assert eps.shape == ((3, N) if f.encoding == "local" else (3,)) and exc.shape == (3,) and vxc.shape == (3, N)
    for b in range(3):
        if ref is None:
            assert not np.any(eps[b]) and exc[b] == 0.0 and not np.any(vxc[b])
            continue
        _close(eps[b], O.xc_energy_density_out(ref, n[b]))
        _close(exc[b], O.xc_energy(ref, n[b], DX))
        _close(vxc[b], O.xc_potential(ref, n[b], DX))


@pytest.mark.parametrize("name", ["lda", "mlp_mfma", "mlp_lpp8", "power_global", "global_mlp"])
def test_xc_potential_is_derivative_of_device_energy(name):
    f, _ = _cases()[name]
    n = _densities(1)[0]
    _, _, v = engine.xc_evaluate(f, n, DX)
    v = v.cpu().numpy()
    h = 1e-6  # truncation and rounding both below 2e-7 relative for these densities
    idx = [0, 13, 50, 77, N - 1]
    plus, minus = np.repeat(n[None], len(idx), 0), np.repeat(n[None], len(idx), 0)
    for r, i in enumerate(idx):
        plus[r, i] += h
        minus[r, i] -= h
    ep = engine.xc_evaluate(f, plus, DX)[1].cpu().numpy()
    em = engine.xc_evaluate(f, minus, DX)[1].cpu().numpy()
    fd = (ep - em) / (2 * h)
    if f.encoding == "local":
        fd = fd / DX
    np.testing.assert_allclose(fd, v[idx], rtol=1e-6, atol=1e-8)


def test_functional_call_and_scf_helpers_run_on_device():
    cases = _cases()
    n = _densities(1)[0]
    grids = torch.as_tensor(GRIDS)
    f, ref = cases["mlp_mfma"]
    e = f(n)
    assert isinstance(e, torch.Tensor) and e.is_cuda
    _close(e.cpu().numpy(), ref.energy_density(n))
    _close(float(scf.get_xc_energy(n, f, grids)), O.xc_energy(ref, n, DX))
    _close(scf.get_xc_potential(n, f, grids).cpu().numpy(), ref.potential(n))
    g, gref = cases["global_mlp"]
    _close(float(g(n)), gref.value(n))
    _close(float(scf.get_xc_energy_amplitude_encoded(n, g, grids)), O.xc_energy(gref, n, DX))
    _close(scf.get_xc_potential_amplitude_encoded(n, g, grids).cpu().numpy(), gref.grad(n) * DX)
    # a local PowerLaw in the amplitude-encoded helpers is its global form: dx sum_i f(n_i)
    p, pref = cases["power_local"]
    _close(float(scf.get_xc_energy_amplitude_encoded(n, p, grids)), O.PowerGlobal(-0.7, 0.4, 1e-3).value(n) * DX)
    # a local functional summed by the amplitude-encoded formula: dx sum_i e_i
    lda, lref = cases["lda"]
    _close(float(scf.get_xc_energy_amplitude_encoded(n, lda, grids)), lref.energy_density(n).sum() * DX)
    with pytest.raises(TypeError, match="global"):
        scf.get_xc_potential_amplitude_encoded(n, lda, grids)


def test_xc_evaluate_argument_checks():
    from qex_amd import _lib

    f, _ = _cases()["global_mlp"]
    with pytest.raises(ValueError, match="grid points"):
        engine.xc_evaluate(f, np.ones((2, N - 1)), DX)
    with pytest.raises(ValueError, match=r"\[N\] or \[B, N\]"):
        engine.xc_evaluate(xc.LDAExp(), np.ones((2, 2, N)), DX)
    with pytest.raises(_lib.QexError, match="dx"):
        engine.xc_evaluate(xc.LDAExp(), np.ones(N), 0.0)


# ---- opaque xc_energy_density_fn: the SCF one device pass at a time (qex1d_kohn_sham_step) ----
def _state_close(st, ref, dx, tol=1e-9):
    s = st.numpy()
    for t in range(ref.total_energy.shape[0]):
        assert abs(s.total_energy[t] - ref.total_energy[t]) < tol, (t, s.total_energy[t], ref.total_energy[t])
        assert np.sqrt(dx * np.sum((s.density[t] - ref.density[t]) ** 2)) < tol, t
        assert np.sqrt(dx * np.sum((s.xc_potential[t] - ref.xc_potential[t]) ** 2)) < tol, t
    np.testing.assert_allclose(s.xc_energy_density, ref.xc_energy_density, rtol=0, atol=tol)
    np.testing.assert_array_equal(s.converged.astype(bool), ref.converged.astype(bool))


@pytest.mark.parametrize("loop,reflect,tol", [("python", True, 1e-4), ("scan", True, -1.0), ("python", False, -1.0)])
def test_stepwise_passes_equal_one_call_bitwise(golden_dir, loop, reflect, tol):
    """T + 1 calls of qex1d_kohn_sham_step give qex1d_kohn_sham's outputs bit for bit (LDA and a
    LocalMLP, with reflection and with early convergence freezing states)."""
    import os

    from qex_amd.od import utils

    z = np.load(os.path.join(golden_dir, "od_h2_rows.npz"))
    grids = z["grids"]
    locs = [z["locations"][0], z["locations"][1]]
    Zs = [z["nuclear_charges"][0], z["nuclear_charges"][1]]
    W, b = _mlp([5, 16, 16, 1], 9)
    for f in (xc.LDAExp(), xc.LocalMLPXC(list(zip(W, b)), window=5, activation="silu", norm=2.0)):
        prob = engine.Problem1D(grids, locs, Zs, [2, 2], utils.exponential_coulomb, f, 8,
                                density_mse_converge_tolerance=tol, enforce_reflection_symmetry=reflect, loop=loop)
        n0 = engine.solve_noninteracting(prob.vext, prob.num_electrons, prob.grids).density
        a = prob.run(n0)
        a = {k: v.clone() for k, v in a.__dict__.items() if isinstance(v, torch.Tensor) and k != "stats"}
        bres = prob.run_stepwise(n0)
        for k, v in a.items():
            assert torch.equal(v, getattr(bres, k)), k


def test_opaque_local_callable_matches_oracle(golden_dir):
    """scf.kohn_sham with a plain torch function (no qex functional): python loop, reflection about
    off-centre nuclei, against the oracle's SCF with the same functional."""
    import os

    from qex_amd.od import utils

    z = np.load(os.path.join(golden_dir, "od_h2_rows.npz"))
    grids, loc, Z = z["grids"], z["locations"][1], z["nuclear_charges"][1]
    fn = lambda n: -0.7 * (n + 1e-3) ** 0.4  # noqa: E731
    ref = O.kohn_sham(loc, Z, 2, 8, grids, O.PowerLocal(-0.7, 0.4, 1e-3), O.exponential_coulomb, None,
                      enforce_reflection_symmetry=True)
    st = scf.kohn_sham(loc, Z, 2, 8, grids, fn, utils.exponential_coulomb, enforce_reflection_symmetry=True)
    _state_close(st, ref, O.get_dx(grids))
    # and against the built-in kind of the same functional
    bi = scf.kohn_sham(loc, Z, 2, 8, grids, xc.PowerLaw(-0.7, 0.4, 1e-3), utils.exponential_coulomb,
                       enforce_reflection_symmetry=True)
    np.testing.assert_allclose(st.numpy().total_energy, bi.numpy().total_energy, rtol=0, atol=1e-12)


def test_opaque_global_callable_matches_oracle():
    """scf.kohn_sham_amplitude_encoded with a plain torch function: E = dx sum fn(n),
    v = dE/dn by torch autograd between passes; against the oracle's PowerGlobal SCF."""
    from qex_amd.od import utils

    loc, Z = np.array([-0.8, 0.8]), np.array([1, 1])
    grids = np.linspace(-8, 8, 129)
    n0 = O.gaussian_initial_density(grids, loc, Z, 2)
    fn = lambda n: -0.7 * (n + 1e-2) ** 0.4  # noqa: E731
    ref = O.kohn_sham(loc, Z, 2, 4, grids, O.PowerGlobal(-0.7, 0.4, 1e-2), O.exponential_coulomb, n0)
    st = scf.kohn_sham_amplitude_encoded(loc, Z, 2, 4, grids, fn, utils.exponential_coulomb, initial_density=n0)
    _state_close(st, ref, O.get_dx(grids))


def test_opaque_callable_scan_loop_matches_builtin_lda():
    """jit_scf.kohn_sham (scan semantics) with the LDA written in torch and passed as an opaque
    callable agrees with the built-in LDA kind (c1 shape: 513 points, reflection, T = 15)."""
    from qex_amd.od import jit_scf, utils

    grids = O.h2_config_grid()
    loc, Z = np.array([-0.7, 0.7]), np.array([1, 1])
    n0 = O.gaussian_initial_density(grids, loc, Z, 2)
    lda = xc.LDAExp()
    fn = lambda n: lda.energy_density(n)  # noqa: E731
    kw = dict(alpha=0.3, alpha_decay=1.0, enforce_reflection_symmetry=True, num_mixing_iterations=1)
    st = jit_scf.kohn_sham(loc, Z, 2, 15, grids, fn, utils.exponential_coulomb, n0, **kw)
    bi = jit_scf.kohn_sham(loc, Z, 2, 15, grids, lda, utils.exponential_coulomb, n0, **kw)
    s, b = st.numpy(), bi.numpy()
    np.testing.assert_allclose(s.total_energy, b.total_energy, rtol=0, atol=1e-11)
    np.testing.assert_allclose(s.density, b.density, rtol=0, atol=1e-11)
    ref = O.jit_kohn_sham(loc, Z, 2, 15, grids, O.LDAExp(), O.exponential_coulomb, n0, **kw)
    np.testing.assert_allclose(s.total_energy, ref.total_energy, rtol=0, atol=1e-8)
