"""FP64 NumPy/SciPy restatement of the reference's 1D Kohn-Sham path.

TEST INFRASTRUCTURE ONLY (the checker): imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``; never by
the product package ``qex_amd``.

What it follows (reference = pasqal-io/qex, package ``qedft``; ``jax_dft`` is the
un-vendored dependency pinned at google-research rev c27ed001, ``pyproject.toml:21``):

* interaction / external potential: ``jax_dft.utils.exponential_coulomb``,
  ``soft_coulomb``, ``get_atomic_chain_potential`` (call sites ``od/scf.py:490-495``,
  ``od/jit_scf.py:219-224``);
* kinetic operator: ``od/scf.py:185-195`` (``get_kinetic_matrix_no_jit``) with the
  5-point ``jax_dft.scf.discrete_laplacian``;
* non-interacting solve + density: ``od/scf.py:198-268`` (dense ``eigh`` of
  T + diag(v), first num_electrons states, repeat-x2 occupation);
* Hartree potential / energy: ``od/scf.py:271-325``;
* local XC energy / potential: ``jax_dft.scf.get_xc_energy`` = dx * e(n) . n and
  ``get_xc_potential`` = grad / dx; global ("amplitude-encoded") XC: ``od/scf.py:32-76``
  (E = dx * sum f(n), v = dE/dn with no 1/dx);
* total energy: ``od/scf.py:137-159`` / ``od/jit_scf.py:46-72``;
* the python SCF loop: ``od/scf.py:434-560`` (and the identical jax_dft ``scf.kohn_sham``);
* the ``lax.scan`` SCF loop: ``od/jit_scf.py:143-244`` (and jax_dft ``jit_scf.kohn_sham``);
* 1D LDA for the exponential interaction: ``jax_dft.xc`` (used at ``od/eval.py:78-81``);
* the per-point MLP XC: ``models/classical/classical_models.py:123-174``.

Pins (see tests/test_oracle_od.py): the closed-form external potential against
``data/od/*/external_potentials.npy``; the kinetic stencil + eigh + density against
``qedft/models/initial_density.pkl``; the global-XC known answer of
``tests/test_scf_od.py:52-61`` and the closed-form potential of ``:64-86``; Hartree
potential/energy against erfc closed forms for Gaussian densities; eigenvalues against
the harmonic oscillator; the energy assembly against an independent T_s + E_ext + E_H +
E_xc evaluation; XC potentials against central finite differences of the energies;
convergence-flag patterns of ``tests/test_scf_od.py:184-215`` and
``tests/test_jit_scf_od.py:186-229``.

Parity unpinned (no fixture or closed form fixes them; only an import of jax_dft
could): the LDA *correlation* constants, and the off-centre branch of
``jax_dft.utils.flip_and_average``.
"""

from __future__ import annotations

import math
from typing import NamedTuple

import numpy as np
import scipy.linalg

# jax_dft.constants (verified against data/od/h2/external_potentials.npy to 2e-15).
EXPONENTIAL_COULOMB_AMPLITUDE = 1.071295
EXPONENTIAL_COULOMB_KAPPA = 1.0 / 2.385345


# ----------------------------------------------------------------------------
# jax_dft.utils
# ----------------------------------------------------------------------------
def exponential_coulomb(displacements, amplitude=EXPONENTIAL_COULOMB_AMPLITUDE,
                        kappa=EXPONENTIAL_COULOMB_KAPPA):
    """u(r) = A exp(-kappa |r|)  (jax_dft.utils.exponential_coulomb)."""
    return amplitude * np.exp(-np.abs(displacements) * kappa)


def soft_coulomb(displacements, soften_factor=1.0):
    """u(r) = 1 / sqrt(r^2 + 1)  (jax_dft.utils.soft_coulomb)."""
    return 1.0 / np.sqrt(np.asarray(displacements) ** 2 + soften_factor)


def get_dx(grids):
    grids = np.asarray(grids)
    return float((grids[-1] - grids[0]) / (len(grids) - 1))


def gaussian(grids, center, sigma=1.0):
    return np.exp(-0.5 * ((np.asarray(grids) - center) / sigma) ** 2) / (math.sqrt(2 * math.pi) * sigma)


def get_atomic_chain_potential(grids, locations, nuclear_charges, interaction_fn):
    """v_ext(x) = -sum_a Z_a u(x - R_a)."""
    grids = np.asarray(grids, dtype=np.float64)
    locations = np.asarray(locations, dtype=np.float64)
    nuclear_charges = np.asarray(nuclear_charges, dtype=np.float64)
    return -np.sum(nuclear_charges[:, None] * interaction_fn(grids[None, :] - locations[:, None]), axis=0)


def flip_and_average_on_center(array):
    """(a + a[::-1]) / 2 -- jax_dft.jit_scf._flip_and_average_on_center (od/jit_scf.py:76-77)."""
    return (array + array[::-1]) / 2


def reflection_window(locations, grids):
    """Index window [lo, hi] symmetric about the centre of ``locations``.

    Restates jax_dft.utils.flip_and_average (od/scf.py:162-167): the centre is the mean
    location; if it is a grid point the window is centred on it, otherwise on the two
    grid points around it; the radius is the largest that stays on the grid.
    """
    grids = np.asarray(grids, dtype=np.float64)
    locations = np.asarray(locations, dtype=np.float64)
    center = float(np.mean(locations))
    on_grid = np.isclose(grids, center, rtol=0.0, atol=1e-9)
    if np.any(on_grid):
        left = right = int(np.argmax(on_grid))
    else:
        dist = np.abs(grids - center)
        left = int(np.argmin(np.where(grids < center, dist, np.inf)))
        right = int(np.argmin(np.where(grids > center, dist, np.inf)))
    radius = min(left, len(grids) - right - 1)
    return left - radius, right + radius


def flip_and_average(locations, grids, array):
    lo, hi = reflection_window(locations, grids)
    out = np.array(array, dtype=np.float64, copy=True)
    seg = out[lo:hi + 1].copy()
    out[lo:hi + 1] = (seg + seg[::-1]) / 2
    return out


# ----------------------------------------------------------------------------
# kinetic operator and the non-interacting solve (od/scf.py:185-268)
# ----------------------------------------------------------------------------
def discrete_laplacian(num_grids):
    """5-point second difference (-1/12, 4/3, -5/2, 4/3, -1/12), Dirichlet truncation."""
    n = num_grids
    return (np.diag(-2.5 * np.ones(n)) + np.diag(4.0 / 3 * np.ones(n - 1), 1) + np.diag(4.0 / 3 * np.ones(n - 1), -1)
            + np.diag(-1.0 / 12 * np.ones(n - 2), 2) + np.diag(-1.0 / 12 * np.ones(n - 2), -2))


def get_kinetic_matrix(grids):
    dx = get_dx(grids)
    return -0.5 * discrete_laplacian(len(grids)) / (dx * dx)


def wavefunctions_to_density(num_electrons, wavefunctions, grids):
    """od/scf.py:198-213: normalise with dx, occupy each state twice, sum the first N_e."""
    wavefunctions = wavefunctions[:num_electrons]
    wavefunctions = wavefunctions / np.sqrt(np.sum(wavefunctions ** 2, axis=1, keepdims=True) * get_dx(grids))
    intensities = np.repeat(wavefunctions ** 2, 2, axis=0)
    return np.sum(intensities[:num_electrons], axis=0)


def get_total_eigen_energies(num_electrons, eigen_energies):
    return float(np.sum(np.repeat(eigen_energies, 2)[:num_electrons]))


def get_gap(num_electrons, eigen_energies):
    rep = np.repeat(eigen_energies, 2)
    return float(rep[num_electrons] - rep[num_electrons - 1])


def solve_noninteracting_system(external_potential, num_electrons, grids, return_eigs=False):
    """od/scf.py:233-250: dense FP64 eigh of T + diag(v) (LAPACK dsyevd)."""
    h = get_kinetic_matrix(grids) + np.diag(external_potential)
    eigen_energies, wavefunctions_transpose = scipy.linalg.eigh(h, driver="evd")
    density = wavefunctions_to_density(num_electrons, wavefunctions_transpose.T, grids)
    out = (density, get_total_eigen_energies(num_electrons, eigen_energies), get_gap(num_electrons, eigen_energies))
    if return_eigs:
        return out + (eigen_energies, wavefunctions_transpose.T)
    return out


# ----------------------------------------------------------------------------
# Hartree and external terms (od/scf.py:271-325; jax_dft scf.get_external_potential_energy)
# ----------------------------------------------------------------------------
def get_hartree_energy(density, grids, interaction_fn):
    n1 = density[None, :]
    n2 = density[:, None]
    r1 = np.asarray(grids)[None, :]
    r2 = np.asarray(grids)[:, None]
    return float(0.5 * np.sum(n1 * n2 * interaction_fn(r1 - r2)) * get_dx(grids) ** 2)


def get_hartree_potential(density, grids, interaction_fn):
    r1 = np.asarray(grids)[None, :]
    r2 = np.asarray(grids)[:, None]
    return np.sum(density[None, :] * interaction_fn(r1 - r2), axis=1) * get_dx(grids)


def get_external_potential_energy(external_potential, density, grids):
    return float(np.dot(density, external_potential) * get_dx(grids))


# ----------------------------------------------------------------------------
# XC functionals
# ----------------------------------------------------------------------------
class LocalXC:
    """Local XC: per-point energy density e(n) (jax_dft scf.get_xc_energy / get_xc_potential).

    E_xc = dx * e(n) . n ;  v_xc = (1/dx) dE_xc/dn = e + J_e(n)^T n.
    """

    kind = "local"

    def energy_density(self, n):  # (N,) -> (N,)
        raise NotImplementedError

    def potential(self, n):  # (N,) -> (N,)
        raise NotImplementedError


class GlobalXC:
    """Global ("amplitude-encoded") XC: f(n) summed to a scalar (od/scf.py:32-76).

    E_xc = dx * sum(f(n)) ;  v_xc = dE_xc/dn = dx * grad f  (no 1/dx).
    """

    kind = "global"

    def value(self, n):  # scalar sum(f(n))
        raise NotImplementedError

    def grad(self, n):  # d value / dn
        raise NotImplementedError


def lda_exchange(density, amplitude=EXPONENTIAL_COULOMB_AMPLITUDE, kappa=EXPONENTIAL_COULOMB_KAPPA, epsilon=1e-15):
    """Exchange energy per electron of the 1D uniform gas with u = A exp(-kappa|r|).

    e_x = A/(2 pi) [ln(1+y^2)/y - 2 atan(y)],  y = pi n / kappa; the series A/(2 pi)(-y + y^3/6)
    below epsilon (jax_dft.xc.exponential_coulomb_uniform_exchange_density).  ln(1+y^2) is
    evaluated as written there (not log1p), so tiny y round the same way.
    Returns (e_x, de_x/dn).
    """
    n = np.asarray(density, dtype=np.float64)
    y = np.pi * n / kappa
    c = amplitude / (2 * np.pi)
    big = y > epsilon
    ys = np.where(big, y, 1.0)
    ex = np.where(big, c * (np.log(1 + ys ** 2) / ys - 2 * np.arctan(ys)), c * (-y + y ** 3 / 6))
    # d/dy [ln(1+y^2)/y - 2 atan y] = 2/(1+y^2) - ln(1+y^2)/y^2 - 2/(1+y^2) = -ln(1+y^2)/y^2
    dex_dy = np.where(big, c * (-np.log(1 + ys ** 2) / ys ** 2), c * (-1 + y ** 2 / 2))
    return ex, dex_dy * np.pi / kappa


LDA_CORRELATION_PARAMS = dict(alpha=2.0, beta=-1.00077, gamma=6.26099, delta=-11.9041, eta=9.62614,
                              sigma=-1.48334, nu=1.0)


def lda_correlation(density, amplitude=EXPONENTIAL_COULOMB_AMPLITUDE, kappa=EXPONENTIAL_COULOMB_KAPPA):
    """Correlation energy per electron (Baker et al., PRB 91, 235141 (2015), Eq. 24 as used by jax_dft.xc).

    e_c = -A y / (pi P(y)),  P = alpha + beta y^1/2 + gamma y + delta y^3/2 + eta y^2 + sigma y^5/2
    + nu pi kappa^2 / A y^3 ; at y = 0: -A y / (pi alpha).  PARITY UNPINNED (recalled constants).
    Returns (e_c, de_c/dn).
    """
    p = LDA_CORRELATION_PARAMS
    n = np.asarray(density, dtype=np.float64)
    y = np.pi * n / kappa
    zero = y == 0.0
    fy = np.where(zero, 1.0, y)
    s = np.sqrt(fy)
    c3 = p["nu"] * np.pi * kappa ** 2 / amplitude
    P = (p["alpha"] + p["beta"] * s + p["gamma"] * fy + p["delta"] * fy * s + p["eta"] * fy ** 2
         + p["sigma"] * fy ** 2 * s + c3 * fy ** 3)
    dP = (0.5 * p["beta"] / s + p["gamma"] + 1.5 * p["delta"] * s + 2 * p["eta"] * fy
          + 2.5 * p["sigma"] * fy * s + 3 * c3 * fy ** 2)
    ec = np.where(zero, -amplitude * y / np.pi / p["alpha"], -amplitude * fy / np.pi / P)
    dec_dy = np.where(zero, -amplitude / np.pi / p["alpha"], -amplitude / np.pi * (P - fy * dP) / P ** 2)
    return ec, dec_dy * np.pi / kappa


class LDAExp(LocalXC):
    """jax_dft.xc.get_lda_xc_energy_density_fn(): exchange + correlation per electron."""

    def energy_density(self, n):
        return lda_exchange(n)[0] + lda_correlation(n)[0]

    def potential(self, n):
        ex, dex = lda_exchange(n)
        ec, dec = lda_correlation(n)
        return ex + ec + n * (dex + dec)


class PowerLocal(LocalXC):
    """e(n) = c (n + eps)^p (the regularised test functional of tests/test_scf_od.py:127-129)."""

    def __init__(self, c=-0.73855, p=1.0 / 3, eps=0.0):
        self.c, self.p, self.eps = c, p, eps

    def energy_density(self, n):
        return self.c * (n + self.eps) ** self.p

    def potential(self, n):
        return self.energy_density(n) + n * self.c * self.p * (n + self.eps) ** (self.p - 1)


class PowerGlobal(GlobalXC):
    """f(n) = sum c n^p (tests/test_scf_od.py:46-48, tests/test_jit_scf_od.py:206-209)."""

    def __init__(self, c=-0.73855, p=1.0 / 3, eps=0.0):
        self.c, self.p, self.eps = c, p, eps

    def value(self, n):
        return float(np.sum(self.c * (n + self.eps) ** self.p))

    def grad(self, n):
        return self.c * self.p * (n + self.eps) ** (self.p - 1)


# ---- MLPs ------------------------------------------------------------------
def _act(name, z):
    if name == "silu":
        s = 1 / (1 + np.exp(-z))
        return z * s, s * (1 + z * (1 - s))
    if name == "tanh":
        t = np.tanh(z)
        return t, 1 - t * t
    if name == "softplus":
        return np.logaddexp(0, z), 1 / (1 + np.exp(-z))
    if name == "sigmoid":
        s = 1 / (1 + np.exp(-z))
        return s, s * (1 - s)
    if name == "relu":
        return np.maximum(z, 0), (z > 0).astype(z.dtype)
    if name == "gelu":  # tanh approximation (jax.nn.gelu default)
        k = math.sqrt(2 / math.pi)
        u = k * (z + 0.044715 * z ** 3)
        t = np.tanh(u)
        return 0.5 * z * (1 + t), 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * k * (1 + 3 * 0.044715 * z * z)
    raise ValueError(f"unknown activation {name}")


def output_transform(kind, y):
    """(g(y), g'(y)) of the network-output transform: "linear", or "negativity" = -swish(y), the
    jax_dft neural_xc.negativity_transform that qedft/models/wrappers.py:201-203 appends
    (range (-inf, 0.278])."""
    if kind == "linear":
        return y, np.ones_like(y)
    if kind == "negativity":
        s = 1 / (1 + np.exp(-y))
        return -y * s, -(s * (1 + y * (1 - s)))
    raise ValueError(f"unknown output transform {kind!r}")


def _mlp_forward(weights, biases, activation, x):
    """x (P, D0) -> out (P, 1); returns activations for the backward pass."""
    cache = []
    h = x
    for li, (w, b) in enumerate(zip(weights, biases)):
        z = h @ w + b
        if li < len(weights) - 1:
            a, da = _act(activation, z)
            cache.append((h, da))
            h = a
        else:
            cache.append((h, None))
            h = z
    return h, cache


def _mlp_input_grad_full(weights, cache, gout):
    g = gout
    for li in range(len(weights) - 1, -1, -1):
        g = g @ weights[li].T
        if li > 0:
            g = g * cache[li - 1][1]
    return g


class LocalMLP(LocalXC):
    """Per-point MLP on a (2h+1)-point density window.

    e_i = MLP([n_{i-h}, ..., n_{i+h}] / norm) with zero padding off the grid; with window=1
    it is ``build_local_mlp`` (classical_models.py:123-174): Dense+act repeated, final
    Dense(1), inputs divided by density_normalization_factor, vmapped over points.
    ``weights[l]`` has shape (fan_in, fan_out) (stax Dense convention: x @ W + b).
    """

    def __init__(self, weights, biases, activation="silu", window=17, norm=2.0, output="linear"):
        self.weights = [np.asarray(w, dtype=np.float64) for w in weights]
        self.biases = [np.asarray(b, dtype=np.float64) for b in biases]
        self.activation = activation
        self.window = window
        self.norm = norm
        self.output = output

    def _windows(self, n):
        h = self.window // 2
        pad = np.concatenate([np.zeros(h), n, np.zeros(h)])
        idx = np.arange(len(n))[:, None] + np.arange(self.window)[None, :]
        return pad[idx] / self.norm

    def energy_density(self, n):
        out, _ = _mlp_forward(self.weights, self.biases, self.activation, self._windows(n))
        return output_transform(self.output, out[:, 0])[0]

    def potential(self, n):
        N = len(n)
        out, cache = _mlp_forward(self.weights, self.biases, self.activation, self._windows(n))
        e, de = output_transform(self.output, out[:, 0])
        gx = _mlp_input_grad_full(self.weights, cache, (n * de)[:, None])  # (N, W): d(e.n)/d window input
        h = self.window // 2
        v = e.copy()
        for o in range(self.window):
            j = np.arange(N) + o - h  # point fed by window slot o of point i
            ok = (j >= 0) & (j < N)
            np.add.at(v, j[ok], gx[ok, o] / self.norm)
        return v


class GlobalMLP(GlobalXC):
    """``build_global_mlp`` (classical_models.py:177-224): whole density / norm -> scalar."""

    def __init__(self, weights, biases, activation="tanh", norm=2.0, output="linear"):
        self.weights = [np.asarray(w, dtype=np.float64) for w in weights]
        self.biases = [np.asarray(b, dtype=np.float64) for b in biases]
        self.activation = activation
        self.norm = norm
        self.output = output

    def value(self, n):
        out, _ = _mlp_forward(self.weights, self.biases, self.activation, (n / self.norm)[None, :])
        return float(output_transform(self.output, out[0, 0])[0])

    def grad(self, n):
        out, cache = _mlp_forward(self.weights, self.biases, self.activation, (n / self.norm)[None, :])
        _, dg = output_transform(self.output, out[0, 0])
        return _mlp_input_grad_full(self.weights, cache, np.full((1, 1), dg))[0] / self.norm


class SymmetrizedLocalXC(LocalXC):
    """e_sym(n) = F e(F n) for the reflection-symmetric loop, F = flip-and-average.

    (jax_dft scf/jit_scf wrap the XC function this way when enforce_reflection_symmetry is on.)
    F is symmetric, so v_sym(n) = F v(F n).
    """

    def __init__(self, base, flip):
        self.base, self.flip = base, flip

    def energy_density(self, n):
        return self.flip(self.base.energy_density(self.flip(n)))

    def potential(self, n):
        return self.flip(self.base.potential(self.flip(n)))


def xc_energy(xc, n, dx):
    if xc.kind == "local":
        return float(np.dot(xc.energy_density(n), n) * dx)
    return float(xc.value(n) * dx)


def xc_potential(xc, n, dx):
    if xc.kind == "local":
        return xc.potential(n)
    return xc.grad(n) * dx


def xc_energy_density_out(xc, n):
    if xc.kind == "local":
        return xc.energy_density(n)
    return np.float64(xc.value(n))


# ----------------------------------------------------------------------------
# SCF
# ----------------------------------------------------------------------------
class KohnShamState(NamedTuple):
    density: np.ndarray
    total_energy: np.ndarray
    locations: np.ndarray
    nuclear_charges: np.ndarray
    external_potential: np.ndarray
    grids: np.ndarray
    num_electrons: np.ndarray
    hartree_potential: np.ndarray | None = None
    xc_potential: np.ndarray | None = None
    xc_energy_density: np.ndarray | None = None
    gap: np.ndarray | None = None
    converged: np.ndarray | None = False


def kohn_sham_iteration(density, external_potential, grids, num_electrons, xc, interaction_fn, flip):
    """One KS iteration (od/scf.py:79-177, od/jit_scf.py:14-86).

    ``flip`` is None or the flip-and-average map applied to the output density (and, for
    local XC, used to symmetrise the functional by the caller).
    Returns density_out (flipped if flip), total_energy, v_H, v_xc, e_xc, gap, eigen_energies.
    """
    dx = get_dx(grids)
    hartree_potential = get_hartree_potential(density, grids, interaction_fn)
    v_xc = xc_potential(xc, density, dx)
    ks_potential = hartree_potential + v_xc + external_potential
    e_xc = xc_energy_density_out(xc, density)
    n_out, e_eig, gap, eigs, _ = solve_noninteracting_system(ks_potential, num_electrons, grids, return_eigs=True)
    total_energy = (e_eig - get_external_potential_energy(ks_potential, n_out, grids)
                    + get_hartree_energy(n_out, grids, interaction_fn) + xc_energy(xc, n_out, dx)
                    + get_external_potential_energy(external_potential, n_out, grids))
    if flip is not None:
        n_out = flip(n_out)
    return n_out, total_energy, hartree_potential, v_xc, e_xc, gap


def _prepare(locations, nuclear_charges, grids, xc, interaction_fn, enforce_reflection_symmetry, center_flip):
    grids = np.asarray(grids, dtype=np.float64)
    locations = np.asarray(locations, dtype=np.float64)
    nuclear_charges = np.asarray(nuclear_charges)
    v_ext = get_atomic_chain_potential(grids, locations, nuclear_charges, interaction_fn)
    flip = None
    if enforce_reflection_symmetry:
        flip = flip_and_average_on_center if center_flip else (lambda a: flip_and_average(locations, grids, a))
        if xc.kind == "local":
            xc = SymmetrizedLocalXC(xc, flip)
    return grids, locations, nuclear_charges, v_ext, flip, xc


def kohn_sham(locations, nuclear_charges, num_electrons, num_iterations, grids, xc, interaction_fn,
              initial_density=None, alpha=0.5, alpha_decay=0.9, enforce_reflection_symmetry=False,
              num_mixing_iterations=2, density_mse_converge_tolerance=-1.0):
    """Python-loop SCF (od/scf.py:434-560; jax_dft scf.kohn_sham for local XC).

    Convergence: MSE(n_out - n_in) < tol on the same iteration; the state keeps the
    *mixed* density and the energy of the unmixed output; once converged, later
    iterations repeat that state.  Reflection is about the centre of the locations.
    """
    grids, locations, nuclear_charges, v_ext, flip, xc = _prepare(
        locations, nuclear_charges, grids, xc, interaction_fn, enforce_reflection_symmetry, center_flip=False)
    if initial_density is None:
        initial_density = solve_noninteracting_system(v_ext, num_electrons, grids)[0]
    density = np.asarray(initial_density, dtype=np.float64)
    rows = []
    differences = []
    converged = False
    state = None
    for _ in range(num_iterations):
        if converged:
            rows.append(state)
            continue
        n_out, e, vh, vxc, exc, gap = kohn_sham_iteration(density, v_ext, grids, num_electrons, xc, interaction_fn, flip)
        differences.append(n_out - density)
        if np.mean(np.square(n_out - density)) < density_mse_converge_tolerance:
            converged = True
        mixed = density + alpha * np.mean(np.array(differences[-num_mixing_iterations:]), axis=0)
        state = dict(density=mixed, total_energy=e, hartree_potential=vh, xc_potential=vxc,
                     xc_energy_density=exc, gap=gap, converged=converged)
        rows.append(state)
        density = mixed
        alpha *= alpha_decay
    return _stack(rows, locations, nuclear_charges, v_ext, grids, num_electrons)


def connection_weights(num_iterations, num_mixing_iterations):
    """jax_dft.jit_scf._connection_weights: row-normalised band of the last m differences."""
    mask = np.triu(np.tril(np.ones((num_iterations, num_iterations))), k=-num_mixing_iterations + 1)
    return mask / np.sum(mask, axis=1, keepdims=True)


def jit_kohn_sham(locations, nuclear_charges, num_electrons, num_iterations, grids, xc, interaction_fn,
                  initial_density, alpha=0.5, alpha_decay=0.9, enforce_reflection_symmetry=False,
                  num_mixing_iterations=2, density_mse_converge_tolerance=-1.0, stop_gradient_step=-1):
    """lax.scan SCF (od/jit_scf.py:143-244; jax_dft jit_scf.kohn_sham for local XC).

    Convergence: MSE(n_mixed - n_in) < tol, visible from the next iteration, which then
    passes the old state through with converged=True.  Reflection about the grid centre.
    ``stop_gradient_step`` only affects gradients and is accepted for signature parity.
    """
    del stop_gradient_step
    grids, locations, nuclear_charges, v_ext, flip, xc = _prepare(
        locations, nuclear_charges, grids, xc, interaction_fn, enforce_reflection_symmetry, center_flip=True)
    N = len(grids)
    weights = connection_weights(num_iterations, num_mixing_iterations)
    differences = np.zeros((num_iterations, N))
    state = dict(density=np.asarray(initial_density, dtype=np.float64), total_energy=np.inf,
                 hartree_potential=np.zeros(N), xc_potential=np.zeros(N),
                 xc_energy_density=np.zeros(N) if xc.kind == "local" else np.float64(0.0), gap=0.0, converged=False)
    converged = False
    rows = []
    for idx in range(num_iterations):
        old = state
        if converged:
            state = dict(old, converged=True)
        else:
            n_out, e, vh, vxc, exc, gap = kohn_sham_iteration(old["density"], v_ext, grids, num_electrons, xc,
                                                              interaction_fn, flip)
            differences[idx] = n_out - old["density"]
            mixed = old["density"] + alpha * weights[idx] @ differences
            state = dict(density=mixed, total_energy=e, hartree_potential=vh, xc_potential=vxc,
                         xc_energy_density=exc, gap=gap, converged=old["converged"])
        converged = bool(np.mean(np.square(state["density"] - old["density"])) < density_mse_converge_tolerance)
        rows.append(state)
        alpha *= alpha_decay
    return _stack(rows, locations, nuclear_charges, v_ext, grids, num_electrons)


def _stack(rows, locations, nuclear_charges, v_ext, grids, num_electrons):
    T = len(rows)
    return KohnShamState(
        density=np.stack([r["density"] for r in rows]),
        total_energy=np.array([r["total_energy"] for r in rows], dtype=np.float64),
        locations=np.stack([locations] * T),
        nuclear_charges=np.stack([nuclear_charges] * T),
        external_potential=np.stack([v_ext] * T),
        grids=np.stack([grids] * T),
        num_electrons=np.array([num_electrons] * T),
        hartree_potential=np.stack([r["hartree_potential"] for r in rows]),
        xc_potential=np.stack([r["xc_potential"] for r in rows]),
        xc_energy_density=np.stack([np.asarray(r["xc_energy_density"]) for r in rows]),
        gap=np.array([r["gap"] for r in rows], dtype=np.float64),
        converged=np.array([bool(r["converged"]) for r in rows]),
    )


# ----------------------------------------------------------------------------
# Synthetic workloads
# ----------------------------------------------------------------------------
def h2_config_grid():
    """513 points on [-25.6, 25.6] bohr (dx = 0.1)."""
    return np.linspace(-25.6, 25.6, 513)


def gaussian_initial_density(grids, locations, nuclear_charges, num_electrons, weighted=False):
    """Sum of unit-norm Gaussians (sigma = 1) on the nuclei; weighted by Z and renormalised to N_e
    when ``weighted``."""
    grids = np.asarray(grids)
    if weighted:
        n = sum(float(z) * gaussian(grids, float(r), 1.0) for r, z in zip(locations, nuclear_charges))
        return n * (num_electrons / (np.sum(n) * get_dx(grids)))
    return sum(gaussian(grids, float(r), 1.0) for r in locations)


def random_molecules(seed, count, max_nuclei=4, max_electrons=8, charge_choices=(1, 2, 3, 4),
                     box=4.0, min_sep=0.4, grids=None):
    """Seeded uneven 1D systems: 1-4 nuclei, Z in {1..4} with sum(Z) <= 8, positions U[-4, 4] with
    >= 0.4 bohr separation, N_e = sum(Z).  Positions are snapped to the grid when ``grids`` is given
    (so the reflection window of the python loop is well defined)."""
    rng = np.random.default_rng(seed)
    mols = []
    while len(mols) < count:
        a = int(rng.integers(1, max_nuclei + 1))
        z = rng.choice(np.asarray(charge_choices), size=a)
        if int(z.sum()) > max_electrons:
            continue
        r = np.sort(rng.uniform(-box, box, size=a))
        if a > 1 and np.min(np.diff(r)) < min_sep:
            continue
        if grids is not None:
            dx = get_dx(grids)
            r = np.round((r - grids[0]) / dx) * dx + grids[0]
            if a > 1 and np.min(np.diff(r)) < min_sep - 1e-9:
                continue
        mols.append((r.astype(np.float64), z.astype(np.int64), int(z.sum())))
    return mols


def glorot_normal_mlp(seed, sizes):
    """Glorot-normal weights (std = sqrt(2/(fan_in+fan_out))), zero biases; weights[l] is (in, out)."""
    rng = np.random.default_rng(seed)
    weights, biases = [], []
    for fi, fo in zip(sizes[:-1], sizes[1:]):
        weights.append(rng.normal(0.0, math.sqrt(2.0 / (fi + fo)), size=(fi, fo)))
        biases.append(np.zeros(fo))
    return weights, biases
