"""FP64 NumPy/SciPy restatement of the reference's 3D RKS cycle (test infrastructure only).

Follows the pure-function template of ``qedft/train/td/scf_functions_masked.py:143-303, 846-967``
and the pyscf-based loop ``qedft/train/td/hf_legacy.py:51-116`` (get_fock with CDIIS
``:379-427``, ``jax_diis.py`` error FDS - SDF with the -1 border), ``rks_legacy.py:36-127``
(Veff = J + Vxc), ``numint_legacy.py:122-482`` (eval_rho / nr_rks), ``hf.py:303-308``
(make_rdm1) and ``generalized_eigensolver.py:264-330`` (eigh(F, S)).

The compiled pieces the reference takes from pyscf/libcint/libxc are restated from their
published mathematics:
* Cartesian s/p Gaussian integrals (S, T, V, ERI) by McMurchie-Davidson with the Boys function
  from the regularised incomplete gamma function (scipy);
* molecular grid: Treutler-Ahlrichs M4 radial (Chebyshev 2nd kind) x Gauss-Legendre(cos theta) x
  uniform-phi angular, Becke partition (3 smoothing iterations);
* LDA: Slater exchange + VWN5 correlation (spin-unpolarised).

Pins (tests/test_oracle_td.py): Szabo & Ostlund's H2/STO-3G integrals and RHF energy, the
reference notebook's PySCF H2/6-31G RHF energies (notebooks/03_notebook_td_dataset.ipynb:40),
finite differences of the XC potential, grid integrals of known functions, and scipy's
eigh(F, S).  Parity unpinned: pyscf's level-0 grid, libxc's "lda" alias and the minao initial
guess (a core-Hamiltonian guess is used).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg
from scipy.special import gamma as gamma_fn
from scipy.special import gammainc

BOHR = 0.52917721092  # Angstrom per bohr (pyscf param.BOHR)

# ---------------------------------------------------------------------------
# basis sets (public data)
# ---------------------------------------------------------------------------
STO3G_H = [(0, [3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])]


def sto3g_shells(zeta_1s, zeta_2sp=None):
    """STO-3G shells of a first-row atom (Hehre, Stewart, Pople 1969): the exponent-1 fits of the 1s
    and 2sp Slater functions, exponents scaled by zeta^2 (2s and 2p share the 2sp exponents)."""
    e1 = np.array([2.227660584, 0.4057711562, 0.1098175104])
    shells = [(0, zeta_1s ** 2 * e1, [0.1543289673, 0.5353281423, 0.4446345422])]
    if zeta_2sp is not None:
        e2 = zeta_2sp ** 2 * np.array([0.9942027, 0.2310313, 0.0751386])
        shells += [(0, e2, [-0.09996722919, 0.3995128261, 0.7001154689]),
                   (1, e2, [0.1559162750, 0.6076837186, 0.3919573931])]
    return shells


def water_molecule(shells_o, shells_h, r_oh=0.9572, angle=104.52):
    """H2O at the experimental geometry (Angstrom, degrees), O at the origin."""
    a = math.radians(angle)
    r = r_oh / BOHR
    coords = np.array([[0, 0, 0], [r * math.sin(a / 2), 0, r * math.cos(a / 2)],
                       [-r * math.sin(a / 2), 0, r * math.cos(a / 2)]])
    return build_molecule(coords, [8, 1, 1], [shells_o, shells_h, shells_h])
B631G_H = [(0, [18.7311370, 2.8253937, 0.6401217], [0.03349460, 0.23472695, 0.81375733]),
           (0, [0.1612778], [1.0])]


@dataclass
class Shell:
    center: int
    l: int
    exps: np.ndarray
    coefs: np.ndarray  # multiply normalised primitives; the contraction is normalised


@dataclass
class Molecule:
    coords: np.ndarray          # (A, 3) bohr
    charges: np.ndarray         # (A,)
    shells: list = field(default_factory=list)

    @property
    def nelectron(self):
        return int(round(float(np.sum(self.charges))))

    def ao_list(self):
        """[(shell index, (lx, ly, lz))] in shell order; p shells as px, py, pz."""
        aos = []
        for si, sh in enumerate(self.shells):
            if sh.l == 0:
                aos.append((si, (0, 0, 0)))
            else:
                aos += [(si, (1, 0, 0)), (si, (0, 1, 0)), (si, (0, 0, 1))]
        return aos

    @property
    def nao(self):
        return len(self.ao_list())

    def energy_nuc(self):
        e = 0.0
        for a in range(len(self.charges)):
            for b in range(a):
                e += self.charges[a] * self.charges[b] / np.linalg.norm(self.coords[a] - self.coords[b])
        return e


def prim_norm(alpha, l):
    """Normalisation of x^l exp(-alpha r^2) (Cartesian, l <= 1)."""
    return (2 * alpha / math.pi) ** 0.75 * (2 * math.sqrt(alpha)) ** l


def build_molecule(coords_bohr, charges, shells_per_atom):
    """shells_per_atom: list (per atom) of [(l, exps, coefs)] with coefficients of normalised
    primitives; contractions are renormalised to unit self-overlap."""
    mol = Molecule(np.asarray(coords_bohr, dtype=np.float64), np.asarray(charges, dtype=np.float64))
    for a, shells in enumerate(shells_per_atom):
        for l, exps, coefs in shells:
            exps = np.asarray(exps, dtype=np.float64)
            coefs = np.asarray(coefs, dtype=np.float64)
            # self overlap of sum_i c_i N_i g_i (same center, same l)
            ss = 0.0
            for i in range(len(exps)):
                for j in range(len(exps)):
                    p = exps[i] + exps[j]
                    sij = (math.pi / p) ** 1.5 * (1.0 / (2 * p)) ** l
                    ss += coefs[i] * coefs[j] * prim_norm(exps[i], l) * prim_norm(exps[j], l) * sij
            mol.shells.append(Shell(a, l, exps, coefs / math.sqrt(ss)))
    return mol


def h2_molecule(r_angstrom, basis=B631G_H):
    coords = np.array([[0, 0, 0], [0, 0, r_angstrom / BOHR]])
    return build_molecule(coords, [1, 1], [basis, basis])


def synthetic_chain(n_atoms, seed):
    """Hydrogen chain along z, bonds U[0.5, 1.5] Angstrom; per atom 2 contracted s shells and 1 p
    shell of 3 primitives each, exponents log-uniform in [0.1, 20] (sorted, descending) and
    normalised random-positive coefficients, all from the seed (5 AOs per atom)."""
    rng = np.random.default_rng(seed)
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n_atoms - 1))]) / BOHR
    coords = np.stack([np.zeros(n_atoms), np.zeros(n_atoms), z - z.mean()], axis=1)
    shells = []
    for _ in range(n_atoms):
        per = []
        for l in (0, 0, 1):
            exps = np.sort(np.exp(rng.uniform(np.log(0.1), np.log(20.0), 3)))[::-1].copy()
            coefs = rng.uniform(0.2, 1.0, 3)
            per.append((l, exps, coefs))
        shells.append(per)
    return build_molecule(coords, np.ones(n_atoms), shells)


# ---------------------------------------------------------------------------
# McMurchie-Davidson integrals
# ---------------------------------------------------------------------------
def boys(m, T):
    """F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt (vectorised)."""
    T = np.asarray(T, dtype=np.float64)
    out = np.empty_like(T)
    small = T < 1e-10
    out[small] = 1.0 / (2 * m + 1) - T[small] / (2 * m + 3)
    Ts = T[~small]
    out[~small] = gamma_fn(m + 0.5) * gammainc(m + 0.5, Ts) / (2 * Ts ** (m + 0.5))
    return out


def hermite_E(i, j, t, Qx, a, b):
    """E_t^{ij} of the 1D Hermite expansion (recursive; arrays broadcast)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return np.zeros_like(np.asarray(Qx, dtype=np.float64) * a)
    if i == j == t == 0:
        return np.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p)) * hermite_E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * hermite_E(i - 1, j, t, Qx, a, b) \
            + (t + 1) * hermite_E(i - 1, j, t + 1, Qx, a, b)
    return (1 / (2 * p)) * hermite_E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * hermite_E(i, j - 1, t, Qx, a, b) \
        + (t + 1) * hermite_E(i, j - 1, t + 1, Qx, a, b)


def hermite_R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    """R^n_{tuv} Hermite Coulomb integrals (recursive)."""
    T = p * RPC * RPC
    if t == u == v == 0:
        return (-2 * p) ** n * boys(n, T)
    if t == u == 0:
        val = 0.0
        if v > 1:
            val = val + (v - 1) * hermite_R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCz * hermite_R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC)
    if t == 0:
        val = 0.0
        if u > 1:
            val = val + (u - 1) * hermite_R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCy * hermite_R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC)
    val = 0.0
    if t > 1:
        val = val + (t - 1) * hermite_R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    return val + PCx * hermite_R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC)


def _prims(mol, ao):
    si, lmn = ao
    sh = mol.shells[si]
    norms = np.array([prim_norm(a, sh.l) for a in sh.exps])
    return mol.coords[sh.center], lmn, sh.exps, sh.coefs * norms


def one_electron(mol):
    """S, T, V (nuclear attraction, all nuclei) in the AO basis."""
    aos = mol.ao_list()
    n = len(aos)
    S = np.zeros((n, n))
    Tk = np.zeros((n, n))
    V = np.zeros((n, n))
    info = [_prims(mol, ao) for ao in aos]
    for i in range(n):
        A, la, ea, ca = info[i]
        for j in range(i + 1):
            B, lb, eb, cb = info[j]
            a = ea[:, None]
            b = eb[None, :]
            p = a + b
            cc = ca[:, None] * cb[None, :]
            Pc = (a[..., None] * A + b[..., None] * B) / p[..., None]
            s1 = [hermite_E(la[d], lb[d], 0, A[d] - B[d], a, b) * np.sqrt(np.pi / p) for d in range(3)]
            S[i, j] = np.sum(cc * s1[0] * s1[1] * s1[2])
            # kinetic: -1/2 sum_d D_d prod_{e != d} s_e
            kin = 0.0
            for d in range(3):
                j_ = lb[d]
                s_m2 = hermite_E(la[d], j_ - 2, 0, A[d] - B[d], a, b) * np.sqrt(np.pi / p) if j_ >= 2 else 0.0
                s_0 = s1[d]
                s_p2 = hermite_E(la[d], j_ + 2, 0, A[d] - B[d], a, b) * np.sqrt(np.pi / p)
                Dd = j_ * (j_ - 1) * s_m2 - 2 * b * (2 * j_ + 1) * s_0 + 4 * b * b * s_p2
                others = np.prod([s1[e] for e in range(3) if e != d], axis=0)
                kin = kin + Dd * others
            Tk[i, j] = -0.5 * np.sum(cc * kin)
            # nuclear attraction
            v = 0.0
            for C, Z in zip(mol.coords, mol.charges):
                PC = Pc - C
                RPC = np.linalg.norm(PC, axis=-1)
                acc = 0.0
                for t in range(la[0] + lb[0] + 1):
                    Ex = hermite_E(la[0], lb[0], t, A[0] - B[0], a, b)
                    for u in range(la[1] + lb[1] + 1):
                        Ey = hermite_E(la[1], lb[1], u, A[1] - B[1], a, b)
                        for w in range(la[2] + lb[2] + 1):
                            Ez = hermite_E(la[2], lb[2], w, A[2] - B[2], a, b)
                            acc = acc + Ex * Ey * Ez * hermite_R(t, u, w, 0, p, PC[..., 0], PC[..., 1], PC[..., 2], RPC)
                v = v - Z * 2 * np.pi / p * acc
            V[i, j] = np.sum(cc * v)
            S[j, i], Tk[j, i], V[j, i] = S[i, j], Tk[i, j], V[i, j]
    return S, Tk, V


def _eri_pair_data(info, i, j):
    A, la, ea, ca = info[i]
    B, lb, eb, cb = info[j]
    a = ea[:, None]
    b = eb[None, :]
    p = (a + b).reshape(-1)
    P = ((a[..., None] * A + b[..., None] * B) / (a + b)[..., None]).reshape(-1, 3)
    cc = (ca[:, None] * cb[None, :]).reshape(-1)
    E = [[hermite_E(la[d], lb[d], t, A[d] - B[d], a, b).reshape(-1) for t in range(la[d] + lb[d] + 1)]
         for d in range(3)]
    return p, P, cc, E, [la[d] + lb[d] for d in range(3)]


def _eri_quartet(d1, d2):
    p, P, c1, E1, L1 = d1
    q, Q, c2, E2, L2 = d2
    pp = p[:, None]
    qq = q[None, :]
    alpha = pp * qq / (pp + qq)
    PQ = P[:, None, :] - Q[None, :, :]
    RPQ = np.linalg.norm(PQ, axis=-1)
    acc = 0.0
    for t in range(L1[0] + 1):
        for u in range(L1[1] + 1):
            for v in range(L1[2] + 1):
                e1 = (E1[0][t] * E1[1][u] * E1[2][v])[:, None]
                for tau in range(L2[0] + 1):
                    for nu in range(L2[1] + 1):
                        for phi in range(L2[2] + 1):
                            e2 = (E2[0][tau] * E2[1][nu] * E2[2][phi])[None, :]
                            sign = (-1) ** (tau + nu + phi)
                            R = hermite_R(t + tau, u + nu, v + phi, 0, alpha, PQ[..., 0], PQ[..., 1], PQ[..., 2], RPQ)
                            acc = acc + sign * e1 * e2 * R
    return float(np.sum(c1[:, None] * c2[None, :] * 2 * np.pi ** 2.5 / (pp * qq * np.sqrt(pp + qq)) * acc))


def _eri_rows(args):
    mol, rows = args
    aos = mol.ao_list()
    n = len(aos)
    info = [_prims(mol, ao) for ao in aos]
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    pdat = {pr: _eri_pair_data(info, *pr) for pr in pairs}
    return [(x, [_eri_quartet(pdat[pairs[x]], pdat[pairs[y]]) for y in range(x + 1)]) for x in rows]


def eri(mol, procs=1):
    """(ij|kl) in chemists' notation, full n^4 array (8-fold symmetry used in the loops); the unique
    pair-pair quartets are optionally spread over `procs` host processes."""
    n = mol.nao
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    npair = len(pairs)
    if procs > 1:
        import multiprocessing as mp
        chunks = [list(range(r, npair, procs * 4)) for r in range(procs * 4)]
        with mp.get_context("spawn").Pool(procs) as pool:
            parts = pool.map(_eri_rows, [(mol, c) for c in chunks])
        rows = [r for part in parts for r in part]
    else:
        rows = _eri_rows((mol, range(npair)))
    out = np.zeros((n, n, n, n))
    for x, vals in rows:
        i, j = pairs[x]
        for y, val in enumerate(vals):
            k, l = pairs[y]
            for (a_, b_) in ((i, j), (j, i)):
                for (c_, d_) in ((k, l), (l, k)):
                    out[a_, b_, c_, d_] = val
                    out[c_, d_, a_, b_] = val
    return out


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------
def radial_treutler(n, xi=0.8):
    """Treutler-Ahlrichs M4 mapping on Chebyshev nodes of the second kind: r, w (w includes r^2)."""
    i = np.arange(1, n + 1)
    th = i * np.pi / (n + 1)
    x = np.cos(th)
    wx = np.pi / (n + 1) * np.sin(th)  # int_{-1}^{1} f(x) dx ~ sum wx f(x)
    ln2 = math.log(2.0)
    r = xi / ln2 * (1 + x) ** 0.6 * np.log(2 / (1 - x))
    drdx = xi / ln2 * (0.6 * (1 + x) ** (-0.4) * np.log(2 / (1 - x)) + (1 + x) ** 0.6 / (1 - x))
    return r, wx * drdx * r * r


def angular_product(n_theta=10, n_phi=11):
    """Gauss-Legendre in cos(theta) x uniform phi: unit vectors (n_theta*n_phi, 3) and weights (sum 4 pi)."""
    ct, wt = np.polynomial.legendre.leggauss(n_theta)
    phi = 2 * np.pi * (np.arange(n_phi) + 0.5) / n_phi
    st = np.sqrt(1 - ct * ct)
    dirs = np.stack([np.outer(st, np.cos(phi)).ravel(), np.outer(st, np.sin(phi)).ravel(),
                     np.outer(ct, np.ones(n_phi)).ravel()], axis=1)
    w = np.outer(wt, np.full(n_phi, 2 * np.pi / n_phi)).ravel()
    return dirs, w


# Lebedev-Laikov octahedral rules: orbit weights (summing to 1 over the points) and the orbits'
# parameters; a1: (1,0,0) x6, a2: (0,1,1)/sqrt2 x12, a3: (1,1,1)/sqrt3 x8, b(l): (l,l,m) x24,
# c(p): (p,q,0) x24.  n points -> (degree, [(orbit, parameter, weight)]).
LEBEDEV = {
    6: (3, [("a1", None, 1 / 6)]),
    14: (5, [("a1", None, 1 / 15), ("a3", None, 3 / 40)]),
    26: (7, [("a1", None, 1 / 21), ("a2", None, 4 / 105), ("a3", None, 9 / 280)]),
    38: (9, [("a1", None, 1 / 105), ("a3", None, 9 / 280), ("c", 0.4597008433809831, 1 / 35)]),
    50: (11, [("a1", None, 0.012698412698412698), ("a2", None, 0.022574955908289243),
              ("a3", None, 0.02109375), ("b", 0.30151134457776363, 0.020173335537918871)]),
    110: (17, [("a1", None, 0.3828270494937162e-2), ("a3", None, 0.9793737512487512e-2),
               ("b", 0.1851156353447362, 0.8211737283191111e-2), ("b", 0.6904210483822922, 0.9942814891178103e-2),
               ("b", 0.3956894730559419, 0.9595471336070963e-2), ("c", 0.4783690288121502, 0.9694996361663028e-2)]),
}


def _orbit(kind, par):
    import itertools

    pts = set()
    if kind == "a1":
        base = [(1.0, 0.0, 0.0)]
    elif kind == "a2":
        h = 1 / np.sqrt(2)
        base = [(0.0, h, h)]
    elif kind == "a3":
        t = 1 / np.sqrt(3)
        base = [(t, t, t)]
    elif kind == "b":
        base = [(par, par, np.sqrt(1 - 2 * par * par))]
    else:  # c
        base = [(par, np.sqrt(1 - par * par), 0.0)]
    for v in base:
        for perm in itertools.permutations(v):
            for sg in itertools.product((1, -1), repeat=3):
                pts.add(tuple(round(s_ * c, 15) + 0.0 for s_, c in zip(sg, perm)))
    return np.array(sorted(pts))


def angular_lebedev(n):
    """Lebedev-Laikov rule with n points: unit vectors (n, 3) and weights summing to 4 pi (the angular
    grid pyscf's dft.Grids uses; level 0 takes 50 points for hydrogen)."""
    if n not in LEBEDEV:
        raise ValueError(f"Lebedev rules available: {sorted(LEBEDEV)}")
    dirs, w = [], []
    for kind, par, wt in LEBEDEV[n][1]:
        o = _orbit(kind, par)
        dirs.append(o)
        w.append(np.full(len(o), 4 * np.pi * wt))
    dirs = np.concatenate(dirs)
    if len(dirs) != n:
        raise AssertionError("orbit sizes do not add up")
    return dirs, np.concatenate(w)


def stratmann_step(mu, a=0.64):
    """Stratmann-Scuseria-Frisch cell step (CPL 257, 213 (1996), eq. 14; pyscf gen_grid.stratmann,
    selected at td/trainer_legacy_no_jit.py:250): (35 z - 35 z^3 + 21 z^5 - 5 z^7) / 16 with z = mu / a,
    and sign(mu) for |mu| >= a."""
    mu = np.asarray(mu, dtype=np.float64)
    z = mu / a
    g = (35 * z - 35 * z ** 3 + 21 * z ** 5 - 5 * z ** 7) / 16
    return np.where(mu <= -a, -1.0, np.where(mu >= a, 1.0, g))


def becke_weights(coords_pts, atom_coords, owner, scheme="becke"):
    """Fuzzy-cell weight of each point's owning atom: Becke (3 iterations of p(mu) = 1.5 mu - 0.5 mu^3)
    or Stratmann (``stratmann_step``)."""
    A = len(atom_coords)
    d = np.linalg.norm(coords_pts[:, None, :] - atom_coords[None, :, :], axis=-1)  # (G, A)
    P = np.ones((len(coords_pts), A))
    for a in range(A):
        for b in range(A):
            if a == b:
                continue
            rab = np.linalg.norm(atom_coords[a] - atom_coords[b])
            mu = (d[:, a] - d[:, b]) / rab
            if scheme == "stratmann":
                mu = stratmann_step(mu)
            else:
                for _ in range(3):
                    mu = 1.5 * mu - 0.5 * mu ** 3
            P[:, a] *= 0.5 * (1 - mu)
    return P[np.arange(len(coords_pts)), owner] / np.sum(P, axis=1)


def molecular_grid(mol, n_rad=50, n_theta=10, n_phi=11, scheme="becke", n_ang=None):
    """Atom-centred grid: Treutler radial x (product, or Lebedev with n_ang points) angular, fuzzy
    cells by ``scheme``."""
    r, wr = radial_treutler(n_rad)
    dirs, wa = angular_product(n_theta, n_phi) if n_ang is None else angular_lebedev(n_ang)
    pts, wts, own = [], [], []
    for a, C in enumerate(mol.coords):
        p = C[None, None, :] + r[:, None, None] * dirs[None, :, :]
        pts.append(p.reshape(-1, 3))
        wts.append((wr[:, None] * wa[None, :]).ravel())
        own.append(np.full(n_rad * len(wa), a))
    pts = np.concatenate(pts)
    wts = np.concatenate(wts)
    own = np.concatenate(own)
    return pts, wts * becke_weights(pts, mol.coords, own, scheme)


def eval_ao(mol, pts):
    """phi[g, mu] for Cartesian s/p contracted Gaussians."""
    aos = mol.ao_list()
    out = np.zeros((len(pts), len(aos)))
    for mu, (si, lmn) in enumerate(aos):
        sh = mol.shells[si]
        d = pts - mol.coords[sh.center]
        r2 = np.sum(d * d, axis=1)
        rad = sum(c * prim_norm(e, sh.l) * np.exp(-e * r2) for e, c in zip(sh.exps, sh.coefs))
        ang = np.prod([d[:, k] ** lmn[k] for k in range(3)], axis=0)
        out[:, mu] = rad * ang
    return out


def eval_rho(ao, dm):
    """rho_g = sum_{mu nu} phi_g,mu D_mu,nu phi_g,nu (numint_legacy.py:351-373)."""
    return np.einsum("gi,ij,gj->g", ao, dm, ao)


# ---------------------------------------------------------------------------
# LDA: Slater exchange + VWN5 correlation (spin-unpolarised)
# ---------------------------------------------------------------------------
VWN5 = dict(A=0.0310907, x0=-0.10498, b=3.72744, c=12.9352)


def lda_exc_vxc(rho, functional="svwn5"):
    """Energy per electron and potential d(rho e)/d rho; zero where rho <= 1e-30.
    functional "svwn5": Slater exchange + VWN5 correlation (pyscf "lda,vwn"); "slater": Slater
    exchange alone (pyscf's bare "lda")."""
    rho = np.asarray(rho, dtype=np.float64)
    ok = rho > 1e-30
    r = np.where(ok, rho, 1.0)
    cx = -0.75 * (3.0 / np.pi) ** (1.0 / 3.0)
    ex = cx * r ** (1.0 / 3.0)
    vx = 4.0 / 3.0 * ex
    if functional == "slater":
        return np.where(ok, ex, 0.0), np.where(ok, vx, 0.0)
    if functional != "svwn5":
        raise ValueError(f"unknown LDA functional {functional!r}")
    A, x0, b, c = VWN5["A"], VWN5["x0"], VWN5["b"], VWN5["c"]
    rs = (3.0 / (4.0 * np.pi * r)) ** (1.0 / 3.0)
    x = np.sqrt(rs)
    X = x * x + b * x + c
    X0 = x0 * x0 + b * x0 + c
    Q = math.sqrt(4 * c - b * b)
    at = np.arctan(Q / (2 * x + b))
    ec = A * (np.log(x * x / X) + 2 * b / Q * at
              - b * x0 / X0 * (np.log((x - x0) ** 2 / X) + 2 * (b + 2 * x0) / Q * at))
    den = (2 * x + b) ** 2 + Q * Q
    dec = A * (2 / x - (2 * x + b) / X - 4 * b / den
               - b * x0 / X0 * (2 / (x - x0) - (2 * x + b) / X - 4 * (b + 2 * x0) / den))
    vc = ec - x / 6.0 * dec
    return np.where(ok, ex + ec, 0.0), np.where(ok, vx + vc, 0.0)


def nr_rks_lda(ao, weights, dm, functional="svwn5"):
    """(nelec, excsum, vmat) of numint.nr_rks for LDA (td/numint.py:42-64, 264-265)."""
    rho = eval_rho(ao, dm)
    exc, vrho = lda_exc_vxc(rho, functional)
    nelec = float(np.dot(rho, weights))
    excsum = float(np.dot(rho * weights, exc))
    aow = ao * (0.5 * weights * vrho)[:, None]
    vmat = ao.T @ aow
    return nelec, excsum, vmat + vmat.T


# ---------------------------------------------------------------------------
# Global (amplitude-encoded) neural XC on the full grid density
# ---------------------------------------------------------------------------
def gelu_tanh(x):
    """flax.linen.gelu with approximate=True (its default)."""
    return 0.5 * x * (1.0 + np.tanh(math.sqrt(2.0 / math.pi) * (x + 0.044715 * x ** 3)))


def gelu_tanh_grad(x):
    k = math.sqrt(2.0 / math.pi)
    t = np.tanh(k * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * k * (1.0 + 3 * 0.044715 * x * x)


def _sigmoid(x):
    return 0.5 * (1.0 + np.tanh(0.5 * x))


class GlobalMLP3D:
    """The MLP of ``trainer_legacy_no_jit.py:96-107`` applied to the whole grid density (the
    amplitude encoding, ``exc_and_vrho_global`` at :46-53): x = rho[n_g] -> Dense -> gelu -> ... ->
    Dense(1) -> -scale * swish; excsum = sum of the output; vrho = d excsum / d rho.
    weights[l] has shape (fan_in, fan_out) as flax Dense kernels."""

    def __init__(self, weights, biases, scale=1e-2):
        self.W = [np.asarray(w, dtype=np.float64) for w in weights]
        self.b = [np.asarray(b, dtype=np.float64) for b in biases]
        self.scale = float(scale)

    def exc_vrho(self, rho):
        h = np.asarray(rho, dtype=np.float64)
        zs = []
        for W, b in zip(self.W[:-1], self.b[:-1]):
            z = h @ W + b
            zs.append(z)
            h = gelu_tanh(z)
        z = h @ self.W[-1] + self.b[-1]
        s = _sigmoid(z)
        out = -self.scale * z * s
        g = -self.scale * (s + z * s * (1.0 - s))  # d out / d z
        gh = self.W[-1] @ g
        for W, z in zip(reversed(self.W[:-1]), reversed(zs)):
            gh = W @ (gh * gelu_tanh_grad(z))
        return float(np.sum(out)), gh


class LocalMLP3D(GlobalMLP3D):
    """The same MLP applied to every grid point's density (``exc_and_vrho_local`` at
    ``trainer_legacy_no_jit.py:56-63``: network.apply(params, rho[:, None]).squeeze(), then the VJP
    with a unit cotangent): exc_g = net(rho_g), vrho_g = d exc_g / d rho_g (points are independent)."""

    def exc_vrho(self, rho):
        h = np.asarray(rho, dtype=np.float64)[:, None]
        zs = []
        for W, b in zip(self.W[:-1], self.b[:-1]):
            z = h @ W + b
            zs.append(z)
            h = gelu_tanh(z)
        z = (h @ self.W[-1] + self.b[-1])[:, 0]
        s = _sigmoid(z)
        exc = -self.scale * z * s
        g = (-self.scale * (s + z * s * (1.0 - s)))[:, None] * self.W[-1][:, 0][None, :]
        for W, zz in zip(reversed(self.W[:-1]), reversed(zs)):
            g = (g * gelu_tanh_grad(zz)) @ W.T
        return exc, g[:, 0]


def nr_rks_nn_local(ao, weights, dm, net):
    """(nelec, excsum, vmat) of nr_rks's "NN" branch (td/numint.py:179-204): excsum = sum rho w exc and
    V = ao^T diag(w vrho / 2) ao with vrho = d exc / d rho (not d(rho exc) / d rho), then V + V^T."""
    rho = eval_rho(ao, dm)
    exc, vrho = net.exc_vrho(rho)
    den = rho * weights
    aow = ao * (0.5 * weights * vrho)[:, None]
    vmat = ao.T @ aow
    return float(den.sum()), float(den @ exc), vmat + vmat.T


def init_global_mlp(seed, n_in, hidden=(512, 512)):
    """Seeded weights: kernels N(0, 1/fan_in) (the variance of flax's lecun_normal), zero biases."""
    rng = np.random.default_rng(seed)
    sizes = [n_in, *hidden, 1]
    W = [rng.standard_normal((a, b)) / math.sqrt(a) for a, b in zip(sizes[:-1], sizes[1:])]
    return W, [np.zeros(b) for b in sizes[1:]]


def nr_rks_nn_global(ao, weights, dm, net):
    """(nelec, excsum, vmat) of nr_rks's "NN-AmplitudeEncoding" branch (td/numint.py:206-231,
    264-265): excsum is the network's scalar and vrho is scaled by the weights once more."""
    rho = eval_rho(ao, dm)
    excsum, vrho = net.exc_vrho(rho)
    nelec = float(np.dot(rho, weights))
    aow = ao * (0.5 * weights * vrho)[:, None]
    vmat = ao.T @ aow
    return nelec, excsum, vmat + vmat.T


# ---------------------------------------------------------------------------
# SCF (RKS-LDA or RHF) with CDIIS
# ---------------------------------------------------------------------------
def get_j(eri_, dm):
    return np.einsum("ijkl,kl->ij", eri_, dm)


def get_k(eri_, dm):
    return np.einsum("ikjl,kl->ij", eri_, dm)


def gen_eigh(F, S):
    """F C = S C e (generalized_eigensolver.py:264-330 semantics: symmetrise, eigh(F, S))."""
    F = 0.5 * (F + F.T)
    return scipy.linalg.eigh(F, S)


def degen_eigh_bwd(eival, eivec, g_eival=None, g_eivec=None, exact=False):
    """Reverse pass of A = V diag(w) V^T with degenerate pairs masked
    (qedft/train/td/generalized_eigensolver.py:158-219): M = F o (U gV) + diag(gw) with
    F_ij = 1/(w_j - w_i), zero where |w_j - w_i| < eps^0.6, result sym(V M U).  The reference's code
    takes U = jnp.transpose(eivec, axes=(-2, -1)), the identity on a 2-D array, so U = V (exact=False
    reproduces it); exact=True takes U = V^T, the derivative of the decomposition (pinned by finite
    differences in tests/test_oracle_td.py)."""
    w = np.asarray(eival, dtype=np.float64)
    V = np.asarray(eivec, dtype=np.float64)
    U = V.T if exact else V
    res = np.zeros_like(V)
    if g_eivec is not None:
        F = w[None, :] - w[:, None]
        F = np.where(np.abs(F) < np.finfo(np.float64).eps ** 0.6, np.inf, F)
        res = V @ ((1.0 / F) * (U @ np.asarray(g_eivec, dtype=np.float64))) @ U
    if g_eival is not None:
        res = res + V @ (np.asarray(g_eival, dtype=np.float64)[:, None] * U)
    return 0.5 * (res + res.T)


def make_rdm1(C, nocc):
    """D = 2 C_occ C_occ^T (hf.py:303-308, aufbau occupation of the N_e/2 lowest)."""
    Co = C[:, :nocc]
    return 2.0 * Co @ Co.T


class CDIIS:
    """Commutator DIIS on F with error F D S - S D F, at most `space` vectors (jax_diis.py)."""

    def __init__(self, space=8):
        self.space = space
        self.F, self.E = [], []

    def update(self, S, D, F):
        err = F @ D @ S - S @ D @ F
        self.F.append(F.copy())
        self.E.append(err.ravel())
        if len(self.F) > self.space:
            self.F.pop(0)
            self.E.pop(0)
        m = len(self.F)
        B = np.zeros((m + 1, m + 1))
        B[0, 1:] = B[1:, 0] = -1.0
        for i in range(m):
            for j in range(m):
                B[i + 1, j + 1] = self.E[i] @ self.E[j]
        rhs = np.zeros(m + 1)
        rhs[0] = -1.0
        c = np.linalg.solve(B, rhs)[1:]
        return sum(ci * Fi for ci, Fi in zip(c, self.F))


@dataclass
class SCFResult:
    e_tot: float
    energies: np.ndarray  # per cycle
    mo_energy: np.ndarray
    mo_coeff: np.ndarray
    dm: np.ndarray
    converged: bool
    cycles: int
    nelec_grid: float = float("nan")


def scf(mol, method="rks", grid=None, max_cycle=30, conv_tol=1e-8, diis_space=8, ints=None, net=None,
        functional="svwn5"):
    """pyscf-style SCF (hf_legacy.py:51-116): core-Hamiltonian guess; DIIS from cycle 1;
    converged when |dE| < conv_tol and |g|/sqrt(size) < sqrt(conv_tol).  method: "rks" (the LDA
    `functional`, "svwn5" or "slater"), "rks_nn" (GlobalMLP3D `net` on the grid density) or "rhf"."""
    S, T, V = ints[:3] if ints is not None else one_electron(mol)
    g = ints[3] if ints is not None else eri(mol)
    H = T + V
    nocc = mol.nelectron // 2
    if method in ("rks", "rks_nn", "rks_nn_local"):
        pts, w = grid if grid is not None else molecular_grid(mol)
        ao = eval_ao(mol, pts)

    def veff(D):
        J = get_j(g, D)
        if method == "rhf":
            K = get_k(g, D)
            return J - 0.5 * K, 0.5 * np.sum(D * J) - 0.25 * np.sum(D * K), np.nan
        if method == "rks_nn":
            nel, exc, vxc = nr_rks_nn_global(ao, w, D, net)
        elif method == "rks_nn_local":
            nel, exc, vxc = nr_rks_nn_local(ao, w, D, net)
        else:
            nel, exc, vxc = nr_rks_lda(ao, w, D, functional)
        return J + vxc, 0.5 * np.sum(D * J) + exc, nel

    e, C = gen_eigh(H, S)
    D = make_rdm1(C, nocc)
    vhf, e2, nel = veff(D)
    e_tot = np.sum(D * H) + e2 + mol.energy_nuc()
    diis = CDIIS(diis_space)
    energies = []
    conv = False
    cycle = 0
    for cycle in range(max_cycle):
        e_last = e_tot
        F = H + vhf
        if cycle >= 1:
            F = diis.update(S, D, F)
        e, C = gen_eigh(F, S)
        D = make_rdm1(C, nocc)
        vhf, e2, nel = veff(D)
        e_tot = np.sum(D * H) + e2 + mol.energy_nuc()
        energies.append(e_tot)
        Fn = H + vhf
        gorb = 2 * C[:, nocc:].T @ Fn @ C[:, :nocc]
        ngorb = np.linalg.norm(gorb) / math.sqrt(gorb.size)
        if abs(e_tot - e_last) < conv_tol and ngorb < math.sqrt(conv_tol):
            conv = True
            break
    return SCFResult(e_tot, np.array(energies), e, C, D, conv, cycle + 1, nel)
