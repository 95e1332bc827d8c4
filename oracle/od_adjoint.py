"""Reverse-mode gradient of a loss on the 1D scan-semantics Kohn-Sham trajectory with respect to the
parameters of a windowed LocalMLP XC functional, or of a GlobalMLP one (amplitude encoded: F = f,
v_xc = dx grad f; under reflection only the density passed on is flipped, not the functional's
inputs) (test infrastructure only: the FP64 reference that the device adjoint is checked against).

What the reference differentiates (``qedft/train/od/train.py:58-126, 250-323``): ``jax.grad`` of a
loss built from the states of ``jit_scf.kohn_sham`` (``od/jit_scf.py:143-244``), unrolled over the
scan, with the XC network parameters as the only leaves.  This module writes that reverse pass by
hand (for ``density_mse_converge_tolerance < 0``, i.e. no frozen iterations, and no reflection):

  forward, iteration k (input density n_k, mixing weight alpha_k, connection weights W):
    v_KS = v_ext + v_H(n_k) + v_xc(n_k; theta)        v_H = dx U n, v_xc = grad_n F, F = sum e(n) n
    (eps_j, psi_j) = eigh(T + diag v_KS);  n_out = sum_j occ_j psi_j^2 / dx
    E_k = sum occ eps - dx v_KS.n_out + E_H(n_out) + dx F(n_out) + dx v_ext.n_out
    d_k = n_out - n_k;  n_{k+1} = n_k + alpha_k sum_j W[k, j] d_j  (stored density of iteration k)

  reverse, iteration k:
    dE/dn_out = dx (-v_KS + v_H(n_out) + v_xc(n_out) + v_ext);  dE/dtheta = dx dF(n_out)/dtheta;
    dE/dv_KS = 0 at fixed n_out (Hellmann-Feynman cancels -dx n_out);
    n_out -> v_KS by first-order perturbation theory over the full spectrum:
      vbar = sum_j occ_j (2/dx) psi_j * y_j,  y_j = -sum_{a != j} psi_a (psi_a . (nbar * psi_j)) / (eps_a - eps_j)
      (pairs of occupied orbitals with equal occupation cancel exactly and are skipped);
    v_KS -> n_k: nbar += dx U vbar + H_F vbar, thetabar += d(vbar . grad_n F)/dtheta, where the
      second-order terms come from reverse mode over the MLP's forward and forward-tangent passes.
"""

from __future__ import annotations

import math

import numpy as np

from . import od as O


# ---- activations with first and second derivatives ---------------------------------------
def act3(name, z):
    """(sigma, sigma', sigma'') elementwise."""
    if name == "silu":
        s = 1 / (1 + np.exp(-z))
        d1 = s * (1 + z * (1 - s))
        d2 = s * (1 - s) * (2 + z * (1 - 2 * s))
        return z * s, d1, d2
    if name == "tanh":
        t = np.tanh(z)
        return t, 1 - t * t, -2 * t * (1 - t * t)
    if name == "softplus":
        s = 1 / (1 + np.exp(-z))
        return np.logaddexp(0, z), s, s * (1 - s)
    if name == "sigmoid":
        s = 1 / (1 + np.exp(-z))
        return s, s * (1 - s), s * (1 - s) * (1 - 2 * s)
    if name == "relu":
        return np.maximum(z, 0), (z > 0).astype(z.dtype), np.zeros_like(z)
    if name == "gelu":
        k = math.sqrt(2 / math.pi)
        c = 0.044715
        u = k * (z + c * z ** 3)
        t = np.tanh(u)
        du = k * (1 + 3 * c * z * z)
        sech2 = 1 - t * t
        d1 = 0.5 * (1 + t) + 0.5 * z * sech2 * du
        d2 = sech2 * du + 0.5 * z * (-2 * t * sech2 * du * du + sech2 * k * 6 * c * z)
        return 0.5 * z * (1 + t), d1, d2
    raise ValueError(f"unknown activation {name}")


def output3(kind, y):
    """(g, g', g'') of the network-output transform (``od.output_transform``)."""
    if kind == "linear":
        return y, np.ones_like(y), np.zeros_like(y)
    if kind == "negativity":  # -swish
        s = 1 / (1 + np.exp(-y))
        return -y * s, -(s * (1 + y * (1 - s))), -(s * (1 - s) * (2 + y * (1 - 2 * s)))
    raise ValueError(f"unknown output transform {kind!r}")


def mlp_second_order(weights, biases, act, X, Xd, fbar, fdbar, output="linear"):
    """Reverse mode over the forward and forward-tangent passes of an MLP.

    X (P, D0) inputs, Xd (P, D0) input tangents; per point the scalar Phi = fbar f + fdbar fdot
    (f = g(MLP(X)) with the output transform g, fdot = grad_x f . Xd).  Returns (f, fdot, Xbar (P, D0),
    dW list, db list) with Xbar = dPhi/dX (fbar grad f + fdbar H_f Xd) and dW, db summed over points.
    """
    L = len(weights)
    hs, hds, zs, d1s, d2s = [X], [Xd], [], [], []
    h, hd = X, Xd
    for li in range(L - 1):
        z = h @ weights[li] + biases[li]
        zd = hd @ weights[li]
        a, d1, d2 = act3(act, z)
        zs.append(zd)  # tangent of z
        d1s.append(d1)
        d2s.append(d2)
        h, hd = a, d1 * zd
        hs.append(h)
        hds.append(hd)
    y = (h @ weights[-1] + biases[-1])[:, 0]
    yd = (hd @ weights[-1])[:, 0]
    g0, g1, g2 = output3(output, y)
    f, fd = g0, g1 * yd
    # reverse through the transform: (f, fd) = (g(y), g'(y) yd)
    fbar, fdbar = fbar * g1 + fdbar * g2 * yd, fdbar * g1
    dW = [np.zeros_like(w) for w in weights]
    db = [np.zeros_like(b) for b in biases]
    # last layer: f = h W + b, fd = hd W
    hb = fbar[:, None] * weights[-1][:, 0][None, :]
    hdb = fdbar[:, None] * weights[-1][:, 0][None, :]
    dW[-1] += (hs[-1] * fbar[:, None]).sum(0)[:, None] + (hds[-1] * fdbar[:, None]).sum(0)[:, None]
    db[-1] += np.array([fbar.sum()])
    for li in range(L - 2, -1, -1):
        # h = sigma(z), hd = sigma'(z) zd
        zdb = hdb * d1s[li]
        zb = hb * d1s[li] + hdb * d2s[li] * zs[li]
        dW[li] += hs[li].T @ zb + hds[li].T @ zdb
        db[li] += zb.sum(0)
        hb = zb @ weights[li].T
        hdb = zdb @ weights[li].T
    return f, fd, hb, dW, db


class LocalMLPAdjoint:
    """First- and second-order pieces of F(n) = sum_i e_i(n) n_i for ``od.LocalMLP``."""

    def __init__(self, xc):
        self.xc = xc
        self.h = xc.window // 2

    def _scatter(self, Xbar):
        """Adjoint of the window map x_i[o] = n[i + o - h] / norm."""
        N, Wd = Xbar.shape
        out = np.zeros(N)
        for o in range(Wd):
            j = np.arange(N) + o - self.h
            ok = (j >= 0) & (j < N)
            np.add.at(out, j[ok], Xbar[ok, o] / self.xc.norm)
        return out

    def energy_and_grads(self, n):
        """F, grad_n F (= v_xc), dF/dtheta (list of (dW, db))."""
        X = self.xc._windows(n)
        z0 = np.zeros_like(X)
        f, _, Xbar, dW, db = mlp_second_order(self.xc.weights, self.xc.biases, self.xc.activation, X, z0, n,
                                              np.zeros(len(n)), getattr(self.xc, "output", "linear"))
        return float(f @ n), f + self._scatter(Xbar), list(zip(dW, db))

    def potential(self, n):
        """v_xc = grad_n F."""
        return self.energy_and_grads(n)[1]

    def vxc_vjp(self, n, vbar):
        """(nbar, thetabar) of vbar . v_xc(n): H_F vbar and d(vbar . grad F)/dtheta."""
        X = self.xc._windows(n)
        Xd = self.xc._windows(vbar)
        f, fd, Xbar, dW, db = mlp_second_order(self.xc.weights, self.xc.biases, self.xc.activation, X, Xd, vbar, n,
                                               getattr(self.xc, "output", "linear"))
        return fd + self._scatter(Xbar), list(zip(dW, db))


class GlobalMLPAdjoint:
    """The same pieces for ``od.GlobalMLP`` (amplitude encoded, ``od/scf.py:32-76``): E_xc = dx f(n),
    v_xc = dx grad_n f, so F = f in the energy and the potential carries the factor dx."""

    def __init__(self, xc, dx):
        self.xc, self.dx = xc, dx

    def _x(self, n):
        return (np.asarray(n, dtype=np.float64) / self.xc.norm)[None, :]

    def energy_and_grads(self, n):
        """f, grad_n f, df/dtheta."""
        X = self._x(n)
        f, _, Xbar, dW, db = mlp_second_order(self.xc.weights, self.xc.biases, self.xc.activation, X,
                                              np.zeros_like(X), np.ones(1), np.zeros(1), self.xc.output)
        return float(f[0]), Xbar[0] / self.xc.norm, list(zip(dW, db))

    def potential(self, n):
        return self.dx * self.energy_and_grads(n)[1]

    def vxc_vjp(self, n, vbar):
        """(nbar, thetabar) of vbar . v_xc(n) = dx grad f . vbar: dx H_f vbar and its theta-gradient."""
        X = self._x(n)
        Xd = self._x(vbar)
        _, _, Xbar, dW, db = mlp_second_order(self.xc.weights, self.xc.biases, self.xc.activation, X, Xd,
                                              np.zeros(1), np.full(1, self.dx), self.xc.output)
        return Xbar[0] / self.xc.norm, list(zip(dW, db))


def flatten(params):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in params])


def unflatten(flat, like):
    out, o = [], 0
    for w, b in like:
        W = flat[o:o + w.size].reshape(w.shape)
        o += w.size
        B = flat[o:o + b.size].reshape(b.shape)
        o += b.size
        out.append((W, B))
    return out


def occupations(num_electrons):
    k = (num_electrons + 1) // 2
    occ = np.full(k, 2.0)
    if num_electrons % 2:
        occ[-1] = 1.0
    return occ


def forward_trajectory(locations, charges, num_electrons, num_iterations, grids, xc, interaction_fn,
                       initial_density, alpha, alpha_decay, num_mixing_iterations, reflect=False, flip=None):
    """Scan-semantics forward (jit_kohn_sham without convergence freezing) keeping what the reverse
    pass needs.  reflect: enforce_reflection_symmetry about the grid centre (od/jit_scf.py:76-77),
    S = flip-and-average: the XC functional is the wrapped e_S(n) = S e(S n), so F_S(n) = e_S(n).n =
    F(S n) and v_xc = S grad F(S n); the energy is that of the raw output density n_raw (with the XC
    part F(S n_raw)), and the density passed on is n_out = S n_raw (jax_dft flips after the energy).
    ``flip``: another flip-and-average map (the python loop's, about the centre of the nuclei,
    od/scf.py:162-167); with a negative convergence tolerance that loop is otherwise this one."""
    S = (flip or O.flip_and_average_on_center) if reflect else (lambda a: a)
    grids = np.asarray(grids, dtype=np.float64)
    dx = O.get_dx(grids)
    v_ext = O.get_atomic_chain_potential(grids, np.asarray(locations, dtype=np.float64), np.asarray(charges),
                                         interaction_fn)
    Wc = O.connection_weights(num_iterations, num_mixing_iterations)
    T = O.get_kinetic_matrix(grids)
    glob = isinstance(xc, O.GlobalMLP)
    adj = GlobalMLPAdjoint(xc, dx) if glob else LocalMLPAdjoint(xc)
    # the XC functional sees flipped densities only when local (od/jit_scf.py:15-78 applies no flip to
    # the amplitude-encoded functional: its potential is at n, its energy at the raw output)
    Sx = (lambda a: a) if glob else S
    K = (num_electrons + 1) // 2
    occ = occupations(num_electrons)
    n = np.asarray(initial_density, dtype=np.float64).copy()
    diffs = np.zeros((num_iterations, len(grids)))
    rec = []
    a = alpha
    for k in range(num_iterations):
        vh = O.get_hartree_potential(n, grids, interaction_fn)
        vxc = Sx(adj.potential(Sx(n)))
        vks = v_ext + vh + vxc
        eps, U = np.linalg.eigh(T + np.diag(vks))
        psi = U[:, :K].T  # (K, N), unit 2-norm
        n_raw = (occ[:, None] * psi ** 2).sum(0) / dx
        n_out = S(n_raw)
        Fo, vxc_out, _ = adj.energy_and_grads(Sx(n_raw))
        E = (float(occ @ eps[:K]) - dx * vks @ n_raw + O.get_hartree_energy(n_raw, grids, interaction_fn)
             + dx * Fo + dx * v_ext @ n_raw)
        diffs[k] = n_out - n
        mixed = n + a * (Wc[k] @ diffs)
        rec.append(dict(n=n, vks=vks, eps=eps, U=U, n_raw=n_raw, n_out=n_out, E=E, alpha=a, mixed=mixed))
        n = mixed
        a *= alpha_decay
    return dict(rec=rec, v_ext=v_ext, Wc=Wc, dx=dx, occ=occ, K=K, adj=adj, grids=grids, S=S, Sx=Sx,
                interaction_fn=interaction_fn,
                energies=np.array([r["E"] for r in rec]), densities=np.stack([r["mixed"] for r in rec]))


def eig_adjoint(eps, U, occ, nbar, dx):
    """vbar of nbar . n_out with n_out = sum_j occ_j psi_j^2 / dx (full-spectrum perturbation theory)."""
    K = len(occ)
    N = len(eps)
    vbar = np.zeros(N)
    for j in range(K):
        r = nbar * U[:, j]
        c = U.T @ r  # psi_a . r
        den = eps - eps[j]
        w = np.zeros(N)
        for a in range(N):
            if a == j or (a < K and occ[a] == occ[j]):
                continue
            w[a] = c[a] / den[a]
        y = -(U @ w)
        vbar += occ[j] * (2.0 / dx) * U[:, j] * y
    return vbar


def backward(fw, dE, dn, stop_gradient_step=-1):
    """Gradient of sum_k dE[k] E_k + sum_k dn[k] . density_k w.r.t. the flat MLP parameters.

    stop_gradient_step: the states of iterations k <= it are constants (jax.lax.stop_gradient on the
    scan state in jax_dft jit_scf.kohn_sham): neither the loss's cotangents on them nor those flowing
    back from later iterations reach them; the differences they contribute to later mixes still do."""
    rec, dx, adj, S, Sx = fw["rec"], fw["dx"], fw["adj"], fw["S"], fw["Sx"]
    grids, ifn, v_ext, Wc = fw["grids"], fw["interaction_fn"], fw["v_ext"], fw["Wc"]
    Tn = len(rec)
    N = len(grids)
    theta = np.zeros_like(flatten(list(zip(adj.xc.weights, adj.xc.biases))))
    dbar = np.zeros((Tn, N))
    nbar_next = np.zeros(N)
    for k in range(Tn - 1, -1, -1):
        r = rec[k]
        g = nbar_next + dn[k] if k > stop_gradient_step else np.zeros(N)
        nbar = g.copy()
        for j in range(k + 1):
            dbar[j] += r["alpha"] * Wc[k, j] * g
        nraw_bar = S(dbar[k])  # cotangent of n_raw (n_out = S n_raw)
        nbar -= dbar[k]
        if dE[k] != 0.0 and k > stop_gradient_step:
            _, vxc_out, dth = adj.energy_and_grads(Sx(r["n_raw"]))
            vh_out = O.get_hartree_potential(r["n_raw"], grids, ifn)
            nraw_bar += dE[k] * dx * (-r["vks"] + vh_out + Sx(vxc_out) + v_ext)
            theta += dE[k] * dx * flatten(dth)
        vbar = eig_adjoint(r["eps"], r["U"], fw["occ"], nraw_bar, dx)
        nbar += O.get_hartree_potential(vbar, grids, ifn)
        nb2, dth2 = adj.vxc_vjp(Sx(r["n"]), Sx(vbar))
        nbar += Sx(nb2)
        theta += flatten(dth2)
        nbar_next = nbar
    return theta


# ---- losses ---------------------------------------------------------------------------------
def c3_loss_and_cotangents(energies, densities, target_energy, target_density, dx, energy_weight=1.0,
                           density_weight=1e-2):
    """Per system: (E_T - E*)^2 + 1e-2 dx sum (n_T - n*)^2 (the batch loss is the mean over systems)."""
    Tn = len(energies)
    dE = np.zeros(Tn)
    dn = np.zeros_like(densities)
    e = energies[-1] - target_energy
    d = densities[-1] - target_density
    loss = energy_weight * e * e + density_weight * dx * float(d @ d)
    dE[-1] = 2 * energy_weight * e
    dn[-1] = 2 * density_weight * dx * d
    return loss, dE, dn


def trajectory_loss_and_cotangents(energies, densities, target_energy, target_density, num_electrons, dx,
                                   discount=0.9, start=10):
    """The reference trainer's loss for one system (od/train.py:250-323 with jax_dft
    losses.trajectory_mse): mean_t discount^(S-1-t) (E_t - E*)^2 over t >= start, / N_e, plus
    mean_i (n_T - n*)^2 * dx * N / N_e."""
    Tn, N = densities.shape
    S = Tn - start
    disc = discount ** np.arange(S)[::-1]
    e = energies[start:] - target_energy
    loss = float(np.mean(disc * e * e)) / num_electrons
    dE = np.zeros(Tn)
    dE[start:] = 2 * disc * e / S / num_electrons
    d = densities[-1] - target_density
    loss += float(np.mean(d * d)) * dx * N / num_electrons
    dn = np.zeros_like(densities)
    dn[-1] = 2 * d / N * dx * N / num_electrons
    return loss, dE, dn
