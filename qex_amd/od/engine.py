"""Batched 1D SCF driver over the C ABI (``qex1d_*`` in include/qex.h).

All arrays live on the device as torch tensors; the library enqueues its kernels on torch's
current stream.  There is no host computation path: without a GPU or without libqex.so
every entry point raises.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from . import xc as xcmod
from .utils import check_uniform, get_dx, reflection_window


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.QexError("no GPU visible: the qex engine runs only on the device")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def as_f64(x, dev) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev).contiguous()


def as_i32(x, dev) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.int32).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.int32), device=dev).contiguous()


def upload(x, dtype, dev) -> torch.Tensor:
    """Host array -> device tensor without waiting for the stream: staged through pinned memory
    (torch's caching host allocator keeps the staging block until the copy has run).  A copy from
    pageable memory would block the host until the work queued before it on the stream is done."""
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=dtype))
    return t.pin_memory().to(dev, non_blocking=True)


def workspace(nbytes: int, dev) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)


def interaction_table(grids: torch.Tensor, interaction_fn) -> torch.Tensor:
    """u(k dx), k = 0..N-1: row 0 of the Toeplitz interaction matrix (u must be even)."""
    n = grids.shape[0]
    dx = get_dx(grids)
    d = dx * torch.arange(n, dtype=torch.float64, device=grids.device)
    return as_f64(interaction_fn(d), grids.device)


def external_potential(grids, locations, nuclear_charges, interaction_fn) -> torch.Tensor:
    """v_ext = -sum_a Z_a u(x - R_a); batched when ``locations`` is 2D ([B, A])."""
    dev = device()
    g = as_f64(grids, dev)
    loc = as_f64(locations, dev)
    z = as_f64(nuclear_charges, dev)
    single = loc.dim() == 1
    if single:
        loc, z = loc[None], z[None]
    B, A = loc.shape
    N = g.shape[0]
    kind = getattr(interaction_fn, "kind", None)
    if kind in (_lib.QEX_INTERACTION_EXPONENTIAL, _lib.QEX_INTERACTION_SOFT):
        out = torch.empty((B, N), dtype=torch.float64, device=dev)
        p0, p1 = interaction_fn.params
        _lib.check(_lib.load().qex1d_external_potential(B, N, _ptr(g), A, _ptr(loc), _ptr(z), kind, p0, p1,
                                                        _ptr(out), _stream()), "qex1d_external_potential")
    else:  # arbitrary even interaction: evaluated on the device through torch
        out = -(z[:, :, None] * interaction_fn(g[None, None, :] - loc[:, :, None])).sum(dim=1)
        out = as_f64(out, dev)
    return out[0] if single else out


def hartree_potential(density, grids, interaction_fn) -> torch.Tensor:
    dev = device()
    g = as_f64(grids, dev)
    n = as_f64(density, dev)
    single = n.dim() == 1
    if single:
        n = n[None]
    B, N = n.shape
    out = torch.empty_like(n)
    table = interaction_table(g, interaction_fn)
    _lib.check(_lib.load().qex1d_hartree_potential(B, N, get_dx(g), _ptr(table), _ptr(n), _ptr(out), _stream()),
               "qex1d_hartree_potential")
    return out[0] if single else out


def xc_evaluate(xc, density, dx: float):
    """The functional ``xc`` on densities [N] or [B, N] with grid spacing ``dx`` through
    ``qex1d_xc_evaluate``: returns (energy density, E_xc, v_xc).  Energy density: [.., N] for local
    functionals, [..] for global (amplitude-encoded) ones; E_xc [..]; v_xc [.., N] (local:
    (dE/dn)/dx, global: dE/dn)."""
    dev = device()
    f = xcmod.resolve(xc)
    n = as_f64(density, dev)
    single = n.dim() == 1
    if single:
        n = n[None]
    if n.dim() != 2:
        raise ValueError("density must be [N] or [B, N]")
    B, N = n.shape
    if N > _lib.QEX1D_MAX_GRIDS:
        raise ValueError(f"num_grids {N} exceeds {_lib.QEX1D_MAX_GRIDS}")
    p = _lib.Qex1dProblem()
    p.batch, p.num_grids, p.dx = B, N, float(dx)
    f.configure(p)
    params = f.device_params(dev)
    lib = _lib.load()
    nbytes = lib.qex1d_xc_workspace_bytes(ctypes.byref(p))
    if nbytes == 0:
        _lib.check(-1, "qex1d_xc_workspace_bytes")
    ws = workspace(nbytes, dev)
    eps = torch.empty((B, N) if f.encoding == "local" else (B,), dtype=torch.float64, device=dev)
    exc = torch.empty(B, dtype=torch.float64, device=dev)
    vxc = torch.empty((B, N), dtype=torch.float64, device=dev)
    _lib.check(lib.qex1d_xc_evaluate(ctypes.byref(p), _ptr(params), _ptr(n), _ptr(eps), _ptr(exc), _ptr(vxc),
                                     _ptr(ws), ws.numel(), _stream()), "qex1d_xc_evaluate")
    if single:
        return eps[0], exc[0], vxc[0]
    return eps, exc, vxc


@dataclass
class NonInteracting:
    density: torch.Tensor
    total_eigen_energies: torch.Tensor
    gap: torch.Tensor
    eigenvalues: torch.Tensor | None = None
    wavefunctions: torch.Tensor | None = None


def solve_noninteracting(potential, num_electrons, grids, with_states=False, max_electrons=None) -> NonInteracting:
    """Lowest eigenpairs of T + diag(v) for a batch [B, N] of potentials."""
    shape = tuple(potential.shape) if hasattr(potential, "shape") else np.shape(potential)
    ne = np.broadcast_to(np.asarray(num_electrons, dtype=np.int64).reshape(-1),
                         (1 if len(shape) == 1 else shape[0],)).copy()
    if np.any(ne < 1):
        raise ValueError("num_electrons must be >= 1")
    me = int(max_electrons or ne.max())
    if ne.max() > me:
        raise ValueError(f"num_electrons {int(ne.max())} exceeds max_electrons {me}")
    if me > _lib.QEX1D_MAX_ELECTRONS:
        raise ValueError(f"num_electrons must be <= {_lib.QEX1D_MAX_ELECTRONS}")
    dev = device()
    v = as_f64(potential, dev)
    single = v.dim() == 1
    if single:
        v = v[None]
    B, N = v.shape
    K = me // 2 + 1
    ne_t = as_i32(ne, dev)
    lib = _lib.load()
    ws = workspace(lib.qex1d_noninteracting_workspace_bytes(B, N, me), dev)
    dens = torch.empty((B, N), dtype=torch.float64, device=dev)
    ee = torch.empty(B, dtype=torch.float64, device=dev)
    gap = torch.empty(B, dtype=torch.float64, device=dev)
    eig = torch.empty((B, K), dtype=torch.float64, device=dev) if with_states else None
    wf = torch.empty((B, K, N), dtype=torch.float64, device=dev) if with_states else None
    _lib.check(lib.qex1d_solve_noninteracting(B, N, get_dx(as_f64(grids, dev)), me, _ptr(v), _ptr(ne_t), _ptr(dens),
                                              _ptr(ee), _ptr(gap), _ptr(eig), _ptr(wf), _ptr(ws), ws.numel(),
                                              _stream()), "qex1d_solve_noninteracting")
    if single:
        return NonInteracting(dens[0], ee[0], gap[0], None if eig is None else eig[0], None if wf is None else wf[0])
    return NonInteracting(dens, ee, gap, eig, wf)


@dataclass
class BatchResult:
    density: torch.Tensor            # [B, T, N]
    total_energy: torch.Tensor       # [B, T]
    hartree_potential: torch.Tensor  # [B, T, N]
    xc_potential: torch.Tensor       # [B, T, N]
    xc_energy_density: torch.Tensor  # [B, T, N] (local) or [B, T] (global)
    gap: torch.Tensor                # [B, T]
    converged: torch.Tensor          # [B, T] bool
    external_potential: torch.Tensor  # [B, N]
    eigenvalues: torch.Tensor | None = None  # [B, T, K]
    stats: torch.Tensor | None = None        # [2] int64: Sturm rounds, eigenpairs (accumulated)
    orbitals: torch.Tensor | None = None     # [B, T, K, N] (the reverse pass needs them)


class Problem1D:
    """Device-resident description of a batch of systems on one grid (reused across solves)."""

    def __init__(self, grids, locations, nuclear_charges, num_electrons, interaction_fn, xc,
                 num_iterations, alpha=0.5, alpha_decay=0.9, num_mixing_iterations=2,
                 density_mse_converge_tolerance=-1.0, enforce_reflection_symmetry=False, loop="python",
                 external_potential_override=None, want_eigenvalues=False, want_orbitals=False,
                 stop_gradient_step=-1):
        dev = device()
        check_uniform(grids)
        self.dev = dev
        self.grids = as_f64(grids, dev)
        N = self.grids.shape[0]
        if N > _lib.QEX1D_MAX_GRIDS:
            raise ValueError(f"num_grids {N} exceeds {_lib.QEX1D_MAX_GRIDS}")
        locs = [np.asarray(l, dtype=np.float64).reshape(-1) for l in locations]
        zs = [np.asarray(z, dtype=np.float64).reshape(-1) for z in nuclear_charges]
        B = len(locs)
        A = max(1, max(len(l) for l in locs))
        loc_pad = np.zeros((B, A))
        z_pad = np.zeros((B, A))
        for b in range(B):
            loc_pad[b, :len(locs[b])] = locs[b]
            z_pad[b, :len(zs[b])] = zs[b]
        self.locations, self.nuclear_charges = locs, zs
        self.B, self.N, self.T = B, N, int(num_iterations)
        ne = np.broadcast_to(np.asarray(num_electrons, dtype=np.int64).reshape(-1), (B,)).copy()
        if np.any(ne < 1) or np.any(ne > _lib.QEX1D_MAX_ELECTRONS):
            raise ValueError(f"num_electrons must be in [1, {_lib.QEX1D_MAX_ELECTRONS}]")
        self.num_electrons = ne
        self.ne_t = as_i32(ne, dev)
        self.xc = xcmod.resolve(xc)
        self.table = interaction_table(self.grids, interaction_fn)
        if external_potential_override is not None:
            v = as_f64(external_potential_override, dev)
            self.vext = v[None].contiguous() if v.dim() == 1 else v
        else:
            self.vext = external_potential(self.grids, loc_pad, z_pad, interaction_fn).reshape(B, N).contiguous()
        p = _lib.Qex1dProblem()
        p.batch, p.num_grids, p.dx = B, N, get_dx(self.grids)
        p.num_iterations = self.T
        p.max_electrons = int(ne.max())
        p.alpha, p.alpha_decay = float(alpha), float(alpha_decay)
        p.num_mixing_iterations = int(num_mixing_iterations)
        p.density_mse_converge_tolerance = float(density_mse_converge_tolerance)
        p.loop = _lib.QEX1D_LOOP_PYTHON if loop == "python" else _lib.QEX1D_LOOP_SCAN
        p.reflect = 1 if enforce_reflection_symmetry else 0
        if stop_gradient_step < -1:
            raise ValueError("stop_gradient_step must be >= -1")
        p.stop_gradient_iterations = int(stop_gradient_step) + 1  # reverse pass only
        self.xc.configure(p)
        self.problem = p
        self.K = p.max_electrons // 2 + 1
        self.want_eigenvalues = want_eigenvalues
        self.want_orbitals = want_orbitals
        if enforce_reflection_symmetry:
            if loop == "python":
                win = np.array([reflection_window(l, self.grids) for l in locs], dtype=np.int32)
            else:
                win = np.tile(np.array([0, N - 1], dtype=np.int32), (B, 1))
            self.win = as_i32(win, dev)
        else:
            self.win = None
        self.xc_params = self.xc.device_params(dev)
        lib = _lib.load()
        nbytes = lib.qex1d_workspace_bytes(ctypes.byref(p))
        if nbytes == 0:
            _lib.check(-1, "qex1d_workspace_bytes")
        self.ws = workspace(nbytes, dev)

    def allocate_outputs(self) -> BatchResult:
        B, T, N, dev = self.B, self.T, self.N, self.dev
        f = lambda *s: torch.empty(s, dtype=torch.float64, device=dev)
        exc = f(B, T, N) if self.xc.encoding == "local" else f(B, T)
        return BatchResult(
            density=f(B, T, N), total_energy=f(B, T), hartree_potential=f(B, T, N), xc_potential=f(B, T, N),
            xc_energy_density=exc, gap=f(B, T), converged=torch.empty((B, T), dtype=torch.uint8, device=dev),
            external_potential=self.vext, eigenvalues=f(B, T, self.K) if self.want_eigenvalues else None,
            stats=torch.zeros(2, dtype=torch.int64, device=dev),
            orbitals=f(B, T, self.K, N) if self.want_orbitals else None)

    def run(self, initial_density, out: BatchResult | None = None) -> BatchResult:
        n0 = as_f64(initial_density, self.dev)
        if n0.dim() == 1:
            n0 = n0[None]
        if tuple(n0.shape) != (self.B, self.N):
            raise ValueError(f"initial_density must have shape {(self.B, self.N)}, got {tuple(n0.shape)}")
        n0 = n0.contiguous()
        out = out or self.allocate_outputs()
        external = xcmod.is_external(self.xc)
        exch = torch.zeros((3, self.B, self.N), dtype=torch.float64, device=self.dev) if external else None
        ins = _lib.Qex1dInputs(_ptr(self.table), _ptr(self.vext), _ptr(n0), _ptr(self.ne_t), _ptr(self.win),
                               _ptr(exch if external else self.xc_params))
        outs = _lib.Qex1dOutputs(_ptr(out.density), _ptr(out.total_energy), _ptr(out.hartree_potential),
                                 _ptr(out.xc_potential), _ptr(out.xc_energy_density), _ptr(out.gap),
                                 _ptr(out.converged), _ptr(out.eigenvalues), _ptr(out.stats), _ptr(out.orbitals))
        self._keep = (n0, ins, outs, exch)  # keep alive until the stream has consumed them
        self._last_n0 = n0
        lib = _lib.load()
        if not external:
            _lib.check(lib.qex1d_kohn_sham(ctypes.byref(self.problem), ctypes.byref(ins), ctypes.byref(outs),
                                           _ptr(self.ws), self.ws.numel(), _stream()), "qex1d_kohn_sham")
            return out
        # an opaque functional: one device pass per iteration, the callable evaluated between passes
        # on the densities the device hands over (qex1d_external_xc_densities)
        d_in = torch.empty((self.B, self.N), dtype=torch.float64, device=self.dev)
        d_out = torch.empty_like(d_in)
        dx = float(self.problem.dx)
        for k in range(self.T + 1):
            _lib.check(lib.qex1d_external_xc_densities(ctypes.byref(self.problem), ctypes.byref(ins),
                                                       ctypes.byref(outs), k, _ptr(self.ws), self.ws.numel(),
                                                       _ptr(d_in), _ptr(d_out), _stream()),
                       "qex1d_external_xc_densities")
            exch.copy_(self.xc.exchange(d_in, d_out, dx))
            _lib.check(lib.qex1d_kohn_sham_step(ctypes.byref(self.problem), ctypes.byref(ins), ctypes.byref(outs), k,
                                                _ptr(self.ws), self.ws.numel(), _stream()), "qex1d_kohn_sham_step")
        return out

    def run_stepwise(self, initial_density, out: BatchResult | None = None) -> BatchResult:
        """``run`` as T + 1 calls of ``qex1d_kohn_sham_step`` (built-in functionals): the same
        results as the one-call loop, bit for bit."""
        n0 = as_f64(initial_density, self.dev)
        if n0.dim() == 1:
            n0 = n0[None]
        n0 = n0.contiguous()
        out = out or self.allocate_outputs()
        ins = _lib.Qex1dInputs(_ptr(self.table), _ptr(self.vext), _ptr(n0), _ptr(self.ne_t), _ptr(self.win),
                               _ptr(self.xc_params))
        outs = _lib.Qex1dOutputs(_ptr(out.density), _ptr(out.total_energy), _ptr(out.hartree_potential),
                                 _ptr(out.xc_potential), _ptr(out.xc_energy_density), _ptr(out.gap),
                                 _ptr(out.converged), _ptr(out.eigenvalues), _ptr(out.stats), _ptr(out.orbitals))
        self._keep = (n0, ins, outs)
        lib = _lib.load()
        for k in range(self.T + 1):
            _lib.check(lib.qex1d_kohn_sham_step(ctypes.byref(self.problem), ctypes.byref(ins), ctypes.byref(outs), k,
                                                _ptr(self.ws), self.ws.numel(), _stream()), "qex1d_kohn_sham_step")
        return out


class BatchRunner:
    """Batched solver for a stream of systems that share the grid, interaction, XC and loop settings
    (the serving path).  The interaction table, XC parameters and a workspace sized for
    ``max_batch`` systems are built once; each ``run`` places b <= max_batch systems in the first
    b slots and enqueues one ``qex1d_kohn_sham`` on the current stream.  A system's result does not
    depend on the other systems in its batch (fixed-order reductions per system)."""

    def __init__(self, grids, interaction_fn, xc, num_iterations, max_batch, max_electrons=8, alpha=0.5,
                 alpha_decay=0.9, num_mixing_iterations=2, density_mse_converge_tolerance=-1.0, loop="scan"):
        dev = device()
        check_uniform(grids)
        self.dev = dev
        self.grids = as_f64(grids, dev)
        self.N = self.grids.shape[0]
        self.T = int(num_iterations)
        self.max_batch = int(max_batch)
        self.interaction_fn = interaction_fn
        kind = getattr(interaction_fn, "kind", None)
        if kind not in (_lib.QEX_INTERACTION_EXPONENTIAL, _lib.QEX_INTERACTION_SOFT):
            raise ValueError("the serving path needs a built-in interaction (exponential or soft Coulomb)")
        self.xc = xcmod.resolve(xc)
        self.table = interaction_table(self.grids, interaction_fn)
        p = _lib.Qex1dProblem()
        p.batch, p.num_grids, p.dx = self.max_batch, self.N, get_dx(self.grids)
        p.num_iterations = self.T
        p.max_electrons = int(max_electrons)
        p.alpha, p.alpha_decay = float(alpha), float(alpha_decay)
        p.num_mixing_iterations = int(num_mixing_iterations)
        p.density_mse_converge_tolerance = float(density_mse_converge_tolerance)
        p.loop = _lib.QEX1D_LOOP_PYTHON if loop == "python" else _lib.QEX1D_LOOP_SCAN
        p.reflect = 0
        self.xc.configure(p)
        self.problem = p
        self.xc_params = self.xc.device_params(dev)
        lib = _lib.load()
        nbytes = lib.qex1d_workspace_bytes(ctypes.byref(p))
        if nbytes == 0:
            _lib.check(-1, "qex1d_workspace_bytes")
        self.ws = workspace(nbytes, dev)

    def run(self, locations, nuclear_charges, num_electrons, initial_density) -> BatchResult:
        """locations / nuclear_charges: [b, A] (zero charges pad), num_electrons [b], initial_density [b, N]."""
        loc = np.atleast_2d(np.asarray(locations, dtype=np.float64))
        z = np.atleast_2d(np.asarray(nuclear_charges, dtype=np.float64))
        ne = np.asarray(num_electrons, dtype=np.int32).reshape(-1)
        b = ne.shape[0]
        if not 1 <= b <= self.max_batch or loc.shape != z.shape or loc.shape[0] != b:
            raise ValueError("bad batch shapes")
        if np.any(ne < 1) or np.any(ne > self.problem.max_electrons):
            raise ValueError(f"num_electrons must be in [1, {self.problem.max_electrons}]")
        dev, N, T = self.dev, self.N, self.T
        loc_d, z_d, ne_d = upload(loc, np.float64, dev), upload(z, np.float64, dev), upload(ne, np.int32, dev)
        if isinstance(initial_density, torch.Tensor):
            n0 = as_f64(initial_density, dev).reshape(b, N).contiguous()
        else:
            n0 = upload(np.asarray(initial_density, dtype=np.float64).reshape(b, N), np.float64, dev)
        lib = _lib.load()
        vext = torch.empty((b, N), dtype=torch.float64, device=dev)
        p0, p1 = self.interaction_fn.params
        _lib.check(lib.qex1d_external_potential(b, N, _ptr(self.grids), loc.shape[1], _ptr(loc_d), _ptr(z_d),
                                                self.interaction_fn.kind, p0, p1, _ptr(vext), _stream()),
                   "qex1d_external_potential")
        p = _lib.Qex1dProblem.from_buffer_copy(self.problem)
        p.batch = b
        f = lambda *s: torch.empty(s, dtype=torch.float64, device=dev)
        out = BatchResult(density=f(b, T, N), total_energy=f(b, T), hartree_potential=f(b, T, N),
                          xc_potential=f(b, T, N),
                          xc_energy_density=f(b, T, N) if self.xc.encoding == "local" else f(b, T), gap=f(b, T),
                          converged=torch.empty((b, T), dtype=torch.uint8, device=dev), external_potential=vext,
                          stats=torch.zeros(2, dtype=torch.int64, device=dev))
        ins = _lib.Qex1dInputs(_ptr(self.table), _ptr(vext), _ptr(n0), _ptr(ne_d), None, _ptr(self.xc_params))
        outs = _lib.Qex1dOutputs(_ptr(out.density), _ptr(out.total_energy), _ptr(out.hartree_potential),
                                 _ptr(out.xc_potential), _ptr(out.xc_energy_density), _ptr(out.gap),
                                 _ptr(out.converged), None, _ptr(out.stats))
        _lib.check(lib.qex1d_kohn_sham(ctypes.byref(p), ctypes.byref(ins), ctypes.byref(outs), _ptr(self.ws),
                                       self.ws.numel(), _stream()), "qex1d_kohn_sham")
        out._keep = (loc_d, z_d, ne_d, n0)  # inputs stay alive until the stream has consumed them
        return out


def kohn_sham_vjp(problem: Problem1D, result: BatchResult, d_total_energy=None, d_density=None) -> torch.Tensor:
    """Gradient of sum(d_total_energy * total_energy) + sum(d_density * density) with respect to the
    packed parameters of the problem's LocalMLP or GlobalMLP functional (summed over the batch), by the device
    reverse pass (``qex1d_kohn_sham_vjp``).  ``result`` must come from ``problem.run`` with
    ``want_orbitals=True`` on the same initial densities."""
    if result.orbitals is None:
        raise ValueError("the forward result has no orbitals: build the problem with want_orbitals=True")
    dev = problem.dev
    B, T, N = problem.B, problem.T, problem.N
    dE = None if d_total_energy is None else as_f64(d_total_energy, dev).reshape(B, T).contiguous()
    dn = None if d_density is None else as_f64(d_density, dev).reshape(B, T, N).contiguous()
    lib = _lib.load()
    nbytes = lib.qex1d_kohn_sham_vjp_workspace_bytes(ctypes.byref(problem.problem))
    if nbytes == 0:
        _lib.check_vjp(-4, "qex1d_kohn_sham_vjp_workspace_bytes")
    ws = workspace(nbytes, dev)
    grad = torch.empty(problem.xc_params.numel(), dtype=torch.float64, device=dev)
    ins = _lib.Qex1dInputs(_ptr(problem.table), _ptr(problem.vext), _ptr(problem._last_n0), _ptr(problem.ne_t),
                           _ptr(problem.win), _ptr(problem.xc_params))
    outs = _lib.Qex1dOutputs(_ptr(result.density), _ptr(result.total_energy), _ptr(result.hartree_potential),
                             _ptr(result.xc_potential), _ptr(result.xc_energy_density), _ptr(result.gap),
                             _ptr(result.converged), None, None, _ptr(result.orbitals))
    _lib.check_vjp(lib.qex1d_kohn_sham_vjp(ctypes.byref(problem.problem), ctypes.byref(ins), ctypes.byref(outs),
                                           _ptr(dE), _ptr(dn), _ptr(grad), _ptr(ws), ws.numel(), _stream()),
                   "qex1d_kohn_sham_vjp")
    problem._keep_vjp = (dE, dn, ws, ins, outs)
    return grad


class SlotRunner:
    """Continuous batching of 1D solves (``qex1d_slots_step``): a table of ``slots`` systems on one
    grid, each running its own SCF iteration, so a request is admitted into the running batch as
    soon as a slot frees up instead of waiting for the next batched launch.

    ``admit(requests)`` places up to ``free_slots()`` requests (locations, charges, N_e, initial
    density) into free slots; ``step()`` enqueues ``iterations_per_step`` iterations for every busy
    slot on the current stream (a slot whose solve ends inside the group idles for the rest of it),
    then a device-to-host copy of the slot states and final energies into pinned memory with an
    event; ``poll()`` returns the requests whose solve completed in the steps whose events have
    fired, as (tag, total energy, converged).  A request's result equals its solo
    ``qex1d_kohn_sham`` result bit for bit.  Nothing here synchronises the host with the device."""

    def __init__(self, grids, interaction_fn, xc, num_iterations, slots=64, max_electrons=8, alpha=0.5,
                 alpha_decay=0.9, num_mixing_iterations=2, density_mse_converge_tolerance=-1.0, loop="scan",
                 depth=3, max_nuclei=4, iterations_per_step=1):
        dev = device()
        check_uniform(grids)
        self.dev = dev
        self.grids = as_f64(grids, dev)
        self.N = self.grids.shape[0]
        self.T = int(num_iterations)
        self.S = int(slots)
        self.A = int(max_nuclei)
        kind = getattr(interaction_fn, "kind", None)
        if kind not in (_lib.QEX_INTERACTION_EXPONENTIAL, _lib.QEX_INTERACTION_SOFT):
            raise ValueError("the serving path needs a built-in interaction (exponential or soft Coulomb)")
        self.interaction_fn = interaction_fn
        self.xc = xcmod.resolve(xc)
        self.table = interaction_table(self.grids, interaction_fn)
        p = _lib.Qex1dProblem()
        p.batch, p.num_grids, p.dx = self.S, self.N, get_dx(self.grids)
        p.num_iterations = self.T
        p.max_electrons = int(max_electrons)
        p.alpha, p.alpha_decay = float(alpha), float(alpha_decay)
        p.num_mixing_iterations = int(num_mixing_iterations)
        p.density_mse_converge_tolerance = float(density_mse_converge_tolerance)
        p.loop = _lib.QEX1D_LOOP_PYTHON if loop == "python" else _lib.QEX1D_LOOP_SCAN
        p.reflect = 0
        self.xc.configure(p)
        self.problem = p
        self.xc_params = self.xc.device_params(dev)
        lib = _lib.load()
        nbytes = lib.qex1d_workspace_bytes(ctypes.byref(p))
        if nbytes == 0:
            _lib.check(-1, "qex1d_workspace_bytes")
        self.ws = workspace(nbytes, dev)
        S, N, T = self.S, self.N, self.T
        f = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        self.vext, self.n0 = f(S, N), f(S, N)
        self.ne = torch.ones(S, dtype=torch.int32, device=dev)
        self.kslot = torch.full((S,), -1, dtype=torch.int32, device=dev)
        self.admit_flags = torch.zeros(S, dtype=torch.int32, device=dev)
        self.out = BatchResult(density=f(S, T, N), total_energy=f(S, T), hartree_potential=f(S, T, N),
                               xc_potential=f(S, T, N),
                               xc_energy_density=f(S, T, N) if self.xc.encoding == "local" else f(S, T),
                               gap=f(S, T), converged=torch.zeros((S, T), dtype=torch.uint8, device=dev),
                               external_potential=self.vext, stats=torch.zeros(2, dtype=torch.int64, device=dev))
        self.ins = _lib.Qex1dInputs(_ptr(self.table), _ptr(self.vext), _ptr(self.n0), _ptr(self.ne), None,
                                    _ptr(self.xc_params))
        o = self.out
        self.outs = _lib.Qex1dOutputs(_ptr(o.density), _ptr(o.total_energy), _ptr(o.hartree_potential),
                                      _ptr(o.xc_potential), _ptr(o.xc_energy_density), _ptr(o.gap), _ptr(o.converged),
                                      None, _ptr(o.stats))
        self.free = list(range(S - 1, -1, -1))
        self.busy = {}  # slot -> (tag, index of the slot iteration call that starts it)
        self.iters_per_step = max(1, int(iterations_per_step))
        self.calls = 0  # slot iteration calls enqueued so far
        self.depth = int(depth)
        self.ring = [(torch.empty(S, dtype=torch.int32).pin_memory(), torch.empty(S, dtype=torch.float64).pin_memory(),
                      torch.empty(S, dtype=torch.uint8).pin_memory(), torch.cuda.Event()) for _ in range(self.depth)]
        self.inflight = []  # ring indices, oldest first
        self.next_ring = 0
        self.steps = 0

    def free_slots(self) -> int:
        return len(self.free)

    def busy_slots(self) -> int:
        return len(self.busy)

    def can_step(self) -> bool:
        return len(self.inflight) < self.depth

    def admit(self, requests):
        """requests: list of (tag, locations, charges, num_electrons, initial_density); at most
        ``free_slots()`` of them.  Enqueues the slots' inputs on the current stream."""
        if not requests:
            return []
        if len(requests) > len(self.free):
            raise ValueError("more requests than free slots")
        a = len(requests)
        loc = np.zeros((a, self.A))
        z = np.zeros((a, self.A))
        ne = np.zeros(a, dtype=np.int32)
        n0 = np.empty((a, self.N))
        slots = []
        for i, (tag, l, c, e, d) in enumerate(requests):
            l, c = np.asarray(l, dtype=np.float64).reshape(-1), np.asarray(c, dtype=np.float64).reshape(-1)
            if len(l) > self.A or not 1 <= int(e) <= self.problem.max_electrons:
                raise ValueError("request exceeds the runner's nuclei or electron limits")
            loc[i, :len(l)], z[i, :len(c)], ne[i] = l, c, int(e)
            n0[i] = np.asarray(d, dtype=np.float64).reshape(self.N)
            s = self.free.pop()
            self.busy[s] = (tag, self.calls)
            slots.append(s)
        dev = self.dev
        idx = upload(np.asarray(slots, dtype=np.int64), np.int64, dev)
        loc_d, z_d = upload(loc, np.float64, dev), upload(z, np.float64, dev)
        tmp = torch.empty((a, self.N), dtype=torch.float64, device=dev)
        p0, p1 = self.interaction_fn.params
        _lib.check(_lib.load().qex1d_external_potential(a, self.N, _ptr(self.grids), self.A, _ptr(loc_d), _ptr(z_d),
                                                        self.interaction_fn.kind, p0, p1, _ptr(tmp), _stream()),
                   "qex1d_external_potential")
        self.vext.index_copy_(0, idx, tmp)
        self.n0.index_copy_(0, idx, upload(n0, np.float64, dev))
        self.ne.index_copy_(0, idx, upload(ne, np.int32, dev))
        self.admit_flags.index_fill_(0, idx, 1)
        self._keep_admit = (idx, loc_d, z_d, tmp)
        return slots

    def step(self):
        """``iterations_per_step`` iterations for every busy slot, then the read-back of the slot states
        (non-blocking)."""
        if not self.can_step():
            raise RuntimeError("too many steps in flight: poll() first")
        lib = _lib.load()
        for _ in range(self.iters_per_step):
            _lib.check(lib.qex1d_slots_step(ctypes.byref(self.problem), ctypes.byref(self.ins), ctypes.byref(self.outs),
                                            _ptr(self.kslot), _ptr(self.admit_flags), _ptr(self.ws), self.ws.numel(),
                                            _stream()), "qex1d_slots_step")
            if _ == 0:
                self.admit_flags.zero_()  # admissions take effect in the step's first iteration
            self.calls += 1
        r = self.next_ring
        self.next_ring = (r + 1) % self.depth
        hk, he, hc, ev = self.ring[r]
        hk.copy_(self.kslot, non_blocking=True)
        he.copy_(self.out.total_energy[:, self.T - 1], non_blocking=True)
        hc.copy_(self.out.converged[:, self.T - 1], non_blocking=True)
        ev.record()
        self.inflight.append((r, self.calls))
        self.steps += 1

    def poll(self, wait=False):
        """[(tag, energy, converged)] of the requests completed by the steps that have finished."""
        done = []
        while self.inflight:
            r, j = self.inflight[0]
            hk, he, hc, ev = self.ring[r]
            if wait:
                ev.synchronize()
            elif not ev.query():
                break
            self.inflight.pop(0)
            # a request first processed by call a finishes with call a + T (T iterations, then its
            # energy-only pass); the polled step ended with call j - 1
            finished = [s for s, (_, a) in self.busy.items() if a + self.T < j]
            ks = hk.numpy()
            for s in finished:
                if ks[s] != -2:
                    raise _lib.QexError(f"slot {s}: expected a finished solve after step {j}, state {int(ks[s])}")
                tag, _ = self.busy.pop(s)
                done.append((tag, float(he[s]), bool(hc[s])))
                self.free.append(s)
        return done
