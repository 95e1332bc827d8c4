"""Serving loop for a mixed stream of 1D and 3D Kohn-Sham requests on one GPU.

Requests arrive over time.  Each 1D request is a system on the shared grid with the shared XC
functional and loop settings.  Each 3D request is a molecule (``qex_amd.td.gto.Mole``) solved by
RKS-LDA.

The dispatcher runs in one host thread and never blocks on the device:
  * 1D, continuous batching (``runner_1d`` an ``engine.SlotRunner``): a table of slots on stream A,
    each running its own SCF iteration (``qex1d_slots_step``); a waiting request is admitted into a
    free slot at the next step, so requests join and leave the running batch between iterations,
    and their work overlaps on the device with the solves already running;
  * 1D, batch per launch (``runner_1d`` an ``engine.BatchRunner``): every released request still
    waiting is packed into the next batched launch (up to ``max_batch``), one ``qex1d_kohn_sham``
    per batch, on stream A (at most ``depth_1d`` batches in flight);
  * 3D: the 3D requests waiting (up to ``batch_3d``) are enqueued together on one of
    ``streams_3d`` streams (the least loaded): integrals, grid and AO values per molecule, then
    one batched SCF (``qex3d_scf_batch``: every cycle stage is one launch over the molecules, so the
    one-workgroup Fock/eigensolver/energy kernels of the batch run side by side), without host
    synchronisation; at most ``depth_3d`` batches in flight per stream;
  * completions are detected with events (``query``), then results are read back.
Admission holds a request back while the device memory its solve needs, by the estimate
n_grid * n_ao * 8 + n_ao^4 / 4 * 8 + 16 * n_ao^2 * 8 bytes (3D) or T * N * 8 * 6 bytes (1D),
would exceed ``memory_budget`` bytes in flight.

Latency is completion time minus arrival time on the host clock.
"""

from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class Request:
    kind: str                 # "1d" or "3d"
    arrival: float            # seconds after the start of the run
    payload: object           # 1D: (locations, charges, num_electrons, initial_density); 3D: Mole
    rid: int = 0


@dataclass
class Completion:
    rid: int
    kind: str
    arrival: float
    done: float
    energy: float
    converged: bool

    @property
    def latency(self):
        return self.done - self.arrival


@dataclass
class _Inflight:
    event: torch.cuda.Event
    reqs: list
    handle: object
    nbytes: int


@dataclass
class ServeStats:
    completions: list = field(default_factory=list)
    wall: float = 0.0
    batches_1d: int = 0
    batch_sizes: list = field(default_factory=list)
    batch_sizes_3d: list = field(default_factory=list)
    work_1d: list = field(default_factory=list)  # per completed 1D batch: (electron counts, iterations)
    work_3d: list = field(default_factory=list)  # per completed 3D solve: (n_ao, grid points, cycles, n_occ)
    host_s: dict = field(default_factory=lambda: {"dispatch_1d": 0.0, "dispatch_3d": 0.0, "complete_1d": 0.0,
                                                  "complete_3d": 0.0})

    def summary(self):
        lat = np.array([c.latency for c in self.completions]) * 1e3
        out = {"requests": len(self.completions), "wall_s": self.wall,
               "throughput_per_s": len(self.completions) / self.wall if self.wall > 0 else 0.0,
               "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
               "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
               "batches_1d": self.batches_1d,
               "mean_batch_1d": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
               "batches_3d": len(self.batch_sizes_3d),
               "mean_batch_3d": float(np.mean(self.batch_sizes_3d)) if self.batch_sizes_3d else 0.0,
               "host_busy_s": {k: round(v, 4) for k, v in self.host_s.items()}}
        for kind in ("1d", "3d"):
            lk = np.array([c.latency for c in self.completions if c.kind == kind]) * 1e3
            if len(lk):
                out[f"p50_ms_{kind}"] = float(np.percentile(lk, 50))
                out[f"p99_ms_{kind}"] = float(np.percentile(lk, 99))
        return out


def _bytes_3d(mol, n_grid):
    n = mol.nao
    return int(n_grid * n * 8 + (n * (n + 1) // 2) ** 2 * 8 + 16 * n * n * 8)


class Server:
    """runner_1d: an ``engine.BatchRunner``; solver_3d(mol) -> SCF object with ``kernel_async`` /
    ``collect`` (e.g. ``lambda m: qex_amd.td.scf.RKS(m)`` with its settings); batch_async_3d(list of
    such objects) enqueues them as one batch (default ``qex_amd.td.scf.kernel_batch_async``)."""

    def __init__(self, runner_1d, solver_3d, max_batch=64, depth_1d=2, depth_3d=2, streams_3d=3,
                 memory_budget=64 << 30, grid_points_per_atom=5500, batch_3d=16, batch_async_3d=None):
        self.runner = runner_1d
        self.slots = runner_1d if hasattr(runner_1d, "step") else None  # continuous batching
        self.solver_3d = solver_3d
        self.max_batch = (runner_1d.S if self.slots is not None else min(max_batch, runner_1d.max_batch)) \
            if runner_1d is not None else max_batch
        self.depth_1d, self.depth_3d = depth_1d, depth_3d
        self.batch_3d = max(1, int(batch_3d))
        if batch_async_3d is None:
            from .td.scf import kernel_batch_async as batch_async_3d
        self.batch_async_3d = batch_async_3d
        self.memory_budget = memory_budget
        self.grid_points_per_atom = grid_points_per_atom
        self.stream_1d = torch.cuda.Stream()
        self.streams_3d = [torch.cuda.Stream() for _ in range(max(1, streams_3d))]

    def _dispatch_1d(self, reqs):
        loc = np.zeros((len(reqs), max(len(r.payload[0]) for r in reqs)))
        z = np.zeros_like(loc)
        ne = np.zeros(len(reqs), dtype=np.int32)
        n0 = np.stack([r.payload[3] for r in reqs])
        for i, r in enumerate(reqs):
            loc[i, :len(r.payload[0])] = r.payload[0]
            z[i, :len(r.payload[1])] = r.payload[1]
            ne[i] = r.payload[2]
        with torch.cuda.stream(self.stream_1d):
            out = self.runner.run(loc, z, ne, n0)
            ev = torch.cuda.Event()
            ev.record(self.stream_1d)
        return out, ev

    def _dispatch_3d(self, reqs, stream):
        with torch.cuda.stream(stream):
            mfs = [self.solver_3d(r.payload) for r in reqs]
            if len(mfs) == 1:
                mfs[0].kernel_async()
            else:
                self.batch_async_3d(mfs)
            ev = torch.cuda.Event()
            ev.record(stream)
        return mfs, ev

    def run(self, requests, poll_sleep=50e-6) -> ServeStats:
        reqs = sorted(requests, key=lambda r: r.arrival)
        pend_1d, pend_3d = deque(), deque()
        inflight_1d = deque()
        inflight_3d = [deque() for _ in self.streams_3d]
        stats = ServeStats()
        i = 0
        used = 0
        t0 = time.perf_counter()
        total = len(reqs)
        while len(stats.completions) < total:
            now = time.perf_counter() - t0
            while i < total and reqs[i].arrival <= now:
                (pend_1d if reqs[i].kind == "1d" else pend_3d).append(reqs[i])
                i += 1
            progressed = False
            # completions (in order per stream)
            for q, kind in [(inflight_1d, "1d")] + [(q3, "3d") for q3 in inflight_3d]:
                while q and q[0].event.query():
                    th = time.perf_counter()
                    f = q.popleft()
                    done = time.perf_counter() - t0
                    used -= f.nbytes
                    if kind == "1d":
                        e = f.handle.total_energy[:, -1].cpu().numpy()
                        conv = f.handle.converged[:, -1].cpu().numpy()
                        for k, r in enumerate(f.reqs):
                            stats.completions.append(Completion(r.rid, "1d", r.arrival, done, float(e[k]),
                                                                bool(conv[k])))
                        stats.work_1d.append(([int(r.payload[2]) for r in f.reqs], self.runner.T))
                    else:
                        for r, mf in zip(f.reqs, f.handle):
                            e = mf.collect()
                            stats.completions.append(Completion(r.rid, "3d", r.arrival, done, e, mf.converged))
                            grids = getattr(mf, "grids", None)
                            stats.work_3d.append((mf.mol.nao, grids.size if grids is not None else 0,
                                                  getattr(mf, "cycles", 0), getattr(mf.mol, "nelectron", 0) // 2))
                    stats.host_s["complete_" + kind] += time.perf_counter() - th
                    progressed = True
            if self.slots is not None:
                # continuous batching: completions, admission into free slots, one step when busy
                th = time.perf_counter()
                fin = self.slots.poll()
                if fin:
                    done = time.perf_counter() - t0
                    for r, e, conv in fin:
                        stats.completions.append(Completion(r.rid, "1d", r.arrival, done, e, conv))
                        stats.work_1d.append(([int(r.payload[2])], self.slots.T))
                        used -= self.slots.T * self.slots.N * 8 * 6
                    stats.host_s["complete_1d"] += time.perf_counter() - th
                    progressed = True
                th = time.perf_counter()
                take = []
                need = self.slots.T * self.slots.N * 8 * 6
                while pend_1d and len(take) < self.slots.free_slots() and (
                        used + need <= self.memory_budget or not (take or self.slots.busy_slots())):
                    r = pend_1d.popleft()
                    take.append((r, r.payload[0], r.payload[1], r.payload[2], r.payload[3]))
                    used += need
                with torch.cuda.stream(self.stream_1d):
                    if take:
                        self.slots.admit(take)
                        progressed = True
                    if self.slots.busy_slots() and self.slots.can_step():
                        self.slots.step()
                        stats.batches_1d += 1
                        stats.batch_sizes.append(self.slots.busy_slots())
                        progressed = True
                stats.host_s["dispatch_1d"] += time.perf_counter() - th
            # dispatch
            if self.slots is None and pend_1d and len(inflight_1d) < self.depth_1d:
                batch = []
                nb = 0
                while pend_1d and len(batch) < self.max_batch:
                    need = self.runner.T * self.runner.N * 8 * 6
                    if used + nb + need > self.memory_budget and batch:
                        break
                    batch.append(pend_1d.popleft())
                    nb += need
                th = time.perf_counter()
                out, ev = self._dispatch_1d(batch)
                stats.host_s["dispatch_1d"] += time.perf_counter() - th
                used += nb
                inflight_1d.append(_Inflight(ev, batch, out, nb))
                stats.batches_1d += 1
                stats.batch_sizes.append(len(batch))
                progressed = True
            k3 = min(range(len(inflight_3d)), key=lambda q: len(inflight_3d[q]))
            if pend_3d and len(inflight_3d[k3]) < self.depth_3d:
                busy = inflight_1d or any(inflight_3d) or (self.slots is not None and self.slots.busy_slots() > 0)
                batch, nb = [], 0
                while pend_3d and len(batch) < self.batch_3d:
                    r = pend_3d[0]
                    need = _bytes_3d(r.payload, r.payload.natm * self.grid_points_per_atom)
                    if used + nb + need > self.memory_budget and (batch or busy):
                        break
                    batch.append(pend_3d.popleft())
                    nb += need
                if batch:
                    th = time.perf_counter()
                    mfs, ev = self._dispatch_3d(batch, self.streams_3d[k3])
                    stats.host_s["dispatch_3d"] += time.perf_counter() - th
                    used += nb
                    inflight_3d[k3].append(_Inflight(ev, batch, mfs, nb))
                    stats.batch_sizes_3d.append(len(batch))
                    progressed = True
            slots_busy = self.slots is not None and self.slots.busy_slots() > 0
            if not progressed:
                if i < total and not (pend_1d or pend_3d or inflight_1d or any(inflight_3d) or slots_busy):
                    wait = reqs[i].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                else:
                    time.sleep(poll_sleep)
        torch.cuda.synchronize()
        stats.wall = time.perf_counter() - t0
        return stats


def run_one_at_a_time(requests, runner_1d, solver_3d) -> ServeStats:
    """The plain path: every request solved alone (1D batch of one, one 3D solve), one after the
    other, each waited for before the next starts."""
    stats = ServeStats()
    t0 = time.perf_counter()
    for r in requests:
        if r.kind == "1d":
            loc, z, ne, n0 = r.payload
            out = runner_1d.run(np.asarray(loc)[None], np.asarray(z)[None], [ne], np.asarray(n0)[None])
            e = float(out.total_energy[0, -1].item())
            conv = bool(out.converged[0, -1].item())
        else:
            mf = solver_3d(r.payload)
            e = mf.kernel()
            conv = mf.converged
        done = time.perf_counter() - t0
        stats.completions.append(Completion(r.rid, r.kind, 0.0, done, e, conv))
    stats.wall = time.perf_counter() - t0
    return stats


def poisson_arrivals(rate, count, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))

