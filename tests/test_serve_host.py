"""Host logic of the serving loop (qex_amd.serve) with stand-in solvers: every request completes
once, 1D requests are batched up to max_batch, arrival order is respected, latencies are
non-negative.  Device streams and events are replaced by host stand-ins."""

import contextlib

import numpy as np
import pytest
import torch

from qex_amd import serve


class _Event:
    def record(self, stream=None):
        pass

    def query(self):
        return True


@pytest.fixture()
def fake_cuda(monkeypatch):
    monkeypatch.setattr(torch.cuda, "Stream", lambda *a, **k: object())
    monkeypatch.setattr(torch.cuda, "Event", lambda *a, **k: _Event())
    monkeypatch.setattr(torch.cuda, "stream", lambda s: contextlib.nullcontext())
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)


class _Out:
    def __init__(self, e):
        self.total_energy = torch.tensor(e, dtype=torch.float64)[:, None].repeat(1, 3)
        self.converged = torch.ones((len(e), 3), dtype=torch.uint8)


class _Runner:
    max_batch, T, N = 8, 3, 5

    def __init__(self):
        self.batches = []

    def run(self, loc, z, ne, n0):
        self.batches.append(len(ne))
        return _Out([float(np.sum(zz)) for zz in z])


class _SCF:
    def __init__(self, mol):
        self.mol = mol
        self.converged = True

    def kernel_async(self):
        return self

    def collect(self):
        return -float(self.mol.natm)

    def kernel(self):
        return self.collect()


def _batch_async(mfs):
    _batch_async.sizes.append(len(mfs))
    for mf in mfs:
        mf.kernel_async()
    return mfs


_batch_async.sizes = []


class _Mol:
    def __init__(self, natm):
        self.natm, self.nao = natm, 5 * natm


def _requests(n, seed, rate):
    rng = np.random.default_rng(seed)
    t = serve.poisson_arrivals(rate, n, seed)
    out = []
    for i in range(n):
        if rng.random() < 0.7:
            z = rng.integers(1, 4, rng.integers(1, 4)).astype(float)
            out.append(serve.Request("1d", float(t[i]), (np.zeros(len(z)), z, int(z.sum()), np.zeros(5)), rid=i))
        else:
            out.append(serve.Request("3d", float(t[i]), _Mol(int(rng.integers(2, 5))), rid=i))
    return out


def test_every_request_completes_once_with_its_own_result(fake_cuda):
    runner = _Runner()
    _batch_async.sizes.clear()
    srv = serve.Server(runner, _SCF, max_batch=8, streams_3d=2, batch_3d=4, batch_async_3d=_batch_async)
    reqs = _requests(200, 1, rate=5000.0)
    st = srv.run(reqs)
    assert sorted(c.rid for c in st.completions) == list(range(200))
    by = {r.rid: r for r in reqs}
    for c in st.completions:
        r = by[c.rid]
        expect = float(np.sum(r.payload[1])) if r.kind == "1d" else -float(r.payload.natm)
        assert c.energy == expect and c.kind == r.kind
        assert c.latency >= 0.0
    assert max(runner.batches) <= 8
    assert sum(runner.batches) == sum(1 for r in reqs if r.kind == "1d")
    s = st.summary()
    assert s["requests"] == 200 and s["p99_ms"] >= s["p50_ms"]
    # 3D requests are grouped (up to batch_3d) and each is solved exactly once
    assert max(st.batch_sizes_3d) <= 4 and sum(st.batch_sizes_3d) == sum(1 for r in reqs if r.kind == "3d")
    assert all(1 < b <= 4 for b in _batch_async.sizes)


def test_fast_arrivals_are_batched_and_slow_ones_are_not_delayed(fake_cuda):
    runner = _Runner()
    srv = serve.Server(runner, _SCF, max_batch=8, batch_async_3d=_batch_async)
    burst = [serve.Request("1d", 0.0, (np.zeros(1), np.ones(1), 1, np.zeros(5)), rid=i) for i in range(32)]
    st = srv.run(burst)
    assert len(st.completions) == 32
    assert max(runner.batches) == 8  # a burst fills whole batches


def test_3d_burst_is_one_batch(fake_cuda):
    _batch_async.sizes.clear()
    srv = serve.Server(_Runner(), _SCF, batch_3d=16, batch_async_3d=_batch_async)
    burst = [serve.Request("3d", 0.0, _Mol(2 + i % 3), rid=i) for i in range(40)]
    st = srv.run(burst)
    assert sorted(c.rid for c in st.completions) == list(range(40))
    assert st.batch_sizes_3d[0] == 16 and sum(st.batch_sizes_3d) == 40


def test_one_at_a_time_path(fake_cuda):
    reqs = _requests(20, 2, rate=1000.0)
    st = serve.run_one_at_a_time(reqs, _Runner(), _SCF)
    assert [c.rid for c in st.completions] == [r.rid for r in reqs]


class _Slots:
    """Host stand-in of engine.SlotRunner: a request admitted before step a finishes at step a + T."""

    T, N = 3, 5

    def __init__(self, S=4, depth=2):
        self.S, self.depth = S, depth
        self.free = list(range(S))
        self.busy = {}
        self.steps = 0
        self.inflight = []
        self.max_busy = 0
        self.admitted_at = {}

    def free_slots(self):
        return len(self.free)

    def busy_slots(self):
        return len(self.busy)

    def can_step(self):
        return len(self.inflight) < self.depth

    def admit(self, reqs):
        assert len(reqs) <= len(self.free)
        for r in reqs:
            s = self.free.pop()
            self.busy[s] = (r[0], self.steps)
            self.admitted_at[r[0].rid] = self.steps

    def step(self):
        assert self.can_step()
        self.inflight.append(self.steps)
        self.steps += 1
        self.max_busy = max(self.max_busy, len(self.busy))

    def poll(self):
        done = []
        while self.inflight:
            j = self.inflight.pop(0)
            for s, (r, a) in list(self.busy.items()):
                if a + self.T == j:
                    del self.busy[s]
                    self.free.append(s)
                    done.append((r, float(np.sum(r.payload[1])), True))
        return done


def test_continuous_batching_admits_into_free_slots_between_steps(fake_cuda):
    """With a slot table, a request that arrives while others run joins at the next step once a slot
    is free (no waiting for a whole batch to drain); every request completes once with its result."""
    slots = _Slots(S=4)
    srv = serve.Server(slots, _SCF, batch_async_3d=_batch_async)
    burst = [serve.Request("1d", 0.0, (np.zeros(1), np.full(1, float(i % 3 + 1)), 1, np.zeros(5)), rid=i)
             for i in range(10)]
    st = srv.run(burst)
    assert sorted(c.rid for c in st.completions) == list(range(10))
    for c in st.completions:
        assert c.energy == float(burst[c.rid].payload[1][0])
    assert slots.max_busy == 4
    # requests 4.. are admitted as earlier ones finish, at different steps (not all at once)
    starts = sorted(slots.admitted_at[i] for i in range(10))
    assert starts[:4] == [0, 0, 0, 0] and len(set(starts[4:])) >= 2 and starts[4] <= slots.T + 1
