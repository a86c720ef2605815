"""Serving (continuous batching of independent steady solves).

CPU: the server's admission and completion logic with a stand-in solver
whose residual history and step time are fixed per request -- admission in
arrival order, never before arrival, within the slots and the HBM budget;
the sequential baseline (one slot) runs the requests back to back; a request
stops at the first step whose residual has dropped by the target, or at the
step cap.

GPU: each request's result is the same alone and in the mix -- residual
history and final fields bit for bit, served six at a time (each step on one
stream, contexts created while the others run) or one at a time (the
step's inviscid fluxes on a second stream).
"""
import time
from types import SimpleNamespace

import numpy as np
import pytest

from orion_amd import serving


class _FakeSolver:
    """Residual rms_k = r0 * q**k, one chunk takes ``dt`` seconds."""

    def __init__(self, case, log, q=0.9, dt=0.01):
        self.case, self.q, self.dt, self.k = case, q, dt, 0
        self.log = log
        self.fs = SimpleNamespace(close=lambda: None)
        log.append(("open", case.name, time.perf_counter()))

    def solve(self, n):
        time.sleep(self.dt)
        out = [SimpleNamespace(rms_residual=1e-3 * self.q ** (self.k + i)) for i in range(n)]
        self.k += n
        return out


def _stream(n=6):
    return serving.channel_stream(n, seed=5, rate=200.0, nodes=(2.0e3, 4.0e3))


def test_served_within_slots_budget_and_arrivals():
    reqs = _stream(6)
    log = []
    budget = 2.5 * max(serving.footprint(r.case) for r in reqs)
    srv = serving.Server(slots=3, hbm_budget=budget, chunk=10, drop=1e-2, max_steps=200,
                         make_solver=lambda c, overlap: _FakeSolver(c, log))
    res = srv.serve(reqs)
    assert sorted(r.rid for r in res) == list(range(6))
    assert srv.peak_resident <= 3 and srv.peak_bytes <= budget
    for r in res:
        assert r.start >= r.arrival - 1e-6
        # 0.9**k <= 1e-2 first at k = 44: 45 steps, in chunks of 10
        assert r.converged and r.steps == 50
        assert int(np.nonzero(r.rms <= 1e-2 * r.rms[0])[0][0]) == 44


def test_sequential_baseline_runs_back_to_back():
    reqs = _stream(4)
    srv = serving.Server(slots=1, chunk=25, drop=1e-30, max_steps=60,
                         make_solver=lambda c, overlap: _FakeSolver(c, [], dt=0.002))
    res = sorted(srv.serve(reqs), key=lambda r: r.rid)
    for a, b in zip(res, res[1:]):
        assert b.start >= a.finish - 1e-6
    for r in res:
        assert not r.converged and r.steps == 60
    s = serving.summarize(res)
    assert s["requests"] == 4 and s["converged"] == 0 and s["p99_latency_s"] >= s["p50_latency_s"]


class _FakeBatch:
    """Stand-in batch: request k's residual is r0_k * q_k**s at its own step
    s, which the next batch takes over from this one (``(batch, k)``), so a
    request's history does not depend on the batch it runs in."""

    def __init__(self, cases, states, device, log, dt=0.003):
        self.cases, self.dt = list(cases), dt
        self.steps = [st[0].steps[st[1]] if st is not None else 0 for st in states]
        log.append([c.name for c in self.cases])

    @staticmethod
    def law(case):
        n = case.interior_nodes()
        return 1.0e-3 * (1 + n % 7), 0.80 + 0.01 * (n % 13)

    def run(self, n):
        time.sleep(self.dt)
        out = np.empty((n, len(self.cases)))
        for k, c in enumerate(self.cases):
            r0, q = self.law(c)
            out[:, k] = r0 * q ** (self.steps[k] + np.arange(n))
            self.steps[k] += n
        return out

    def prim(self, k):
        return [np.array([float(self.steps[k])])]

    def close(self):
        pass


def test_batched_histories_carry_across_rebuilds():
    """Requests join and leave the batch at chunk boundaries; each one's
    history is its own law from step 0, it stops at the first chunk end
    past the drop, and the batch never holds more than the slots."""
    reqs = serving.channel_stream(8, seed=7, rate=300.0, nodes=(2.0e3, 6.0e3))
    log = []
    srv = serving.BatchServer(slots=3, chunk=5, drop=1e-4, max_steps=200,
                              make_batch=lambda cs, st, dev: _FakeBatch(cs, st, dev, log))
    res = {r.rid: r for r in srv.serve(reqs)}
    assert sorted(res) == list(range(8))
    assert max(len(b) for b in log) <= 3 and srv.peak_resident <= 3
    assert srv.rebuilds == len(log) and len(log) >= 4
    for r in reqs:
        got = res[r.rid]
        r0, q = _FakeBatch.law(r.case)
        k = int(np.ceil(np.log(1e-4) / np.log(q) - 1e-9))  # first step with q**k <= 1e-4
        assert got.converged and got.steps == 5 * (k // 5 + 1)
        assert np.array_equal(got.rms, r0 * q ** np.arange(got.steps))
        assert got.start >= got.arrival - 1e-6


def test_batched_server_refuses_a_bad_request_and_serves_the_rest():
    """A request the device refuses (its batch cannot be built) comes back
    with the error and no steps; the others are served as usual."""
    reqs = serving.channel_stream(5, seed=9, rate=300.0, nodes=(2.0e3, 6.0e3))
    bad = reqs[2].case.name

    def make(cs, st, dev):
        if any(c.name == bad for c in cs):
            raise RuntimeError("refused " + bad)
        return _FakeBatch(cs, st, dev, [])

    srv = serving.BatchServer(slots=3, chunk=5, drop=1e-4, max_steps=200, make_batch=make)
    res = {r.rid: r for r in srv.serve(reqs)}
    assert sorted(res) == list(range(5))
    assert res[2].error and "refused" in res[2].error and res[2].steps == 0 and not res[2].converged
    for r in reqs:
        if r.rid == 2:
            continue
        r0, q = _FakeBatch.law(r.case)
        got = res[r.rid]
        assert got.error is None and got.converged
        assert np.array_equal(got.rms, r0 * q ** np.arange(got.steps))
    s = serving.summarize(list(res.values()))
    assert s["refused"] == 1 and s["requests"] == 4 and s["converged"] == 4


def test_batched_server_checks_its_arguments():
    with pytest.raises(ValueError):
        serving.BatchServer(chunk=30, max_steps=100)
    with pytest.raises(ValueError):
        serving.BatchServer(slots=1, hbm_budget=1.0).serve(_stream(1))


def test_batch_create_checks_requests_on_the_host():
    """orion_hip_batch_create refuses, before any device work (so on the
    CPU), requests without blocks or whose block counts do not add up,
    requests that differ in an input shared by the batch, and an interface
    that joins two requests' blocks."""
    import ctypes as C
    from orion_amd import _lib
    from orion_amd.solver import case_inputs, region_descs
    L = _lib.load()
    cs = [r.case for r in serving.channel_stream(2, seed=3, nodes=(2.0e3, 3.0e3))]
    dims = np.ascontiguousarray(np.concatenate([np.asarray(c.dims, dtype=np.int32).reshape(-1) for c in cs]))
    xyz = np.ascontiguousarray(np.concatenate([np.asarray(c.xyz_flat(), dtype=np.float64) for c in cs]))
    regs, nregs = region_descs(cs, [0, 1])

    def call(nreq, counts, inps, regions=regs):
        h = C.c_void_p()
        rc = L.orion_hip_batch_create(nreq, (C.c_int * 2)(*counts), 2, dims.ctypes.data_as(C.POINTER(C.c_int)),
                                      xyz.ctypes.data_as(C.POINTER(C.c_double)), nregs, regions, inps, 0,
                                      C.byref(h))
        assert h.value is None
        return rc, L.orion_hip_last_error().decode()

    good = (_lib.Inputs * 2)(*[case_inputs(c) for c in cs])
    rc, msg = call(0, (1, 1), good)
    assert rc != 0 and "at least one request" in msg
    rc, msg = call(2, (1, 2), good)
    assert rc != 0 and "do not add up" in msg
    rc, msg = call(2, (0, 2), good)
    assert rc != 0 and "has no blocks" in msg
    mixed = (_lib.Inputs * 2)(*[case_inputs(c) for c in cs])
    mixed[1].nvis = 0
    rc, msg = call(2, (1, 1), mixed)
    assert rc != 0 and "must share" in msg
    cross, _ = region_descs(cs, [0, 1])
    q = next(i for i in range(nregs) if cross[i].block == 1 and cross[i].bctype < 0)
    cross[q].nbt = 1  # request 1's periodic interface pointed at request 0's block
    rc, msg = call(2, (1, 1), good, cross)
    assert rc != 0 and "joins blocks of requests 1 and 0" in msg


def test_oversized_request_is_refused():
    reqs = _stream(1)
    with pytest.raises(ValueError):
        serving.Server(slots=1, hbm_budget=1.0).serve(reqs)


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix():
    reqs = serving.channel_stream(6, seed=11, rate=1000.0, nodes=(6.0e3, 2.0e4))
    print("\n[gpu] serving: 6 requests alone and in the mix", flush=True)
    kw = dict(chunk=20, drop=1e-30, max_steps=80, keep_fields=True)
    alone = {r.rid: r for r in serving.Server(slots=1, **kw).serve(reqs)}
    mixed = {r.rid: r for r in serving.Server(slots=6, **kw).serve(reqs)}
    for rid in alone:
        a, m = alone[rid], mixed[rid]
        assert a.steps == m.steps == 80
        assert np.array_equal(a.rms, m.rms), rid
        for pa, pm in zip(a.prim, m.prim):
            assert np.array_equal(pa, pm), rid


@pytest.mark.gpu
def test_same_result_alone_and_batched():
    """Six requests of different sizes and inputs through the batch server
    with four slots: they join at chunk boundaries while others run, and the
    last two wait for departures, so batches are rebuilt around requests
    that carry their state over.  Each result is its result alone."""
    reqs = serving.channel_stream(6, seed=13, rate=60.0, nodes=(6.0e3, 2.0e4))
    print("\n[gpu] serving: 6 requests alone and batched", flush=True)
    kw = dict(chunk=10, drop=1e-30, max_steps=60, keep_fields=True)
    alone = {r.rid: r for r in serving.Server(slots=1, **kw).serve(reqs)}
    srv = serving.BatchServer(slots=4, **kw)
    batched = {r.rid: r for r in srv.serve(reqs)}
    assert srv.rebuilds >= 3 and srv.peak_resident <= 4
    for rid in alone:
        a, m = alone[rid], batched[rid]
        assert a.steps == m.steps == 60
        assert np.array_equal(a.rms, m.rms), rid
        for pa, pm in zip(a.prim, m.prim):
            assert np.array_equal(pa, pm), rid


@pytest.mark.gpu
def test_batch_refuses_per_request_calls_it_cannot_answer():
    """A batch reports residuals per request only through orion_hip_run, and
    a state copy needs blocks of equal sizes."""
    import ctypes as C
    from orion_amd import _lib
    cs = [r.case for r in serving.channel_stream(2, seed=17, nodes=(5.0e3, 9.0e3))]
    print("\n[gpu] serving: batch refusals", flush=True)
    L = _lib.load()
    b = serving.Batch(cs)
    try:
        assert L.orion_hip_requests(b.ctx) == 2
        st = _lib.ResidualStats()
        assert L.orion_hip_residual(b.ctx, C.byref(st)) != 0
        assert "through orion_hip_run" in L.orion_hip_last_error().decode()
        assert L.orion_hip_calculate_time_step(b.ctx, None) != 0
        assert b.run(3).shape == (3, 2)
        same = tuple(cs[0].dims[0]) == tuple(cs[1].dims[0])
        rc = L.orion_hip_copy_block_state(b.ctx, 0, b.ctx, 1)
        assert (rc == 0) == same
    finally:
        b.close()


@pytest.mark.gpu
def test_batch_of_multiblock_requests_equals_each_alone():
    """Three requests in one batch -- a 4-block lattice (interfaces, halo and
    averaging maps), a channel, and a 4-block supersonic lattice at another
    Mach number -- give each request's own statistics (residuals, maximum
    and its location numbered within the request, time step) and fields, bit
    for bit."""
    from tests import _cases
    from orion_amd.solver import OrionSolver
    cs = [_cases.lattice_small(nsteps=6), _cases.channel_small(nsteps=6), _cases.lattice_supersonic(nsteps=6)]
    print("\n[gpu] serving: batch of multi-block requests", flush=True)
    alone = []
    for c in cs:
        s = OrionSolver(c)
        h = s.solve(6)
        alone.append((h, [s.fs.field(b, "prim") for b in range(c.nblocks)]))
        s.fs.close()
    b = serving.Batch(cs)
    try:
        st = b.run(6, stats=True)
        for k, c in enumerate(cs):
            h, prims = alone[k]
            for n in range(6):
                x, y = st[n][k], h[n]
                assert (x.rms_residual, x.max_residual, x.max_loc, x.phydtime) == \
                    (y.rms_residual, y.max_residual, y.max_loc, y.phydtime), (k, n)
            for blk, p in enumerate(b.prim(k)):
                assert np.array_equal(p, prims[blk]), (k, blk)
    finally:
        b.close()


@pytest.mark.gpu
def test_batch_with_a_million_node_request_equals_each_alone():
    """A batch holding a channel-bump request of about 1.2M nodes in 6
    blocks (the serving stream's shape), a 3-block bump request at another
    Mach and Reynolds number and a single-block channel: each request's
    residuals, maximum, its location and time step per step, and its final
    fields, equal its run alone bit for bit."""
    from orion_amd import cases as C
    from orion_amd.solver import OrionSolver
    cs = [C.bump_request(11, 1.2e6, 6, 0.8, 3.0e5, 0.06, False, nsteps=4),
          C.bump_request(12, 5.0e4, 3, 0.4, 2.0e3, 0.1, True, nsteps=4),
          C.bump_request(13, 2.0e4, 1, 0.6, 1.0e4, 0.0, False, nsteps=4)]
    assert cs[0].interior_nodes() >= 1_000_000
    print("\n[gpu] serving: batch with a %d-node request" % cs[0].interior_nodes(), flush=True)
    alone = []
    for c in cs:
        s = OrionSolver(c)
        h = s.solve(4)
        alone.append((h, [s.fs.field(b, "prim") for b in range(c.nblocks)]))
        s.fs.close()
    b = serving.Batch(cs)
    try:
        st = b.run(4, stats=True)
        for k, c in enumerate(cs):
            h, prims = alone[k]
            for n in range(4):
                x, y = st[n][k], h[n]
                assert (x.rms_residual, x.max_residual, x.max_loc, x.phydtime) == \
                    (y.rms_residual, y.max_residual, y.max_loc, y.phydtime), (k, n)
            for blk, p in enumerate(b.prim(k)):
                assert np.array_equal(p, prims[blk]), (k, blk)
    finally:
        b.close()
