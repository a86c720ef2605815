"""Serving: continuous batching of independent steady solves on one GPU.

A request is one steady solve (a case) that arrives at a given time and is
done when its RMS residual has dropped by ``drop`` relative to its first
step (the reference's monitor, ResidualMonitor::compute,
src/solver/ResidualMonitor.cpp:11-138), or after ``max_steps`` steps.
Requests are admitted in arrival order as soon as one has arrived, a slot
is free and its HBM footprint fits the budget beside the requests already
resident.  A request's result does not depend on what else is resident.

``BatchServer`` (the served form): the resident requests share one batch
context (``Batch``, orion_hip_batch_create), so every kernel launch of a
step covers all of them.  The batch advances ``chunk`` pseudo-steps at a
time; at each chunk boundary finished requests leave, arrived ones join,
and the batch is rebuilt with the staying requests' state copied over on
the device.  Each request's arithmetic and reduction order are those of a
context of its own, so its results are too, bit for bit.

``Server``: one context per resident request, each advanced by a host
thread; their steps overlap on the GPU.  It is the simpler form and,
with one slot, the sequential baseline: the same requests solved one after
another in arrival order.  With several slots each context's step is one
stream of kernels (orion_hip_set_overlap off), and HIP maps a process's
streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), kernels of
streams that share a queue running one after another; a process using it
sets GPU_MAX_HW_QUEUES to HW_QUEUES before HIP starts (``prepare_process``,
as ``python -m orion_amd.serving`` does).  Many small contexts share the
device's launch throughput, which is why the batch is the served form.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import threading
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import cases

# HBM held by one context per padded point: the arena (fields, scratch,
# static face geometry, the inviscid flux array) plus the LU-SGS sweep
# arrays.  Measured (orion_hip_device_bytes) on the channel requests of
# channel_stream: 1.33-1.46 KB, and 1.39 KB for a batch of sixteen;
# admission counts 2 KB
BYTES_PER_PADDED_POINT = 2048

# hardware queues asked for by a serving process: one per resident request
# up to this many (HIP accepts at most 32 here)
HW_QUEUES = 32


def prepare_process(hw_queues: int = HW_QUEUES) -> None:
    """Ask HIP for ``hw_queues`` hardware queues (1 to 32), replacing what
    the environment holds; effective only before the process's first HIP
    call."""
    if not 1 <= int(hw_queues) <= 32:
        raise ValueError("hw_queues must lie in 1..32")
    os.environ["GPU_MAX_HW_QUEUES"] = str(int(hw_queues))


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    case: cases.Case


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    start: float
    finish: float
    steps: int
    converged: bool
    nodes: int
    rms: np.ndarray  # residual history
    prim: Optional[List[np.ndarray]] = None  # final fields, when asked for
    error: Optional[str] = None  # why the request was refused (it then ran no step)

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def footprint(case: cases.Case) -> int:
    """HBM bytes a request is admitted for."""
    return BYTES_PER_PADDED_POINT * case.padded_points()


def channel_stream(n: int, seed: int = 2024, rate: float = 20.0, nodes=(1.0e4, 1.0e5), moo=(0.5, 0.85),
                   reynolds=(1.0e3, 1.0e6)) -> List[Request]:
    """A seeded Poisson stream of ``n`` steady channel solves: arrival rate
    ``rate`` per second; per request log-uniform interior nodes and Reynolds
    number, uniform Mach number, upper boundary a wall or far-field."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    out = []
    for r in range(n):
        nn = float(np.exp(rng.uniform(np.log(nodes[0]), np.log(nodes[1]))))
        m = float(rng.uniform(*moo))
        re = float(np.exp(rng.uniform(np.log(reynolds[0]), np.log(reynolds[1]))))
        upper = bool(rng.integers(0, 2))
        c = cases.bump_channel_case(seed * 1000 + r, nn, m, re, 0.0, upper_wall=upper)
        out.append(Request(r, float(t[r]), c))
    return out


def bump_stream(n: int, seed: int = 4242, rate: float = 20.0, nodes=(2.0e5, 8.0e6), blocks=(1, 16),
                moo=(0.3, 0.85), reynolds=(1.0e3, 1.0e6), bump=(0.0, 0.1)) -> List[Request]:
    """A seeded Poisson stream of ``n`` steady channel-bump solves in the
    shape of the surveyed serving workload: arrival rate ``rate`` per
    second; per request log-uniform interior nodes in ``nodes``, a block
    count uniform in ``blocks`` (the channel cut along i, each piece at least
    9 nodes long), uniform Mach number, log-uniform Reynolds number, a
    cosine bump of uniform height on the lower wall, the upper boundary a
    wall or far-field (cases.bump_request)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    out = []
    for r in range(n):
        nn = float(np.exp(rng.uniform(np.log(nodes[0]), np.log(nodes[1]))))
        nb = int(rng.integers(blocks[0], blocks[1] + 1))
        m = float(rng.uniform(*moo))
        re = float(np.exp(rng.uniform(np.log(reynolds[0]), np.log(reynolds[1]))))
        h = float(rng.uniform(*bump))
        upper = bool(rng.integers(0, 2))
        out.append(Request(r, float(t[r]), cases.bump_request(seed * 1000 + r, nn, nb, m, re, h, upper)))
    return out


def _device_solver(case: cases.Case, overlap: bool = True, per_block: bool = False):
    from .solver import OrionSolver
    s = OrionSolver(case)
    s.fs.set_overlap(overlap)
    if per_block:
        s.fs.set_per_block(True)
    return s


def _plain_solver(case: cases.Case, overlap: bool = True):
    """The plain per-block form: a launch per block per kernel, one stream,
    no captured step."""
    return _device_solver(case, False, per_block=True)


class Server:
    def __init__(self, slots: int = 8, hbm_budget: float = 200.0e9, chunk: int = 50, drop: float = 1.0e-8,
                 max_steps: int = 3000, make_solver: Callable = _device_solver, keep_fields: bool = False,
                 overlap: Optional[bool] = None, progress: bool = False):
        """overlap: the side stream of each context (orion_hip_set_overlap);
        None: on for one slot, off for several.  progress: a line on stderr
        per finished request."""
        self.progress = progress
        if slots < 1 or chunk < 1:
            raise ValueError("slots and chunk must be positive")
        self.overlap = (slots == 1) if overlap is None else bool(overlap)
        self.slots, self.hbm_budget, self.chunk = slots, float(hbm_budget), chunk
        self.drop, self.max_steps = drop, max_steps
        self.make_solver, self.keep_fields = make_solver, keep_fields
        self.peak_resident = 0
        self.peak_bytes = 0

    def _solve(self, req: Request, t0: float) -> Result:
        start = time.perf_counter() - t0
        s = self.make_solver(req.case, self.overlap)
        rms: List[float] = []
        converged = False
        try:
            while len(rms) < self.max_steps:
                n = min(self.chunk, self.max_steps - len(rms))
                hist = s.solve(n)
                rms.extend(h.rms_residual for h in hist)
                r = np.asarray(rms)
                hit = np.nonzero(r <= self.drop * r[0])[0]
                if hit.size:
                    converged = True
                    break
            prim = None
            if self.keep_fields:
                prim = [s.fs.field(b, "prim") for b in range(req.case.nblocks)]
        finally:
            s.fs.close()
        finish = time.perf_counter() - t0
        return Result(req.rid, req.arrival, start, finish, len(rms), converged, req.case.interior_nodes(),
                      np.asarray(rms), prim)

    def serve(self, requests: Sequence[Request]) -> List[Result]:
        """Run the stream in real time: a request is not admitted before its
        arrival time.  Returns the results in completion order."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if footprint(r.case) > self.hbm_budget:
                raise ValueError("request %d needs %.1f GB, more than the budget" % (r.rid, footprint(r.case) / 1e9))
        results: List[Result] = []
        active = {}
        used = 0
        lock = threading.Lock()
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=self.slots) as pool:
            while pending or active:
                now = time.perf_counter() - t0
                while (pending and pending[0].arrival <= now and len(active) < self.slots
                       and used + footprint(pending[0].case) <= self.hbm_budget):
                    req = pending.pop(0)
                    with lock:
                        used += footprint(req.case)
                        self.peak_bytes = max(self.peak_bytes, used)
                    active[pool.submit(self._solve, req, t0)] = req
                    self.peak_resident = max(self.peak_resident, len(active))
                timeout = None
                if pending and len(active) < self.slots and used + footprint(pending[0].case) <= self.hbm_budget:
                    timeout = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                if not active:
                    time.sleep(timeout or 0.0)
                    continue
                done, _ = wait(list(active), timeout=timeout, return_when=FIRST_COMPLETED)
                for f in done:
                    req = active.pop(f)
                    with lock:
                        used -= footprint(req.case)
                    results.append(f.result())
                    if self.progress:
                        r = results[-1]
                        print("[serving]   request %d done at %.1f s: %d steps, %d nodes" % (
                            r.rid, r.finish, r.steps, r.nodes), file=sys.stderr, flush=True)
        return results


class Batch:
    """Resident requests batched in one device context
    (orion_hip_batch_create): each kernel launch of a step covers all of
    them, and each request's results are those it has in a context of its
    own.  ``states[k]`` for a request carried over from an earlier batch:
    ``(batch, j)``, request j of that (still open) batch, whose state is
    copied on the device, or its prim per block as host arrays; None for a
    request starting at step 0."""

    def __init__(self, case_list: Sequence[cases.Case], states: Optional[Sequence] = None, device: int = 0):
        import ctypes as C
        from . import _lib
        from .solver import case_inputs, region_descs
        L = _lib.load()
        t_in = time.perf_counter()
        _lib.check(L.orion_hip_backend_init(int(device)), "backend_init")
        self.cases = list(case_list)
        n = len(self.cases)
        if n == 0:
            raise ValueError("a batch needs at least one request")
        nbl = np.array([c.nblocks for c in self.cases], dtype=np.int32)
        self.first = [int(v) for v in np.concatenate([[0], np.cumsum(nbl)[:-1]])]
        dims = np.ascontiguousarray(np.concatenate([np.asarray(c.dims, dtype=np.int32).reshape(-1)
                                                    for c in self.cases]))
        # each request's grid where its case holds it (no gathered copy)
        grids = [np.ascontiguousarray(c.xyz_flat(), dtype=np.float64) for c in self.cases]
        gptr = (C.POINTER(C.c_double) * n)(*[g.ctypes.data_as(C.POINTER(C.c_double)) for g in grids])
        regs, nreg = region_descs(self.cases, self.first)
        inps = (_lib.Inputs * n)(*[case_inputs(c) for c in self.cases])
        h = C.c_void_p()
        t_cr = time.perf_counter()
        _lib.check(L.orion_hip_batch_create_grids(n, nbl.ctypes.data_as(C.POINTER(C.c_int)), int(nbl.sum()),
                                                  dims.ctypes.data_as(C.POINTER(C.c_int)), gptr, nreg, regs, inps,
                                                  0, C.byref(h)), "batch_create_grids")
        del grids
        self.ctx = h
        t_su = time.perf_counter()
        try:
            self._set_up(states, L)
        except Exception:
            self.close()
            raise
        self.timing = {"inputs": t_cr - t_in, "create": t_su - t_cr, "setup": time.perf_counter() - t_su}

    def _set_up(self, states, L) -> None:
        from . import _lib
        n = len(self.cases)
        states = list(states) if states is not None else [None] * n
        if len(states) != n:
            raise ValueError("%d states for %d requests" % (len(states), n))
        # requests starting at step 0: their initial field, then the
        # temperature, as a context of their own sets them up; carried-over
        # requests: their saved prim (temperature included) after that
        fresh = False
        for k, (c, st) in enumerate(zip(self.cases, states)):
            if st is None and c.init_rho is not None:
                for b in range(c.nblocks):
                    self._prim_io(k, b, c.initial_prim(b), upload=True)
                fresh = True
        if fresh:
            _lib.check(L.orion_hip_update_temperature(self.ctx), "update_temperature")
        for k, st in enumerate(states):
            if isinstance(st, tuple):
                src, j = st
                for b in range(self.cases[k].nblocks):
                    _lib.check(L.orion_hip_copy_block_state(self.ctx, self.first[k] + b, src.ctx, src.first[j] + b),
                               "copy_block_state")
            elif st is not None:
                for b, a in enumerate(st):
                    self._prim_io(k, b, a, upload=True)

    def _prim_io(self, k: int, b: int, arr: Optional[np.ndarray], upload: bool) -> np.ndarray:
        import ctypes as C
        from . import _lib
        nx, ny, nz = self.cases[k].padded(b)
        fid, nv = _lib.FIELD_IDS["prim"]
        if upload:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(nv, nz, ny, nx)
            _lib.check(_lib.load().orion_hip_upload(self.ctx, self.first[k] + b, fid,
                                                    a.ctypes.data_as(C.POINTER(C.c_double)), a.size), "upload")
            return a
        a = np.empty((nv, nz, ny, nx))
        _lib.check(_lib.load().orion_hip_download(self.ctx, self.first[k] + b, fid,
                                                  a.ctypes.data_as(C.POINTER(C.c_double)), a.size), "download")
        return a

    def prim(self, k: int) -> List[np.ndarray]:
        """Request k's prim, per block."""
        return [self._prim_io(k, b, None, upload=False) for b in range(self.cases[k].nblocks)]

    def run(self, nsteps: int, stats: bool = False):
        """nsteps pseudo-steps of every request: the RMS residuals, (nsteps,
        requests); with ``stats``, each step's statistics per request
        instead (a list of steps, each a list of ResidualStatsPy, max_loc's
        block numbered within the request)."""
        from . import _lib
        from .solver import _stats
        n = len(self.cases)
        hist = (_lib.ResidualStats * (nsteps * n))()
        _lib.check(_lib.load().orion_hip_run(self.ctx, nsteps, hist), "run")
        if stats:
            return [[_stats(hist[s * n + k]) for k in range(n)] for s in range(nsteps)]
        return np.array([h.rms_residual for h in hist]).reshape(nsteps, n)

    def close(self) -> None:
        if getattr(self, "ctx", None):
            from . import _lib
            _lib.load().orion_hip_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchServer:
    """Continuous batching: the resident requests share one batch context
    (``Batch``) that advances them all ``chunk`` pseudo-steps at a time.  At
    each chunk boundary finished requests leave and arrived ones are
    admitted (arrival order, within ``slots`` and the HBM budget); the batch
    is then rebuilt, the staying requests' state copied over on the device.  Every request
    advances in whole chunks from its step 0, as under ``Server``, so it
    stops at the same step with the same history."""

    def __init__(self, slots: int = 16, hbm_budget: float = 200.0e9, chunk: int = 50, drop: float = 1.0e-8,
                 max_steps: int = 3000, keep_fields: bool = False, device: int = 0,
                 make_batch: Callable = None, progress: bool = False, refuse_oversize: bool = False):
        """make_batch(cases, states, device): the batch factory (``Batch``);
        progress: a line on stderr per batch rebuild; refuse_oversize: a
        request larger than the whole budget is refused (a result with its
        error) instead of failing the stream."""
        self.progress = progress
        self.refuse_oversize = refuse_oversize
        self.delayed_by_budget = 0  # admissions that waited for HBM (a slot was free)
        self.host_stashes = 0  # rebuilds whose staying state went through host memory
        if slots < 1 or chunk < 1:
            raise ValueError("slots and chunk must be positive")
        if max_steps % chunk:
            raise ValueError("max_steps must be a multiple of chunk")
        self.slots, self.hbm_budget, self.chunk = slots, float(hbm_budget), chunk
        self.drop, self.max_steps, self.keep_fields, self.device = drop, max_steps, keep_fields, device
        self.make_batch = make_batch or Batch
        self.peak_resident = 0
        self.peak_bytes = 0
        self.rebuilds = 0
        self.rebuild_s = 0.0  # host time spent rebuilding batches
        # where the rebuild time goes: host inputs, context creation, state set-up, old batch release
        self.rebuild_parts = {"inputs": 0.0, "create": 0.0, "setup": 0.0, "release": 0.0}

    def serve(self, requests: Sequence[Request]) -> List[Result]:
        """Run the stream in real time; returns the results in completion order."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        results: List[Result] = []
        for r in list(pending):
            if footprint(r.case) > self.hbm_budget:
                msg = "request %d needs %.1f GB, more than the budget" % (r.rid, footprint(r.case) / 1e9)
                if not self.refuse_oversize:
                    raise ValueError(msg)
                pending.remove(r)
                results.append(Result(r.rid, r.arrival, r.arrival, r.arrival, 0, False, r.case.interior_nodes(),
                                      np.zeros(0), None, msg))
        waited: set = set()
        resident: List[dict] = []  # in batch order: {req, start, rms}
        leaving: set = set()  # indices into resident that finished at the last boundary
        batch: Optional[Batch] = None
        t0 = time.perf_counter()
        try:
            while pending or len(resident) > len(leaving):
                now = time.perf_counter() - t0
                stay = [k for k in range(len(resident)) if k not in leaving]
                used = sum(footprint(resident[k]["req"].case) for k in stay)
                admitted = []
                while (pending and pending[0].arrival <= now and len(stay) + len(admitted) < self.slots
                       and used + footprint(pending[0].case) <= self.hbm_budget):
                    req = pending.pop(0)
                    used += footprint(req.case)
                    admitted.append({"req": req, "start": now, "rms": []})
                if (pending and pending[0].arrival <= now and len(stay) + len(admitted) < self.slots
                        and pending[0].rid not in waited):
                    waited.add(pending[0].rid)  # arrived, a slot free, but its HBM does not fit yet
                    self.delayed_by_budget += 1
                if admitted or leaving:
                    # one rebuild per boundary: the staying requests' state
                    # copied on the device from the old batch, then the
                    # newcomers.  The old batch and the new one are both on the
                    # device until the copies are done; when the two would not
                    # fit the budget together, the staying requests' state
                    # goes through host memory and the old batch is released
                    # first.
                    old_bytes = sum(footprint(m["req"].case) for m in resident) if batch is not None else 0
                    if batch is not None and old_bytes + used > self.hbm_budget:
                        states = [batch.prim(k) for k in stay]
                        self.host_stashes += 1
                        batch.close()
                        batch = None
                    else:
                        states = [(batch, k) for k in stay] if batch is not None else []
                    resident = [resident[k] for k in stay] + admitted
                    leaving = set()
                    new = None
                    tb = time.perf_counter()
                    if resident:
                        try:
                            new = self.make_batch([m["req"].case for m in resident],
                                                  states + [None] * len(admitted), self.device)
                        except Exception:
                            if not admitted:
                                raise
                            # a newcomer the device refuses must not take the others down:
                            # admit them one at a time, refusing those that fail alone
                            keep = resident[:len(stay)]
                            for m in admitted:
                                try:
                                    self.make_batch([x["req"].case for x in keep + [m]],
                                                    states + [None] * (len(keep) - len(stay) + 1), self.device).close()
                                    keep.append(m)
                                except Exception as e:  # noqa: BLE001 -- reported in the result
                                    now = time.perf_counter() - t0
                                    results.append(Result(m["req"].rid, m["req"].arrival, now, now, 0, False,
                                                          m["req"].case.interior_nodes(), np.zeros(0), None, str(e)))
                            resident = keep
                            used = sum(footprint(m["req"].case) for m in resident)
                            new = (self.make_batch([m["req"].case for m in resident],
                                                   states + [None] * (len(resident) - len(stay)), self.device)
                                   if resident else None)
                        self.rebuilds += 1
                        self.peak_resident = max(self.peak_resident, len(resident))
                        self.peak_bytes = max(self.peak_bytes, used)
                        if self.progress:
                            print("[serving]   batch %d at %.1f s: %d resident, %.1f GB admitted, %d pending" % (
                                self.rebuilds, time.perf_counter() - t0, len(resident), used / 1e9, len(pending)),
                                file=sys.stderr, flush=True)
                    if new is not None and hasattr(new, "timing"):
                        for k, v in new.timing.items():
                            self.rebuild_parts[k] += v
                    tr = time.perf_counter()
                    if batch is not None:
                        batch.close()
                    self.rebuild_parts["release"] += time.perf_counter() - tr
                    batch = new
                    self.rebuild_s += time.perf_counter() - tb
                if not resident:
                    if pending:
                        time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                    continue
                hist = batch.run(self.chunk)
                finish = time.perf_counter() - t0
                for k, m in enumerate(resident):
                    m["rms"].extend(hist[:, k].tolist())
                    r = np.asarray(m["rms"])
                    converged = bool(np.any(r <= self.drop * r[0]))
                    if converged or len(r) >= self.max_steps:
                        leaving.add(k)
                        prim = batch.prim(k) if self.keep_fields else None
                        results.append(Result(m["req"].rid, m["req"].arrival, m["start"], finish, len(r), converged,
                                              m["req"].case.interior_nodes(), r, prim))
        finally:
            if batch is not None:
                batch.close()
        return results


def summarize(results: Sequence[Result]) -> dict:
    """solves/s = completed / (last completion - first arrival); latency =
    completion - arrival, p50 and p99; refused requests are counted apart."""
    refused = sum(r.error is not None for r in results)
    results = [r for r in results if r.error is None]
    if not results:
        return {"requests": 0, "refused": refused}
    lat = np.array([r.latency for r in results])
    span = max(r.finish for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "refused": refused, "converged": int(sum(r.converged for r in results)),
            "solves_per_s": len(results) / span if span > 0 else float("nan"),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_steps": float(np.mean([r.steps for r in results])),
            "cell_steps_per_s": float(sum(r.steps * r.nodes for r in results) / span) if span > 0 else float("nan"),
            "span_s": span}


def measure(nreq: int = 16, slots: int = 16, seed: int = 2024, sequential: bool = True) -> dict:
    """The serving measurement: the seeded stream ``channel_stream(nreq,
    seed)`` solved one request at a time (the sequential baseline) and then
    served ``slots`` at a time; with the check that every request's
    residual history is the same in both runs.  sequential=False: the
    served run alone."""
    reqs = channel_stream(nreq, seed=seed)
    lo = min(r.case.interior_nodes() for r in reqs)
    hi = max(r.case.interior_nodes() for r in reqs)
    print("[serving] %d requests, %d-%d nodes, GPU_MAX_HW_QUEUES=%s" % (
        nreq, lo, hi, os.environ.get("GPU_MAX_HW_QUEUES", "(HIP default)")), file=sys.stderr, flush=True)
    seq = None
    if sequential:
        t0 = time.perf_counter()
        seq = Server(slots=1).serve(reqs)
        print("[serving] sequential %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    srv_ = BatchServer(slots=slots)
    srv = srv_.serve(reqs)
    print("[serving] batched, %d slots: %.1f s, %d batch rebuilds (%.2f s)" % (
        slots, time.perf_counter() - t0, srv_.rebuilds, srv_.rebuild_s), file=sys.stderr, flush=True)
    a = summarize(srv)
    b = summarize(seq) if seq is not None else None
    same = None
    if seq is not None:
        by = {r.rid: r for r in seq}
        same = all(np.array_equal(by[r.rid].rms, r.rms) for r in srv)
    return {"workload": "%d steady channel solves (1 block, %d-%d nodes, M 0.5-0.85, Re 1e3-1e6), Poisson arrivals "
                        "20/s, done at a 1e-8 RMS drop or 3000 steps" % (nreq, lo, hi),
            "slots": slots, "form": "one batch context of the resident requests, rebuilt at 50-step boundaries",
            "batch_rebuilds": srv_.rebuilds, "rebuild_s": srv_.rebuild_s, "peak_resident": srv_.peak_resident,
            "served": a, "sequential": b,
            "solves_per_s_vs_sequential": a["solves_per_s"] / b["solves_per_s"] if b else None,
            "same_histories_alone_and_mixed": same}


def measure_bumps(nreq: int = 32, slots: int = 16, seed: int = 4242, max_steps: int = 200,
                  hbm_budget: float = 200.0e9, plain: bool = True, batched_sequential: bool = False) -> dict:
    """The serving measurement at the surveyed shape: ``bump_stream(nreq,
    seed)`` (0.2M-8M nodes, 1-16 blocks, M 0.3-0.85, Re 1e3-1e6, bumps up
    to 0.1) served ``slots`` at a time from one batch context, against the
    same requests solved one after another in the plain per-block form
    (the sequential plain GPU solves) and, with ``batched_sequential``, one
    after another in the batched form.  A request stops at a 1e-8 RMS drop
    relative to its first step or after ``max_steps`` steps, whichever
    comes first; every request's residual history must be the same in
    every run."""
    reqs = bump_stream(nreq, seed=seed)
    nodes = [r.case.interior_nodes() for r in reqs]
    print("[serving] %d bump requests, %d-%d nodes, %d-%d blocks, stop at 1e-8 drop or %d steps" % (
        nreq, min(nodes), max(nodes), min(r.case.nblocks for r in reqs), max(r.case.nblocks for r in reqs),
        max_steps), file=sys.stderr, flush=True)
    out = {"workload": "%d steady channel-bump solves (%d-%d nodes, 1-16 blocks cut along i, M 0.3-0.85, "
                       "Re 1e3-1e6 log-uniform, cosine bumps of height 0-0.1, upper wall or far-field), Poisson "
                       "arrivals 20/s, seed %d" % (nreq, min(nodes), max(nodes), seed),
           "stop_rule": "RMS residual drop of 1e-8 relative to the first step, or %d steps, whichever comes first"
                        % max_steps,
           "slots": slots, "hbm_budget_bytes": hbm_budget,
           "form": "one batch context of the resident requests, rebuilt at 50-step boundaries"}
    t0 = time.perf_counter()
    srv_ = BatchServer(slots=slots, max_steps=max_steps, hbm_budget=hbm_budget, progress=True)
    srv = srv_.serve(reqs)
    print("[serving] batched, %d slots: %.1f s, %d batch rebuilds (%.2f s)" % (
        slots, time.perf_counter() - t0, srv_.rebuilds, srv_.rebuild_s), file=sys.stderr, flush=True)
    out.update(batch_rebuilds=srv_.rebuilds, rebuild_s=srv_.rebuild_s, peak_resident=srv_.peak_resident,
               peak_admitted_bytes=srv_.peak_bytes, served=summarize(srv))
    out["rebuild_breakdown_s"] = dict(srv_.rebuild_parts)
    by = {r.rid: r for r in srv}
    out["requests_final_rms_not_finite"] = int(sum(not np.all(np.isfinite(r.rms)) for r in srv))
    same = True
    differ = {}
    for key, on, make in (("sequential_plain", plain, _plain_solver),
                          ("sequential_batched_form", batched_sequential, _device_solver)):
        if not on:
            continue
        t0 = time.perf_counter()
        seq = Server(slots=1, max_steps=max_steps, make_solver=make, hbm_budget=hbm_budget, progress=True).serve(reqs)
        print("[serving] %s %.1f s" % (key, time.perf_counter() - t0), file=sys.stderr, flush=True)
        out[key] = summarize(seq)
        out["solves_per_s_vs_" + key] = out["served"]["solves_per_s"] / out[key]["solves_per_s"]
        # histories bit for bit (a diverged request's non-finite entries compare equal to themselves)
        bad = [r.rid for r in seq if not np.array_equal(by[r.rid].rms, r.rms, equal_nan=True)]
        if bad:
            first = {}
            for rid in bad[:4]:
                a, b = by[rid].rms, next(x.rms for x in seq if x.rid == rid)
                n = min(len(a), len(b))
                d = np.nonzero(~((a[:n] == b[:n]) | (np.isnan(a[:n]) & np.isnan(b[:n]))))[0]
                first[rid] = int(d[0]) if d.size else -1
            differ[key] = {"requests": bad, "first_differing_step": first}
        same = same and not bad
    out["same_histories_alone_and_mixed"] = same if (plain or batched_sequential) else None
    if differ:
        out["differing"] = differ
    return out


def device_memory(device: int = 0):
    """(free, total) bytes of the device, from the HIP runtime."""
    import ctypes as C
    from . import _lib
    _lib.load()
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipSetDevice(int(device)) != 0 or hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return int(free.value), int(total.value)


STEP_BYTES = 1240  # algorithmic bytes per interior cell-step (the step's roofline)
HBM_PEAK = 8.0e12


def measure_capacity(nreq: int = 32, slots: int = 32, seed: int = 4242, max_steps: int = 200, alone_steps: int = 10,
                     chunk: int = 10, margin: float = 8.0e9) -> dict:
    """The capacity workload (cases.capacity_case: 48 blocks of 257x129x21,
    33.4M nodes, transonic bump channel, CFL 20) resident beside the served
    bump stream (measure_bumps' requests).  The capacity case runs on its
    own host thread and stream: first ``alone_steps`` steps alone, then in
    chunks of ``chunk`` steps for as long as the stream is served; the
    stream's admission budget is the device memory left beside it less
    ``margin``.  Reports the HBM held, the admissions the budget delayed or
    refused, the capacity case's step time and step roofline alone and under
    contention, and the stream's solves/s and latencies."""
    from .solver import OrionSolver
    cap = cases.capacity_case()
    print("[capacity] %s: %d nodes" % (cap.name, cap.interior_nodes()), file=sys.stderr, flush=True)
    free0, total = device_memory()
    s = OrionSolver(cap)
    cap_bytes = s.fs.device_bytes()
    s.solve(2)
    s.fs.synchronize()
    t0 = time.perf_counter()
    s.solve(alone_steps)
    s.fs.synchronize()
    alone_ms = 1e3 * (time.perf_counter() - t0) / alone_steps
    print("[capacity] alone: %.2f ms/step, %.1f GB on the device" % (alone_ms, cap_bytes / 1e9), file=sys.stderr,
          flush=True)
    free1, _ = device_memory()
    budget = float(free1 - margin)
    reqs = bump_stream(nreq, seed=seed)
    stop = threading.Event()
    chunks: List[float] = []
    err: List[str] = []

    def drive():
        try:
            while not stop.is_set():
                ta = time.perf_counter()
                s.solve(chunk)
                s.fs.synchronize()
                chunks.append(1e3 * (time.perf_counter() - ta) / chunk)
        except Exception as e:  # noqa: BLE001 -- reported
            err.append(str(e))

    th = threading.Thread(target=drive, daemon=True)
    th.start()
    srv_ = BatchServer(slots=slots, max_steps=max_steps, hbm_budget=budget, progress=True, refuse_oversize=True)
    t1 = time.perf_counter()
    res = srv_.serve(reqs)
    span = time.perf_counter() - t1
    stop.set()
    th.join()
    n = cap.interior_nodes()
    cont = float(np.mean(chunks)) if chunks else float("nan")
    out = {"workload": "%s (%d nodes, M 0.85, Re 1e6, CFL 20) resident beside %d channel-bump requests "
                       "(bump_stream seed %d, stop at a 1e-8 drop or %d steps)" % (cap.name, n, nreq, seed, max_steps),
           "device_total_bytes": total, "device_free_before_bytes": free0, "capacity_device_bytes": cap_bytes,
           "serving_budget_bytes": budget, "peak_admitted_bytes": srv_.peak_bytes,
           "peak_resident_requests": srv_.peak_resident,
           "peak_hbm_fraction": (cap_bytes + srv_.peak_bytes) / total,
           "admissions_delayed_by_budget": srv_.delayed_by_budget, "refused": sum(r.error is not None for r in res),
           "rebuilds": srv_.rebuilds, "rebuilds_through_host": srv_.host_stashes, "rebuild_s": srv_.rebuild_s,
           "served": summarize(res), "serving_span_s": span,
           "capacity_alone_ms_per_step": alone_ms, "capacity_contended_ms_per_step": cont,
           "capacity_steps_under_contention": chunk * len(chunks),
           "capacity_step_roofline_alone": n * STEP_BYTES / (alone_ms * 1e-3) / HBM_PEAK,
           "capacity_step_roofline_contended": n * STEP_BYTES / (cont * 1e-3) / HBM_PEAK if chunks else None,
           "capacity_error": err[0] if err else None}
    s.fs.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="serve a seeded stream of steady channel solves; prints one JSON line")
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--hw-queues", type=int, default=HW_QUEUES)
    ap.add_argument("--batched-only", action="store_true", help="skip the sequential baseline")
    ap.add_argument("--stream", choices=("channels", "bumps"), default="channels",
                    help="channels: small single-block channels to a 1e-8 drop; bumps: the surveyed shape "
                         "(0.2M-8M nodes, 1-16 blocks)")
    ap.add_argument("--max-steps", type=int, default=200, help="bumps: the step cap of the stop rule")
    ap.add_argument("--capacity", action="store_true",
                    help="the 33.4M-node capacity case resident beside the bump stream (measure_capacity)")
    ap.add_argument("--batched-sequential", action="store_true",
                    help="bumps: also solve the requests one after another in the batched form")
    a = ap.parse_args(argv)
    prepare_process(a.hw_queues)
    if a.capacity:
        res = measure_capacity(a.requests, a.slots, max_steps=a.max_steps)
    elif a.stream == "bumps":
        res = measure_bumps(a.requests, a.slots, seed=a.seed if a.seed != 2024 else 4242, max_steps=a.max_steps,
                            plain=not a.batched_only, batched_sequential=a.batched_sequential)
    else:
        res = measure(a.requests, a.slots, a.seed, sequential=not a.batched_only)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
