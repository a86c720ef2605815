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


def _device_solver(case: cases.Case, overlap: bool = True):
    from .solver import OrionSolver
    s = OrionSolver(case)
    s.fs.set_overlap(overlap)
    return s


class Server:
    def __init__(self, slots: int = 8, hbm_budget: float = 200.0e9, chunk: int = 50, drop: float = 1.0e-8,
                 max_steps: int = 3000, make_solver: Callable = _device_solver, keep_fields: bool = False,
                 overlap: Optional[bool] = None):
        """overlap: the side stream of each context (orion_hip_set_overlap);
        None: on for one slot, off for several."""
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
        _lib.check(L.orion_hip_backend_init(int(device)), "backend_init")
        self.cases = list(case_list)
        n = len(self.cases)
        if n == 0:
            raise ValueError("a batch needs at least one request")
        nbl = np.array([c.nblocks for c in self.cases], dtype=np.int32)
        self.first = [int(v) for v in np.concatenate([[0], np.cumsum(nbl)[:-1]])]
        dims = np.ascontiguousarray(np.concatenate([np.asarray(c.dims, dtype=np.int32).reshape(-1)
                                                    for c in self.cases]))
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(c.xyz_flat(), dtype=np.float64) for c in self.cases]))
        regs, nreg = region_descs(self.cases, self.first)
        inps = (_lib.Inputs * n)(*[case_inputs(c) for c in self.cases])
        h = C.c_void_p()
        _lib.check(L.orion_hip_batch_create(n, nbl.ctypes.data_as(C.POINTER(C.c_int)), int(nbl.sum()),
                                            dims.ctypes.data_as(C.POINTER(C.c_int)),
                                            xyz.ctypes.data_as(C.POINTER(C.c_double)), nreg, regs, inps, 0,
                                            C.byref(h)), "batch_create")
        self.ctx = h
        try:
            self._set_up(states, L)
        except Exception:
            self.close()
            raise

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
                 make_batch: Callable = None):
        """make_batch(cases, states, device): the batch factory (``Batch``)."""
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

    def serve(self, requests: Sequence[Request]) -> List[Result]:
        """Run the stream in real time; returns the results in completion order."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if footprint(r.case) > self.hbm_budget:
                raise ValueError("request %d needs %.1f GB, more than the budget" % (r.rid, footprint(r.case) / 1e9))
        results: List[Result] = []
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
                if admitted or leaving:
                    # one rebuild per boundary: the staying requests' state
                    # copied on the device from the old batch, then the newcomers
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
                    if batch is not None:
                        batch.close()
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="serve a seeded stream of steady channel solves; prints one JSON line")
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--hw-queues", type=int, default=HW_QUEUES)
    ap.add_argument("--batched-only", action="store_true", help="skip the sequential baseline")
    a = ap.parse_args(argv)
    prepare_process(a.hw_queues)
    print(json.dumps(measure(a.requests, a.slots, a.seed, sequential=not a.batched_only)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
