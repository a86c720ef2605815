#!/usr/bin/env python3
"""Benchmark: FP64 cell-steps/s of ORION's pseudo-time step on one MI355X.

Workload (default): the uneven 64-block mesh of about 8M interior nodes
(4x4x4 lattice, per-axis block sizes log-uniform in [17,129], seeded), M=0.5,
Re=5e3, CFL 5, WCNS-E-5 + Roe inviscid flux, 4th/6th-order viscous flux,
LU-SGS with two extra symmetric Gauss-Seidel sweeps, FP64, synthetic
seeded initial field.  ``--workload eight_block`` runs 8 blocks of 129x65x65
(M=0.8, Re=1e5, CFL 10) instead.

A step is one full pseudo-time step over every block (halo, boundary
conditions, flow states, time step, viscous and inviscid right-hand sides
with interface averaging, LU-SGS update, residual).  value = interior nodes x
timed steps / wall time of the timed steps, with all state resident in HBM.

Prints one JSON line.  The roofline entry is the dominant kernel family,
timed live with HIP events on the solver's stream; the CPU baseline is the
C oracle (a port of the reference) timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-steps/s at 8M cells vs CPU; served solves/s + p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# compulsory FP64 words per interior cell for each stage: prim, geometry
# (9 metrics + vol), right-hand side and increments read or written once
WORDS = {"lu_fwd": 25, "lu_bwd": 25, "gs_fwd": 30, "gs_bwd": 30, "rhs": 20, "update": 25}


def make_case(name, nsteps):
    from orion_amd import cases
    if name == "uneven64":
        return cases.uneven_lattice_case(seed=11, nsteps=nsteps)
    if name == "eight_block":
        return cases.eight_block_case(seed=7, nsteps=nsteps)
    if name == "channel":
        return cases.channel_wall_case(nsteps=nsteps)
    raise SystemExit("unknown workload " + name)


def roofline(fs, case, steps):
    import ctypes as C
    from orion_amd import _lib
    L = _lib.load()
    n = len(_lib.PHASES)
    ms = (C.c_double * n)()
    nl = (C.c_longlong * n)()
    _lib.check(L.orion_hip_profile_run(fs.ctx, steps, ms, nl), "profile_run")
    phase = {_lib.PHASES[i]: (ms[i] / steps, nl[i] / steps) for i in range(n)}
    N = case.interior_nodes()
    # dominant kernel family with a defined algorithmic byte count
    cand = {k: v for k, v in phase.items() if k in WORDS}
    dom = max(cand, key=lambda k: cand[k][0])
    t_ms, launches = cand[dom]
    bytes_step = WORDS[dom] * 8.0 * N
    achieved = bytes_step / (t_ms * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "kernel": "lusgs_plane_kernel (%s)" % dom if dom.startswith(("lu", "gs")) else dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "launch_avg_us": round(1e3 * t_ms / max(launches, 1), 3),
        "launches_per_step": launches,
        "algorithmic_bytes_per_launch": bytes_step / max(launches, 1),
    }, {k: round(v[0], 4) for k, v in phase.items()}


def cpu_baseline(case, threads):
    from oracle.binding import OracleCase
    o = OracleCase(case, threads=threads)
    t0 = time.perf_counter()
    o.step()
    dt = time.perf_counter() - t0
    return {"value": case.interior_nodes() / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": "1 full pseudo-step of the same %d-block, %d-node case on %d OpenMP threads (%.1f s)" % (
                case.nblocks, case.interior_nodes(), threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="uneven64")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by block across GPUs, which is "
              "designed but not built here); run with --gpus 1", file=sys.stderr)
        return 2
    from orion_amd.solver import OrionSolver
    case = make_case(a.workload, a.steps)
    N = case.interior_nodes()
    t_setup = time.perf_counter()
    s = OrionSolver(case, device=0)
    t_setup = time.perf_counter() - t_setup
    print("[bench] %s: %d blocks, %d interior nodes, setup %.1f s" % (case.name, case.nblocks, N, t_setup),
          file=sys.stderr, flush=True)
    if a.warmup > 0:
        s.solve(a.warmup)
    s.fs.synchronize()
    t0 = time.perf_counter()
    hist = s.solve(a.steps)
    s.fs.synchronize()
    t = time.perf_counter() - t0
    value = N * a.steps / t
    print("[bench] %d steps in %.3f s, rms %.6e -> %.6e" % (a.steps, t, hist[0].rms_residual, hist[-1].rms_residual),
          file=sys.stderr, flush=True)
    roof, phases = roofline(s.fs, case, a.profile_steps)
    print("[bench] phase ms/step: " + json.dumps(phases), file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": case.name, "blocks": case.nblocks, "interior_nodes": N,
                   "scheme": "WCNS-E-5+Roe, 6th-order viscous, LU-SGS+2 SGS sweeps", "precision": "fp64",
                   "moo": case.inputs["moo"], "reynolds": case.inputs["reynolds"], "cfl": case.inputs["cfl"]},
        "roofline": roof,
        "phase_ms_per_step": phases,
        "residual_first_last": [hist[0].rms_residual, hist[-1].rms_residual],
    }
    s.fs.close()
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(case, threads)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
