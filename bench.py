#!/usr/bin/env python3
"""Benchmark: FP64 cell-steps/s of ORION's pseudo-time step on one MI355X.

Workload (default): the 64-block mesh of about 8.05M interior nodes on a
seeded block-adjacency graph with partial-face interface windows
(cases.column_case(seed=3): 4x4 columns of 4 blocks, staggered block
edges), M=0.5, Re=5e3, CFL 5, WCNS-E-5 + Roe inviscid flux, 4th/6th-order
viscous flux, LU-SGS with two extra symmetric Gauss-Seidel sweeps, FP64,
synthetic seeded initial field.  ``--workload uneven64`` runs the 4x4x4
lattice of uneven blocks (7.85M nodes), ``--workload eight_block`` 8 blocks
of 129x65x65 (M=0.8, Re=1e5, CFL 10).

A step is one full pseudo-time step over every block (halo, boundary
conditions, flow states, time step, viscous and inviscid right-hand sides
with interface averaging, LU-SGS update, residual).  value = interior nodes x
timed steps / wall time of the timed steps, with all state resident in HBM.

value is the median of 5 timed runs of --steps steps, each from the same
state, in the batched form (every kernel over all blocks, the step replayed
as a captured graph); the plain per-block form (a launch per block per
kernel) is timed beside it from the same state and its results compared.

Prints one JSON line.  The roofline entry is the dominant kernel family,
timed live with HIP events on the solver's stream.  The CPU baseline is the
reference's own MPI CPU path (the reference solver built from its sources,
oracle/_ref) on as many ranks of this host as the mesh allows (up to 16;
the reference's interface-residual messages are tagged by source block, so
a rank may hold at most one interface partner of any block of another
rank), for a few steps from the same start state, its residuals compared
with the GPU's; the C oracle (a port of the reference, OpenMP) is timed
beside it as cpu_port.  Both samples are bounded.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-steps/s at 8M cells vs CPU; served solves/s + p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# Compulsory FP64 words per interior cell for each stage of the step (prim,
# geometry = 9 metrics + vol, right-hand side and increments read or written
# once; derived per-cell scalars recomputed; O(surface) work excluded).  The
# stages sum to 155 words = 1240 B per cell-step.  Each stage lists the
# kernels that implement it and the profiled phases they run in.
STAGES = {
    "rhs": (20, ["visc_deriv", "visc_face", "visc_diff", "inv_face", "inv_closure", "inv_diff"],
            "visc_deriv_kernel+visc_face_queue_kernel+diff_kernel<false,true>+inv_face_pair_kernel+inv_closure_kernel"
            "+diff_kernel<true,false>"),
    # a layout copy: no algorithmic words of its own.  Its state part runs
    # inside state_pack_kernel with the flow states (the state phase); the
    # rhs part is lu_pack_kernel<2>
    "lu_pack": (0, ["lu_pack_state", "lu_pack"], "lu_pack_kernel<2>"),
    # the four sweeps (LU forward 25, LU backward 25, GS forward 30, GS
    # backward 30 words): one chained launch, or one launch per sweep
    "lu_sweeps": (110, ["lu_fwd", "lu_bwd", "gs_fwd", "gs_bwd", "lu_chain"], "lusgs_chain_kernel<true>"),
    "update": (25, ["update", "residual"], "update_kernel+residual_final_kernel"),
}
SWEEP_KERNELS = {True: "lusgs_chain_kernel<true>",
                 False: "lusgs_tile_kernel<0,true>+lusgs_tile_kernel<1,true>+lusgs_tile_kernel<2,true>"
                        "+lusgs_tile_kernel<3,true>"}
STEP_WORDS = sum(v[0] for v in STAGES.values())
# The dominant kernel: the LU-SGS sweeps (LU forward / backward, GS forward
# / backward), one chained launch per step (or four, one per sweep, under
# the per-sweep schedules), which take the largest share of the step of any
# kernel and run alone (the captured step overlaps only the face-flux and
# pack kernels).  Its algorithmic words are those of the four sweep stages
# (25 + 25 + 30 + 30); lu_pack, a layout copy, is given none.
DOMINANT = ("lusgs_chain_kernel", 110, STAGES["lu_sweeps"][1], STAGES["lu_sweeps"][2].split("+"))


def use_sweep_kernels(chained):
    """Name the sweep kernels of the schedule in use in the stage and
    dominant-kernel tables."""
    global DOMINANT
    k = SWEEP_KERNELS[bool(chained)]
    STAGES["lu_sweeps"] = STAGES["lu_sweeps"][:2] + (k,)
    DOMINANT = ("lusgs_chain_kernel" if chained else "lusgs_tile_kernel",) + DOMINANT[1:3] + (k.split("+"),)


def use_visc_kernels(fused):
    """Name the viscous kernels of the form in use in the stage table."""
    faces = "visc_fused_kernel<true>" if fused else "visc_face_queue_kernel+diff_kernel<false,true>"
    STAGES["rhs"] = STAGES["rhs"][:2] + ("visc_deriv_kernel+" + faces + "+inv_face_pair_kernel+inv_closure_kernel"
                                         "+diff_kernel<true,false>",)


def make_case(name, nsteps):
    from orion_amd import cases
    if name == "uneven64":
        return cases.uneven_lattice_case(seed=11, nsteps=nsteps)
    if name == "columns64":
        return cases.column_case(seed=3, nsteps=nsteps)
    if name == "eight_block":
        return cases.eight_block_case(seed=7, nsteps=nsteps)
    if name == "channel":
        return cases.channel_wall_case(nsteps=nsteps)
    raise SystemExit("unknown workload " + name)


def profile_phases(fs, steps):
    """Per-phase ms/step and launches/step, timed with HIP events recorded on
    the solver's own stream around each phase of ``steps`` eager steps."""
    import ctypes as C
    from orion_amd import _lib
    L = _lib.load()
    n = len(_lib.PHASES)
    ms = (C.c_double * n)()
    nl = (C.c_longlong * n)()
    _lib.check(L.orion_hip_profile_run(fs.ctx, steps, ms, nl), "profile_run")
    return {_lib.PHASES[i]: (ms[i] / steps, nl[i] / steps) for i in range(n)}


def roofline(phase, N, ms_per_step):
    """Stage rooflines: algorithmic bytes (stage words x 8 B x interior
    cells) over the stage's measured time.  The dominant kernel is reported
    as ``roofline``; every stage and the whole step are reported beside it."""
    stages = {}
    for name, (words, phases, kernels) in STAGES.items():
        t_ms = sum(phase[p][0] for p in phases)
        launches = sum(phase[p][1] for p in phases)
        nbytes = words * 8.0 * N
        gbs = nbytes / (t_ms * 1e-3) / 1e9 if t_ms > 0 else 0.0
        stages[name] = {"kernels": kernels, "ms": round(t_ms, 4), "launches": launches,
                        "algorithmic_bytes": nbytes, "achieved_GBs": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 5)}
    # the four sweeps: one launch per sweep phase, or one chained launch
    # (phase lu_chain) that runs all four
    kname, words, kphases, _ = DOMINANT
    t_ms = sum(phase[p][0] for p in kphases)
    launches = int(round(sum(phase[p][1] for p in kphases))) or len(kphases)
    chained = phase.get("lu_chain", (0.0, 0))[1] > 0
    per_launch = words * 8.0 * N / launches
    avg_s = t_ms * 1e-3 / launches
    gbs = per_launch / avg_s / 1e9
    roof = {
        "bound": "hbm",
        "kernel": ("%s (LU forward, LU backward, GS forward and GS backward sweeps as one launch)" % kname
                   if chained else
                   "%s (LU forward, LU backward, GS forward, GS backward sweeps)" % kname),
        "achieved": round(gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 5),
        "traffic": None,
        "algorithmic_bytes_per_launch": per_launch,
        "launch_avg_us": round(1e6 * avg_s, 3),
        "launches_per_step": launches,
    }
    step_gbs = STEP_WORDS * 8.0 * N / (ms_per_step * 1e-3) / 1e9
    step = {"bytes_per_cell_step": STEP_WORDS * 8, "achieved_GBs": round(step_gbs, 1),
            "frac": round(step_gbs / HBM_PEAK_GBS, 5)}
    return roof, stages, step


def copy_calibration(mib):
    """A plain device copy of ``mib`` MiB (read mib, write mib) repeated 5
    times: the counter scale for FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 5}), flush=True)
    return 0


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_{workload}.txt")
DEFAULT_WORKLOAD = "columns64"


def _kernel_key(name):
    """Short kernel name of a rocprofv3 Kernel_Name ("void ns::k<2>(args)" -> "k<2>")."""
    n = name.split("(")[0].replace("void ", "").strip()
    return n.split("::")[-1].replace(" ", "")


def summarize_counters(fetch_csv, write_csv, cfetch_csv, cwrite_csv, steps, cmd, commit=""):
    """Per-kernel HBM bytes per step from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE, KiB per dispatch), each scaled by what the same
    counter reads for a plain device copy of known size."""
    import csv
    from collections import defaultdict

    def load(path):
        agg = defaultdict(lambda: [0, 0.0])
        for r in csv.DictReader(open(path)):
            a = agg[_kernel_key(r["Kernel_Name"])]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
        return agg

    def copy_kib(path):
        rows = [r for r in csv.DictReader(open(path)) if "copyBuffer" in r["Kernel_Name"]]
        return sum(float(r["Counter_Value"]) for r in rows) / len(rows)

    copy_bytes = 1024 * (1 << 20)
    sf = copy_bytes / (copy_kib(cfetch_csv) * 1024.0)
    sw = copy_bytes / (copy_kib(cwrite_csv) * 1024.0)
    f, w = load(fetch_csv), load(write_csv)
    lines = ["# HBM traffic per kernel (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)",
             "# command: " + cmd,
             "# commit: " + (commit or "unknown"),
             "# calibration: a %d-byte device copy reads FETCH_SIZE x %.4f and WRITE_SIZE x %.4f bytes" % (
                 copy_bytes, sf, sw),
             "# columns: kernel dispatches_per_step fetch_bytes_per_step write_bytes_per_step",
             "scale_fetch %.6f" % sf, "scale_write %.6f" % sw]
    for k in sorted(f, key=lambda k: -(f[k][1] * sf + w[k][1] * sw)):
        lines.append("%s %g %.0f %.0f" % (k, f[k][0] / steps, f[k][1] * 1024 * sf / steps,
                                          w.get(k, [0, 0.0])[1] * 1024 * sw / steps))
    return "\n".join(lines) + "\n"


def stage_traffic(workload, groups):
    """HBM bytes per launch of each kernel group (name -> list of kernels)
    from the committed counter summary (empty when there is none for this
    workload)."""
    path = TRAFFIC_SUMMARY.format(workload=workload)
    if not os.path.exists(path):
        return {}
    per = {}
    commit = "unknown"
    for line in open(path):
        if line.startswith("# commit: "):
            commit = line.split(":", 1)[1].strip()
        parts = line.split()
        if line.startswith("#") or len(parts) != 4:
            continue
        per[parts[0]] = (float(parts[1]), float(parts[2]) + float(parts[3]))
    out = {}
    for name, ks in groups.items():
        if not all(k in per for k in ks):
            continue
        nbytes = sum(per[k][1] for k in ks)
        launches = sum(per[k][0] for k in ks)
        out[name] = {"bytes_per_step": nbytes, "bytes_per_launch": nbytes / max(launches, 1),
                     "source": "%s (counters of the tree at commit %s)" % (os.path.relpath(path, ROOT), commit)}
    return out


def serving_run(nreq, slots, stream="bumps", max_steps=200):
    """The serving form (orion_amd.serving): by default the surveyed shape
    (measure_bumps): a seeded Poisson stream of channel-bump solves of
    0.2M-8M nodes in 1-16 blocks, each stopped at a 1e-8 RMS residual drop
    relative to its first step or after ``max_steps`` steps, served
    ``slots`` at a time with continuous admission from one batch context,
    and the same stream solved one request at a time in the plain per-block
    form (the sequential plain-GPU baseline); with the check that every
    request's residual history is the same in both runs.  stream
    "channels": small single-block channels to a 1e-8 drop (measure).
    It runs as a child process,
    because the hardware queue count it asks HIP for (one per resident
    request) must be set before HIP starts in a process."""
    cmd = [sys.executable, "-u", "-m", "orion_amd.serving", "--requests", str(nreq), "--slots", str(slots),
           "--stream", stream, "--max-steps", str(max_steps)]
    try:
        p = subprocess.run(cmd, cwd=os.path.dirname(os.path.abspath(__file__)), stdout=subprocess.PIPE, timeout=1200)
        if p.returncode != 0:
            raise RuntimeError("exit status %d" % p.returncode)
        return json.loads(p.stdout.decode().strip().splitlines()[-1])
    except Exception as e:  # noqa: BLE001 -- reported in the line; the plain measurement stands
        print("[bench] serving measurement failed: %s" % e, file=sys.stderr, flush=True)
        return {"error": str(e)}


def host_info():
    """Host cores as the bench sees them and the CPU model."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"nproc": avail, "cpu_count": os.cpu_count(), "cpu_model": model}


def binding_max_ranks(case, most):
    from oracle import binding
    return binding.max_ranks(case, most)


def cpu_reference(case, ranks, nsteps, gpu_hist=None):
    """The reference's MPI CPU path (its own solver built from its sources,
    oracle/_ref) on ``ranks`` processes of this host, the blocks spread
    over them: ``nsteps`` whole pseudo-steps of the same case from the same
    seeded start state, set-up excluded.  With ``gpu_hist`` (the GPU's
    residuals of its first steps from that state) the two histories are
    compared.  None without the build."""
    import tempfile
    from oracle import binding
    if not (binding.have_reference() and os.path.exists(binding.MPIEXEC)):
        return None
    try:
        with tempfile.TemporaryDirectory() as d:
            steps, secs, nproc, hist = binding.time_reference(case, ranks, nsteps, d)
    except Exception as e:  # noqa: BLE001 -- the port then stands as the CPU baseline
        print("[bench] reference cpu path unavailable (%s); the port is the cpu baseline" % e, file=sys.stderr,
              flush=True)
        return None
    print("[bench] reference cpu path: %d steps in %.1f s on %d MPI ranks" % (steps, secs, nproc), file=sys.stderr,
          flush=True)
    out = {"value": case.interior_nodes() * steps / secs, "unit": "cell-steps/s", "cores": nproc, "kind": "reference",
           "sample": "%d pseudo-steps of the same %d-block, %d-node case from the same seeded start state, by the "
                     "reference solver built from its own sources on %d MPI ranks (blocks spread over them, "
                     "largest block %.1f%% of the nodes), step loop timed between barriers (%.1f s)" % (
                         steps, case.nblocks, case.interior_nodes(), nproc,
                         100.0 * float(max(case.dims.prod(axis=1))) / case.interior_nodes(), secs)}
    out.update(host_info())
    if gpu_hist:
        n = min(len(gpu_hist), steps)
        rel = [abs(gpu_hist[i] / hist[i, 0] - 1.0) for i in range(n)]
        out["history_check"] = {"steps": n, "reference_rms": [float(v) for v in hist[:n, 0]],
                                "gpu_rms": [float(v) for v in gpu_hist[:n]], "max_rel_diff": max(rel) if rel else None}
    return out


def cpu_baseline(case, threads, seconds):
    """The C oracle (a restatement of the reference path, OpenMP over blocks)
    on the same case: whole pseudo-steps until ``seconds`` have passed."""
    from oracle.binding import OracleCase
    o = OracleCase(case, threads=threads)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    print("[bench] cpu baseline: %d steps in %.1f s on %d threads" % (n, dt, threads), file=sys.stderr, flush=True)
    out = {"value": case.interior_nodes() * n / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
           "sample": "%d full pseudo-steps of the same %d-block, %d-node case on %d OpenMP threads (%.1f s)" % (
               n, case.nblocks, case.interior_nodes(), threads, dt)}
    out.update(host_info())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--runs", type=int, default=5, help="timed runs of --steps steps each; value is their median")
    ap.add_argument("--per-block-steps", type=int, default=5,
                    help="steps of the plain per-block form timed beside the batched form (0: none)")
    ap.add_argument("--tune-steps", type=int, default=3,
                    help="steps per trial of the launch-shape autotuner after the warmup (0: no tuning)")
    ap.add_argument("--schedule", type=int, default=-1,
                    help="fix the sweep schedule (0 diagonal, 1 critical path, 2 chained) instead of tuning")
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--ref-steps", type=int, default=2, help="steps of the reference CPU baseline (0: port only)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-stream", choices=("bumps", "channels"), default="bumps")
    ap.add_argument("--serve-max-steps", type=int, default=200)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--summarize-counters", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="write the per-kernel HBM traffic summary from rocprofv3 counter CSVs and exit")
    ap.add_argument("--counter-steps", type=int, default=2, help="steps covered by the counter run")
    ap.add_argument("--counter-command", default="", help="command recorded in the summary")
    ap.add_argument("--counter-commit", default="", help="commit of the tree the counters were measured on")
    a = ap.parse_args()
    if a.profile_steps < 1:
        ap.error("--profile-steps must be at least 1: the rooflines come from the per-phase timings")
    if a.summarize_counters:
        sys.stdout.write(summarize_counters(*a.summarize_counters, a.counter_steps, a.counter_command,
                                            a.counter_commit))
        return 0
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by block across GPUs, which is "
              "designed but not built here); run with --gpus 1", file=sys.stderr)
        return 2
    if a.copy_calibration_mib:
        return copy_calibration(a.copy_calibration_mib)
    from orion_amd.solver import OrionSolver
    case = make_case(a.workload, a.steps)
    N = case.interior_nodes()
    t_setup = time.perf_counter()
    s = OrionSolver(case, device=0)
    t_setup = time.perf_counter() - t_setup
    print("[bench] %s: %d blocks, %d interior nodes, setup %.1f s" % (case.name, case.nblocks, N, t_setup),
          file=sys.stderr, flush=True)
    warm_hist = []
    if a.warmup > 0:
        warm_hist = [h.rms_residual for h in s.solve(a.warmup)]
    tuning = None
    if a.schedule >= 0:
        s.fs.set_schedule(a.schedule)
    elif a.tune_steps > 0:
        # untimed, like the warmup; the state is restored after each trial:
        # the sweep schedule, then the viscous stage's tiling
        tuning = s.fs.tune(a.tune_steps)
        print("[bench] launch-shape tuning: %s" % json.dumps(tuning), file=sys.stderr, flush=True)
    # a.runs timed runs of exactly a.steps steps, each from the same state
    # (prim alone fixes the next step); value is their median
    start = [s.fs.field(b, "prim") for b in range(case.nblocks)] if a.runs > 1 else None
    run_t = []
    for r in range(a.runs):
        if r > 0:
            for b in range(case.nblocks):
                s.fs.set_field(b, "prim", start[b])
        s.fs.synchronize()
        t0 = time.perf_counter()
        hist = s.solve(a.steps)
        s.fs.synchronize()
        run_t.append(time.perf_counter() - t0)
        print("[bench] run %d: %d steps in %.4f s, rms %.6e -> %.6e" % (
            r, a.steps, run_t[-1], hist[0].rms_residual, hist[-1].rms_residual), file=sys.stderr, flush=True)
    t = sorted(run_t)[len(run_t) // 2]
    value = N * a.steps / t
    plain = None
    if a.per_block_steps > 0 and start is not None:
        # the plain per-block form (a launch per block per kernel, no captured
        # step) from the same state, against the batched run from that state
        n = min(a.per_block_steps, a.steps)
        for b in range(case.nblocks):
            s.fs.set_field(b, "prim", start[b])
        ref = s.solve(n)
        end_batched = [s.fs.field(b, "prim") for b in range(case.nblocks)]
        for b in range(case.nblocks):
            s.fs.set_field(b, "prim", start[b])
        s.fs.set_per_block(True)
        s.solve(1)  # untimed: first launches of the form
        for b in range(case.nblocks):
            s.fs.set_field(b, "prim", start[b])
        s.fs.synchronize()
        t0 = time.perf_counter()
        hp = s.solve(n)
        s.fs.synchronize()
        tp = time.perf_counter() - t0
        same = ([h.rms_residual for h in hp] == [h.rms_residual for h in ref]
                and all(np.array_equal(s.fs.field(b, "prim"), end_batched[b]) for b in range(case.nblocks)))
        s.fs.set_per_block(False)
        for b in range(case.nblocks):
            s.fs.set_field(b, "prim", start[b])
        plain = {"form": "plain per-block: every kernel launched once per block, in block order, one stream, "
                         "no captured step", "steps": n, "ms_per_step": 1e3 * tp / n, "value": N * n / tp,
                 "unit": "cell-steps/s", "batched_over_plain": (N * a.steps / t) / (N * n / tp),
                 "same_results_as_batched": bool(same)}
        print("[bench] plain per-block form: %.3f ms/step (batched %.3f), same results: %s" % (
            1e3 * tp / n, 1e3 * t / a.steps, same), file=sys.stderr, flush=True)
    use_sweep_kernels(s.fs.schedule() == 2)
    use_visc_kernels(s.fs.visc_form())
    phase = profile_phases(s.fs, a.profile_steps)
    phases = {k: round(v[0], 4) for k, v in phase.items()}
    print("[bench] phase ms/step: " + json.dumps(phases), file=sys.stderr, flush=True)
    roof, stages, step_roof = roofline(phase, N, 1e3 * t / a.steps)
    traffic = stage_traffic(a.workload, {k: v[2].split("+") for k, v in STAGES.items()})
    for name, tr in traffic.items():
        stages[name]["traffic_bytes_per_launch"] = tr["bytes_per_launch"]
    dom = stage_traffic(a.workload, {"dominant": DOMINANT[3]}) if DOMINANT[3] else {}
    if dom:
        roof["traffic"] = dom["dominant"]["bytes_per_launch"]
        roof["traffic_source"] = dom["dominant"]["source"]
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
                   "device_bytes": s.fs.device_bytes(),
                   "scheme": "WCNS-E-5+Roe, 6th-order viscous, LU-SGS+2 SGS sweeps", "precision": "fp64",
                   "moo": case.inputs["moo"], "reynolds": case.inputs["reynolds"], "cfl": case.inputs["cfl"],
                   "sweep_schedule": s.fs.SCHEDULES[s.fs.schedule()],
                   "visc_tiles": dict(zip(("segment", "grid"), s.fs.visc_tiles()))},
        "schedule_tuning": tuning,
        "roofline": roof,
        "stage_rooflines": stages,
        "step_roofline": step_roof,
        "phase_ms_per_step": phases,
        "residual_first_last": [hist[0].rms_residual, hist[-1].rms_residual],
        "runs_ms_per_step": [round(1e3 * v / a.steps, 4) for v in run_t],
        "timing": "median of %d timed runs of %d steps, each from the same state" % (a.runs, a.steps),
        "plain_per_block": plain,
    }
    s.fs.close()
    if not a.no_serve:
        out["serving"] = serving_run(a.serve_requests, a.serve_slots, a.serve_stream, a.serve_max_steps)
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        port = cpu_baseline(case, threads, a.cpu_seconds)
        ref = cpu_reference(case, binding_max_ranks(case, threads), a.ref_steps, warm_hist) if a.ref_steps > 0 else None
        out["cpu_baseline"] = ref or port
        if ref:
            out["cpu_port"] = port
            out["gpu_over_cpu_port"] = value / port["value"]
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
