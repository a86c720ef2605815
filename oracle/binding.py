"""ctypes binding of the C oracle (oracle/_build/liborion_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the orion_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborion_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

FIELDS = {"prim": (0, 6), "metrics": (1, 9), "vol": (2, 1), "q": (3, 5), "dq": (4, 5), "dt": (5, 1),
          "c": (6, 1), "mu": (7, 1), "spec": (8, 3), "spec_visc": (9, 3)}


class _Inputs(C.Structure):
    _fields_ = [("moo", C.c_double), ("reynolds", C.c_double), ("tref", C.c_double),
                ("twall", C.c_double), ("cfl", C.c_double), ("attack", C.c_double),
                ("sideslip", C.c_double), ("csrv", C.c_int), ("nvis", C.c_int), ("method", C.c_int),
                ("ntmst", C.c_int), ("gcl", C.c_int)]


class _Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gama", "prl", "moo", "reynolds", "visc", "tref", "twall",
                                           "cfl", "roo", "uoo", "voo", "woo", "poo", "too")] + \
               [(n, C.c_int) for n in ("csrv", "nvis", "method", "ntmst", "gcl")] + \
               [("phydtime", C.c_double)]


class _Residual(C.Structure):
    _fields_ = [("rms", C.c_double), ("max", C.c_double), ("max_loc", C.c_int * 5)]


def build(quiet: bool = True) -> str:
    subprocess.run(["make", "-C", HERE, "oracle"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ob_create.restype = C.c_void_p
        L.ob_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int,
                                C.POINTER(C.c_int), C.POINTER(_Inputs)]
        L.ob_destroy.argtypes = [C.c_void_p]
        L.ob_last_error.restype = C.c_char_p
        L.ob_field.restype = C.POINTER(C.c_double)
        L.ob_field.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ob_get_params.restype = C.POINTER(_Params)
        L.ob_get_params.argtypes = [C.c_void_p]
        for n in ("ob_update_temperature", "ob_update_flow_states", "ob_calculate_time_step",
                  "ob_compute_viscous_rhs", "ob_compute_inviscid_rhs", "ob_update_solution"):
            getattr(L, n).argtypes = [C.c_void_p]
            getattr(L, n).restype = None
        for n in ("ob_exchange_bc", "ob_apply_physical_bc", "ob_average_interface_residuals"):
            getattr(L, n).argtypes = [C.c_void_p]
            getattr(L, n).restype = C.c_int
        L.ob_residual_compute.restype = _Residual
        L.ob_residual_compute.argtypes = [C.c_void_p]
        L.ob_step.argtypes = [C.c_void_p, C.POINTER(_Residual)]
        L.ob_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.ob_set_threads.argtypes = [C.c_int]
        L.ob_lusgs_sweeps.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.ob_output_fields.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        L.ob_output_fields.restype = None
        L.ob_guard_counts.argtypes = [C.POINTER(C.c_long)]
        L.ob_guard_counts.restype = None
        L.ob_guard_reset.restype = None
        L.ob_set_accurate_sum.argtypes = [C.c_int]
        L.ob_set_accurate_sum.restype = None
        _lib = L
    return _lib


class OracleCase:
    """One case held by the C oracle; fields are exposed as numpy views in
    OrionArray layout: (nvar, nz, ny, nx), i fastest."""

    def __init__(self, case, threads: int = 1, apply_init: bool = True, accurate_sum: bool = False):
        """accurate_sum: form the RMS residual with a compensated sum (about
        the exact sum) rather than the reference's loop-order sum, whose
        rounding error reaches 1e-12 relative at millions of nodes; the
        setting is process-wide and lasts until the next OracleCase."""
        L = lib()
        L.ob_set_threads(int(threads))
        L.ob_set_accurate_sum(1 if accurate_sum else 0)
        self.case = case
        dims = np.ascontiguousarray(case.dims, dtype=np.int32).reshape(-1)
        xyz = np.ascontiguousarray(case.xyz_flat())
        regs = np.ascontiguousarray(case.region_table(), dtype=np.int32)
        inp = _Inputs(**{k: case.inputs[k] for k in (n for n, _ in _Inputs._fields_)})
        self._h = L.ob_create(case.nblocks, dims.ctypes.data_as(C.POINTER(C.c_int)),
                              xyz.ctypes.data_as(C.POINTER(C.c_double)), regs.shape[0],
                              regs.ctypes.data_as(C.POINTER(C.c_int)), C.byref(inp))
        if not self._h:
            raise RuntimeError("oracle: " + L.ob_last_error().decode())
        if apply_init and case.init_rho is not None:
            for b in range(case.nblocks):
                self.field(b, "prim")[...] = case.initial_prim(b)
            L.ob_update_temperature(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ob_destroy(self._h)
            self._h = None

    def field(self, b: int, name: str) -> np.ndarray:
        fid, nv = FIELDS[name]
        nx, ny, nz = self.case.padded(b)
        ptr = lib().ob_field(self._h, b, fid)
        arr = np.ctypeslib.as_array(ptr, shape=(nv * nz * ny * nx,))
        return arr.reshape(nv, nz, ny, nx)

    def params(self):
        p = lib().ob_get_params(self._h).contents
        return {n: getattr(p, n) for n, _ in _Params._fields_}

    def op(self, name: str):
        fn = getattr(lib(), "ob_" + name)
        rc = fn(self._h)
        if isinstance(rc, int) and rc != 0:
            raise RuntimeError("oracle %s: %s" % (name, lib().ob_last_error().decode()))

    def residual(self):
        r = lib().ob_residual_compute(self._h)
        return r.rms, r.max, tuple(r.max_loc)

    def step(self):
        r = _Residual()
        if lib().ob_step(self._h, C.byref(r)) != 0:
            raise RuntimeError("oracle step: " + lib().ob_last_error().decode())
        return r.rms, r.max, tuple(r.max_loc)

    def run(self, n: int) -> np.ndarray:
        hist = np.zeros(2 * n)
        if lib().ob_run(self._h, n, hist.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise RuntimeError("oracle run: " + lib().ob_last_error().decode())
        return hist.reshape(n, 2)

    def output_fields(self, b: int) -> np.ndarray:
        """The VTS point fields of block b (PostProcess.cpp:183-240), float32."""
        i, j, k = (int(v) for v in self.case.dims[b])
        out = np.empty(7 * i * j * k, dtype=np.float32)
        lib().ob_output_fields(self._h, b, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def lusgs_sweeps(self, b: int, rhs: np.ndarray) -> None:
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        lib().ob_lusgs_sweeps(self._h, b, rhs.ctypes.data_as(C.POINTER(C.c_double)))


def guard_counts(reset: bool = False):
    """Times the reference's guard branches were taken in the oracle since
    the last reset: (left fallback, right fallback, damped update, kept
    state)."""
    out = (C.c_long * 4)()
    lib().ob_guard_counts(out)
    if reset:
        lib().ob_guard_reset()
    return tuple(out)


# --------------------------------------------------------------------------
# reference harness (oracle/_ref/ref_harness, this container only)
# --------------------------------------------------------------------------

def have_reference() -> bool:
    return os.path.exists(REF_HARNESS)


# ref_harness history columns: rms, max, max_loc[5], phydtime, compensated rms
HIST_COLS = 9
MPIEXEC = os.environ.get("ORION_MPIEXEC", "/opt/conda/bin/mpiexec")  # the MPI the reference build links


def max_ranks(case, most: int) -> int:
    """The most MPI ranks (at most ``most``) the reference can run ``case``
    on: a placement of its blocks must exist in which no block sends two
    interface messages to one other rank (Case.block_pids)."""
    for r in range(max(1, min(int(most), case.nblocks)), 0, -1):
        try:
            case.block_pids(r)
            return r
        except ValueError:
            continue
    return 1


def _harness_cmd(ranks: int):
    ranks = max(1, int(ranks))
    if ranks == 1:
        return [REF_HARNESS]
    return [MPIEXEC, "-launcher", "fork", "-n", str(ranks), REF_HARNESS]


def time_reference(case, ranks: int, nsteps: int, workdir: str):
    """The reference solver's own CPU path (oracle/_ref/ref_harness in its
    timing mode) on ``ranks`` MPI processes of this host, the blocks spread
    over the ranks (Case.block_pids): ``nsteps`` pseudo-steps of ``case``
    from its seeded start state, timed between barriers around the step
    loop (set-up excluded).  Returns (steps, seconds, ranks, hist[n, 9])."""
    casedir, outdir = os.path.join(workdir, "case"), os.path.join(workdir, "out")
    case.write(casedir, nprocs=ranks)
    os.makedirs(outdir, exist_ok=True)
    env = dict(os.environ)
    env.pop("LD_LIBRARY_PATH", None)
    subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(max(1, int(ranks))), REF_HARNESS, casedir,
                    str(int(nsteps)), outdir, "0", "time"],
                   check=True, stdout=subprocess.DEVNULL, env=env, timeout=1800)
    with open(os.path.join(outdir, "time.txt")) as f:
        steps, secs, nproc = f.read().split()
    hist = np.fromfile(os.path.join(outdir, "hist.bin")).reshape(int(steps), HIST_COLS)
    return int(steps), float(secs), int(nproc), hist


def run_reference(case, casedir: str, outdir: str, nsteps: int, dump_steps: int = 0, vts: bool = False,
                  ranks: int = 1):
    """Write ``case`` in the reference's formats and run the reference
    solver's operators for ``nsteps`` (on ``ranks`` MPI processes, the
    blocks spread over them).  Returns (hist[n,9], reader) where
    reader(stage_file) -> list of per-block dicts of fields ("final.bin"
    also reads the per-block files of a run on several ranks).  With
    ``vts`` the reference's PostProcess::write_solution also runs after the
    last step, writing into the case's flowname directory under casedir.
    Stage dumps and ``vts`` need one rank."""
    ranks = max(1, int(ranks))
    if ranks > 1 and (dump_steps or vts):
        raise ValueError("stage dumps and VTS output need a single-rank run")
    case.write(casedir, nprocs=ranks)
    os.makedirs(outdir, exist_ok=True)
    env = dict(os.environ)
    env.pop("LD_LIBRARY_PATH", None)
    subprocess.run(_harness_cmd(ranks) + [casedir, str(nsteps), outdir, str(dump_steps)] + (["vts"] if vts else []),
                   check=True, stdout=subprocess.DEVNULL, env=env)
    hist = np.fromfile(os.path.join(outdir, "hist.bin")).reshape(nsteps, HIST_COLS)

    def block(raw, pos, b):
        nx, ny, nz = case.padded(b)
        P = nx * ny * nz
        blk = {}
        for key, nv in (("prim", 6), ("dq", 5), ("c", 1), ("mu", 1), ("dt", 1)):
            blk[key] = raw[pos:pos + nv * P].reshape(nv, nz, ny, nx)
            pos += nv * P
        return blk, pos

    def read(name: str):
        if ranks > 1 and name == "final.bin":
            out = []
            for b in range(case.nblocks):
                raw = np.fromfile(os.path.join(outdir, "final_b%d.bin" % b))
                blk, pos = block(raw, 0, b)
                assert pos == raw.size
                out.append(blk)
            return out
        raw = np.fromfile(os.path.join(outdir, name))
        out, pos = [], 0
        for b in range(case.nblocks):
            blk, pos = block(raw, pos, b)
            out.append(blk)
        assert pos == raw.size
        return out

    return hist, read
