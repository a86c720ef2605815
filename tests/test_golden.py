"""The oracle and the HIP path against the golden fixtures in tests/golden/,
which hold outputs of the reference solver built from its own sources
(tests/golden/make_golden.py writes them; the inputs come from the seeded
generators named in each fixture).

The oracle must reproduce every fixture bit for bit (it uses glibc's pow,
as the reference does).  The HIP path is bit-exact where no far-field
boundary calls pow, except the RMS residual, a parallel sum taken in another
order than the reference's loop (1e-13 relative); with far-field boundaries it is held to 1e-12 relative
on every step's residual (RMS) and 1e-10 relative on the fields and on each
step's maximum residual.
"""
import os

import numpy as np
import pytest

from oracle.binding import OracleCase
from orion_amd import cases
from tests import _cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NG = 3
STAGE_OPS = [("halo", "exchange_bc"), ("bc", "apply_physical_bc"), ("state", "update_flow_states"),
             ("dt", "calculate_time_step"), ("visc", "compute_viscous_rhs"),
             ("avg1", "average_interface_residuals"), ("inv", "compute_inviscid_rhs"),
             ("avg2", "average_interface_residuals"), ("upd", "update_solution")]


GENERATORS = {
    "channel_65x33x9": ("cases.channel_wall_case(seed=1, nsteps=200)",
                        lambda: cases.channel_wall_case(seed=1, nsteps=200)),
    "lattice_small": ("tests._cases.lattice_small(nsteps=10)", lambda: _cases.lattice_small(nsteps=10)),
    "stages_17x9x7": ("cases.channel_wall_case(seed=5, nsteps=1, dims=(17, 9, 7))",
                      lambda: cases.channel_wall_case(seed=5, nsteps=1, dims=(17, 9, 7))),
    "bc_variety": ("tests._cases.bc_variety(nvis=0, nsteps=10)", lambda: _cases.bc_variety(nvis=0, nsteps=10)),
}


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    text, make = GENERATORS[name]
    assert str(d["generator"]) == text, "fixture %s was written from %s" % (name, d["generator"])
    return make(), d


def interior(a):
    return a[..., NG:-NG, NG:-NG, NG:-NG]


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _uses_pow(case):
    return any(r.bctype == 4 for r in case.regions)


class _Checker:
    def __init__(self, exact):
        self.exact = exact

    def fields(self, got, want, what):
        if self.exact:
            bad = int((got != want).sum())
            assert bad == 0, "%s: %d values differ, max rel %.3e" % (what, bad, _rel(got, want))
        else:
            assert _rel(got, want) < 1e-10, "%s: rel %.3e" % (what, _rel(got, want))

    def history(self, rms, mx, loc, hist, rms_exact=True):
        if self.exact:
            if rms_exact:
                assert np.array_equal(rms, hist[:, 0]), np.abs(rms / hist[:, 0] - 1).max()
            else:
                # a parallel sum of the same terms in another order
                assert np.all(np.abs(rms - hist[:, 0]) <= 1e-13 * hist[:, 0]), np.abs(rms / hist[:, 0] - 1).max()
            assert np.array_equal(mx, hist[:, 1])
            assert np.array_equal(loc, hist[:, 2:7].astype(np.int64))
        else:
            # the per-step residual (RMS) to 1e-12; the pointwise maximum
            # carries the drift of single nodes next to far-field ghosts
            # (about 2e-12 after 200 steps) and is held to the field bar
            assert np.all(np.abs(rms - hist[:, 0]) <= 1e-12 * hist[:, 0]), np.abs(rms / hist[:, 0] - 1).max()
            assert np.all(np.abs(mx - hist[:, 1]) <= 1e-10 * hist[:, 1]), np.abs(mx / hist[:, 1] - 1).max()


# ---------------------------------------------------------------- runners

class _OracleRun:
    def __init__(self, case):
        self.o = OracleCase(case, threads=4)

    def op(self, name):
        self.o.op(name)

    def field(self, b, name):
        return self.o.field(b, name)

    def steps(self, n):
        out = [self.o.step() for _ in range(n)]
        return (np.array([r[0] for r in out]), np.array([r[1] for r in out]),
                np.array([r[2] for r in out], dtype=np.int64))


class _GpuRun:
    def __init__(self, case):
        from orion_amd.solver import OrionSolver
        self.s = OrionSolver(case)

    def op(self, name):
        from tests.test_gpu_parity import _run_op
        _run_op(self.s.fs, name)

    def field(self, b, name):
        return self.s.fs.field(b, name)

    def steps(self, n):
        h = self.s.solve(n)
        return (np.array([r.rms_residual for r in h]), np.array([r.max_residual for r in h]),
                np.array([r.max_loc for r in h], dtype=np.int64))


# ---------------------------------------------------------------- checks

def check_stages(runner_cls):
    case, d = load("stages_17x9x7")
    r = runner_cls(case)
    chk = _Checker(exact=False) if runner_cls is _GpuRun and _uses_pow(case) else _Checker(exact=True)
    for stage, op in STAGE_OPS:
        r.op(op)
        for key in ("prim", "dq", "c", "mu", "dt"):
            chk.fields(interior(r.field(0, key)), d["%s_%s" % (stage, key)], "%s %s" % (stage, key))
        if stage in ("halo", "bc"):
            # ghost layers too; corners of the padded box are never written
            chk.fields(r.field(0, "prim"), d["%s_prim_full" % stage], "%s prim with ghosts" % stage)


def check_channel(runner_cls):
    case, d = load("channel_65x33x9")
    r = runner_cls(case)
    chk = _Checker(exact=runner_cls is _OracleRun or not _uses_pow(case))
    rms, mx, loc = r.steps(200)
    chk.history(rms, mx, loc, d["hist"], rms_exact=runner_cls is _OracleRun)
    prim = interior(r.field(0, "prim"))
    dq = interior(r.field(0, "dq"))
    k = int(d["kmid"])
    chk.fields(prim[:, k], d["prim_kmid"], "prim k-plane")
    chk.fields(dq[:, k], d["dq_kmid"], "dq k-plane")
    chk.fields(prim.sum(axis=(1, 2, 3)), d["prim_sum"], "prim sums")
    chk.fields(np.abs(prim).sum(axis=(1, 2, 3)), d["prim_abs_sum"], "|prim| sums")
    chk.fields(np.abs(dq).sum(axis=(1, 2, 3)), d["dq_abs_sum"], "|dq| sums")


def check_run(runner_cls, name, nsteps):
    case, d = load(name)
    r = runner_cls(case)
    chk = _Checker(exact=runner_cls is _OracleRun or not _uses_pow(case))
    rms, mx, loc = r.steps(nsteps)
    chk.history(rms, mx, loc, d["hist"], rms_exact=runner_cls is _OracleRun)
    for b in range(case.nblocks):
        sfx = "_%d" % b if case.nblocks > 1 else ""
        chk.fields(interior(r.field(b, "prim")), d["prim" + sfx], "block %d prim" % b)
        chk.fields(interior(r.field(b, "dq")), d["dq" + sfx], "block %d dq" % b)


def test_fixture_inputs_regenerate():
    """Each fixture's generator call rebuilds a case of the recorded size."""
    case, d = load("channel_65x33x9")
    assert tuple(case.dims[0]) == (65, 33, 9) and d["hist"].shape == (200, 8)
    case, d = load("lattice_small")
    assert case.nblocks == 4 and all(("prim_%d" % b) in d.files for b in range(4))


def test_oracle_stages_golden():
    check_stages(_OracleRun)


def test_oracle_channel_golden():
    check_channel(_OracleRun)


@pytest.mark.parametrize("name", ["lattice_small", "bc_variety"])
def test_oracle_run_golden(name):
    check_run(_OracleRun, name, 10)


@pytest.mark.gpu
def test_gpu_stages_golden():
    print("\n[gpu] golden stages 17x9x7", flush=True)
    check_stages(_GpuRun)


@pytest.mark.gpu
def test_gpu_channel_golden():
    print("\n[gpu] golden channel 65x33x9, 200 steps", flush=True)
    check_channel(_GpuRun)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lattice_small", "bc_variety"])
def test_gpu_run_golden(name):
    print("\n[gpu] golden run", name, flush=True)
    check_run(_GpuRun, name, 10)
