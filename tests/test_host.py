"""Host-side checks that need no GPU: the C ABI library loads and exports
every function include/orion_hip.h declares, the product path refuses to
run without it, the case files round-trip, and bench.py refuses N>1."""
import os
import subprocess
import sys

import numpy as np
import pytest

from orion_amd import _lib, cases
from tests import _cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _built():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liborion_hip.so not built (python orion_amd/build.py)")


def test_header_declares_operator_surface():
    names = _lib.declared_symbols()
    for op in ("exchange_bc", "apply_physical_bc", "update_flow_states", "calculate_time_step",
               "compute_viscous_rhs", "average_interface_residuals", "compute_inviscid_rhs", "update_solution",
               "residual", "backend_init", "ctx_create", "ctx_destroy", "last_error"):
        assert "orion_hip_" + op in names, op


def test_library_exports_every_declared_symbol():
    _built()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in _lib.declared_symbols() if s not in exported]
    assert not missing, missing
    # exported C symbols are unmangled and nothing else leaks under the prefix
    extra = sorted(s for s in exported if s.startswith("orion_hip_") and s not in _lib.declared_symbols())
    assert not extra, extra


def test_ctypes_table_covers_header():
    """load() gives every declared function an explicit signature (loading
    the library touches no device)."""
    _built()
    L = _lib.load()
    for s in _lib.declared_symbols():
        fn = getattr(L, s)
        assert fn.argtypes is not None, s
    assert L.orion_hip_last_error() is not None


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.OrionHipError, match="not built"):
        _lib.load()


def test_plot3d_round_trip(tmp_path):
    case = _cases.lattice_small()
    path = str(tmp_path / "grid.grd")
    cases.write_plot3d(path, case)
    dims, xyz = cases.read_plot3d(path)
    assert np.array_equal(dims, case.dims)
    for b in range(case.nblocks):
        assert np.array_equal(xyz[b], case.xyz[b])


def test_region_table_layout():
    """16 ints per region, in the order of orion_region_desc."""
    case = _cases.bc_variety()
    t = case.region_table()
    assert t.shape == (len(case.regions), 16) and t.dtype == np.int32
    r = case.regions[4]
    assert list(t[4, :2]) == [r.block, r.bctype]
    assert list(t[4, 2:5]) == list(r.s_st) and list(t[4, 5:8]) == list(r.s_ed)
    assert list(t[4, 8:11]) == list(r.t_st) and list(t[4, 11:14]) == list(r.t_ed)
    assert list(t[4, 14:]) == [r.nbt, r.ibcwin]


def test_lattice_interfaces_pair_up():
    """Every interface window has a partner window of the same node count
    pointing back at it."""
    case = cases.lattice_case([13, 11], [9, 12], [10], seed=5)
    ifs = [r for r in case.regions if r.bctype < 0]
    assert ifs
    for r in ifs:
        n_s = np.prod([abs(e - s) + 1 for s, e in zip(r.s_st, r.s_ed)])
        n_t = np.prod([abs(e - s) + 1 for s, e in zip(r.t_st, r.t_ed)])
        assert n_s == n_t
        back = [q for q in case.regions if q.block == r.nbt - 1 and q.bctype < 0 and tuple(q.s_st) == tuple(r.t_st)
                and tuple(q.t_st) == tuple(r.s_st)]
        assert len(back) == 1


def test_uneven_workload_size():
    case = cases.uneven_lattice_case(seed=11, nsteps=1)
    assert case.nblocks == 64
    assert 7.0e6 < case.interior_nodes() < 9.0e6
    d = np.asarray(case.dims)
    assert d.min() >= 17 and d.max() <= 129


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_stage_kernels_are_in_the_traffic_summary():
    """Every kernel the bench's stages and its dominant entry name appears in
    the committed counter summary of the default workload, so no stage
    silently loses its traffic."""
    sys.path.insert(0, ROOT)
    import bench
    groups = {k: v[2].split("+") for k, v in bench.STAGES.items()}
    groups["dominant"] = bench.DOMINANT[3]
    t = bench.stage_traffic(bench.DEFAULT_WORKLOAD, groups)
    assert sorted(t) == sorted(groups), sorted(set(groups) - set(t))
    for name, v in t.items():
        assert v["bytes_per_launch"] > 0, name


def test_new_entry_points_refuse_a_null_context():
    """The entry points without a reference counterpart check their context
    before any device work, so a null one fails with a message (no GPU)."""
    import ctypes as C
    from orion_amd import _lib
    L = _lib.load()
    ms = (C.c_double * 2)()
    chosen = C.c_int()
    for name, call in [
        ("set_schedule", lambda: L.orion_hip_set_schedule(None, 0)),
        ("tune_schedule", lambda: L.orion_hip_tune_schedule(None, 3, C.byref(chosen), ms)),
        ("set_overlap", lambda: L.orion_hip_set_overlap(None, 1)),
        ("copy_block_state", lambda: L.orion_hip_copy_block_state(None, 0, None, 0)),
    ]:
        assert call() != 0, name
        assert L.orion_hip_last_error().decode(), name
    assert L.orion_hip_get_schedule(None) == -1 and L.orion_hip_requests(None) == 0
