"""The LU-SGS tile launch orders (a launch per sweep, or the four sweeps
chained in one launch), their autotuner, and the plain per-block launch
form.

Every schedule launches a tile after the tiles it waits for, and each
cell's arithmetic does not depend on the order, so the results must be the
same bit for bit under every schedule; the tuner runs trial steps and must
leave the solution state as it found it.
"""
import numpy as np
import pytest

from orion_amd import cases
from tests import _cases


def _case():
    # 64 blocks of 17-26 nodes per axis, up to four tiles per block: the orders differ
    return cases.uneven_lattice_case(seed=11, nsteps=4, target=6.0e5, lo=17, hi=26)


@pytest.mark.gpu
def test_schedules_give_identical_results():
    from orion_amd.solver import OrionSolver
    case = _case()
    print("\n[gpu] sweep schedules: %d blocks, %d nodes" % (case.nblocks, case.interior_nodes()), flush=True)
    runs = []
    for k in (0, 1, 2):
        s = OrionSolver(case)
        s.fs.set_schedule(k)
        assert s.fs.schedule() == k
        h = s.solve(4)
        runs.append(([x.rms_residual for x in h], [x.max_loc for x in h],
                     [s.fs.field(b, "prim") for b in range(case.nblocks)]))
        s.fs.close()
    for k in (1, 2):
        assert runs[0][0] == runs[k][0] and runs[0][1] == runs[k][1], k
        for a, b in zip(runs[0][2], runs[k][2]):
            assert np.array_equal(a, b), k


@pytest.mark.gpu
def test_tuning_restores_the_state():
    from orion_amd.solver import OrionSolver
    case = _case()
    print("\n[gpu] sweep schedule tuning", flush=True)
    a = OrionSolver(case)
    a.solve(2)
    before = [a.fs.field(b, "prim") for b in range(case.nblocks)]
    chosen, ms = a.fs.tune_schedule(2)
    assert chosen in (0, 1, 2) and a.fs.schedule() == chosen and all(m > 0 for m in ms)
    for b in range(case.nblocks):
        assert np.array_equal(a.fs.field(b, "prim"), before[b]), b
    # and the solve goes on as one that was never tuned
    ref = OrionSolver(case)
    ref.solve(2)
    ha, hr = a.solve(2), ref.solve(2)
    assert [x.rms_residual for x in ha] == [x.rms_residual for x in hr]
    for s in (a, ref):
        s.fs.close()


def test_uneven_sizes_search_is_bounded():
    sx, sy, sz, _ = cases.uneven_sizes(11, lo=17, hi=26, target=6.0e5)
    assert abs(sum(sx) * sum(sy) * sum(sz) - 6.0e5) <= 0.02 * 6.0e5
    assert min(sx + sy + sz) >= 17 and max(sx + sy + sz) <= 26
    with pytest.raises(ValueError):  # unreachable from sizes 17-129
        cases.uneven_sizes(11, target=6.0e5, tries=500)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["columns_small", "lattice_small", "bc_variety_inviscid"])
def test_per_block_form_equals_batched(name):
    """The plain per-block form (a launch per block per kernel, no captured
    step) and the batched form (all blocks per launch, captured step) give
    the same residual histories and fields bit for bit, from the same state
    and through a switch mid-run."""
    from orion_amd.solver import OrionSolver
    case = _cases.ALL[name](nsteps=6)
    print("\n[gpu] per-block vs batched:", name, flush=True)
    a, b = OrionSolver(case), OrionSolver(case)
    b.fs.set_per_block(True)
    ha, hb = a.solve(3), b.solve(3)
    b.fs.set_per_block(False)
    a.fs.set_per_block(True)
    ha += a.solve(3)
    hb += b.solve(3)
    assert [h.rms_residual for h in ha] == [h.rms_residual for h in hb]
    assert [h.max_loc for h in ha] == [h.max_loc for h in hb]
    for blk in range(case.nblocks):
        for f in ("prim", "dq", "dt"):
            assert np.array_equal(a.fs.field(blk, f), b.fs.field(blk, f)), (blk, f)
    a.fs.close()
    b.fs.close()


@pytest.mark.gpu
def test_viscous_tiling_gives_identical_results():
    """The viscous face-flux work queue and the fused viscous form under
    other tilings (faces or planes per item, persistent grid) give the same
    histories and fields bit for bit, and the launch-shape tuner leaves the
    state as it found it."""
    from orion_amd.solver import OrionSolver
    case = _cases.columns_small(nsteps=4)
    print("\n[gpu] viscous tilings", flush=True)
    ref = OrionSolver(case)
    hr = ref.solve(4)
    for fused, seg, grid in ((False, 8, 64), (False, 64, 4096), (False, 1, 1), (True, 1, 1), (True, 5, 7),
                             (True, 64, 4096)):
        s = OrionSolver(case)
        s.fs.set_visc_form(fused)
        s.fs.set_visc_tiles(seg, grid)
        assert s.fs.visc_tiles() == (seg, grid)
        hs = s.solve(4)
        assert [h.rms_residual for h in hs] == [h.rms_residual for h in hr], (fused, seg, grid)
        for b in range(case.nblocks):
            assert np.array_equal(s.fs.field(b, "prim"), ref.fs.field(b, "prim")), (fused, seg, grid, b)
        s.fs.close()
    t = OrionSolver(case)
    t.solve(2)
    before = [t.fs.field(b, "prim") for b in range(case.nblocks)]
    res = t.fs.tune(1)
    assert res["trials"] >= 2
    for b in range(case.nblocks):
        assert np.array_equal(t.fs.field(b, "prim"), before[b])
    t.fs.close()
    ref.fs.close()


VISCOUS = [n for n in sorted(_cases.ALL) if n not in ("bc_variety_inviscid", "lattice_supersonic")]


@pytest.mark.gpu
@pytest.mark.parametrize("name", VISCOUS)
def test_fused_viscous_form_equals_face_queue(name):
    """The fused viscous form (face fluxes of the three directions and their
    difference in one pass) and the face-flux queue with its difference
    kernel give the same residual histories and fields bit for bit: in the
    captured step, through a switch mid-run, and operator by operator (the
    accumulating, not-first form of the kernel)."""
    from orion_amd.solver import OrionSolver
    case = _cases.ALL[name](nsteps=6)
    if int(case.inputs["nvis"]) != 1:
        pytest.skip("inviscid case")
    print("\n[gpu] fused viscous vs face queue:", name, flush=True)
    a, b = OrionSolver(case), OrionSolver(case)
    assert not a.fs.visc_form()
    a.fs.set_visc_form(True)
    assert a.fs.visc_form()
    b.fs.set_visc_form(False)
    assert not b.fs.visc_form()
    ha, hb = a.solve(3), b.solve(3)
    b.fs.set_visc_form(True)
    a.fs.set_visc_form(False)
    ha += a.solve(3)
    hb += b.solve(3)
    assert [h.rms_residual for h in ha] == [h.rms_residual for h in hb]
    assert [h.max_loc for h in ha] == [h.max_loc for h in hb]
    for blk in range(case.nblocks):
        for f in ("prim", "dq", "dt"):
            assert np.array_equal(a.fs.field(blk, f), b.fs.field(blk, f)), (blk, f)
    # one viscous operator call on top of a non-zero dq (the inviscid one
    # first), in both forms, on the same state
    from orion_amd.solver import FluxComputer, InviscidFluxComputer
    for s, fused in ((a, True), (b, False)):
        s.fs.set_visc_form(fused)
        InviscidFluxComputer.compute_inviscid_rhs(s.fs, s.params)
        FluxComputer.compute_viscous_rhs(s.fs, s.params)
    for blk in range(case.nblocks):
        assert np.array_equal(a.fs.field(blk, "dq"), b.fs.field(blk, "dq")), blk
    a.fs.close()
    b.fs.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rotated_pair", "columns_small", "bc_variety_visc"])
def test_halo_field_pairs_equal_whole_points(name, monkeypatch):
    """The halo moves whole points (five fields per entry) for blocks whose
    interface windows pair equal point counts, and one field value per
    entry otherwise; a context built with ORION_HIP_HALO_FIELDS=1 keeps
    every block on the per-field pairs.  Both give the same residual
    histories and fields bit for bit."""
    from orion_amd.solver import OrionSolver
    case = _cases.ALL[name](nsteps=4)
    print("\n[gpu] halo per-field pairs vs whole points:", name, flush=True)
    a = OrionSolver(case)
    monkeypatch.setenv("ORION_HIP_HALO_FIELDS", "1")
    b = OrionSolver(case)
    monkeypatch.delenv("ORION_HIP_HALO_FIELDS")
    # the per-field list holds five entries per ghost point: on the 12-block
    # mesh both lists exceed the small-allocation pool, so the context that
    # keeps them is larger (the switch took effect)
    assert b.fs.device_bytes() >= a.fs.device_bytes()
    if name == "columns_small":
        assert b.fs.device_bytes() > a.fs.device_bytes()
    ha, hb = a.solve(4), b.solve(4)
    assert [h.rms_residual for h in ha] == [h.rms_residual for h in hb]
    assert [h.max_loc for h in ha] == [h.max_loc for h in hb]
    for blk in range(case.nblocks):
        for f in ("prim", "dq"):
            assert np.array_equal(a.fs.field(blk, f), b.fs.field(blk, f)), (blk, f)
    a.fs.close()
    b.fs.close()
