"""Solution output (postprocess::PostProcess) and restart.

The VTS/VTM files are held to the reference's bytes: the committed fixture
tests/golden/vts_bc_variety.npz holds what the reference's own
PostProcess::write_solution (src/postprocess/PostProcess.cpp:44-283) wrote
after 3 steps of the viscous bc_variety block (tests/golden/make_golden.py),
and where the reference build exists (this container) a multi-block case is
compared against it live.  The CPU tests feed the oracle's fields
(ob_output_fields, a restatement of the extraction loops at :174-240) to the
library's host-only writers; the GPU tests write from the device state.

Restart has no counterpart in the reference (it always starts from the
freestream): a state saved after n steps and loaded into a fresh context
must continue bit for bit like the run that saved it.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from oracle import binding
from oracle.binding import OracleCase
from tests import _cases

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "vts_bc_variety.npz")


def _lib():
    from orion_amd import _lib as L
    return L.load()


def _fixture_files():
    d = np.load(FIXTURE)
    return {str(n): bytes(d["file_%d" % i]) for i, n in enumerate(d["names"])}


def _write_from_fields(case, fields_of_block, outdir, step):
    """The library's host-only writers fed with given fields, named as
    write_solution names them for flowname 'output/' (empty stem)."""
    L = _lib()
    off = 0
    xyz = np.ascontiguousarray(case.xyz_flat(), dtype=np.float64)
    for b in range(case.nblocks):
        i, j, k = (int(v) for v in case.dims[b])
        n = i * j * k
        f = np.ascontiguousarray(fields_of_block(b), dtype=np.float32)
        assert f.size == 7 * n
        path = os.path.join(outdir, "_%06d_block_%d.vts" % (step, b)).encode()
        P = C.POINTER(C.c_double)
        rc = L.orion_hip_write_vts(path, i, j, k, xyz[off:].ctypes.data_as(P), xyz[off + n:].ctypes.data_as(P),
                                   xyz[off + 2 * n:].ctypes.data_as(P), f.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0, L.orion_hip_last_error()
        off += 3 * n
    assert L.orion_hip_write_vtm(outdir.encode(), b"", case.nblocks, step) == 0, L.orion_hip_last_error()


def _read_dir(d):
    return {n: open(os.path.join(d, n), "rb").read() for n in sorted(os.listdir(d))}


def test_vts_from_oracle_fields_matches_reference_fixture(tmp_path):
    case = _cases.bc_variety(nvis=1, nsteps=3)
    o = OracleCase(case)
    o.run(3)
    _write_from_fields(case, o.output_fields, str(tmp_path), 3)
    want = _fixture_files()
    got = _read_dir(str(tmp_path))
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name] == want[name], name


@pytest.mark.skipif(not binding.have_reference(), reason="reference build (oracle/_ref) not present")
def test_vts_multiblock_matches_live_reference(tmp_path):
    """Four blocks with interfaces, wall and far-field: the reference writes
    its files, the oracle's fields go through the library's writers."""
    case = _cases.lattice_small(nsteps=3)
    casedir = str(tmp_path / "case")
    binding.run_reference(case, casedir, str(tmp_path / "ref"), 3, vts=True)
    o = OracleCase(case)
    o.run(3)
    mine = tmp_path / "mine"
    mine.mkdir()
    _write_from_fields(case, o.output_fields, str(mine), 3)
    want = _read_dir(os.path.join(casedir, "output"))
    got = _read_dir(str(mine))
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name] == want[name], name


def test_vts_layout_decodes_to_its_inputs(tmp_path):
    """The file's own structure (independent of the reference): the header's
    offsets address UInt32-counted float32 arrays that decode to the points
    and fields written."""
    case = _cases.bc_variety(nvis=1, nsteps=1)
    rng = np.random.default_rng(3)
    i, j, k = (int(v) for v in case.dims[0])
    n = i * j * k
    fields = rng.standard_normal(7 * n).astype(np.float32)
    _write_from_fields(case, lambda b: fields, str(tmp_path), 12)
    raw = open(tmp_path / "_000012_block_0.vts", "rb").read()
    head, data = raw.split(b"<AppendedData encoding=\"raw\">\n    _", 1)
    text = head.decode()
    assert 'WholeExtent="0 %d 0 %d 0 %d"' % (i - 1, j - 1, k - 1) in text
    offs = [int(t.split('"')[0]) for t in text.split('offset="')[1:]]
    assert len(offs) == 6
    sizes = [3 * n, n, n, n, n, 3 * n]
    arrays = []
    for o_, m in zip(offs, sizes):
        (nbytes,) = struct.unpack_from("<I", data, o_)
        assert nbytes == 4 * m
        arrays.append(np.frombuffer(data, dtype="<f4", count=m, offset=o_ + 4))
    assert data.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
    x, y, z = np.split(np.asarray(case.xyz_flat(), dtype=np.float64), 3)
    pts = arrays[0].reshape(n, 3)
    assert np.array_equal(pts[:, 0], x.astype(np.float32)) and np.array_equal(pts[:, 2], z.astype(np.float32))
    for a in range(4):
        assert np.array_equal(arrays[1 + a], fields[a * n:(a + 1) * n])
    assert np.array_equal(arrays[5], fields[4 * n:])
    vtm = open(tmp_path / "_000012.vtm").read()
    assert '<DataSet index="0" file="_000012_block_0.vts"/>' in vtm


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

@pytest.mark.gpu
def test_device_output_matches_reference_fixture(tmp_path):
    from orion_amd.solver import OrionSolver, PostProcess
    case = _cases.bc_variety(nvis=1, nsteps=3)
    print("\n[gpu] output fields and write_solution, bc_variety 3 steps", flush=True)
    s = OrionSolver(case)
    s.solve(3)
    o = OracleCase(case)
    o.run(3)
    assert np.array_equal(s.fs.output_fields(0), o.output_fields(0))
    s.params.filename.flowname = str(tmp_path / "out") + "/"
    PostProcess.write_solution(s.fs, case, s.params, 3)
    got = _read_dir(str(tmp_path / "out"))
    want = _fixture_files()
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name] == want[name], name
    s.fs.close()


@pytest.mark.gpu
def test_restart_continues_bit_for_bit(tmp_path):
    from orion_amd import _lib
    from orion_amd.solver import OrionSolver
    case = _cases.lattice_small(nsteps=5)
    print("\n[gpu] restart: lattice_small 3 + 2 steps", flush=True)
    a = OrionSolver(case)
    a.solve(3)
    path = str(tmp_path / "state.bin")
    a.fs.save_state(path, 3)
    ha = a.solve(2)
    b = OrionSolver(case)
    assert b.fs.load_state(path) == 3
    hb = b.solve(2)
    assert [h.rms_residual for h in ha] == [h.rms_residual for h in hb]
    assert [h.max_loc for h in ha] == [h.max_loc for h in hb]
    for blk in range(case.nblocks):
        assert np.array_equal(a.fs.field(blk, "prim"), b.fs.field(blk, "prim")), blk
    # a state of another grid is refused
    other = OrionSolver(_cases.short_duct(nsteps=1))
    with pytest.raises(_lib.OrionHipError):
        other.fs.load_state(path)
    for s in (a, b, other):
        s.fs.close()
