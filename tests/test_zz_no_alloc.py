"""orion_hip_residual_noalloc -- the residual monitor over caller-held
device fields with no device allocation -- against orion_hip_residual, the
entry point whose kernels it shares, through the C ABI: two calls in a row
on one GPU, the second on a smaller input (fewer blocks, fewer nodes), each
equal to the context's own residual (same kernels, same partial order:
bit for bit).  Sizes beyond its maximum are refused before any device
work."""
import ctypes as C

import numpy as np
import pytest

from orion_amd import _lib
from tests import _cases


def _declare(L):
    L.orion_hip_residual_noalloc.restype = C.c_int
    L.orion_hip_residual_noalloc.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                             C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p,
                                             C.POINTER(_lib.ResidualStats)]
    return L


def _noalloc(L, fs, case):
    nb = case.nblocks
    dims = (C.c_int * (3 * nb))(*[int(v) for v in np.asarray(case.dims).reshape(-1)])
    ptr = {}
    for name, fid in (("dq", 4), ("dt", 5), ("vol", 2)):
        arr = (C.c_void_p * nb)()
        for b in range(nb):
            p = C.c_void_p()
            n = C.c_longlong()
            _lib.check(L.orion_hip_field(fs.ctx, b, fid, C.byref(p), C.byref(n)), "field")
            arr[b] = p.value
        ptr[name] = arr
    out = _lib.ResidualStats()
    rc = L.orion_hip_residual_noalloc(nb, dims, ptr["dq"], ptr["dt"], ptr["vol"], None, C.byref(out))
    _lib.check(rc, "residual_noalloc")
    return out


def test_refuses_sizes_beyond_its_maximum():
    L = _declare(_lib.load())
    out = _lib.ResidualStats()
    nb = 257
    dims = (C.c_int * (3 * nb))(*([9] * (3 * nb)))
    nulls = (C.c_void_p * nb)()
    assert L.orion_hip_residual_noalloc(nb, dims, nulls, nulls, nulls, None, C.byref(out)) != 0
    assert b"blocks" in L.orion_hip_last_error()
    assert L.orion_hip_residual_noalloc(0, dims, nulls, nulls, nulls, None, C.byref(out)) != 0


@pytest.mark.gpu
def test_noalloc_equals_the_context_residual():
    from orion_amd.solver import OrionSolver, ResidualMonitor
    L = _declare(_lib.load())
    for name, steps in (("lattice_small", 3), ("channel_small", 2)):  # the second input is smaller
        case = _cases.ALL[name](nsteps=steps)
        print("\n[gpu] residual_noalloc on", name, flush=True)
        s = OrionSolver(case)
        s.solve(steps)
        s.fs.synchronize()
        want = ResidualMonitor.compute(s.fs, None)
        got = _noalloc(L, s.fs, case)
        assert got.rms_residual == want.rms_residual, (name, got.rms_residual, want.rms_residual)
        assert got.max_residual == want.max_residual, name
        assert tuple(got.max_loc) == tuple(want.max_loc), (name, tuple(got.max_loc), want.max_loc)
        s.fs.close()
