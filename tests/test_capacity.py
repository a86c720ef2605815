"""The capacity workload at full size: 48 blocks of 257x129x21 (33.4M
nodes), transonic channel over a bump, CFL 20.  Properties checked on the
GPU: its first steps are finite; they are the same bit for bit whether it
runs alone or beside a serving batch working on the device at the same time
(each context's arithmetic is its own); and the LU-SGS tile schedule does
not change them."""
import hashlib
import threading

import numpy as np
import pytest

from orion_amd import cases, serving

pytestmark = pytest.mark.gpu


def _digest(fs, nblocks):
    h = hashlib.sha256()
    finite = True
    for b in range(nblocks):
        a = fs.field(b, "prim")
        finite = finite and bool(np.all(np.isfinite(a)))
        h.update(a.tobytes())
    return h.hexdigest(), finite


def test_capacity_case_alone_beside_a_batch_and_under_both_schedules():
    from orion_amd.solver import OrionSolver
    cap = cases.capacity_case(nsteps=3)
    print("\n[gpu] capacity case %s, %d nodes" % (cap.name, cap.interior_nodes()), flush=True)
    a = OrionSolver(cap)
    ha = a.solve(2)
    da, fa = _digest(a.fs, cap.nblocks)
    assert fa, "non-finite state after two steps"
    assert all(np.isfinite(h.rms_residual) for h in ha)
    small = [cases.bump_request(21, 3.0e5, 4, 0.7, 1.0e5, 0.05, True, nsteps=40),
             cases.bump_request(22, 1.0e5, 1, 0.5, 1.0e4, 0.02, False, nsteps=40)]
    batch = serving.Batch(small)
    done = threading.Event()

    def busy():
        while not done.is_set():
            batch.run(10)

    th = threading.Thread(target=busy)
    th.start()
    try:
        b = OrionSolver(cap)
        hb = b.solve(2)
    finally:
        done.set()
        th.join()
        batch.close()
    assert [h.rms_residual for h in ha] == [h.rms_residual for h in hb]
    assert [h.max_loc for h in ha] == [h.max_loc for h in hb]
    assert _digest(b.fs, cap.nblocks)[0] == da
    a.fs.set_schedule(0)
    b.fs.set_schedule(1)
    ha, hb = a.solve(1), b.solve(1)
    assert ha[0].rms_residual == hb[0].rms_residual and ha[0].max_loc == hb[0].max_loc
    assert _digest(a.fs, cap.nblocks)[0] == _digest(b.fs, cap.nblocks)[0]
    a.fs.close()
    b.fs.close()
