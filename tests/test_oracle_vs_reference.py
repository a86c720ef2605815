"""The oracle against the reference solver compiled from its own sources
(oracle/_ref/ref_harness, built by ``make -C oracle ref`` where the
reference tree is present): every stage of the first steps and a residual
history, bit for bit, on cases covering every boundary type, interfaces,
viscous and inviscid flow and supersonic inflow.

Skipped where the reference build is absent; the committed fixtures in
tests/golden/ (test_golden.py) carry the same check there.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import binding
from oracle.binding import OracleCase
from tests import _cases

pytestmark = pytest.mark.skipif(not binding.have_reference(), reason="reference build (oracle/_ref) not present")

STAGE_OPS = [("halo", "exchange_bc"), ("bc", "apply_physical_bc"), ("state", "update_flow_states"),
             ("dt", "calculate_time_step"), ("visc", "compute_viscous_rhs"),
             ("avg1", "average_interface_residuals"), ("inv", "compute_inviscid_rhs"),
             ("avg2", "average_interface_residuals"), ("upd", "update_solution")]


@pytest.mark.parametrize("name", sorted(_cases.ALL))
def test_stages_bit_exact(name):
    case = _cases.ALL[name](nsteps=2)
    tmp = tempfile.mkdtemp(prefix="orion_ref_")
    _, read = binding.run_reference(case, os.path.join(tmp, "case"), os.path.join(tmp, "out"), 2, dump_steps=2)
    o = OracleCase(case)
    for step in (1, 2):
        for stage, op in STAGE_OPS:
            if stage in ("visc", "avg1") and case.inputs["nvis"] != 1:
                continue  # OrionSolver.cpp:149-153
            o.op(op)
            ref = read("s%d_%s.bin" % (step, stage))
            for b in range(case.nblocks):
                for key in ("prim", "dq", "c", "mu", "dt"):
                    got, want = o.field(b, key), ref[b][key]
                    bad = int((got != want).sum())
                    assert bad == 0, "%s step %d %s block %d %s: %d values differ" % (name, step, stage, b, key, bad)


@pytest.mark.parametrize("name", ["channel_small", "lattice_small", "bc_variety_visc", "columns_small",
                                  "rotated_pair"])
def test_history_bit_exact(name):
    n = 25
    case = _cases.ALL[name](nsteps=n)
    tmp = tempfile.mkdtemp(prefix="orion_ref_")
    hist, read = binding.run_reference(case, os.path.join(tmp, "case"), os.path.join(tmp, "out"), n)
    o = OracleCase(case)
    got = np.array([o.step()[:2] for _ in range(n)])
    assert np.array_equal(got[:, 0], hist[:, 0])
    assert np.array_equal(got[:, 1], hist[:, 1])
    fin = read("final.bin")
    for b in range(case.nblocks):
        assert np.array_equal(o.field(b, "prim"), fin[b]["prim"])
