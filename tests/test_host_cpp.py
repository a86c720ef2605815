"""The C++ host mirror (orion_amd/host) and its orion_run executable: case
files in the reference's formats are read natively, and the step runs from
them match the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle.binding import OracleCase
from orion_amd import build as ob
from tests import _cases


def _runner():
    if not os.path.exists(ob.RUNNER):
        pytest.skip("orion_run not built (python orion_amd/build.py)")
    return ob.RUNNER


def _check(casedir):
    r = subprocess.run([_runner(), casedir, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@pytest.mark.parametrize("name", ["channel_small", "lattice_small", "bc_variety_visc", "isothermal_wall"])
def test_reads_case_files(name, tmp_path):
    case = _cases.ALL[name](nsteps=11)
    case.write(str(tmp_path))
    got = _check(str(tmp_path))
    assert got["blocks"] == case.nblocks
    assert got["nodes"] == case.interior_nodes()
    assert got["regions"] == len(case.regions)
    assert got["steps"] == 11
    for k in ("moo", "reynolds", "tref", "cfl"):
        assert got[k] == case.inputs[k], k
    assert got["nvis"] == case.inputs["nvis"] and got["gcl"] == case.inputs["gcl"]


def test_namelist_forms(tmp_path):
    """Fortran forms the reference's reader accepts (Params.cpp:12-100,
    :238-268): '&' groups, comma-separated items, '!' comments, 'd'
    exponents, quoted strings.  The $force lengths do not scale the
    Reynolds number: init_inflow scales by params.inflow.lref / lfref
    (Inflow.cpp:323-325), which the namelist never sets (the $force values
    go to params.force, Params.cpp:150-153), so by 1 -- checked against
    the reference build where it is present."""
    case = _cases.channel_small(nsteps=3)
    case.write(str(tmp_path))
    text = open(tmp_path / "param.dat").read()
    text = text.replace("$inflow", "&INFLOW  ! comment").replace("reynolds = 5000.0", "reynolds = 5.0d3, attack = 0.0")
    text = text.replace("lfref = 1.0", "lfref = 2.0").replace("lref = 1.0", "lref = 0.5")
    text = text.replace("gridname = 'grid.x'", "gridname = \"grid.x\"  ! quoted")
    open(tmp_path / "param.dat", "w").write(text)
    got = _check(str(tmp_path))
    assert got["reynolds"] == 5.0e3
    from oracle import binding
    if binding.have_reference():
        out = tmp_path / "ref"
        out.mkdir()
        env = dict(os.environ)
        env.pop("LD_LIBRARY_PATH", None)
        subprocess.run([binding.REF_HARNESS, str(tmp_path), "0", str(out), "0"], check=True,
                       stdout=subprocess.DEVNULL, env=env)
        assert np.fromfile(str(out / "params.bin"))[1] == got["reynolds"]


def test_rejects_what_the_reference_rejects(tmp_path):
    case = _cases.channel_small(nsteps=3)
    case.write(str(tmp_path))
    bc = open(tmp_path / "grid.inp").read()
    open(tmp_path / "grid.inp", "w").write(bc.replace("1976", "1975", 1))
    r = subprocess.run([_runner(), str(tmp_path), "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "marker" in r.stderr
    case.write(str(tmp_path))
    p = open(tmp_path / "param.dat").read()
    # pressure, temperature and density all given (Inflow.cpp:367-370)
    open(tmp_path / "param.dat", "w").write(p.replace("pref = -1.0", "pref = 101325.0"))
    r = subprocess.run([_runner(), str(tmp_path), "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot all be given" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["graph", "ops"])
def test_orion_run_matches_oracle(mode, tmp_path):
    """orion_run from case files (with the seeded init.bin) against the
    oracle on the same case: bit-exact residual maxima and locations, RMS to
    a parallel summation order (no far-field boundaries in this case)."""
    n = 15
    case = _cases.bc_variety(nvis=1, nsteps=n)
    case.write(str(tmp_path))
    hist = str(tmp_path / "hist.bin")
    print("\n[gpu] orion_run", mode, flush=True)
    cmd = [_runner(), str(tmp_path), str(n), "--hist", hist] + (["--ops"] if mode == "ops" else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h = np.fromfile(hist).reshape(n, 8)
    o = OracleCase(case)
    ref = [o.step() for _ in range(n)]
    rms = np.array([x[0] for x in ref])
    assert np.all(np.abs(h[:, 0] - rms) <= 1e-13 * rms)
    assert np.array_equal(h[:, 1], np.array([x[1] for x in ref]))
    assert np.array_equal(h[:, 2:7].astype(int), np.array([x[2] for x in ref]))
