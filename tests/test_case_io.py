"""Case input formats beyond the binary grid and the custom reference state:
the ASCII Plot3D form the reference falls back to
(src/mesh/Plot3DReader.cpp:162-224) and the reference state of
preprocess::init_inflow (src/preprocess/Inflow.cpp:54-415) -- the standard
atmosphere at a height and the custom combinations of pressure,
temperature and density -- read by the C++ host mirror (orion_run --check,
no device use).

Pins: the reference build where present (its derived values written by
ref_harness to params.bin, and its run on an ASCII grid), and everywhere the
committed fixture tests/golden/inflow_reference.npz that
tests/golden/make_golden.py writes from it.  The GPU test runs orion_run on
such cases against the oracle given the derived inputs."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import binding
from tests import _cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "orion_amd", "_build", "orion_run")
GOLDEN = os.path.join(ROOT, "tests", "golden", "inflow_reference.npz")

# heights (km) in every layer of the standard atmosphere, on layer tops,
# and above its last layer; then the custom combinations (height < 0)
HEIGHTS = [0.0, 3.5, 11.019, 15.0, 25.0, 40.0, 50.0, 55.0, 70.0, 85.0, 95.0, 105.0, 115.0, 130.0, 155.0, 165.0,
           180.0, 200.0, 260.0, 350.0, 450.0, 550.0, 650.0, 720.0]
CUSTOM = [dict(tref=288.15, pref=-1.0, rref=1.225), dict(tref=250.0, pref=5.0e4, rref=-1.0),
          dict(tref=-1.0, pref=8.0e4, rref=1.0)]


def inflow_cases():
    """(label, inputs) of every reference-state case: heights, then the
    custom combinations."""
    out = [("height %g" % h, dict(height=h, moo=0.6, reynolds=2.0e3, tref=288.15)) for h in HEIGHTS]
    out += [("custom %d" % n, dict(height=-1.0, moo=0.6, reynolds=3.0e3, **c)) for n, c in enumerate(CUSTOM)]
    return out


def case_with(inputs, nsteps=1):
    c = _cases.bc_variety(nvis=1, nsteps=nsteps)
    c.inputs = dict(c.inputs, **inputs)
    return c


def check(casedir):
    r = subprocess.run([RUNNER, casedir, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def reference_params(case, tmp):
    binding.run_reference(case, os.path.join(tmp, "case"), os.path.join(tmp, "out"), 0)
    v = np.fromfile(os.path.join(tmp, "out", "params.bin"))
    return dict(moo=v[0], reynolds=v[1], tref=v[2], visc=v[3], pref=v[8], rref=v[9])


@pytest.mark.parametrize("label,inputs", inflow_cases())
def test_reference_state_matches_fixture(label, inputs, tmp_path):
    d = np.load(GOLDEN, allow_pickle=False)
    n = [str(x) for x in d["labels"]].index(label)
    c = case_with(inputs)
    c.write(str(tmp_path))
    got = check(str(tmp_path))
    for k in ("reynolds", "tref", "pref", "rref"):
        assert got[k] == d[k][n], (label, k, got[k], d[k][n])


@pytest.mark.skipif(not binding.have_reference(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("label,inputs", inflow_cases()[::3])
def test_reference_state_matches_reference(label, inputs, tmp_path):
    c = case_with(inputs)
    ref = reference_params(c, str(tmp_path / "ref"))
    c.write(str(tmp_path / "own"))
    got = check(str(tmp_path / "own"))
    for k in ("reynolds", "tref", "pref", "rref"):
        assert got[k] == ref[k], (label, k, got[k], ref[k])


def test_ascii_grid_reads_as_binary(tmp_path):
    c = _cases.lattice_small(nsteps=1)
    c.write(str(tmp_path / "bin"))
    c.write(str(tmp_path / "asc"), ascii_grid=True)
    a, b = check(str(tmp_path / "asc")), check(str(tmp_path / "bin"))
    assert a == b


@pytest.mark.skipif(not binding.have_reference(), reason="reference build (oracle/_ref) not present")
def test_reference_reads_ascii_grid(tmp_path):
    """The reference's own fallback reads the ASCII grid to the same
    doubles: two steps from it equal two steps from the binary grid."""
    c = _cases.lattice_small(nsteps=2)
    hb, rb = binding.run_reference(c, str(tmp_path / "bc"), str(tmp_path / "bo"), 2)
    c.write(str(tmp_path / "ac"), ascii_grid=True)
    env = dict(os.environ)
    env.pop("LD_LIBRARY_PATH", None)
    os.makedirs(str(tmp_path / "ao"))
    subprocess.run([binding.REF_HARNESS, str(tmp_path / "ac"), "2", str(tmp_path / "ao"), "0"], check=True,
                   stdout=subprocess.DEVNULL, env=env)
    ha = np.fromfile(str(tmp_path / "ao" / "hist.bin")).reshape(2, binding.HIST_COLS)
    assert np.array_equal(ha, hb)


@pytest.mark.gpu
@pytest.mark.parametrize("inputs", [dict(height=11.0), dict(height=-1.0, tref=-1.0, pref=8.0e4, rref=1.0)])
def test_orion_run_from_derived_reference_state(inputs, tmp_path):
    """orion_run from case files whose reference state the host derives
    (standard atmosphere; temperature from pressure and density) and with
    an ASCII grid, against the oracle given the derived inputs: maxima and
    locations bit for bit, the RMS to a parallel summation order (no
    far-field boundary in this case)."""
    from oracle.binding import OracleCase
    n = 4
    c = case_with(inputs, nsteps=n)
    c.write(str(tmp_path), ascii_grid=True)
    derived = check(str(tmp_path))
    print("\n[gpu] orion_run with", inputs, flush=True)
    hist = str(tmp_path / "hist.bin")
    r = subprocess.run([RUNNER, str(tmp_path), str(n), "--hist", hist, "--no-output"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    h = np.fromfile(hist).reshape(n, 8)
    c.inputs = dict(c.inputs, reynolds=derived["reynolds"], tref=derived["tref"])
    o = OracleCase(c)
    ref = [o.step() for _ in range(n)]
    rms = np.array([x[0] for x in ref])
    assert np.all(np.abs(h[:, 0] - rms) <= 1e-13 * rms)
    assert np.array_equal(h[:, 1], np.array([x[1] for x in ref]))
    assert np.array_equal(h[:, 2:7].astype(int), np.array([x[2] for x in ref]))
