"""Particle-state CSV reader / writer (Python and the C++ header), CPU only.

The fixture is the reference's resume file input/resume_data/
particle_state_00500.csv, kept as tests/golden/resume_state_00500.npz (same
columns, same values).  Pins:
* writing the fixture and reading it back gives the fixture's values exactly
  (they have at most 12 significant digits);
* the Python writer (``%.12g``) and the C++ writer (``precision(12)``) write
  the same bytes, and the C++ reader returns the same levels;
* the level rule reduces to a closed form: a particle of size sigma 2^k is at
  level max_level - k (initialization.cpp:660-666);
* where the reference tree is present, reading its own CSV gives the fixture.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lsmps_vpm_v2_amd import state_io
from lsmps_vpm_v2_amd.solver import Particle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SIGMA, MAX_LEVEL = 0.02, 6
REF_CSV = "/root/reference/input/resume_data/particle_state_00500.csv"


def _fixture():
    d = np.load(os.path.join(HERE, "golden", "resume_state_00500.npz"))
    p = Particle(x=d["xp"], y=d["yp"], s=d["sp"], vorticity=d["vor"],
                 isActive=(d["active"] != 0).astype(np.uint8), u=d["up"], v=d["vp"])
    return d, p


def test_level_rule_closed_form():
    k = np.arange(0, 6)
    sp = SIGMA * 2.0 ** k
    assert np.array_equal(state_io.particle_level(sp, MAX_LEVEL, SIGMA), MAX_LEVEL - k)
    # sizes between powers of two truncate like the reference's int division
    assert state_io.particle_level([SIGMA * 3.0], MAX_LEVEL, SIGMA)[0] == MAX_LEVEL - 1
    with pytest.raises(ValueError):
        state_io.particle_level([0.5 * SIGMA], MAX_LEVEL, SIGMA)


def test_python_round_trip(tmp_path):
    d, p = _fixture()
    path = str(tmp_path / "state.csv")
    state_io.save_par_state_2D(p, path, gz=d["gpz"], chi=d["chi"])
    q, extra = state_io.read_2d_particle(path, MAX_LEVEL, SIGMA)
    for a, b in ((q.x, d["xp"]), (q.y, d["yp"]), (q.s, d["sp"]), (q.vorticity, d["vor"]),
                 (q.u, d["up"]), (q.v, d["vp"]), (extra["gz"], d["gpz"]),
                 (extra["chi"], d["chi"])):
        assert np.array_equal(a, b)
    assert np.array_equal(q.isActive, (d["active"] != 0).astype(np.uint8))
    assert q.num == 21728 and int(q.isActive.sum()) == 17522
    lv = extra["level"]
    assert np.array_equal(lv, MAX_LEVEL - np.log2(np.round(d["sp"] / SIGMA)).astype(int))


def test_writer_body_velocity_and_pressure(tmp_path):
    p = Particle(x=np.array([0.5]), y=np.array([-1.25]), s=np.array([0.04]),
                 vorticity=np.array([2.0]), isActive=np.array([1], np.uint8),
                 u=np.array([1.5]), v=np.array([0.25]))
    path = str(tmp_path / "one.csv")
    state_io.save_par_state_2D(p, path, P=np.array([3.0]), ubody=1.0, vbody=0.5)
    lines = open(path).read().splitlines()
    assert lines[0] == "xp,yp,gpz,vor,up,vp,sp,active,chi,pressure"
    assert lines[1] == "0.5,-1.25,0.0032,2,0.5,-0.25,0.04,1,0,3"
    q, extra = state_io.read_2d_particle(path, MAX_LEVEL, SIGMA, pressure=True)
    assert extra["P"][0] == 3.0 and extra["level"][0] == MAX_LEVEL - 1


def test_reader_errors(tmp_path):
    with pytest.raises(RuntimeError, match="not existed"):
        state_io.read_2d_particle(str(tmp_path / "missing.csv"), MAX_LEVEL, SIGMA)
    bad = tmp_path / "bad.csv"
    bad.write_text("xp,yp,gpz,vor,up,vp,sp,active,chi\n1,2,3,4,5,6,0.02,1,0\n\n")
    with pytest.raises(ValueError):
        state_io.read_2d_particle(str(bad), MAX_LEVEL, SIGMA)


@pytest.fixture(scope="module")
def cpp_tool(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("cpp") / "particle_state_roundtrip")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", f"-I{ROOT}/include",
                    os.path.join(HERE, "cpp", "particle_state_roundtrip.cpp"), "-o", exe],
                   check=True)
    return exe


def test_cpp_header_matches_python(tmp_path, cpp_tool):
    d, p = _fixture()
    a = str(tmp_path / "py.csv")
    b = str(tmp_path / "cpp.csv")
    state_io.save_par_state_2D(p, a, gz=d["gpz"], chi=d["chi"])
    r = subprocess.run([cpp_tool, a, b, str(MAX_LEVEL), str(SIGMA)], capture_output=True,
                       text=True, check=True)
    assert open(a, "rb").read() == open(b, "rb").read()
    _, extra = state_io.read_2d_particle(a, MAX_LEVEL, SIGMA)
    assert np.array_equal(np.array(r.stdout.split(), int), extra["level"])


def test_cpp_header_errors(tmp_path, cpp_tool):
    r = subprocess.run([cpp_tool, str(tmp_path / "missing.csv"), str(tmp_path / "o.csv"),
                        "6", "0.02"], capture_output=True, text=True)
    assert r.returncode == 2 and "not existed" in r.stderr
    small = tmp_path / "small.csv"
    small.write_text("xp,yp,gpz,vor,up,vp,sp,active,chi\n1,2,3,4,5,6,0.001,1,0\n")
    r = subprocess.run([cpp_tool, str(small), str(tmp_path / "o.csv"), "6", "0.02"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "below" in r.stderr


@pytest.mark.skipif(not os.path.isfile(REF_CSV), reason="reference tree not present")
def test_reads_the_reference_file():
    d, _ = _fixture()
    q, extra = state_io.read_2d_particle(REF_CSV, MAX_LEVEL, SIGMA)
    assert np.array_equal(q.x, d["xp"]) and np.array_equal(q.vorticity, d["vor"])
    assert np.array_equal(q.u, d["up"]) and np.array_equal(extra["gz"], d["gpz"])
