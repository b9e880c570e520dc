"""The C++ drop-in (include/vpmfmm_velocity_calc.hpp) on the GPU: the
reference's call get_velocity(particle, grid, step) through a compiled C++
program, on the FMM branch, the type-1 branch and (compiled with DIM 3) the
3-D branch, checked against the oracle's velocities to 1e-10."""
import subprocess

import numpy as np
import pytest

from test_abi import build_drop_in

pytestmark = pytest.mark.gpu


def test_cpp_drop_in_velocity_matches_oracle(gpu, oracle, tmp_path):
    exe = build_drop_in(tmp_path)
    out = tmp_path / "uv.bin"
    n = 6000
    r = subprocess.run([str(exe), str(n), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    assert int(raw[0]) == n
    x, y, s, w, a, u, v, gz, u1, v1 = raw[1:].reshape(10, n)
    uo, vo = oracle.velocity(x, y, s, w, a.astype(np.uint8), oracle.opts())
    scale = np.sqrt(np.mean((uo - 1.0) ** 2 + vo ** 2))
    err = np.max(np.hypot(u - uo, v - vo)) / scale
    print(f"drop-in vs oracle: max |d(u,v)| / rms = {err:.2e}")
    assert err <= 1e-10
    # the type-1 branch (VELOCITY_2D_TYPE 1): the reference's type-1 options,
    # every particle a source, plus the free stream (1, 0)
    ut, vt = oracle.t1_velocity(x, y, s, gz, oracle.t1_opts())
    err1 = np.max(np.hypot(u1 - 1.0 - ut, v1 - vt)) / np.sqrt(np.mean(ut ** 2 + vt ** 2))
    print(f"drop-in type-1 vs oracle: max |d(u,v)| / rms = {err1:.2e}")
    assert err1 <= 1e-10


def test_cpp_drop_in_3d_matches_oracle(gpu, oracle, tmp_path):
    exe = build_drop_in(tmp_path, "drop_in_velocity3d")
    out = tmp_path / "uvw.bin"
    n = 8000
    r = subprocess.run([str(exe), str(n), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    assert int(raw[0]) == n
    x, y, z, s, wx, wy, wz, a, u, v, w, u2, v2, w2 = raw[1:].reshape(14, n)
    # the reference's DIM == 3 defaults: critNum 30, free stream (1, 0, 0)
    uo, vo, wo = oracle.o3_velocity(x, y, z, s, wx, wy, wz, a.astype(np.uint8),
                                    oracle.o3_opts(), 1.0, 0.0, 0.0)
    scale = np.sqrt(np.mean((uo - 1.0) ** 2 + vo ** 2 + wo ** 2))
    err = np.max(np.sqrt((u - uo) ** 2 + (v - vo) ** 2 + (w - wo) ** 2)) / scale
    print(f"3-D drop-in vs oracle: max |d(u,v,w)| / rms = {err:.2e}")
    assert err <= 1e-10
    # VELOCITY_3D_TYPE 2 through set_default_fmm3d_options
    uo, vo, wo = oracle.o3_velocity(x, y, z, s, wx, wy, wz, a.astype(np.uint8),
                                    oracle.o3_opts(type=2), 1.0, 0.0, 0.0)
    scale = np.sqrt(np.mean((uo - 1.0) ** 2 + vo ** 2 + wo ** 2))
    err = np.max(np.sqrt((u2 - uo) ** 2 + (v2 - vo) ** 2 + (w2 - wo) ** 2)) / scale
    print(f"3-D drop-in type 2 vs oracle: max |d(u,v,w)| / rms = {err:.2e}")
    assert err <= 1e-10
