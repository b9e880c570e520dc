"""The reference's own time loop driving the GPU VelocityCalc (the drop-in
header in place of velocity_calc.hpp): tests/cpp/reference_loop.cpp, built
by oracle/build_loop.py from the reference's unchanged simulation sources --
initialization (Perlman vortex, the grid-block particle generation and its
multiresolution adaptation), advection, diffusion (LSMPS), remeshing with
adaptation every 5 steps, the domain-boundary update and get_velocity.

The same driver built with the oracle's VelocityCalc (the FMM branch as the
oracle restates it, pinned to the reference build) runs beside it on the
host.  After 12 steps (three adaptations, 54,400 -> 53,344 particles) both
hold the same particles, and positions, sizes, vorticity and velocity agree
to 1e-10 of their scales: the GPU FP64 path changes nothing the loop sees.
The binaries are built where the reference tree exists (build()); without
them the test is skipped."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LOOP = os.path.join(ROOT, "oracle", "_ref", "loop")
STEPS = 12


def _run(exe, tmp, name):
    work = tmp / name
    (work / "output").mkdir(parents=True)
    out = work / "state.bin"
    with open(work / "stdout.txt", "w") as log:
        r = subprocess.run([exe, str(STEPS), str(out)], cwd=work, stdout=log,
                           stderr=subprocess.PIPE, text=True, timeout=240)
    print(name, r.stderr.strip().splitlines()[-1:] if r.stderr else "")
    assert r.returncode == 0, (name, r.returncode, r.stderr[-2000:])
    raw = np.fromfile(out, dtype=np.float64)
    n = int(raw[0])
    return n, raw[1:].reshape(6, n)


@pytest.mark.skipif(not os.path.exists(os.path.join(LOOP, "reference_loop_gpu")),
                    reason="the reference loop driver was not built (no reference tree)")
def test_reference_loop_with_gpu_velocity_matches_oracle(gpu, tmp_path):
    n_o, st_o = _run(os.path.join(LOOP, "reference_loop_oracle"), tmp_path, "oracle")
    n_g, st_g = _run(os.path.join(LOOP, "reference_loop_gpu"), tmp_path, "gpu")
    assert n_g == n_o
    names = ("x", "y", "s", "vorticity", "u", "v")
    for k, name in enumerate(names):
        a, b = st_g[k], st_o[k]
        if name in ("u", "v"):
            scale = np.sqrt(np.mean((st_o[4] - 1.0) ** 2 + st_o[5] ** 2))  # rotational part
        else:
            scale = max(np.max(np.abs(b)), 1e-300)
        err = float(np.max(np.abs(a - b)) / scale)
        print(f"{name}: max |gpu - oracle| / scale = {err:.2e}")
        assert err <= 1e-10, name
