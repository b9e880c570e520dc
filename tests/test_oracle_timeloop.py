"""The oracle's advection sub-step and loop, pinned on the CPU.

* Advection (main_advection.cpp:44-83): with dyadic positions, velocities
  and step, every operation is exact, so k Euler steps equal x + k u dt and
  the reference's "rk2" update equals x + (11/24) u dt (its weights
  (1 + 2/2 + 2/4 + 1/4) / 6).
* Loop: the singular Biot-Savart velocity is a sum of antisymmetric pair
  terms, so the linear impulse sum(G u) and the angular impulse rate
  sum(G x . u) vanish for the exact sum; the oracle FMM at p = 16 must keep
  them at its own error level at every step, with the tree rebuilt around the
  moved particles.  (p = 16 with exact binomials and list 4 by P2L: the
  field's error is 2.5e-8 there, the impulses about 3e-10; the reference's
  list-4 M2L alone leaves 1e-6.)
"""
import numpy as np
import pytest

from lsmps_vpm_v2_amd import workloads as W


def test_advect_closed_forms(oracle):
    rng = np.random.default_rng(3)
    x = rng.integers(-512, 512, 64) / 64.0
    y = rng.integers(-512, 512, 64) / 64.0
    u = rng.integers(-64, 64, 64) / 16.0
    v = rng.integers(-64, 64, 64) / 16.0
    dt = 1.0 / 128
    xe, ye = x, y
    for _ in range(5):
        xe, ye = oracle.advect(xe, ye, u, v, dt)
    assert np.array_equal(xe, x + 5 * u * dt) and np.array_equal(ye, y + 5 * v * dt)
    dt = 3.0 / 64  # (11/24) dt = 11/512: exact for these u
    xr, yr = oracle.advect(x, y, u, v, dt, "rk2")
    np.testing.assert_allclose(xr, x + u * (11.0 / 512), rtol=0, atol=1e-15)
    np.testing.assert_allclose(yr, y + v * (11.0 / 512), rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        oracle.advect(x, y, u, v, dt, "rk4")


def _impulses(f, x, y, u, v):
    g = f["omega"] * f["s"] ** 2
    cx, cy = np.sum(g * x) / np.sum(g), np.sum(g * y) / np.sum(g)
    rx, ry = x - cx, y - cy
    un = np.sqrt(np.sum(u * u + v * v))
    lin = np.hypot(np.sum(g * u), np.sum(g * v)) / (np.linalg.norm(g) * un)
    ang = abs(np.sum(g * (rx * u + ry * v))) / (np.linalg.norm(g * np.hypot(rx, ry)) * un)
    return lin, ang


def test_loop_keeps_the_impulses(oracle):
    f = W.random_clusters(3000, seed=11)
    o = oracle.opts(p=16, max_par=40, max_src=40, binom_mode=0, list4_mode=2)
    x, y = f["x"], f["y"]
    u, v = oracle.velocity(x, y, f["s"], f["omega"], f["active"], o, 0.0, 0.0)
    dt = 0.02 * np.ptp(x) / np.max(np.hypot(u, v))
    for _ in range(3):
        lin, ang = _impulses(f, x, y, u, v)
        assert lin < 1e-8 and ang < 1e-8, (lin, ang)
        x, y = oracle.advect(x, y, u, v, dt)
        u, v = oracle.velocity(x, y, f["s"], f["omega"], f["active"], o, 0.0, 0.0)
    # the loop helper is the same composition
    lx, ly, lu, lv = oracle.inviscid_loop(f["x"], f["y"], f["s"], f["omega"], f["active"], dt,
                                          3, o, 0.0, 0.0)
    assert np.array_equal(lx, x) and np.array_equal(lu, u)
    # a velocity 1 % off on half of the particles (a target-side factor that
    # breaks the pair antisymmetry) fails both bounds by orders of magnitude
    bad = np.where(x > np.median(x), 1.01, 1.0)
    lin_bad, ang_bad = _impulses(f, x, y, u * bad, v * bad)
    assert lin_bad > 1e-5 and ang_bad > 1e-5, (lin_bad, ang_bad)
