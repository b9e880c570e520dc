"""GPU parity of the mixed-precision path (FP32 arithmetic, FP32 MFMA M2L,
FP64 geometry and displacements).

Reference: the oracle run in FP64 with the same semantics the mixed path uses
(exact binomials; list 4 by P2L or by M2L).  The only difference left is FP32
rounding, so fields agree to rel-L2 <= 5e-6 (FP32 accumulation of the near
field dominates: about 3e-6 on clustered inputs; each stage <= 1e-5 of its
own norm).  Against the FP64 direct sum the mixed path is bounded by its FMM
truncation error, stated per test.
"""
import numpy as np
import pytest

from conftest import rel_l2
from lsmps_vpm_v2_amd import FmmSolver, VpmFmmError
from lsmps_vpm_v2_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p", [4, 6, 8, 12, 16])
def test_mixed_matches_oracle(gpu, oracle, p):
    f = W.random_clusters(6000, 51, active_frac=0.9)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"],
                                oracle.opts(p=p, binom_mode=0, list4_mode=2))
    s = FmmSolver(p=p, precision="mixed", binom_mode=0, list4_mode=2)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex, ey, ex0, ey0)
    print(f"p={p} mixed vs oracle {err:.3e}")
    assert err <= 5e-6


@pytest.mark.parametrize("mask", [1, 2, 4])
def test_mixed_stages(gpu, oracle, mask):
    f = W.random_clusters(6000, 52, active_frac=0.9)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"],
                                oracle.opts(p=8, binom_mode=0, list4_mode=2, stage_mask=mask))
    s = FmmSolver(p=8, precision="mixed", list4_mode=2, stage_mask=mask)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex, ey, ex0, ey0)
    print(f"mask={mask}: {err:.3e}")
    assert err <= 1e-5


def test_mixed_list4_by_m2l(gpu, oracle):
    f = W.random_clusters(6000, 53)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"],
                                oracle.opts(p=10, binom_mode=0, list4_mode=1))
    s = FmmSolver(p=10, precision="mixed", list4_mode=1)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    assert rel_l2(ex, ey, ex0, ey0) <= 5e-6


def test_mixed_against_direct(gpu, oracle):
    f = W.random_clusters(20000, 54)
    t = W.subsample(f, 3000)
    dx, dy = oracle.direct_field(f["x"], f["y"], f["q"], targets=t)
    s = FmmSolver(p=8, precision="mixed", list4_mode=2)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex[t], ey[t], dx, dy)
    print(f"mixed p=8 vs FP64 direct: {err:.3e}")
    assert err <= 1e-4


def test_mixed_velocity_on_boundary_layer_sample(gpu, oracle):
    """A 60k-particle slice of the flat-plate field, velocity entry point."""
    f = W.boundary_layer(lx=0.8)
    t = W.subsample(f, 2000)
    dx, dy = oracle.direct_field(f["x"], f["y"], f["q"], targets=t)
    s = FmmSolver(p=8, precision="mixed", list4_mode=2)
    u, v = s.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
    err = rel_l2(u[t] - 1.0, v[t], dy, -dx)
    print(f"n={f['n']} mixed velocity vs direct {err:.3e}")
    assert err <= 1e-4


def test_per_level_orders(gpu, oracle):
    f = W.random_clusters(8000, 55)
    t = W.subsample(f, 2000)
    dx, dy = oracle.direct_field(f["x"], f["y"], f["q"], targets=t)
    errs = {}
    for name, orders in (("p6", [6] * 17), ("p12", [12] * 17), ("mixed", [12, 12, 12, 10, 8, 8, 6, 6] + [6] * 9)):
        s = FmmSolver(p=8, precision="mixed", list4_mode=2)
        s.set_level_orders(orders)
        ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
        errs[name] = rel_l2(ex[t], ey[t], dx, dy)
    print(errs)
    assert errs["p12"] < errs["p6"]
    assert errs["p12"] <= errs["mixed"] <= errs["p6"] * 1.5
    s = FmmSolver(p=8, precision="mixed")
    with pytest.raises(VpmFmmError):
        s.set_level_orders([17])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_results_do_not_depend_on_call_history(gpu, precision):
    """The particle sort covers the key bits the tree's depth needs, guessed
    from the previous call and checked after the leaf split; the bit count is
    a function of the depth alone, so the order of particles inside a leaf
    (and every rounding) depends on the input only.  Fresh solvers, a solver
    whose previous tree was deeper or shallower, and repeated calls agree bit
    for bit (uniform field: depth 5, 16 key bits; field with a tight cluster:
    depth 13, 32 key bits)."""
    rng = np.random.default_rng(64)
    n = 20000
    uni = (rng.random(n), rng.random(n), rng.standard_normal(n), np.ones(n, np.uint8))
    h = n // 2
    deep = (np.concatenate([rng.random(h), 0.3 + 1e-3 * rng.standard_normal(h)]),
            np.concatenate([rng.random(h), 0.6 + 1e-3 * rng.standard_normal(h)]),
            rng.standard_normal(n), np.ones(n, np.uint8))
    kw = dict(p=8, precision=precision, binom_mode=0, list4_mode=2)
    fu, fd = FmmSolver(**kw), FmmSolver(**kw)
    eu, ed = fu.field(*uni), fd.field(*deep)
    du, dd = fu.stats()["max_leaf_level"], fd.stats()["max_leaf_level"]
    print(f"{precision}: depth uniform {du}, clustered {dd}")
    assert du <= 8 and dd >= 13  # 16 and 32 sorted key bits
    s = FmmSolver(**kw)
    runs = [(deep, ed), (uni, eu), (uni, eu), (deep, ed), (deep, ed)]
    for k, (f, ref) in enumerate(runs):
        e = s.field(*f)
        assert np.array_equal(e[0], ref[0]) and np.array_equal(e[1], ref[1]), k
