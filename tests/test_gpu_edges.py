"""Edge cases of the solve on the GPU: tiny and ragged particle counts (around
the 256-particle split blocks and the 1024-key radix tiles), trees one level
deep, no active particle, collinear particles, and a single particle.

FP64 path against the oracle to 1e-10 (rel-L2; absolute 1e-12 where the
field is zero).  Mixed path against the oracle in FP64 to 5e-6 plus twice
the floor that FP32 positions set: the mixed path rounds each particle's
offset from its leaf centre to FP32, and on small uniform random fields a
close pair (r ~ 1e-3) makes that rounding alone worth about 1e-5.  The floor
is the direct sum's change under exactly that rounding, on the oracle tree.
"""
import numpy as np
import pytest

from conftest import rel_l2
from lsmps_vpm_v2_amd import FmmSolver, VpmFmmError

pytestmark = pytest.mark.gpu


def _field(n, seed, active_frac=1.0):
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    q = rng.normal(0, 1, n)
    a = (rng.uniform(0, 1, n) < active_frac).astype(np.uint8)
    return x, y, q, a


def _fp32_floor(oracle, x, y, q, a):
    """rel-L2 change of the direct field when each particle's offset from its
    leaf centre (the oracle tree, reference leaf rule) is rounded to FP32."""
    t = oracle.Tree(x, y, a)
    R, piv = t.geometry()
    cells = t.cells()
    where = {int(i): k for k, i in enumerate(cells["id"])}
    lid, poff, pidx = t.leaves()
    xr, yr = x.copy(), y.copy()
    for li, cid in enumerate(lid):
        k = where[int(cid)]
        side = R / 2.0 ** int(cells["level"][k])
        cx = piv[0] + (cells["ix"][k] + 0.5) * side
        cy = piv[1] + (cells["iy"][k] + 0.5) * side
        idx = pidx[poff[li]:poff[li + 1]]
        xr[idx] = cx + (x[idx] - cx).astype(np.float32).astype(np.float64)
        yr[idx] = cy + (y[idx] - cy).astype(np.float32).astype(np.float64)
    ex, ey = oracle.direct_field(x, y, q)
    er, eyr = oracle.direct_field(xr, yr, q)
    return rel_l2(er, eyr, ex, ey)


def _check(ex, ey, ex0, ey0, tol):
    scale = np.sqrt(np.sum(ex0 ** 2 + ey0 ** 2))
    if scale == 0.0:
        assert np.abs(ex).max() <= 1e-12 and np.abs(ey).max() <= 1e-12
    else:
        assert rel_l2(ex, ey, ex0, ey0) <= tol


@pytest.mark.parametrize("n", [2, 3, 5, 255, 257, 1023, 1025, 4097])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_small_and_ragged_counts(gpu, oracle, n, precision):
    print(f"n={n} {precision}", flush=True)
    x, y, q, a = _field(n, 100 + n)
    if precision == "fp64":
        o, kw, tol = oracle.opts(p=10), dict(p=10), 1e-10
    else:
        o = oracle.opts(p=10, binom_mode=0, list4_mode=2)
        kw = dict(p=10, precision="mixed", binom_mode=0, list4_mode=2)
        floor = _fp32_floor(oracle, x, y, q, a)
        tol = 5e-6 + 2 * floor
        print(f"FP32 position floor {floor:.2e}", flush=True)
    ex0, ey0 = oracle.fmm_field(x, y, q, a, o)
    ex, ey = FmmSolver(**kw).field(x, y, q, a)
    _check(ex, ey, ex0, ey0, tol)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_no_active_particle(gpu, oracle, precision):
    x, y, q, _ = _field(3000, 7)
    a = np.zeros(3000, np.uint8)
    kw = dict(p=10) if precision == "fp64" else dict(p=10, precision="mixed", binom_mode=0,
                                                     list4_mode=2)
    o = oracle.opts(p=10) if precision == "fp64" else oracle.opts(p=10, binom_mode=0,
                                                                  list4_mode=2)
    ex0, ey0 = oracle.fmm_field(x, y, q, a, o)
    ex, ey = FmmSolver(**kw).field(x, y, q, a)
    _check(ex, ey, ex0, ey0, 1e-10)


def test_collinear_particles(gpu, oracle):
    n = 5000
    rng = np.random.default_rng(9)
    x = np.sort(rng.uniform(0, 3, n))
    y = np.full(n, 0.25)
    q = rng.normal(0, 1, n)
    a = np.ones(n, np.uint8)
    ex0, ey0 = oracle.fmm_field(x, y, q, a, oracle.opts(p=12))
    ex, ey = FmmSolver(p=12).field(x, y, q, a)
    _check(ex, ey, ex0, ey0, 1e-10)


def test_single_particle_is_a_degenerate_domain(gpu):
    s = FmmSolver()
    with pytest.raises(VpmFmmError) as e:
        s.field(np.array([0.5]), np.array([0.5]), np.array([1.0]), np.ones(1, np.uint8))
    assert e.value.status == "DEGENERATE_DOMAIN"
    # and the handle still works
    x, y, q, a = _field(100, 3)
    ex, ey = s.field(x, y, q, a)
    assert np.all(np.isfinite(ex)) and np.all(np.isfinite(ey))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_duplicate_points_raise(gpu, oracle, precision):
    """Two active particles at one position: the singular P2P has r = 0 (the
    reference would divide by zero).  Both paths report DUPLICATE_POINTS."""
    x, y, q, a = _field(2000, 21)
    x[7], y[7] = x[3], y[3]
    kw = dict(p=10) if precision == "fp64" else dict(p=10, precision="mixed", binom_mode=0,
                                                     list4_mode=2)
    s = FmmSolver(**kw)
    with pytest.raises(VpmFmmError) as e:
        s.field(x, y, q, a)
    assert e.value.status == "DUPLICATE_POINTS"
    # the handle still works on a clean field
    x2, y2, q2, a2 = _field(2000, 22)
    ex0, ey0 = oracle.fmm_field(x2, y2, q2, a2, oracle.opts(
        p=10, **({} if precision == "fp64" else dict(binom_mode=0, list4_mode=2))))
    ex, ey = s.field(x2, y2, q2, a2)
    _check(ex, ey, ex0, ey0, 1e-10 if precision == "fp64" else
           5e-6 + 2 * _fp32_floor(oracle, x2, y2, q2, a2))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_depth_overflow_raises(gpu, oracle, precision):
    """100 particles inside a 1e-9 square with the 75-particle leaf rule need
    more than 16 levels: TREE_DEPTH_EXCEEDED, as the oracle reports."""
    x, y, q, a = _field(1000, 23)
    rng = np.random.default_rng(24)
    x = np.concatenate([x, 0.3 + 1e-9 * rng.uniform(0, 1, 100)])
    y = np.concatenate([y, 0.2 + 1e-9 * rng.uniform(0, 1, 100)])
    q = np.concatenate([q, rng.normal(0, 1, 100)])
    a = np.ones(len(x), np.uint8)
    with pytest.raises(oracle.OracleError):
        oracle.fmm_field(x, y, q, a, oracle.opts(p=10))
    kw = dict(p=10) if precision == "fp64" else dict(p=10, precision="mixed", binom_mode=0,
                                                     list4_mode=2)
    with pytest.raises(VpmFmmError) as e:
        FmmSolver(**kw).field(x, y, q, a)
    assert e.value.status == "TREE_DEPTH_EXCEEDED"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_level_tree_after_a_large_solve_on_one_handle(gpu, oracle, precision):
    """A tree one level deep (leaves at level 1) has no far field.  The mixed
    path forms local coefficients from level 2 down only, so its L2P must not
    read the level-1 leaves' coefficients: on a handle whose buffers still
    hold a large solve's coefficients those are stale, not zero."""
    from lsmps_vpm_v2_amd import workloads as W
    kw = dict(p=10) if precision == "fp64" else dict(p=10, precision="mixed", binom_mode=0,
                                                      list4_mode=2)
    s = FmmSolver(**kw)
    big = W.random_clusters(50000, seed=61)
    s.field(big["x"], big["y"], big["q"], big["active"])
    for n in (2, 3, 5, 40):
        x, y, q, a = _field(n, 300 + n)
        ex, ey = s.field(x, y, q, a)
        assert s.stats()["max_leaf_level"] == 1
        ex0, ey0 = oracle.direct_field(x, y, q)
        tol = 1e-10 if precision == "fp64" else 5e-6 + 2 * _fp32_floor(oracle, x, y, q, a)
        _check(ex, ey, ex0, ey0, tol)
