"""GPU parity of the plain FP64 path against the oracle (reference semantics).

Tolerance: rel-L2 of (ex, ey) <= 1e-10 against the oracle on the same inputs,
options and binomial table (the oracle matches the reference build to 1e-15,
see test_oracle_reference.py).  Tree and list membership are integer work and
must be bit-identical.
"""
import numpy as np
import pytest

from conftest import rel_l2
from lsmps_vpm_v2_amd import FmmSolver, VpmFmmError
from lsmps_vpm_v2_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _case(kind, n, seed):
    if kind == "uniform":
        rng = np.random.Generator(np.random.PCG64(seed))
        f = dict(x=rng.uniform(-1, 1, n), y=rng.uniform(-1, 1, n), q=rng.normal(0, 1, n),
                 active=np.ones(n, np.uint8))
    elif kind == "clusters":
        f = W.random_clusters(n, seed)
    elif kind == "inactive":
        f = W.random_clusters(n, seed, active_frac=0.7)
    else:
        raise ValueError(kind)
    return f


@pytest.mark.parametrize("kind,n,seed", [("uniform", 3000, 1), ("clusters", 6000, 2),
                                         ("inactive", 6000, 3)])
@pytest.mark.parametrize("p", [8, 10, 12])
def test_field_matches_oracle(gpu, oracle, kind, n, seed, p):
    print(f"case {kind} n={n} seed={seed} p={p}", flush=True)
    f = _case(kind, n, seed)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"], oracle.opts(p=p))
    s = FmmSolver(p=p)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex, ey, ex0, ey0)
    print(f"rel-L2 vs oracle {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("p", [1, 2, 3, 5, 6, 7, 9, 11, 13, 14, 15])
def test_every_order_matches_oracle(gpu, oracle, p):
    """The list-2 M2L runs on the matrix cores with a kernel instance and a
    translation table per order (one or two 16-row tiles, K padded to 4):
    every order up to 16 against the oracle (8, 10, 12 and 16 above)."""
    print(f"case clusters n=4000 seed={40 + p} p={p}", flush=True)
    f = _case("clusters", 4000, 40 + p)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"], oracle.opts(p=p))
    s = FmmSolver(p=p)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex, ey, ex0, ey0)
    print(f"rel-L2 vs oracle {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("p", [17, 24, 40])
def test_high_orders_match_oracle(gpu, oracle, p):
    """Orders 17 .. 40 run the thread-per-(cell, order) M2M, M2L and L2L
    kernels (no matrix-core M2L, no LDS staging): field and potential against
    the oracle up to the largest accepted order."""
    print(f"case clusters n=3000 seed={70 + p} p={p}", flush=True)
    f = _case("clusters", 3000, 70 + p)
    o = oracle.opts(p=p)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"], o)
    s = FmmSolver(p=p)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex, ey, ex0, ey0)
    print(f"field rel-L2 vs oracle {err:.3e}", flush=True)
    assert err <= TOL
    phi0 = np.asarray(oracle.fmm_potential(f["x"], f["y"], f["q"], f["active"], o))
    phi = np.asarray(s.potential(f["x"], f["y"], f["q"], f["active"]))
    perr = float(np.linalg.norm(phi - phi0) / np.linalg.norm(phi0))
    print(f"potential rel-L2 vs oracle {perr:.3e}", flush=True)
    assert perr <= TOL


@pytest.mark.parametrize("binom_mode,list4_mode", [(0, 1), (0, 2), (1, 2)])
def test_field_modes_match_oracle(gpu, oracle, binom_mode, list4_mode):
    f = _case("clusters", 5000, 11)
    o = oracle.opts(p=16, binom_mode=binom_mode, list4_mode=list4_mode)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"], o)
    s = FmmSolver(p=16, binom_mode=binom_mode, list4_mode=list4_mode)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    assert rel_l2(ex, ey, ex0, ey0) <= TOL


@pytest.mark.parametrize("kind,max_par,max_src", [("inactive", 75, 50), ("clusters", 48, 48),
                                                  ("clusters", 64, 64), ("uniform", 32, 32),
                                                  ("uniform", 63, 63)])
def test_tree_and_lists_bit_exact(gpu, oracle, kind, max_par, max_src):
    """Reference leaf rule (75/50, some inactive) and the equal thresholds of
    the learned path on all-active fields."""
    f = _case(kind, 4000, 5)
    t = oracle.Tree(f["x"], f["y"], f["active"], oracle.opts(max_par=max_par, max_src=max_src))
    s = FmmSolver(max_par=max_par, max_src=max_src)
    s.build_tree(f["x"], f["y"], f["active"])
    assert s.tree_counts() == t.counts()
    co, cg = t.cells(), s.tree_cells()
    for k in co:
        assert np.array_equal(co[k], cg[k]), k
    lo = t.leaves()
    lg = s.tree_leaves(len(f["x"]))
    order = np.argsort(lo[0])
    assert np.array_equal(lo[0][order], lg[0])
    for k, li in enumerate(order):
        a = lo[2][lo[1][li]:lo[1][li + 1]]
        b = lg[2][lg[1][k]:lg[1][k + 1]]
        assert np.array_equal(a, b)
    bad = []
    for cid, lev, leaf in zip(co["id"], co["level"], co["is_leaf"]):
        l1, l2, l3, l4 = s.tree_lists(cid)
        if leaf:
            r1, r3 = t.int_list(cid)
            if sorted(l1) != sorted(r1) or sorted(l3) != sorted(r3):
                bad.append(("int", int(cid), sorted(l1), sorted(r1), sorted(l3), sorted(r3)))
        if lev >= 2:
            r2, r4 = t.ext_list(cid)
            if sorted(l2) != sorted(r2) or sorted(l4) != sorted(r4):
                bad.append(("ext", int(cid), sorted(l2), sorted(r2), sorted(l4), sorted(r4)))
    for b in bad[:6]:
        print(b)
    assert not bad, f"{len(bad)} cells with list mismatches"


def test_velocity_entry_point(gpu, oracle):
    f = W.random_clusters(4000, 9)
    u0, v0 = oracle.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"], oracle.opts(),
                             1.0, 0.0)
    s = FmmSolver()
    u, v = s.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
    assert rel_l2(u - 1.0, v, u0 - 1.0, v0) <= TOL


def test_errors(gpu):
    s = FmmSolver()
    x = np.array([0.0, 1.0, np.nan])
    with pytest.raises(VpmFmmError) as e:
        s.field(x, x, x, np.ones(3))
    assert e.value.status == "NONFINITE_INPUT"
    z = np.zeros(4)
    with pytest.raises(VpmFmmError) as e:
        s.field(z, z, np.ones(4), np.ones(4))
    assert e.value.status == "DEGENERATE_DOMAIN"
    # solver still usable after errors
    f = _case("uniform", 500, 3)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    assert np.all(np.isfinite(ex)) and np.all(np.isfinite(ey))


def test_tiny_leaf_threshold_many_cells(gpu, oracle):
    """Leaf thresholds of one particle give about 1.3 cells per particle, more
    than the cell emission has room for before the count is read back on a
    fresh handle (n / 4 + 1024): the build grows its buffers and emits again.
    Tree counts and the field equal the oracle's."""
    f = _case("uniform", 6000, 13)
    t = oracle.Tree(f["x"], f["y"], f["active"], oracle.opts(p=6, max_par=1, max_src=1))
    s = FmmSolver(p=6, max_par=1, max_src=1)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    nc, nl, _, _ = t.counts()
    assert nc > 6000 // 4 + 1024
    assert s.tree_counts() == t.counts()
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"],
                                oracle.opts(p=6, max_par=1, max_src=1))
    assert rel_l2(ex, ey, ex0, ey0) <= TOL
