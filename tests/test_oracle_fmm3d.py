"""Pins of the 3-D oracle (oracle/fmm3d_oracle.c), the restatement of the
reference's default 3-D velocity path (fast3DFMM.cpp, fmm3D.cpp helpers).

The reference's 3-D sources cannot be built here (global.hpp:85 fixes DIM at
2), so the oracle is pinned by what the mathematics fixes:

* the direct sum against a numpy cross-product brute force;
* each M2P multiplier against finite differences of the kernel r / |r|^3;
* M2M against P2M about the parent (the order-2 shift is exact);
* the truncation order of P2M -> M2P (error ~ |d|^3: a wrong zeroth-, first-
  or second-order term gives a lower order);
* the octree's invariants against numpy (root box, octant of every particle
  from the parent centre, leaf counts, active marks);
* the evaluation with every cell opened against the direct sum, including the
  reference's handling of level-1 leaves (their sources are never summed:
  evaluateFMM starts from the children of the level-1 cells);
* the treecode error with the reference's opening test.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _cloud(n, seed, kind="uniform"):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        p = rng.uniform(-1.0, 1.0, (3, n))
    else:
        c = rng.uniform(-1.0, 1.0, (6, 3))
        k = rng.integers(0, 6, n)
        p = (c[k] + 0.08 * rng.standard_normal((n, 3))).T
    a = 1e-3 * rng.standard_normal((3, n))
    return p[0].copy(), p[1].copy(), p[2].copy(), a[0].copy(), a[1].copy(), a[2].copy()


def _brute(x, y, z, ax, ay, az, src=None, targets=None):
    P = np.stack([x, y, z], 1)
    A = np.stack([ax, ay, az], 1)
    n = len(x)
    src = np.ones(n, bool) if src is None else np.asarray(src, bool)
    t = np.arange(n) if targets is None else np.asarray(targets)
    out = np.zeros((len(t), 3))
    for k, i in enumerate(t):
        m = src.copy()
        m[i] = False
        r = P[i] - P[m]
        out[k] = (np.cross(A[m], r) / (np.linalg.norm(r, axis=1) ** 3)[:, None]).sum(0)
    return out[:, 0], out[:, 1], out[:, 2]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def test_direct_sum_brute_force():
    x, y, z, ax, ay, az = _cloud(60, 1)
    act = (np.arange(60) % 5 != 0).astype(np.uint8)
    u, v, w = O.o3_direct(x, y, z, ax, ay, az, act)
    bu, bv, bw = _brute(x, y, z, ax, ay, az, act)
    assert _rel(np.r_[u, v, w], np.r_[bu, bv, bw]) < 1e-13
    t = np.array([3, 17, 59], np.int32)
    tu, tv, tw = O.o3_direct(x, y, z, ax, ay, az, act, targets=t)
    np.testing.assert_allclose(np.c_[tu, tv, tw], np.c_[u[t], v[t], w[t]], rtol=0, atol=0)


@pytest.mark.parametrize("D", [(0.7, -0.4, 1.1), (-2.0, 0.3, 0.2), (0.1, 1.5, -0.9)])
def test_m2p_multipliers_are_kernel_derivatives(D):
    """mul_c[k] = d^k (D_c / |D|^3) for the moment order of P2M_calc:
    [1, x, y, z, xx, yy, zz, xy, yz, xz] (second derivatives, the 1/2 of the
    squares being in the moments)."""
    D = np.array(D)

    def f(p):
        return p / np.linalg.norm(p) ** 3

    h = 1e-4
    e = np.eye(3)
    d1 = [(f(D + h * e[a]) - f(D - h * e[a])) / (2 * h) for a in range(3)]

    def d2(a, b):
        return (f(D + h * e[a] + h * e[b]) - f(D + h * e[a] - h * e[b])
                - f(D - h * e[a] + h * e[b]) + f(D - h * e[a] - h * e[b])) / (4 * h * h)

    want = np.stack([f(D), d1[0], d1[1], d1[2], d2(0, 0), d2(1, 1), d2(2, 2), d2(0, 1),
                     d2(1, 2), d2(0, 2)], 1)  # (component, term)
    mx, my, mz = O.o3_m2p_mul(*D)
    got = np.stack([mx, my, mz], 0)
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * scale)


def test_m2m_equals_p2m_about_parent():
    rng = np.random.default_rng(3)
    cc = np.array([0.3, -0.2, 0.5])
    cp = cc + np.array([0.25, 0.25, -0.25])
    pts = cc + 0.25 * rng.uniform(-1, 1, (40, 3))
    q = rng.standard_normal(40)
    child = np.zeros(10)
    parent_direct = np.zeros(10)
    for p, s in zip(pts, q):
        child = O.o3_p2m(child, s, *(cc - p))
        parent_direct = O.o3_p2m(parent_direct, s, *(cp - p))
    parent = O.o3_m2m(np.zeros(10), child, *(cp - cc))
    np.testing.assert_allclose(parent, parent_direct, rtol=1e-12, atol=1e-13)


def test_p2m_m2p_truncation_order():
    """Error of one source through P2M -> M2P shrinks as |d|^3."""
    A = np.array([0.3, -0.7, 0.5])
    D = np.array([1.3, 0.4, -0.8])
    dir_ = np.array([0.6, -0.3, 0.74]) / np.linalg.norm([0.6, -0.3, 0.74])
    errs = []
    for eps in (0.08, 0.04, 0.02):
        d = eps * dir_                       # centre - particle
        r = D + d                            # target - particle (D = target - centre)
        exact = np.cross(A, r) / np.linalg.norm(r) ** 3
        m = [O.o3_p2m(np.zeros(10), A[k], *d) for k in range(3)]
        mx, my, mz = O.o3_m2p_mul(*D)
        got = np.array([m[1] @ mz - m[2] @ my, m[2] @ mx - m[0] @ mz, m[0] @ my - m[1] @ mx])
        errs.append(np.linalg.norm(got - exact))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 6.5 < r1 < 9.5 and 6.5 < r2 < 9.5, errs


@pytest.mark.parametrize("kind,n,crit", [("uniform", 3000, 50), ("clusters", 6000, 50),
                                         ("clusters", 2000, 7)])
def test_octree_invariants(kind, n, crit):
    x, y, z, *_ = _cloud(n, 5, kind)
    act = (np.random.default_rng(9).random(n) < 0.8).astype(np.uint8)
    o = O.o3_opts(crit_num=crit)
    T = O.O3Tree(x, y, z, act, o)
    P = np.stack([x, y, z], 1)
    # root: the box of the particles and the origin, expanded by exp_tree %
    mn = np.minimum(P.min(0), 0.0)
    mx = np.maximum(P.max(0), 0.0)
    assert T.level[0] == 0 and T.key[0] == 0
    np.testing.assert_array_equal(T.centre[0], 0.5 * (mx + mn))
    assert T.size[0] == max(mx - mn) * (1 + 1.0 / 100)
    # parents by key; child centre = parent centre + (size / 2) (bit - 1/2)
    where = {(int(l), int(k)): c for c, (l, k) in enumerate(zip(T.level, T.key))}
    for c in range(1, len(T.level)):
        p = where[(int(T.level[c]) - 1, int(T.key[c]) >> 3)]
        oc = int(T.key[c]) & 7
        cl = T.size[p] / 2.0
        assert T.size[c] == cl
        bits = np.array([oc & 1, (oc >> 1) & 1, (oc >> 2) & 1], float)
        np.testing.assert_array_equal(T.centre[c], T.centre[p] + cl * (bits - 0.5))
        assert T.count[p] >= crit  # only cells with crit_num particles or more split
    # leaves partition the particles; each particle's octant path from the
    # parent centres (x > xc strictly) leads to its leaf
    leaves = np.flatnonzero((T.count < crit) & (T.level > 0))
    seen = np.zeros(n, int)
    for c in leaves:
        ids = T.particles(c)
        assert len(ids) == T.count[c] and np.all(np.diff(ids) > 0)
        seen[ids] += 1
        assert bool(T.active[c]) == bool(act[ids].any())
        L, k = int(T.level[c]), int(T.key[c])
        for i in ids:
            cell = 0
            for lev in range(1, L + 1):
                oc = int((P[i, 0] > T.centre[cell, 0]) + 2 * (P[i, 1] > T.centre[cell, 1])
                         + 4 * (P[i, 2] > T.centre[cell, 2]))
                assert (k >> (3 * (L - lev))) & 7 == oc
                cell = where[(lev, k >> (3 * (L - lev)))]
            assert cell == c
    assert np.all(seen == 1)
    # counts and active marks of the non-leaf cells from their subtrees
    for c in np.flatnonzero(T.count >= crit):
        L, k = int(T.level[c]), int(T.key[c])
        sub = [d for d in leaves if T.level[d] > L and (int(T.key[d]) >> (3 * (int(T.level[d]) - L))) == k]
        assert sum(int(T.count[d]) for d in sub) == T.count[c]
        if L > 0:
            assert bool(T.active[c]) == any(T.active[d] for d in sub)
    assert T.depth == T.level.max()


def _level1_leaf_particles(T):
    out = []
    for c in np.flatnonzero((T.level == 1) & (T.leaf == 1)):
        out.extend(T.particles(c).tolist())
    return np.array(out, int)


def test_all_open_equals_direct_sum():
    x, y, z, ax, ay, az = _cloud(4000, 11)
    act = np.ones(4000, np.uint8)
    act[::7] = 0
    o = O.o3_opts(support=1e6)
    u, v, w = O.o3_velocity_alpha(x, y, z, ax, ay, az, act, o)
    T = O.O3Tree(x, y, z, act)
    assert len(_level1_leaf_particles(T)) == 0
    du, dv, dw = O.o3_direct(x, y, z, ax, ay, az, act)
    assert _rel(np.r_[u, v, w], np.r_[du, dv, dw]) < 1e-13


def test_level1_leaf_sources_are_not_summed():
    """A level-1 octant with fewer than crit_num particles is a leaf that the
    evaluation never opens (fast3DFMM.cpp:352-364 queues the level-1 cells and
    tests only their children): its sources are missing from every velocity,
    its particles still receive the others'."""
    x, y, z, ax, ay, az = _cloud(3000, 13)
    few = np.flatnonzero((x > 0) & (y > 0) & (z > 0))
    keep = np.setdiff1d(np.arange(3000), few[20:])
    x, y, z, ax, ay, az = (t[keep] for t in (x, y, z, ax, ay, az))
    n = len(x)
    act = np.ones(n, np.uint8)
    T = O.O3Tree(x, y, z, act)
    lost = _level1_leaf_particles(T)
    assert len(lost) == 20
    o = O.o3_opts(support=1e6)
    u, v, w = O.o3_velocity_alpha(x, y, z, ax, ay, az, act, o)
    src = act.astype(bool).copy()
    src[lost] = False
    du, dv, dw = O.o3_direct(x, y, z, ax, ay, az, src.astype(np.uint8))
    assert _rel(np.r_[u, v, w], np.r_[du, dv, dw]) < 1e-13
    # fewer particles than crit_num: every level-1 cell is a leaf, nothing is summed
    u, v, w = O.o3_velocity_alpha(x[:25], y[:25], z[:25], ax[:25], ay[:25], az[:25],
                                  act[:25])
    assert not np.any(u) and not np.any(v) and not np.any(w)


def test_treecode_error_with_reference_opening_test():
    x, y, z, ax, ay, az = _cloud(20000, 17)
    act = np.ones(20000, np.uint8)
    du, dv, dw = O.o3_direct(x, y, z, ax, ay, az, act)
    errs = []
    for sup in (2.0, 3.0, 4.0):
        u, v, w = O.o3_velocity_alpha(x, y, z, ax, ay, az, act, O.o3_opts(support=sup))
        errs.append(_rel(np.r_[u, v, w], np.r_[du, dv, dw]))
    assert errs[0] < 2e-3 and errs[2] < errs[1] < errs[0], errs


def test_velocity_forms_alpha_and_adds_free_stream():
    rng = np.random.default_rng(21)
    x, y, z, *_ = _cloud(2500, 21, "clusters")
    s = 0.01 + 0.01 * rng.random(2500)
    wx, wy, wz = rng.standard_normal((3, 2500))
    act = np.ones(2500, np.uint8)
    u, v, w = O.o3_velocity(x, y, z, s, wx, wy, wz, act, u_inf=1.5, v_inf=-0.25, w_inf=0.125)
    al = [c * s ** 3 / (4 * np.pi) for c in (wx, wy, wz)]
    ru, rv, rw = O.o3_velocity_alpha(x, y, z, *al, act)
    np.testing.assert_allclose(u, ru + 1.5, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(v, rv - 0.25, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(w, rw + 0.125, rtol=1e-14, atol=1e-15)


def test_bad_inputs():
    x, y, z, ax, ay, az = _cloud(200, 23)
    act = np.ones(200, np.uint8)
    x2 = x.copy()
    x2[5] = np.nan
    with pytest.raises(O.OracleError):
        O.o3_velocity_alpha(x2, y, z, ax, ay, az, act)
    # crit_num coincident particles never separate: depth bound
    xs = np.r_[x, np.full(60, 0.5)]
    ys = np.r_[y, np.full(60, 0.5)]
    zs = np.r_[z, np.full(60, 0.5)]
    with pytest.raises(O.OracleError):
        O.O3Tree(xs, ys, zs, np.ones(260, np.uint8), O.o3_opts(max_depth=12))


# ---------------------------------------------------------------------------
# type 2: fmm3D::calcVelocity over the 3-D TreeCell (VELOCITY_3D_TYPE 2)
# ---------------------------------------------------------------------------
def _t2(**kw):
    return O.o3_opts(type=2, **kw)


@pytest.mark.parametrize("kind,n", [("uniform", 4000), ("clusters", 8000)])
def test_type2_treecell_invariants(kind, n):
    """TreeCell::initializeTree in 3-D against numpy: the root box of the
    particles alone, side x 1.01, pivot (max + min - side) / 2; every
    particle's cell index per level (int)((x - pivot) / (side / 2^L)); cell
    centres pivot + (0.5 + index) side / 2^L; the leaf rule (at most 50
    particles and 30 active ones, the root never a leaf); active = an active
    particle inside."""
    x, y, z, *_ = _cloud(n, 31, kind)
    act = (np.random.default_rng(4).random(n) < 0.7).astype(np.uint8)
    T = O.O3Tree(x, y, z, act, _t2())
    P = np.stack([x, y, z], 1)
    mn, mx = P.min(0), P.max(0)
    side = max(np.abs(mx - mn)) * (1.0 + 1.0 / 100.0)
    piv = (mx + mn - side) / 2.0
    assert T.size[0] == side
    leaf = T.leaf.astype(bool)
    assert not leaf[0]
    where = {(int(l), int(k)): c for c, (l, k) in enumerate(zip(T.level, T.key))}
    seen = np.zeros(n, int)
    for c in range(1, len(T.level)):
        L, k = int(T.level[c]), int(T.key[c])
        cs = side / 2.0 ** L
        assert T.size[c] == cs
        idx = np.zeros(3, np.int64)
        for lev in range(1, L + 1):
            oc = (k >> (3 * (L - lev))) & 7
            idx = 2 * idx + np.array([oc & 1, (oc >> 1) & 1, (oc >> 2) & 1])
        np.testing.assert_array_equal(T.centre[c], piv + (0.5 + idx) * cs)
        if leaf[c]:
            ids = T.particles(c)
            seen[ids] += 1
            np.testing.assert_array_equal(((P[ids] - piv) / cs).astype(np.int64),
                                          np.tile(idx, (len(ids), 1)))
            assert len(ids) == T.count[c]
            assert len(ids) <= 50 and act[ids].sum() <= 30
            assert bool(T.active[c]) == bool(act[ids].any())
        par = where[(L - 1, k >> 3)]
        assert not leaf[par]
    assert np.all(seen == 1)
    # a non-leaf cell is over one of the two limits
    for c in np.flatnonzero(~leaf)[1:]:
        L, k = int(T.level[c]), int(T.key[c])
        sub = [d for d in np.flatnonzero(leaf) if T.level[d] > L and
               (int(T.key[d]) >> (3 * (int(T.level[d]) - L))) == k]
        ids = np.concatenate([T.particles(d) for d in sub])
        assert len(ids) == T.count[c]
        assert len(ids) > 50 or act[ids].sum() > 30


def test_type2_all_open_equals_direct_sum_with_inactive_sources():
    """With every cell opened, type 2 sums every particle of the active
    leaves directly, inactive ones included (fmm3D.cpp:1172-1186 reads no
    active mark), and nothing of the inactive leaves."""
    x, y, z, ax, ay, az = _cloud(5000, 37)
    act = np.ones(5000, np.uint8)
    act[np.random.default_rng(2).random(5000) < 0.3] = 0
    act[(x > 0.6) & (y > 0.6)] = 0  # some leaves wholly inactive
    o = _t2(support=1e6)
    u, v, w = O.o3_velocity_alpha(x, y, z, ax, ay, az, act, o)
    T = O.O3Tree(x, y, z, act, o)
    src = np.zeros(5000, np.uint8)
    for c in np.flatnonzero(T.leaf.astype(bool) & T.active.astype(bool)):
        src[T.particles(c)] = 1
    assert src.sum() > act.sum()  # inactive particles of active leaves are sources
    du, dv, dw = O.o3_direct(x, y, z, ax, ay, az, src)
    assert _rel(np.r_[u, v, w], np.r_[du, dv, dw]) < 1e-13


def test_type2_treecode_error_and_level1_leaf():
    x, y, z, ax, ay, az = _cloud(20000, 41)
    act = np.ones(20000, np.uint8)
    du, dv, dw = O.o3_direct(x, y, z, ax, ay, az, act)
    u, v, w = O.o3_velocity_alpha(x, y, z, ax, ay, az, act, _t2())
    e2 = _rel(np.r_[u, v, w], np.r_[du, dv, dw])
    u1, v1, w1 = O.o3_velocity_alpha(x, y, z, ax, ay, az, act, O.o3_opts())
    e1 = _rel(np.r_[u1, v1, w1], np.r_[du, dv, dw])
    assert e2 < 2e-3 and e1 < 2e-3, (e1, e2)
    # a level-1 leaf beside deeper cells: the reference's get_childID throws
    # (treeCell.cpp:384-388) when the walk asks the leaf for its children
    few = np.flatnonzero((x > 0) & (y > 0) & (z > 0))
    keep = np.setdiff1d(np.arange(20000), few[20:])
    with pytest.raises(O.OracleError):
        O.o3_velocity_alpha(*(t[keep] for t in (x, y, z, ax, ay, az)), act[keep], _t2())
    # a tree one level deep: the walk never starts, the rotational part is zero
    u, v, w = O.o3_velocity_alpha(x[:40], y[:40], z[:40], ax[:40], ay[:40], az[:40], act[:40],
                                  _t2())
    assert not np.any(u) and not np.any(v) and not np.any(w)
    # every particle at one point: no root box
    with pytest.raises(O.OracleError):
        O.O3Tree(np.zeros(60), np.zeros(60), np.zeros(60), np.ones(60, np.uint8), _t2())
