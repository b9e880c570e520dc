"""The 3-D treecode on the GPU (vpmfmm_velocity3d[_device|_alpha]: the
reference's VelocityCalc::velocity_fmm_3d_fast -> fmm3D::calcVelocityFast)
against the 3-D oracle (oracle/fmm3d_oracle.c, pinned on the CPU by the
kernel's derivatives, the moment identities and the direct sum).

The octree is compared exactly (every cell's level, key, particle count,
leaf and active marks, centre and side, bit for bit).  Velocities are FP64
sums in another order (and moments shifted from the leaves in one step):
max |d(u, v, w)| / rms(rotational velocity) <= 1e-11."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-11


@pytest.fixture(scope="module")
def solver(gpu):
    from lsmps_vpm_v2_amd import FmmSolver
    s = FmmSolver()
    yield s
    s.close()


def _err(got, ref, free=(0.0, 0.0, 0.0)):
    g = np.stack(got, 1)
    r = np.stack(ref, 1)
    rot = r - np.asarray(free)[None, :]
    sc = np.sqrt(np.mean((rot ** 2).sum(1)))
    return float(np.max(np.linalg.norm(g - r, axis=1)) / sc)


def _cloud(n, seed, kind="uniform", spread=0.08):
    g = np.random.Generator(np.random.PCG64(seed))
    if kind == "uniform":
        p = g.uniform(-1.0, 1.0, (3, n))
    elif kind == "mixed":  # clusters over a uniform background (every root octant dense)
        c = g.uniform(-1.0, 1.0, (5, 3))
        k = g.integers(0, 5, n)
        p = (c[k] + spread * g.standard_normal((n, 3))).T
        bg = g.random(n) < 0.3
        p[:, bg] = g.uniform(-1.0, 1.0, (3, int(bg.sum())))
    else:
        c = g.uniform(-1.0, 1.0, (5, 3))
        k = g.integers(0, 5, n)
        p = (c[k] + spread * g.standard_normal((n, 3))).T
    s = 0.02 * (1 + g.uniform(0, 1, n))
    w = g.standard_normal((3, n))
    return p[0].copy(), p[1].copy(), p[2].copy(), s, w[0].copy(), w[1].copy(), w[2].copy()


FREE = (1.0, 0.0, 0.0)  # the default free stream (global.cpp:633-635)


def _oracle(x, y, z, s, wx, wy, wz, act, crit=30, support=2.0, free=FREE):
    from oracle import oracle as O
    o = O.o3_opts(crit_num=crit, support=support)
    return O.o3_velocity(x, y, z, s, wx, wy, wz, act, o, *free)


CASES = [
    ("uniform_20k", 20000, "uniform", 50, 1.0),
    ("clusters_30k", 30000, "clusters", 50, 0.8),
    ("clusters_8k_crit7", 8000, "clusters", 7, 1.0),
]


@pytest.mark.parametrize("name,n,kind,crit,afrac", CASES, ids=[c[0] for c in CASES])
def test_velocity_matches_oracle(solver, name, n, kind, crit, afrac):
    x, y, z, s, wx, wy, wz = _cloud(n, 3, kind)
    act = (np.random.default_rng(5).random(n) < afrac).astype(np.uint8)
    free = (0.5, -0.25, 0.125)
    print(f"3-D velocity {name} n={n} crit={crit}", flush=True)
    got = solver.velocity3d(x, y, z, s, wx, wy, wz, act, crit_num=crit, u_inf=free[0],
                            v_inf=free[1], w_inf=free[2])
    ref = _oracle(x, y, z, s, wx, wy, wz, act, crit=crit, free=free)
    e = _err(got, ref, free)
    st = solver.fmm3d_stats()
    print(f"  err {e:.2e} cells {st['n_cells']} depth {st['depth']} pairs {st['pairs']}")
    assert e <= TOL


def _cells_by_key(level, key, *cols):
    order = np.lexsort((key, level))
    return tuple(np.asarray(c)[order] for c in (level, key) + cols)


@pytest.mark.parametrize("kind,spread,crit", [("uniform", 0.0, 50), ("clusters", 0.08, 50),
                                              ("clusters", 2e-5, 20)],
                         ids=["uniform", "clusters", "deep_clusters"])
def test_octree_matches_oracle(solver, kind, spread, crit):
    from oracle import oracle as O
    n = 12000
    x, y, z, s, wx, wy, wz = _cloud(n, 7, kind, spread)
    act = (np.random.default_rng(8).random(n) < 0.9).astype(np.uint8)
    print(f"3-D octree {kind} spread={spread} crit={crit}", flush=True)
    solver.velocity3d(x, y, z, s, wx, wy, wz, act, crit_num=crit)
    st = solver.fmm3d_stats()
    G = solver.fmm3d_cells()
    T = O.O3Tree(x, y, z, act, O.o3_opts(crit_num=crit))
    m = T.level > 0  # the GPU keeps the root implicit
    ref = _cells_by_key(T.level[m], T.key[m], T.count[m], T.leaf[m], T.active[m],
                        T.centre[m], T.size[m])
    got = _cells_by_key(G["level"], G["key"], G["count"], G["leaf"], G["active"], G["centre"],
                        G["size"])
    assert st["depth"] == T.depth
    assert (st["sorted_levels"] > 10) == (T.depth > 10)
    assert st["n_cells"] == int(m.sum())
    assert st["n_leaves"] == int((T.leaf[m] == 1).sum())
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(st["root_centre"], T.centre[0])
    assert st["root_side"] == T.size[0]
    # leaves hold the oracle's particles: in depth-first order the leaves
    # cover the tree-ordered particles one after another
    perm = solver.fmm3d_perm()
    idx = {(int(l), int(k)): c for c, (l, k) in enumerate(zip(T.level, T.key))}
    begin = 0
    for j in range(len(G["level"])):
        L, k, c = int(G["level"][j]), int(G["key"][j]), int(G["count"][j])
        if G["leaf"][j]:
            mine = np.sort(perm[begin:begin + c])
            np.testing.assert_array_equal(mine, T.particles(idx[(L, k)]))
            begin += c
    assert begin == n
    if st["depth"] > 10:
        assert st["sorted_levels"] == 20


def test_vortex_ring_matches_oracle(solver):
    from lsmps_vpm_v2_amd import workloads as W
    f = W.vortex_ring(60000, jitter=0.25, active_frac=0.95)
    print(f"3-D vortex ring n={f['n']}", flush=True)
    got = solver.velocity3d(f["x"], f["y"], f["z"], f["s"], f["wx"], f["wy"], f["wz"],
                            f["active"])
    ref = _oracle(f["x"], f["y"], f["z"], f["s"], f["wx"], f["wy"], f["wz"], f["active"])
    e = _err(got, ref, FREE)
    print(f"  err {e:.2e}")
    assert e <= TOL
    # the ring moves along +z: the mean axial velocity over the core is positive
    core = f["wx"] ** 2 + f["wy"] ** 2 > 0.25 * np.max(f["wx"] ** 2 + f["wy"] ** 2)
    assert np.mean(got[2][core]) > 0


def test_alpha_entry_matches_oracle(solver):
    from oracle import oracle as O
    x, y, z, s, wx, wy, wz = _cloud(15000, 11, "clusters")
    al = [c * s ** 3 / (4 * np.pi) for c in (wx, wy, wz)]
    act = np.ones(15000, np.uint8)
    print("3-D alpha entry n=15000", flush=True)
    got = solver.velocity3d_alpha(x, y, z, *al, act, u_inf=3.0)  # free stream not added
    ref = O.o3_velocity_alpha(x, y, z, *al, act)
    assert _err(got, ref) <= TOL
    from lsmps_vpm_v2_amd import fmm3D
    tool = fmm3D(solver)
    tool.calcVelocityFast(np.stack([x, y, z], 1), act, np.ones(15000, bool), *al)
    np.testing.assert_array_equal(tool.get_Field_x(), got[0])
    np.testing.assert_array_equal(tool.get_Field_z(), got[2])


def test_device_entry_and_determinism(solver):
    import torch
    x, y, z, s, wx, wy, wz = _cloud(25000, 13, "clusters")
    act = np.ones(25000, np.uint8)
    print("3-D device entry n=25000", flush=True)
    h = solver.velocity3d(x, y, z, s, wx, wy, wz, act, u_inf=1.0)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z, s, wx, wy, wz)]
    ta = torch.from_numpy(act).cuda()
    u, v, w = solver.velocity3d_device(*dev, ta, u_inf=1.0)
    torch.cuda.synchronize()
    for a, b in zip((u, v, w), h):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    # a deep tree in between does not change the rounding of the next call
    xd, yd, zd, sd, *wd = _cloud(12000, 7, "clusters", 2e-5)
    solver.velocity3d(xd, yd, zd, sd, *wd, np.ones(12000, np.uint8), crit_num=20)
    assert solver.fmm3d_stats()["sorted_levels"] == 20
    h2 = solver.velocity3d(x, y, z, s, wx, wy, wz, act, u_inf=1.0)
    assert solver.fmm3d_stats()["sorted_levels"] == 10
    for a, b in zip(h2, h):
        np.testing.assert_array_equal(a, b)


def test_small_and_level1_leaf_cases(solver):
    from oracle import oracle as O
    x, y, z, s, wx, wy, wz = _cloud(25, 17)
    act = np.ones(25, np.uint8)
    print("3-D fewer than crit_num particles", flush=True)
    u, v, w = solver.velocity3d(x, y, z, s, wx, wy, wz, act, u_inf=0.25, w_inf=-1.0)
    assert np.all(u == 0.25) and np.all(v == 0.0) and np.all(w == -1.0)
    # one root octant with fewer than crit_num particles: its sources are not summed
    x, y, z, s, wx, wy, wz = _cloud(3000, 13)
    few = np.flatnonzero((x > 0) & (y > 0) & (z > 0))
    keep = np.setdiff1d(np.arange(3000), few[20:])
    cols = [t[keep] for t in (x, y, z, s, wx, wy, wz)]
    act = np.ones(len(keep), np.uint8)
    print("3-D level-1 leaf", flush=True)
    got = solver.velocity3d(*cols, act)
    ref = O.o3_velocity(*cols, act, O.o3_opts(), *FREE)
    assert _err(got, ref, FREE) <= TOL
    one = solver.velocity3d(x[:1], y[:1], z[:1], s[:1], wx[:1], wy[:1], wz[:1],
                            np.ones(1, np.uint8), u_inf=0.0, v_inf=2.0)
    assert one[0][0] == 0.0 and one[1][0] == 2.0 and one[2][0] == 0.0


def test_error_paths(solver):
    from lsmps_vpm_v2_amd import VpmFmmError
    x, y, z, s, wx, wy, wz = _cloud(500, 19)
    act = np.ones(500, np.uint8)
    bad = x.copy()
    bad[7] = np.inf
    print("3-D error paths", flush=True)
    with pytest.raises(VpmFmmError) as e:
        solver.velocity3d(bad, y, z, s, wx, wy, wz, act)
    assert e.value.code == 2
    # crit_num coincident particles never separate
    xs, ys, zs = (np.r_[t, np.full(60, 0.3)] for t in (x, y, z))
    ss, w1, w2, w3 = (np.r_[t, t[:60]] for t in (s, wx, wy, wz))
    with pytest.raises(VpmFmmError) as e:
        solver.velocity3d(xs, ys, zs, ss, w1, w2, w3, np.ones(560, np.uint8))
    assert e.value.code == 4
    with pytest.raises(VpmFmmError) as e:
        solver.velocity3d(x, y, z, s, wx, wy, wz, act, crit_num=0)
    assert e.value.code == 1
    # the handle still works
    got = solver.velocity3d(x, y, z, s, wx, wy, wz, act)
    ref = _oracle(x, y, z, s, wx, wy, wz, act)
    assert _err(got, ref, FREE) <= TOL


def test_velocity_calc_dim3(solver):
    from lsmps_vpm_v2_amd import Particle, VelocityCalc
    x, y, z, s, wx, wy, wz = _cloud(5000, 23, "clusters")
    act = np.ones(5000, np.uint8)
    p = Particle(x=x, y=y, s=s, vorticity=np.zeros(5000), isActive=act, z=z, vortx=wx,
                 vorty=wy, vortz=wz)
    vc = VelocityCalc(dim=3, u_inf=1.0, v_inf=0.0)
    print("3-D VelocityCalc", flush=True)
    vc.get_velocity(p)
    ref = _oracle(x, y, z, s, wx, wy, wz, act, free=(1.0, 0.0, 0.0))
    assert _err((p.u, p.v, p.w), ref, (1.0, 0.0, 0.0)) <= TOL


# ---------------------------------------------------------------------------
# type 2: fmm3D::calcVelocity over the 3-D TreeCell (VELOCITY_3D_TYPE 2)
# ---------------------------------------------------------------------------
def _oracle_t2(x, y, z, s, wx, wy, wz, act, free=FREE, **kw):
    from oracle import oracle as O
    return O.o3_velocity(x, y, z, s, wx, wy, wz, act, O.o3_opts(type=2, **kw), *free)


T2_CASES = [
    ("uniform_20k", 20000, "uniform", 1.0),
    ("mixed_30k_inactive", 30000, "mixed", 0.7),
]


@pytest.mark.parametrize("name,n,kind,afrac", T2_CASES, ids=[c[0] for c in T2_CASES])
def test_type2_velocity_matches_oracle(solver, name, n, kind, afrac):
    x, y, z, s, wx, wy, wz = _cloud(n, 43, kind)
    act = (np.random.default_rng(6).random(n) < afrac).astype(np.uint8)
    print(f"3-D type 2 velocity {name} n={n}", flush=True)
    got = solver.velocity3d(x, y, z, s, wx, wy, wz, act, type=2)
    ref = _oracle_t2(x, y, z, s, wx, wy, wz, act)
    e = _err(got, ref, FREE)
    st = solver.fmm3d_stats()
    print(f"  err {e:.2e} cells {st['n_cells']} depth {st['depth']} pairs {st['pairs']}")
    assert e <= TOL
    # other leaf limits
    got = solver.velocity3d(x, y, z, s, wx, wy, wz, act, type=2, max_par=20, max_src=12)
    ref = _oracle_t2(x, y, z, s, wx, wy, wz, act, max_par=20, max_src=12)
    assert _err(got, ref, FREE) <= TOL


@pytest.mark.parametrize("kind,spread", [("uniform", 0.0), ("mixed", 0.08),
                                         ("mixed", 2e-5)],
                         ids=["uniform", "mixed", "deep_mixed"])
def test_type2_octree_matches_oracle(solver, kind, spread):
    from oracle import oracle as O
    n = 12000
    x, y, z, s, wx, wy, wz = _cloud(n, 47, kind, spread)
    act = (np.random.default_rng(9).random(n) < 0.8).astype(np.uint8)
    print(f"3-D type 2 octree {kind} spread={spread}", flush=True)
    solver.velocity3d(x, y, z, s, wx, wy, wz, act, type=2)
    st = solver.fmm3d_stats()
    G = solver.fmm3d_cells()
    T = O.O3Tree(x, y, z, act, O.o3_opts(type=2))
    m = T.level > 0
    ref = _cells_by_key(T.level[m], T.key[m], T.count[m], T.leaf[m], T.active[m],
                        T.centre[m], T.size[m])
    got = _cells_by_key(G["level"], G["key"], G["count"], G["leaf"], G["active"], G["centre"],
                        G["size"])
    assert st["depth"] == T.depth
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(st["root_centre"], T.centre[0])
    assert st["root_side"] == T.size[0]
    perm = solver.fmm3d_perm()
    idx = {(int(l), int(k)): c for c, (l, k) in enumerate(zip(T.level, T.key))}
    begin = 0
    for j in range(len(G["level"])):
        if G["leaf"][j]:
            c = int(G["count"][j])
            np.testing.assert_array_equal(np.sort(perm[begin:begin + c]),
                                          T.particles(idx[(int(G["level"][j]),
                                                           int(G["key"][j]))]))
            begin += c
    assert begin == n


def test_type2_ring_edges_and_mirror(solver):
    from lsmps_vpm_v2_amd import Particle, VelocityCalc, VpmFmmError, workloads as W
    f = W.vortex_ring(60000, jitter=0.25, active_frac=0.9)
    cols = [f[k] for k in ("x", "y", "z", "s", "wx", "wy", "wz")]
    print(f"3-D type 2 vortex ring n={f['n']}", flush=True)
    got = solver.velocity3d(*cols, f["active"], type=2)
    ref = _oracle_t2(*cols, f["active"])
    assert _err(got, ref, FREE) <= TOL
    # a level-1 leaf beside deeper cells: the reference throws; here LEVEL1_LEAF
    x, y, z, s, wx, wy, wz = _cloud(20000, 41)
    few = np.flatnonzero((x > 0) & (y > 0) & (z > 0))
    keep = np.setdiff1d(np.arange(20000), few[20:])
    cols = [t[keep] for t in (x, y, z, s, wx, wy, wz)]
    with pytest.raises(VpmFmmError) as e:
        solver.velocity3d(*cols, np.ones(len(keep), np.uint8), type=2)
    assert e.value.code == 10
    # one level deep: the walk never starts, the free stream alone
    u, v, w = solver.velocity3d(x[:40], y[:40], z[:40], s[:40], wx[:40], wy[:40], wz[:40],
                                np.ones(40, np.uint8), type=2, u_inf=0.5)
    assert np.all(u == 0.5) and np.all(v == 0.0) and np.all(w == 0.0)
    # every particle at one point
    with pytest.raises(VpmFmmError) as e:
        solver.velocity3d(np.zeros(60), np.zeros(60), np.zeros(60), np.full(60, 0.1),
                          np.ones(60), np.ones(60), np.ones(60), np.ones(60, np.uint8), type=2)
    assert e.value.code == 3
    # the mirror's VELOCITY_3D_TYPE 2
    x, y, z, s, wx, wy, wz = _cloud(6000, 53, "mixed")
    act = np.ones(6000, np.uint8)
    p = Particle(x=x, y=y, s=s, vorticity=np.zeros(6000), isActive=act, z=z, vortx=wx,
                 vorty=wy, vortz=wz)
    VelocityCalc(dim=3, velocity_type=2).get_velocity(p)
    ref = _oracle_t2(x, y, z, s, wx, wy, wz, act)
    assert _err((p.u, p.v, p.w), ref, FREE) <= TOL
    # type 1 after type 2 on the same handle is unchanged
    a = solver.velocity3d(x, y, z, s, wx, wy, wz, act)
    b = _oracle(x, y, z, s, wx, wy, wz, act)
    assert _err(a, b, FREE) <= TOL


def test_no_active_particle_and_single_particle(solver):
    """No active particle: every cell is inactive and skipped, the velocity
    is the free stream (both types); one particle: the free stream."""
    x, y, z, s, wx, wy, wz = _cloud(5000, 59, "mixed")
    act = np.zeros(5000, np.uint8)
    print("3-D no active particle", flush=True)
    for t in (1, 2):
        u, v, w = solver.velocity3d(x, y, z, s, wx, wy, wz, act, type=t, u_inf=0.75,
                                    w_inf=-0.5)
        assert np.all(u == 0.75) and np.all(v == 0.0) and np.all(w == -0.5)
        assert solver.fmm3d_stats()["n_active"] == 0
    # one particle: type 1's box holds the domain centre too; type 2's box
    # has no size (the reference divides by a zero cell size there)
    u, v, w = solver.velocity3d(x[:1], y[:1], z[:1], s[:1], wx[:1], wy[:1], wz[:1],
                                np.ones(1, np.uint8), u_inf=0.0)
    assert u[0] == 0.0 and v[0] == 0.0 and w[0] == 0.0
    from lsmps_vpm_v2_amd import VpmFmmError
    with pytest.raises(VpmFmmError) as e:
        solver.velocity3d(x[:1], y[:1], z[:1], s[:1], wx[:1], wy[:1], wz[:1],
                          np.ones(1, np.uint8), type=2)
    assert e.value.code == 3
