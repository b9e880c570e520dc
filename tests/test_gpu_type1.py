"""The type-1 FMM on the GPU (vpmfmm_velocity_type1[_device]: the reference's
VelocityCalc::velocity_old -> biotsavart_fmm_2d) against the reference
build's goldens and the type-1 oracle, which the CPU tests pin bit for bit
to that build.

FP64 arithmetic in another order (and exact binomials where the reference
rounds its own): max |d(u, v)| / rms(rotational velocity) <= 1e-10.  The
tree (boxes, counts, particle order in the leaves) and the four lists are
compared as sets of box positions, exactly."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


def _err(u, v, ur, vr):
    sc = np.sqrt(np.mean(ur ** 2 + vr ** 2))
    return float(np.max(np.hypot(u - ur, v - vr)) / sc)


def _field(n, seed, lo=-1.0, hi=1.0):
    g = np.random.Generator(np.random.PCG64(seed))
    return (g.uniform(lo, hi, n), g.uniform(lo, hi, n), 0.02 * (1 + g.uniform(0, 1, n)),
            1e-3 * g.normal(0, 1, n))


@pytest.mark.parametrize("name", ["type1_uniform_2k", "type1_clusters_6k_far",
                                  "type1_resume_00500"])
def test_type1_matches_reference_goldens(gpu, name):
    from lsmps_vpm_v2_amd import FmmSolver
    d = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    s = FmmSolver(u_inf=0.0, v_inf=0.0)
    for i, (ns, lm, nd, ic) in enumerate(d["configs"]):
        u, v = s.velocity_type1(d["x"], d["y"], d["s"], d["g"], n_s=int(ns), lmax=int(lm),
                                ndp=int(nd), icutoff=int(ic))
        err = _err(u, v, d[f"u{i}"], d[f"v{i}"])
        print(f"{name} n_s {ns} lmax {lm} ndp {nd} icutoff {ic}: vs the reference {err:.2e}")
        assert err <= TOL


def test_type1_matches_oracle_over_options(gpu, oracle):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    c = W.random_clusters(7000, seed=51)
    cases = [("uniform", _field(5000, 50)),
             ("clusters_far", (c["x"] * 11 - 60, c["y"] * 11 + 25,
                               np.full(7000, 0.04), c["q"]))]
    s = FmmSolver(u_inf=0.3, v_inf=-0.2)
    for name, (x, y, sz, g) in cases:
        dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, sz, g)]
        for ns, lm, nd, ic, mode in ((50, 7, 10, 2, 1), (8, 10, 6, 0, 1), (30, 8, 16, 3, 0),
                                     (100, 5, 1, 1, 1), (3, 12, 20, 2, 0), (60, 2, 12, 0, 0)):
            u, v = s.velocity_type1_device(*dev, n_s=ns, lmax=lm, ndp=nd, icutoff=ic,
                                           l2l_mode=mode)
            u, v = u.cpu().numpy() - 0.3, v.cpu().numpy() + 0.2
            ur, vr = oracle.t1_velocity(x, y, sz, g, oracle.t1_opts(n_s=ns, lmax=lm, ndp=nd,
                                                                    icutoff=ic, l2l_mode=mode))
            err = _err(u, v, ur, vr)
            print(f"{name} n_s {ns} lmax {lm} ndp {nd} icutoff {ic} l2l {mode}: {err:.2e}")
            assert err <= TOL


def test_type1_tree_and_lists_match_oracle(gpu, oracle):
    """Boxes (level, position, count), the particles of every leaf in order,
    and lists 1-4 of every box, as the oracle's (the reference's) sets."""
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    c = W.random_clusters(3000, seed=61)
    for x, y in ((c["x"] * 19 + 5, c["y"] * 19 - 11), _field(2000, 62)[:2]):
        sz, g = np.full(len(x), 0.01), np.ones(len(x))
        s = FmmSolver()
        s.velocity_type1(x, y, sz, g, n_s=10, lmax=8)
        T = oracle.T1Tree(x, y, oracle.t1_opts(n_s=10, lmax=8))
        B = T.boxes()
        cells = s.tree_cells()
        keep = cells["level"] > 0  # the type-1 hierarchy has no root box
        gpu_boxes = {(int(L), int(i), int(j)): int(n) for L, i, j, n in zip(
            cells["level"][keep], cells["ix"][keep], cells["iy"][keep], cells["par_num"][keep])}
        ora_boxes = {(int(L), int(i), int(j)): int(n) for L, i, j, n in zip(
            B["level"], B["ix"], B["iy"], B["count"])}
        assert gpu_boxes == ora_boxes
        pos = {b: (int(B["level"][b]), int(B["ix"][b]), int(B["iy"][b])) for b in range(T.nboxes)}
        cid = {}
        for k, key in enumerate(zip(cells["level"], cells["ix"], cells["iy"])):
            cid[tuple(int(v) for v in key)] = int(cells["id"][k])
        id2pos = {v: k for k, v in cid.items()}
        # leaf particles, in the reference's order inside each leaf
        lid, poff, pidx = s.tree_leaves(len(x))
        leaves = {id2pos[int(i)]: pidx[poff[k]:poff[k + 1]].tolist() for k, i in enumerate(lid)}
        for b in range(T.nboxes):
            if B["nchild"][b] == 0:
                assert leaves[pos[b]] == T.particles(b).tolist()
        for b in range(T.nboxes):
            got = s.tree_lists(cid[pos[b]])
            for w in range(4):
                want = {pos[int(e)] for e in T.list(b, w + 1)}
                assert {id2pos[int(e)] for e in got[w]} == want, (pos[b], w + 1)


def test_type1_full_size_flat_plate(gpu, oracle):
    """The 779,280-particle flat plate (gz = omega s^2): the GPU type-1 FMM at
    the reference's n_s 50, ndp 10 and Gaussian kernel, lmax 10 (the default
    lmax 7 leaves thousands of particles per leaf here), against the oracle at
    every particle; the exact shift against the regularized direct sum at
    2,000 seeded targets."""
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    f = W.boundary_layer()
    g = f["omega"] * f["s"] ** 2
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (f["x"], f["y"], f["s"], g)]
    s = FmmSolver(u_inf=0.0, v_inf=0.0, timing=True)
    for mode in (1, 0):
        u, v = s.velocity_type1_device(*dev, n_s=50, lmax=10, ndp=10, icutoff=2, l2l_mode=mode)
        u, v = u.cpu().numpy(), v.cpu().numpy()
        st = s.stats()
        ur, vr = oracle.t1_velocity(f["x"], f["y"], f["s"], g,
                                    oracle.t1_opts(n_s=50, lmax=10, ndp=10, icutoff=2,
                                                   l2l_mode=mode))
        err = _err(u, v, ur, vr)
        t = W.subsample(f, 2000, seed=7)
        ud, vd = oracle.t1_direct(f["x"], f["y"], f["s"], g, icutoff=2, targets=t)
        e_dir = float(np.sqrt(np.mean((u[t] - ud) ** 2 + (v[t] - vd) ** 2) /
                              np.mean(ud ** 2 + vd ** 2)))
        print(f"flat plate type-1 l2l {mode}: vs the oracle {err:.2e}, rel-L2 vs direct "
              f"{e_dir:.2e}, {st['ms_total']:.2f} ms, {st['n_cells']} cells")
        assert err <= TOL
        if mode == 0:
            assert e_dir <= 1e-5


def test_type1_entry_points_and_mirrors(gpu):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, Particle, VelocityCalc, velocity_biot_savart
    x, y, sz, g = _field(3000, 71)
    s = FmmSolver(u_inf=1.0, v_inf=0.0)
    u, v = s.velocity_type1(x, y, sz, g)
    dev = [torch.from_numpy(a).cuda() for a in (x, y, sz, g)]
    ud, vd = s.velocity_type1_device(*dev)
    np.testing.assert_array_equal(ud.cpu().numpy(), u)
    np.testing.assert_array_equal(vd.cpu().numpy(), v)
    # icutoff 0 needs no sizes
    u0, _ = s.velocity_type1(x, y, None, g, icutoff=0)
    u0s, _ = s.velocity_type1(x, y, sz, g, icutoff=0)
    np.testing.assert_array_equal(u0, u0s)
    # VelocityCalc type 1 (gz from vorticity s^2 when absent) and the
    # velocity_biot_savart mirror (rotational part only)
    w = g / sz ** 2
    p = Particle(x=x, y=y, s=sz, vorticity=w, isActive=np.ones(3000, bool))
    vc = VelocityCalc(velocity_type=1, u_inf=1.0, v_inf=0.0)
    vc.get_velocity(p)
    uw, vw = s.velocity_type1(x, y, sz, w * sz ** 2)
    np.testing.assert_array_equal(p.u, uw)
    np.testing.assert_array_equal(p.v, vw)
    pi = Particle(x=x, y=y, s=sz, vorticity=w, isActive=np.ones(3000, bool), gz=g)
    velocity_biot_savart(solver=s).biotsavart_fmm_2d(pi, pi, 2, 50, 1, 10)
    r = FmmSolver(u_inf=0.0, v_inf=0.0).velocity_type1(x, y, sz, g)
    np.testing.assert_array_equal(pi.u, r[0])
    np.testing.assert_array_equal(pi.v, r[1])
    assert s.opts.u_inf == 1.0


def test_type1_then_type2_on_one_handle(gpu):
    """A type-1 call leaves the handle's own path as it was: the FMM velocity
    after it equals a fresh handle's, bit for bit, and back again."""
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    f = W.random_clusters(20000, seed=81, active_frac=0.9)
    for prec in ("fp64", "mixed"):
        a = FmmSolver(precision=prec)
        a.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
        t1 = a.velocity_type1(f["x"], f["y"], f["s"], f["omega"] * f["s"] ** 2, n_s=20, lmax=9)
        u, v = a.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
        u0, v0 = FmmSolver(precision=prec).velocity(f["x"], f["y"], f["s"], f["omega"],
                                                    f["active"])
        np.testing.assert_array_equal(u, u0)
        np.testing.assert_array_equal(v, v0)
        t2 = a.velocity_type1(f["x"], f["y"], f["s"], f["omega"] * f["s"] ** 2, n_s=20, lmax=9)
        np.testing.assert_array_equal(t1[0], t2[0])


def test_type1_edge_cases_and_errors(gpu, oracle):
    from lsmps_vpm_v2_amd import FmmSolver, VpmFmmError
    s = FmmSolver(u_inf=0.5, v_inf=0.25)
    # one particle, every particle at one point: the free stream alone
    for x, y in ((np.array([0.3]), np.array([0.1])), (np.full(9, 0.2), np.full(9, -1.0))):
        u, v = s.velocity_type1(x, y, np.full(len(x), 0.1), np.ones(len(x)))
        assert np.all(u == 0.5) and np.all(v == 0.25)
    # two and three particles
    x, y = np.array([0.0, 1.0, 0.3]), np.array([0.0, 0.5, 0.9])
    for n in (2, 3):
        u, v = s.velocity_type1(x[:n], y[:n], np.full(n, 0.2), np.ones(n))
        ur, vr = oracle.t1_velocity(x[:n], y[:n], np.full(n, 0.2), np.ones(n))
        np.testing.assert_allclose(u - 0.5, ur, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(v - 0.25, vr, rtol=1e-12, atol=1e-15)
    # coincident pairs add nothing (no DUPLICATE_POINTS on this path)
    xd, yd, sd, gd = _field(800, 91)
    xd[10], yd[10] = xd[20], yd[20]
    u, v = s.velocity_type1(xd, yd, sd, gd, icutoff=0, n_s=20)
    ur, vr = oracle.t1_velocity(xd, yd, sd, gd, oracle.t1_opts(icutoff=0, n_s=20))
    assert _err(u - 0.5, v - 0.25, ur, vr) <= TOL
    # a line of particles: the last one lies on the level-1 boxes' upper bound
    with pytest.raises(VpmFmmError) as e:
        s.velocity_type1(np.linspace(0, 1, 50), np.zeros(50), np.full(50, 0.1), np.ones(50))
    assert e.value.status == "OUTSIDE_BOXES"
    for kw in (dict(n_s=0), dict(lmax=1), dict(lmax=17), dict(ndp=0), dict(ndp=41),
               dict(icutoff=4), dict(l2l_mode=2)):
        with pytest.raises(VpmFmmError) as e:
            s.velocity_type1(xd, yd, sd, gd, **kw)
        assert e.value.status == "BAD_ARG", kw
    with pytest.raises(VpmFmmError) as e:
        s.velocity_type1(xd, yd, None, gd, icutoff=2)
    assert e.value.status == "BAD_ARG"
    xn = xd.copy()
    xn[3] = np.nan
    with pytest.raises(VpmFmmError) as e:
        s.velocity_type1(xn, yd, sd, gd)
    assert e.value.status == "NONFINITE_INPUT"
    # the handle stays usable
    u, v = s.velocity_type1(xd, yd, sd, gd)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(v))
