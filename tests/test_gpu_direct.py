"""The GPU FP64 direct sum (vpmfmm_direct_device) against the CPU oracle's
direct sum (oracle/fmm2d_oracle.c: orc_direct_field, the closed form
E_i = sum_{j != i, mask_j} q_j (x_i - x_j) / r_ij^2 of fmm2D.cpp:135-179
summed over every source).

This kernel is the accuracy reference of every <= 1e-4 claim (the
configuration tests, the bench's error against the direct sum and the
selector's training labels), so it is pinned here on uniform and clustered
fields, over all targets and over target subsets, with and without a source
mask.  Both sums add the sources in index order in FP64; they round each term
differently (q dx / r^2 against (q / r^2) dx, and the GPU may contract
q r^-2 dx + a into one FMA), so they agree to a few units in the last place
per term: rel-L2 <= 1e-13 is written below."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _fields():
    from lsmps_vpm_v2_amd import workloads as W
    out = []
    for n in (257, 4097, 30000):
        g = np.random.Generator(np.random.PCG64(900 + n))
        u = dict(x=g.uniform(-1, 1, n), y=g.uniform(-1, 1, n), q=g.normal(0, 1, n))
        out.append((f"uniform_{n}", u))
        c = W.random_clusters(n, seed=n)
        out.append((f"clusters_{n}", dict(x=c["x"], y=c["y"], q=c["q"])))
    return out


def _rel(ax, ay, bx, by):
    return float(np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2) / np.sum(bx ** 2 + by ** 2)))


@pytest.mark.parametrize("case", range(6))
def test_direct_device_matches_oracle(gpu, case):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver
    from oracle import oracle as O

    name, f = _fields()[case]
    n = len(f["x"])
    s = FmmSolver()
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    g = np.random.Generator(np.random.PCG64(17 + case))
    mask = (g.random(n) < 0.7).astype(np.uint8)
    tsub = np.sort(g.choice(n, size=max(1, n // 7), replace=False)).astype(np.int32)
    # shuffled targets with repeats: the output follows the index order given
    trep = g.integers(0, n, size=max(1, n // 11)).astype(np.int32)
    for label, m, t in (("all", None, None), ("mask", mask, None), ("subset", None, tsub),
                        ("subset+mask", mask, tsub), ("repeats", mask, trep)):
        md = None if m is None else torch.from_numpy(m).cuda()
        td = None if t is None else torch.from_numpy(t).cuda()
        ex, ey = s.direct_device(d["x"], d["y"], d["q"], src_mask=md, targets=td)
        ex, ey = ex.cpu().numpy(), ey.cpu().numpy()
        rx, ry = O.direct_field(f["x"], f["y"], f["q"], targets=t, src_mask=m)
        err = _rel(ex, ey, rx, ry)
        # pointwise too, so no target hides behind the norm: each difference
        # against the larger of the target's own field and the field's RMS (a
        # field that cancels at one target keeps the rounding of its terms)
        mag = np.hypot(rx, ry)
        scale = np.maximum(mag, np.sqrt(np.mean(mag ** 2)))
        pt = np.max(np.hypot(ex - rx, ey - ry) / scale)
        print(f"{name} {label}: rel-L2 {err:.2e}, max pointwise {pt:.2e} of the scale")
        assert ex.shape == rx.shape
        assert err <= TOL, (name, label, err)
        assert pt <= 1e-11, (name, label, pt)


def test_direct_device_planted_error_is_caught(gpu):
    """The comparison above fails for a plausible mistake: the oracle with one
    source's strength flipped in sign differs from the GPU sum by far more
    than the tolerance."""
    import torch
    from lsmps_vpm_v2_amd import FmmSolver
    from oracle import oracle as O

    _, f = _fields()[1]
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    ex, ey = FmmSolver().direct_device(d["x"], d["y"], d["q"])
    q = f["q"].copy()
    q[123] = -q[123]
    rx, ry = O.direct_field(f["x"], f["y"], q)
    assert _rel(ex.cpu().numpy(), ey.cpu().numpy(), rx, ry) > 1e3 * TOL
