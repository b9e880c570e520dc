"""Stage-level parity of the FP64 path against the oracle.

Each evaluation part (P2P over list 1, M2P over list 3, L2P of the local
expansion) is compared on its own, and every cell's multipole (P2M + M2M) and
local (M2L + L2L) coefficients are compared cell by cell.  Tolerance: 1e-10
relative to the largest coefficient magnitude of the same order (rows), or
rel-L2 1e-10 for fields.
"""
import numpy as np
import pytest

from conftest import rel_l2
from lsmps_vpm_v2_amd import FmmSolver
from lsmps_vpm_v2_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mask", [1, 2, 4])
def test_stage_fields(gpu, oracle, mask):
    f = W.random_clusters(5000, 21, active_frac=0.8)
    ex0, ey0 = oracle.fmm_field(f["x"], f["y"], f["q"], f["active"],
                                oracle.opts(p=10, stage_mask=mask))
    s = FmmSolver(p=10, stage_mask=mask)
    ex, ey = s.field(f["x"], f["y"], f["q"], f["active"])
    err = rel_l2(ex, ey, ex0, ey0)
    print(f"stage_mask={mask}: rel-L2 {err:.3e}")
    assert err <= 1e-10


@pytest.mark.parametrize("p,binom", [(10, 1), (12, 0)])
def test_expansions(gpu, oracle, p, binom):
    f = W.random_clusters(5000, 22, active_frac=0.8)
    ids0, ak0, bk0 = oracle.fmm_expansions(f["x"], f["y"], f["q"], f["active"],
                                           oracle.opts(p=p, binom_mode=binom))
    s = FmmSolver(p=p, binom_mode=binom)
    s.field(f["x"], f["y"], f["q"], f["active"])
    t = oracle.Tree(f["x"], f["y"], f["active"], oracle.opts(p=p))
    rg, pg = s.tree_geometry()
    ro, po = t.geometry()
    print("geometry gpu", rg, pg, "oracle", ro, po)
    assert rg == ro and np.array_equal(pg, po)
    ids, ak, bk = s.tree_expansions()
    assert np.array_equal(ids, ids0)
    # b_0 is not formed on the device (it never reaches the field)
    for name, a, b in (("ak", ak, ak0), ("bk", bk[:, 1:], bk0[:, 1:])):
        scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
        err = (np.abs(a - b) / scale).max(axis=0)
        print(name, "max err per order", np.array2string(err, precision=2))
        assert err.max() <= 1e-10, name
