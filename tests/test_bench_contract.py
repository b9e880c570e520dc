"""bench.py's command-line contract that holds without a GPU: more than one
GPU is refused with exit status 2 and a message, before torch is imported;
the evaluation kernel's flop count follows the list statistics."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_eval_flops_formula():
    sys.path.insert(0, ROOT)
    import bench
    st = {"p2p_pairs": 10, "m2p_terms": 3, "l2p_terms": 5}
    assert bench.eval_flops(st) == 11 * 10 + 8 * 3 + 8 * 5


class _FakeTree:
    """Root, four level-1 cells (the first one split), four level-2 leaves
    under it; level-2 leaf 0 has two list-2 cells and one list-4 cell."""

    def tree_cells(self):
        import numpy as np
        return {"id": np.arange(9), "level": np.array([0, 1, 1, 1, 1, 2, 2, 2, 2]),
                "is_leaf": np.array([0, 0, 1, 1, 1, 1, 1, 1, 1], np.uint8),
                "is_active": np.array([1, 1, 1, 1, 0, 1, 1, 1, 1], np.uint8),
                "src_num": np.array([30, 12, 5, 6, 0, 3, 3, 3, 3]),
                "par_num": np.array([31, 12, 5, 6, 1, 3, 3, 3, 3])}

    def tree_lists(self, cid, cap=256):
        import numpy as np
        e = np.zeros(0, np.int64)
        return (e, np.array([7, 8]), e, np.array([3])) if cid == 5 else (e, e, e, e)


def test_stage_work_counts_a_small_tree():
    sys.path.insert(0, ROOT)
    import bench
    p = [4, 5, 6] + [0] * 14
    w = bench.stage_work(_FakeTree(), p, fp64=True)
    # P2M over the active leaves: levels 1 (5, 6 particles, p 5), 2 (4 x 3, p 6)
    assert w["p2m_flops"] == 8 * (11 * 6 + 12 * 7)
    # M2M into parents of level >= 1 (FP64): the four level-2 children of cell 1 (p 5)
    assert w["m2m_flops"] == 4 * 8 * 6 * 7 / 2
    # L2L from parents of level >= 2: none
    assert w["l2l_flops"] == 0
    # M2L: two list-2 and (FP64) one list-4 translation at level 2 (p 6)
    assert w["m2l_flops"] == 8 * 3 * 7 ** 2
    assert bench.stage_work(_FakeTree(), p, fp64=False)["m2l_flops"] == 8 * 2 * 7 ** 2
    # the mixed upward pass stops at level 2: no M2M into level 1
    assert bench.stage_work(_FakeTree(), p, fp64=False)["m2m_flops"] == 0


def test_stage_rooflines_time_weighting():
    sys.path.insert(0, ROOT)
    import bench
    st = {"n_list1": 10, "n_list2": 20, "n_list3": 0, "n_list4": 5, "n_cells": 100,
          "p2p_pairs": 1e9, "m2p_terms": 0, "l2p_terms": 0}
    ms = {"ms_tree": 0.1, "ms_lists": 0.05, "ms_upward": 0.05, "ms_m2l": 0.02,
          "ms_downward": 0.03, "ms_eval": 0.25}
    work = {"p2m_flops": 1e8, "m2m_flops": 1e8, "l2l_flops": 1e8, "m2l_flops": 2e8}
    stages, wt = bench.stage_rooflines(st, ms, work, 1000000, fp64=False)
    tree = stages[0]
    assert tree["bound"] == "hbm" and abs(tree["achieved"] - 48e6 / 1e-4 / 1e9) < 1e-9
    ev = stages[-1]
    assert abs(ev["achieved"] - 11e9 / 0.25e-3 / 1e12) < 1e-9
    assert stages[3]["bound"] == "mfma"
    expect = sum(e["ms"] * e["frac"] for e in stages) / sum(e["ms"] for e in stages)
    assert abs(wt - expect) < 1e-15 and 0 < wt < 1
