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
