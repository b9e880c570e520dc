"""GPU parity of the regularized near field (opts.icutoff 1 / 2 / 3).

Reference: the oracle FMM with the same icutoff (its P2P follows the
reference's type-1 direct_sum, subroutine_all_fmm_2d.cpp:594-627, with
regul_func_2d, regularization_function_2d.cpp:11-37), pinned against the
numpy regularized direct sum and the Lamb-Oseen closed form on the CPU.
* FP64 path: rel-L2 <= 1e-10 against the oracle (uniform sizes and sizes
  varying by 4x), for every kernel;
* mixed path: <= 5e-6 against the oracle in FP64 (the FP32 bar of the
  singular kernel);
* learned path: <= 1e-4 against the regularized direct sum;
* a field call (no sizes) with icutoff > 0 is refused; coincident particles
  add nothing instead of raising DUPLICATE_POINTS.
"""
import numpy as np
import pytest

from conftest import rel_l2
from lsmps_vpm_v2_amd import FmmSolver, VpmFmmError
from lsmps_vpm_v2_amd import workloads as W

pytestmark = pytest.mark.gpu


def _multires(n=6000, seed=71):
    f = W.random_clusters(n, seed)
    rng = np.random.default_rng(seed)
    s = f["s"] * rng.uniform(0.5, 2.0, n)
    return dict(f, s=s)


@pytest.mark.parametrize("icut", [1, 2, 3])
@pytest.mark.parametrize("field", ["perlman", "multires"])
def test_fp64_matches_oracle(gpu, oracle, icut, field):
    f = W.perlman() if field == "perlman" else _multires()
    o = oracle.opts(icutoff=icut)
    u0, v0 = oracle.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"], o)
    s = FmmSolver(icutoff=icut)
    u, v = s.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
    err = rel_l2(u - 1.0, v, u0 - 1.0, v0)
    print(f"{field} icutoff {icut}: FP64 vs oracle {err:.2e}")
    assert err <= 1e-10


@pytest.mark.parametrize("icut", [1, 2, 3])
def test_mixed_matches_oracle(gpu, oracle, icut):
    f = _multires(8000, 72)
    o = oracle.opts(p=12, binom_mode=0, list4_mode=2, icutoff=icut)
    u0, v0 = oracle.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"], o)
    s = FmmSolver(p=12, precision="mixed", binom_mode=0, list4_mode=2, icutoff=icut)
    u, v = s.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
    err = rel_l2(u - 1.0, v, u0 - 1.0, v0)
    print(f"icutoff {icut}: mixed vs oracle {err:.2e}")
    assert err <= 5e-6


def test_learned_against_regularized_direct(gpu, oracle):
    f = W.perlman()
    t = W.subsample(f, 2000)
    ud, vd = oracle.regularized_direct_velocity(f["x"], f["y"], f["s"], f["omega"], 2,
                                                targets=t)
    s = FmmSolver(precision="mixed", p=12, max_par=48, max_src=48, binom_mode=0, list4_mode=2,
                  order="learned", icutoff=2, u_inf=0.0, v_inf=0.0)
    u, v = s.velocity(f["x"], f["y"], f["s"], f["omega"], f["active"])
    err = rel_l2(u[t], v[t], ud, vd)
    print(f"learned icutoff 2 vs regularized direct {err:.2e}, source "
          f"{s.last_selection()['source']}")
    assert err <= 1e-4


def test_field_call_refused(gpu):
    f = W.random_clusters(500, 73)
    s = FmmSolver(icutoff=2)
    with pytest.raises(VpmFmmError):
        s.field(f["x"], f["y"], f["q"], f["active"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coincident_particles_add_nothing(gpu, oracle, precision):
    f = W.random_clusters(3000, 74)
    x, y = f["x"].copy(), f["y"].copy()
    x[10:20], y[10:20] = x[0], y[0]  # ten particles on top of particle 0
    kw = dict(binom_mode=0, list4_mode=2, p=12)
    o = oracle.opts(icutoff=2, **kw)
    u0, v0 = oracle.velocity(x, y, f["s"], f["omega"], f["active"], o)
    s = FmmSolver(icutoff=2, precision=precision, **kw)
    u, v = s.velocity(x, y, f["s"], f["omega"], f["active"])
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(v))
    assert rel_l2(u - 1.0, v, u0 - 1.0, v0) <= (1e-10 if precision == "fp64" else 5e-6)
