"""Host-side workload generators: sizes, determinism and source packing."""
import numpy as np

from lsmps_vpm_v2_amd import workloads as W


def test_uniform_vortices_shape_and_seed():
    a, b = W.uniform_vortices(), W.uniform_vortices()
    assert a["n"] == 316 * 316 == 99856
    assert np.array_equal(a["omega"], b["omega"])
    assert not np.array_equal(a["omega"], W.uniform_vortices(seed=2)["omega"])
    h = 0.0632
    assert np.allclose(a["q"], -a["omega"] * h * h / (2 * np.pi), rtol=1e-15)


def test_boundary_layer_is_780k_strip():
    f = W.boundary_layer()
    assert 770_000 <= f["n"] <= 790_000
    assert np.abs(f["y"]).max() < 0.36 and np.abs(f["x"]).max() < 5.1
    assert np.all(np.abs(f["omega"]) >= 0.95e-6)
    assert set(np.unique(f["s"])) == {0.0025, 0.005}
    assert np.all(np.sign(f["omega"]) == np.sign(f["y"]))
    assert np.array_equal(f["x"], W.boundary_layer()["x"])


def test_perlman_closed_form_is_continuous_at_r1():
    f = W.perlman(h=0.02)
    r = np.hypot(f["x"], f["y"])
    m = np.abs(r - 1.0) < 0.02
    u_out = -f["y"][m] / (16 * r[m] ** 2)
    assert np.allclose(f["u_exact"][m], u_out, atol=1e-6)


def test_subsample_is_sorted_unique_and_seeded():
    f = W.random_clusters(5000, 1)
    t = W.subsample(f, 400)
    assert len(t) == 400 and np.all(np.diff(t) > 0)
    assert np.array_equal(t, W.subsample(f, 400))
