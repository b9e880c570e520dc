"""vpmfmm_velocity_device_noalloc (through the C ABI) against
vpmfmm_velocity_device on the same inputs: the same kernels on the same tree,
so the outputs must be bitwise equal.  Two calls in a row, the second with a
smaller input, on the FP64 path (reference options), the mixed fixed-order
path and the learned path; a fresh handle whose first call is the
no-allocation one; the size limit; and that the call allocates no device
memory."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(f):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(f[k])).cuda()
            for k in ("x", "y", "s", "omega", "active")]


def _both(a, b, dev):
    """(u, v) of a.velocity_device_noalloc and b.velocity_device as numpy."""
    import torch
    n = dev[0].numel()
    u1, v1 = torch.empty(n, dtype=torch.float64, device="cuda"), torch.empty(
        n, dtype=torch.float64, device="cuda")
    a.velocity_device_noalloc(*dev, u1, v1)
    st = a.stats()
    u2, v2 = b.velocity_device(*dev)
    torch.cuda.synchronize()
    return u1.cpu().numpy(), v1.cpu().numpy(), u2.cpu().numpy(), v2.cpu().numpy(), st


OPTS = {
    "fp64": dict(),
    "mixed": dict(p=8, max_par=48, max_src=48, precision="mixed", binom_mode=0, list4_mode=2),
    "learned": dict(p=12, max_par=48, max_src=48, precision="mixed", binom_mode=0,
                    list4_mode=2, order="learned"),
}


@pytest.mark.parametrize("path", sorted(OPTS))
def test_noalloc_equals_velocity_device_two_calls(gpu, path):
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    from lsmps_vpm_v2_amd._lib import NOALLOC_SCRATCH_BYTES

    big = W.random_clusters(60000, seed=41, active_frac=0.9 if path == "fp64" else 1.0)
    small = W.random_clusters(9000, seed=42, active_frac=0.9 if path == "fp64" else 1.0)
    a = FmmSolver(**OPTS[path])
    b = FmmSolver(**OPTS[path])
    for label, f in (("first", big), ("second, smaller", small)):
        u1, v1, u2, v2, st = _both(a, b, _dev(f))
        print(f"{path} {label} call n={f['n']}: scratch {st['scratch_bytes'] / 2**20:.1f} MiB, "
              f"max |du| {np.max(np.abs(u1 - u2)):.1e}")
        assert 0 < st["scratch_bytes"] <= NOALLOC_SCRATCH_BYTES
        assert st["n"] == f["n"]
        np.testing.assert_array_equal(u1, u2)
        np.testing.assert_array_equal(v1, v2)
    # the handle's own buffers are still good after the no-allocation calls
    u3, v3 = a.velocity_device(*_dev(big))
    u4, v4 = b.velocity_device(*_dev(big))
    np.testing.assert_array_equal(u3.cpu().numpy(), u4.cpu().numpy())
    np.testing.assert_array_equal(v3.cpu().numpy(), v4.cpu().numpy())


def test_noalloc_fresh_handle_allocates_nothing_780k(gpu):
    """A fresh handle whose first call is the no-allocation one, at the
    bench's 779,280-particle field (learned path), then a smaller call; the
    free device memory does not drop across either call."""
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    from lsmps_vpm_v2_amd._lib import NOALLOC_SCRATCH_BYTES

    f = W.boundary_layer()
    sub = W.subsample(f, 200000, seed=3)
    g = {k: (f[k][sub] if isinstance(f[k], np.ndarray) and f[k].shape == (f["n"],) else f[k])
         for k in f}
    g["n"] = len(sub)
    ref = FmmSolver(**OPTS["learned"])
    for label, fld in (("780k", f), ("200k subsample", g)):
        dev = _dev(fld)
        n = fld["n"]
        u1 = torch.empty(n, dtype=torch.float64, device="cuda")
        v1 = torch.empty_like(u1)
        torch.cuda.synchronize()
        if label == "780k":
            a = FmmSolver(**OPTS["learned"])
        free0 = torch.cuda.mem_get_info()[0]
        a.velocity_device_noalloc(*dev, u1, v1)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        st = a.stats()
        u2, v2 = ref.velocity_device(*dev)
        print(f"{label}: scratch {st['scratch_bytes'] / 2**20:.1f} MiB, free memory change "
              f"{(free0 - free1) / 2**20:.1f} MiB, selection {a.last_selection()['source']}")
        assert st["scratch_bytes"] <= NOALLOC_SCRATCH_BYTES
        assert free0 - free1 < (4 << 20)
        np.testing.assert_array_equal(u1.cpu().numpy(), u2.cpu().numpy())
        np.testing.assert_array_equal(v1.cpu().numpy(), v2.cpu().numpy())


def test_noalloc_rejects_sizes_beyond_the_limit(gpu):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, VpmFmmError, workloads as W
    from lsmps_vpm_v2_amd._lib import NOALLOC_MAX_N

    n = NOALLOC_MAX_N + 1
    g = np.random.Generator(np.random.PCG64(5))
    dev = [torch.from_numpy(a).cuda() for a in
           (g.uniform(-1, 1, n), g.uniform(-1, 1, n), np.full(n, 0.01), g.normal(0, 1, n),
            np.ones(n, np.uint8))]
    u, v = torch.empty_like(dev[0]), torch.empty_like(dev[0])
    s = FmmSolver()
    with pytest.raises(VpmFmmError) as e:
        s.velocity_device_noalloc(*dev, u, v)
    assert e.value.status == "BAD_ARG" and "limit" in str(e.value)
    # the handle stays usable
    f = W.random_clusters(3000, seed=9)
    u1, v1, u2, v2, _ = _both(s, FmmSolver(), _dev(f))
    np.testing.assert_array_equal(u1, u2)
    np.testing.assert_array_equal(v1, v2)
