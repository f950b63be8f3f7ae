"""hpt_trace_closest_device, the entry point that allocates nothing: its
hits equal hpt_trace_closest's bit for bit on two calls in a row (the second
smaller), over a shallow tree and over a 21-level tree whose traversal stack
spills past the LDS short stack into the library's per-GPU scratch array;
a call the scratch array cannot hold is rejected before any work."""
import numpy as np
import pytest

import denoiser
import hpt
import hpt_scenes

pytestmark = pytest.mark.gpu


def _rays(n, seed, lo, hi, target_hi):
    """Origins uniform in [lo, hi], aimed at points uniform in [0, target_hi]^3."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.uniform(0.0, target_hi, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return hpt.make_rays(o, d)


@pytest.mark.parametrize("case", ["cornell", "nested_triangles"])
def test_trace_closest_device_equals_trace_closest(case):
    if case == "cornell":
        scene, lo, hi, target = hpt_scenes.cornell_box(), (-0.9, -0.9, 0.1), (0.9, 0.9, 1.9), 1.0
    else:
        scene, lo, hi, target = hpt_scenes.nested_triangles(200, 0.7), (0.5, -0.2, -0.2), (1.5, 1.2, 1.2), 0.4
    r = hpt.HalaRenderer(16, 16, max_depth=4, rr_depth=2)
    r.set_scene(scene)
    r.commit()
    depth = r.bvh_info()["max_depth"]
    print(f"[gpu] no-alloc trace {case}: {depth}-level tree", flush=True)
    assert (depth > 8) == (case == "nested_triangles")  # the deep tree uses the scratch array
    for n, seed in ((200_000, 1), (37_123, 2)):
        rays = _rays(n, seed, lo, hi, target)
        want = r.trace_closest(rays)
        d_rays = denoiser.DeviceArray.from_numpy(rays.view(np.uint8))
        d_hits = denoiser.DeviceArray(n * hpt.HIT_DTYPE.itemsize, np.uint8)
        r.trace_closest_device(d_rays.ptr, n, d_hits.ptr)
        r.wait_idle()
        got = d_hits.numpy().view(hpt.HIT_DTYPE)
        assert (got["kind"] == hpt.HIT_TRIANGLE).mean() > 0.3
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    if depth > 8:
        big = 8_388_608 // (depth - 8) + 256
        with pytest.raises(hpt.HptError, match="scratch"):
            r.trace_closest_device(d_rays.ptr, big, d_hits.ptr)
    r.close()
