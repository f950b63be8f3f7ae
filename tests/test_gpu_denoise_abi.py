"""The denoiser behind the C ABI (hpt_denoiser_create / hpt_denoise), the
path a host without Python takes in place of denoise.py:26-30: the same
bytes as the Python driver of the same kernels (denoiser.UNet.denoise_renderer)
for host and device destinations, several image sizes (multiples of 16 and
not) and sample counts."""
import numpy as np
import pytest

import denoiser
import hpt
import hpt_scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nets():
    path = denoiser.weights_path()
    return hpt.Denoiser(path), denoiser.UNet.load(path)


@pytest.mark.parametrize("size,spp", [((96, 64), 8), ((100, 60), 3), ((320, 192), 17)])
def test_c_abi_denoiser_equals_python_driver(nets, size, spp):
    c_net, py_net = nets
    w, h = size
    print(f"[gpu] C-ABI denoiser {w}x{h} at {spp} spp", flush=True)
    r = hpt.HalaRenderer(w, h, max_depth=8, rr_depth=3)
    r.set_scene(hpt_scenes.cornell_box())
    r.commit()
    r.render(spp)
    got = c_net.denoise(r)
    want = py_net.denoise_renderer(r).numpy()
    denoiser.synchronize()
    assert got.shape == (h, w, 4) and np.isfinite(got).all()
    assert np.array_equal(got, want)
    # device destination
    dev = denoiser.DeviceArray(h * w * 4, np.float32, (h, w, 4))
    c_net.denoise_into(r, dev.ptr)
    assert np.array_equal(dev.numpy(), want)
    # the denoised image is not the noisy one
    assert not np.allclose(got[..., :3], r.read_image(hpt.IMAGE_ACCUM)[..., :3])
    r.close()


def test_c_abi_denoiser_rejects_wrong_sizes(nets):
    c_net, _ = nets
    r = hpt.HalaRenderer(32, 32, max_depth=4, rr_depth=2)
    r.set_scene(hpt_scenes.cornell_box())
    r.commit()
    out = np.empty((16, 32, 4), np.float32)
    lib = hpt.load_library()
    assert lib.hpt_denoise(c_net._h, r._h, out.ctypes.data, out.nbytes, 0) == 1  # HPT_ERROR_INVALID_ARGUMENT
    assert b"size" in lib.hpt_last_error()
    r.close()


def test_c_abi_denoiser_two_widths_with_one_padded_size(nets):
    """Widths 97 and 100 pad to the same network size (112), so the second
    call reuses the first call's buffers: its host copy must still hold the
    larger image (the buffer is sized by the padded size)."""
    c_net, py_net = nets
    for w in (97, 100):
        print(f"[gpu] C-ABI denoiser {w}x60 after a smaller image", flush=True)
        r = hpt.HalaRenderer(w, 60, max_depth=6, rr_depth=3)
        r.set_scene(hpt_scenes.cornell_box())
        r.commit()
        r.render(4)
        got = c_net.denoise(r)
        want = py_net.denoise_renderer(r).numpy()
        denoiser.synchronize()
        assert np.array_equal(got, want)
        r.close()
