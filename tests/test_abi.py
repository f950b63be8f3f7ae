"""The C-ABI library loads and exports every entry point include/hpt.h
declares, and the Python/numpy mirrors of its structs have the std140 sizes
and offsets of shaders/src/hala-pathtracer/inc/types.glsl.  No compute calls:
these run without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hpt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hpt.h")


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"HPT_API\s+[\w\s\*]+?\b(hpt_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_entry_points()
    for required in ("hpt_create", "hpt_set_scene", "hpt_set_env_map", "hpt_set_ground_color", "hpt_set_sky_color",
                     "hpt_set_env_intensity", "hpt_set_exposure_value", "hpt_set_blue_noise", "hpt_commit",
                     "hpt_render", "hpt_read_image", "hpt_trace_closest", "hpt_trace_any", "hpt_wait_idle",
                     "hpt_destroy", "hpt_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = hpt.load_library()
    for name in declared_entry_points():
        assert hasattr(lib, name), name
    assert set(declared_entry_points()) == set(hpt.SIGNATURES)


def test_last_error_is_empty_string_initially():
    lib = hpt.load_library()
    assert isinstance(lib.hpt_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device():
    lib = hpt.load_library()
    assert lib.hpt_create(None, None) == 1
    assert b"null" in lib.hpt_last_error()
    assert lib.hpt_commit(None) == 1
    lib.hpt_destroy(None)


def test_std140_layouts():
    assert hpt.CAMERA_DTYPE.itemsize == 80
    assert hpt.LIGHT_DTYPE.itemsize == 80
    assert hpt.MATERIAL_DTYPE.itemsize == 144
    assert hpt.MATERIAL_DTYPE.fields["type"][1] == 140
    assert hpt.MATERIAL_DTYPE.fields["ior"][1] == 112
    assert hpt.MATERIAL_DTYPE.fields["clearcoat_tint"][1] == 96
    assert hpt.LIGHT_DTYPE.fields["area"][1] == 64
    assert hpt.CAMERA_DTYPE.fields["type"][1] == 72
    assert C.sizeof(hpt._PrimitiveDesc) == 96
    assert C.sizeof(hpt._CreateInfo) == 48
    assert hpt.RAY_DTYPE.itemsize == 32 and hpt.HIT_DTYPE.itemsize == 24


def test_scene_desc_round_trip_fields():
    import hpt_scenes

    sc = hpt_scenes.cornell_box()
    d, keep = sc.to_desc()
    assert d.camera_count == 1 and d.light_count == 1 and d.material_count == 3
    assert d.primitive_count == 7
    assert sc.triangle_count == 34
    tri = sc.world_triangles()
    assert tri.shape == (34, 3, 3)
    assert np.isclose(tri[:, :, 2].max(), 2.0) and np.isclose(tri[:, :, 2].min(), 0.0, atol=1e-6)


def test_pfm_round_trip(tmp_path):
    img = np.random.default_rng(0).random((5, 7, 3)).astype(np.float32)
    p = tmp_path / "x.pfm"
    hpt.write_pfm(str(p), img)
    assert np.array_equal(hpt.read_pfm(str(p)), img)


NN_HEADER = os.path.join(ROOT, "include", "hpt_nn.h")


def test_denoiser_header_symbols_are_exported_and_typed():
    import denoiser

    names = sorted(set(re.findall(r"HPT_NN_API\s+[\w\s\*]+?\b(hpt_nn_\w+)\s*\(", open(NN_HEADER).read())))
    assert len(names) == 31
    assert set(names) == set(denoiser.NN_SIGNATURES)
    lib = denoiser.nn_library()
    for name in names:
        assert hasattr(lib, name), name


def test_denoiser_entry_points_reject_bad_shapes_without_a_device():
    import denoiser

    lib = denoiser.nn_library()
    # no launch happens: the argument checks come first (1 = hipErrorInvalidValue)
    assert lib.hpt_nn_conv_forward(None, None, None, None, None, None, 1, 8, 8, 16, 32, 32, 3, 1) == 1
    assert lib.hpt_nn_conv_forward(None, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None,
                                   1, 8, 8, 12, 32, 32, 3, 1) == 1   # cin not a multiple of 8
    assert lib.hpt_nn_conv_forward(None, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None,
                                   1, 8, 8, 16, 32, 32, 5, 1) == 1   # only 1x1 and 3x3
    assert lib.hpt_nn_pool_forward(None, C.c_void_p(16), C.c_void_p(16), 1, 7, 8, 16) == 1  # odd height
    assert lib.hpt_nn_adam(None, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 4, 0.1, 0.9,
                           0.999, 1e-8, 0) == 1  # step 0
    assert lib.hpt_nn_wgrad_scratch_floats(1, 32, 32, 16, 32, 3) == 1 * 32 * 9 * 16 + 32


def test_group_and_exr_entry_points_reject_bad_arguments_without_a_device():
    lib = hpt.load_library()
    assert lib.hpt_group_render(None, None, None, None, 0) == 1
    assert b"null" in lib.hpt_last_error()
    assert lib.hpt_group_set_launch_mode(None, 0) == 1
    assert lib.hpt_group_set_shade_sort(None, 0) == 1
    assert lib.hpt_group_wait_idle(None) == 1
    lib.hpt_group_destroy(None)
    assert lib.hpt_set_shade_sort(None, 0) == 1
    assert lib.hpt_exr_info(None, 0, None, None) == 1
    w = C.c_uint32()
    assert lib.hpt_exr_info(b"not an exr file", 15, C.byref(w), C.byref(w)) == 1
    assert b"magic" in lib.hpt_last_error()


def _npz_bytes(tmp_path, compressed=False, **arrays):
    p = tmp_path / ("w.npz")
    (np.savez_compressed if compressed else np.savez)(p, **arrays)
    return p.read_bytes()


def test_denoiser_weights_are_checked_before_the_device(tmp_path):
    """hpt_denoiser_create reads the .npz bytes (stored or deflated zip
    entries, .npy arrays) and checks the layer layout before it touches a
    device: malformed files fail with HPT_ERROR_INVALID_ARGUMENT here, a
    well-formed one gets as far as the device check."""
    lib = hpt.load_library()
    w = os.path.join(ROOT, "hala-pathtracer_amd", "weights", "unet_denoiser.npz")
    good = open(w, "rb").read()
    with np.load(w, allow_pickle=False) as f:
        levels, params, step, residual = f["levels"], f["params"], f["step"], f["residual"]
    h = C.c_void_p()

    def create(data):
        rc = lib.hpt_denoiser_create(data, len(data), 0, C.byref(h))
        return rc, lib.hpt_last_error().decode()

    for data, what in ((b"not a zip at all, just bytes", "zip"), (good[: len(good) // 2], "zip"),
                       (_npz_bytes(tmp_path, levels=levels, params=params[:-1], step=step, residual=residual),
                        "parameters"),
                       (_npz_bytes(tmp_path, levels=levels, step=step, residual=residual), "params"),
                       (_npz_bytes(tmp_path, levels=levels, params=params.astype(np.float64), step=step,
                                   residual=residual), "dtype")):
        rc, msg = create(data)
        assert rc != 0 and what in msg, (rc, msg)
    for data in (good, _npz_bytes(tmp_path, True, levels=levels, params=params, step=step, residual=residual)):
        rc, msg = create(data)
        if rc != 0:  # no GPU here: the file passed and the device check failed
            assert "weights" not in msg, msg
        else:
            lib.hpt_denoiser_destroy(h)
