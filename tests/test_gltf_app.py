"""Host side above the C ABI: the glTF scene loader, the YAML config of the
reference application (src/config/*.rs) and the environment readers.

The glTF conversion lives in the reference's unvendored renderer, so these
tests pin the loader's stated rules: exact round trips through the writer,
the structure of the reference's own assets (tests/golden/gltf), and
closed-form checks of the conversions (light radiance from the exported
power, camera axes from node rotations).
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import app
import gltf_scene
import hpt
import hpt_scenes
import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gltf")


def _world_tris(scene):
    return scene.world_triangles().astype(np.float64)


def test_round_trip_keeps_geometry_materials_lights_cameras(tmp_path):
    src = hpt_scenes.displaced_icosphere_lights_medium(level=2)
    path = str(tmp_path / "scene.gltf")
    gltf_scene.write_gltf(path, src)
    got = gltf_scene.load_gltf(path).scene
    assert got.triangle_count == src.triangle_count
    assert np.allclose(_world_tris(got), _world_tris(src), atol=1e-5)
    n = len(src.materials)
    for k in hpt.MATERIAL_DTYPE.names:
        if not k.endswith("_map_index"):
            assert np.allclose(got.materials[k][:n], src.materials[k]), k
    assert len(got.lights) == len(src.lights) == 5
    for a, b in zip(got.lights, src.lights):
        assert a["type"] == b["type"]
        for k in ("intensity", "position", "u", "v", "radius", "area"):
            assert np.allclose(a[k], b[k], rtol=1e-5, atol=1e-5), (int(a["type"]), k, a[k], b[k])
    for k in ("position", "right", "up", "forward", "yfov", "focal_distance_or_xmag", "aperture_or_ymag", "type"):
        assert np.allclose(got.cameras[0][k], src.cameras[0][k], atol=1e-6), k


def test_glb_container(tmp_path):
    src = hpt_scenes.cornell_box()
    path = str(tmp_path / "c.gltf")
    gltf_scene.write_gltf(path, src)
    doc = json.load(open(path))
    blob = open(str(tmp_path / "c.bin"), "rb").read()
    del doc["buffers"][0]["uri"]
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    blob += b"\0" * (-len(blob) % 4)
    glb = struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js) + 8 + len(blob))
    glb += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
    gpath = str(tmp_path / "c.glb")
    open(gpath, "wb").write(glb)
    got = gltf_scene.load_gltf(gpath).scene
    assert np.allclose(_world_tris(got), _world_tris(src), atol=1e-6)


def test_reference_cornell_asset():
    g = gltf_scene.load_gltf(os.path.join(GOLDEN, "cornell-box.gltf"))
    sc = g.scene
    assert sc.triangle_count == 34 and len(sc.primitives) == 7
    assert g.names["meshes"] == ["Floor", "Ceiling", "BackWall", "RightWall", "LeftWall", "ShortBox", "TallBox"]
    w = _world_tris(sc).reshape(-1, 3)
    assert np.allclose(w.min(0), [-1, -1, 0], atol=1e-5) and np.allclose(w.max(0), [1, 1, 2], atol=1e-5)
    assert (sc.materials["type"][:7] == hpt.MATERIAL_DIFFUSE).all()  # metallic 0, no Disney extensions
    # Camera node: 90 degrees about +x, so its -Z looks along +y and its +Y is +z.
    cam = sc.cameras[0]
    assert np.allclose(cam["position"], [0, -8.784, 1], atol=1e-5)
    assert np.allclose(cam["forward"], [0, 1, 0], atol=1e-6) and np.allclose(cam["up"], [0, 0, 1], atol=1e-6)
    assert math.isclose(cam["yfov"], 0.2563342779540469, rel_tol=1e-6)
    # The "Area" light: a point light in the file with extras {type 1, 0.47 x 0.38}:
    # a rectangle centred on the node, emitting downwards with 4 * I / A.
    L = sc.lights[0]
    assert L["type"] == hpt.LIGHT_QUAD
    u, v = np.asarray(L["u"], np.float64), np.asarray(L["v"], np.float64)
    centre = np.asarray(L["position"], np.float64) + 0.5 * (u + v)
    assert np.allclose(centre, [-0.005, 0.03, 1.999], atol=1e-5)
    assert math.isclose(np.linalg.norm(u), 0.47, rel_tol=1e-6) and math.isclose(np.linalg.norm(v), 0.38, rel_tol=1e-6)
    n = np.cross(v, u)
    assert np.allclose(n / np.linalg.norm(n), [0, 0, -1], atol=1e-6)
    assert math.isclose(L["area"], 0.47 * 0.38, rel_tol=1e-6)
    assert np.allclose(L["intensity"], np.array([1, 0.7083765864372253, 0.23455078899860382]) *
                       4 * 1.5915494309189535 / (0.47 * 0.38), rtol=1e-5)


def test_reference_orthographic_asset():
    g = gltf_scene.load_gltf(os.path.join(GOLDEN, "orthographic_camera.gltf"))
    sc = g.scene
    cam = sc.cameras[0]
    assert cam["type"] == hpt.CAMERA_ORTHOGRAPHIC
    assert math.isclose(cam["focal_distance_or_xmag"], 5.107142925262451, rel_tol=1e-6)
    assert math.isclose(cam["aperture_or_ymag"], 3.8303571939468384, rel_tol=1e-6)
    q = [0.483536034822464, 0.20870360732078552, 0.33687159419059753, 0.7804827094078064]
    R = gltf_scene._quat_matrix(q)
    assert np.allclose(cam["forward"], -R[:, 2], atol=1e-6) and np.allclose(cam["right"], R[:, 0], atol=1e-6)
    L = sc.lights[0]
    assert L["type"] == hpt.LIGHT_DIRECTIONAL
    R = gltf_scene._quat_matrix([0.16907574236392975, 0.27217137813568115, 0.7558803558349609, 0.570947527885437])
    assert np.allclose(L["u"], -R[:, 2], atol=1e-6)  # travels along the node's -Z
    assert sc.materials[0]["type"] == hpt.MATERIAL_DIFFUSE and math.isclose(sc.materials[0]["roughness"], 0.5)
    assert sc.triangle_count == 12


def test_embedded_texture_and_missing_attributes(tmp_path):
    from PIL import Image
    import io

    img = np.zeros((2, 3, 4), np.uint8)
    img[..., 0] = [[10, 20, 30], [40, 50, 60]]
    img[..., 3] = 255
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, format="PNG")
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 1]}],
           "nodes": [{"mesh": 0}, {"camera": 0, "translation": [0, 0, 5]}],
           "cameras": [{"type": "perspective", "perspective": {"yfov": 0.5, "znear": 0.1}}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "material": 0}]}],
           "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 0.0}}],
           "textures": [{"source": 0}],
           "images": [{"uri": "data:image/png;base64," + __import__("base64").b64encode(buf.getvalue()).decode()}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}],
           "bufferViews": [{"buffer": 0, "byteLength": 36}],
           "buffers": [{"byteLength": 36, "uri": "data:application/octet-stream;base64," +
                        __import__("base64").b64encode(pos.tobytes()).decode()}]}
    path = str(tmp_path / "t.gltf")
    json.dump(doc, open(path, "w"))
    g = gltf_scene.load_gltf(path)
    sc = g.scene
    assert len(sc.textures) == 1 and np.array_equal(sc.textures[0], img)
    assert sc.materials[0]["base_color_map_index"] == 0
    v = sc.primitives[0].vertices
    assert np.allclose(v[:, 3:6], [0, 0, 1])        # flat normal of the un-normalled triangle
    assert np.allclose(v[:, 6:9], [0, -1, 0])       # onb tangent: cross(up = +x, N = +z) = -y
    assert np.allclose(sc.cameras[0]["forward"], [0, 0, -1])


def test_config_parsing_and_validation(tmp_path):
    scene = tmp_path / "s.gltf"
    gltf_scene.write_gltf(str(scene), hpt_scenes.cornell_box())
    text = f"""window:
  width: 640
  height: 480
tracer:
  max_depth: 8
  rr_depth: 4
  exposure_value: 0.5
  tonemap:
    enable: true
    enable_aces: true
  max_samples: 0
scene_file: {scene}
hdri_file:
hdri_rotation: 0.0
sky_color: [0.1, 0.2, 0.3]
"""
    c = app.parse_app_config(text)
    assert (c.window.width, c.window.height) == (640, 480)
    assert (c.tracer.max_depth, c.tracer.rr_depth, c.tracer.exposure_value) == (8, 4, 0.5)
    assert c.tracer.tonemap.enable and c.tracer.tonemap.enable_aces and not c.tracer.tonemap.use_simple_aces
    assert c.hdri_file == "" and c.ground_color == (0.5, 0.5, 0.5) and c.sky_color == (0.1, 0.2, 0.3)
    assert c.env_intensity == 1.0
    app.validate_app_config(c)
    c.tracer.rr_depth = 9
    with pytest.raises(app.ConfigError, match="Russian Roulette depth is greater"):
        app.validate_app_config(c)
    c.tracer.rr_depth = 4
    c.window.width = 0
    with pytest.raises(app.ConfigError, match="The width is 0."):
        app.validate_app_config(c)
    c.window.width = 640
    c.scene_file = str(tmp_path / "missing.gltf")
    with pytest.raises(app.ConfigError, match="is not found"):
        app.validate_app_config(c)
    with pytest.raises(app.ConfigError, match="missing field `hdri_rotation`"):
        app.parse_app_config(text.replace("hdri_rotation: 0.0\n", ""))
    # the tracer section's serde defaults (src/config/tracer.rs)
    c2 = app.parse_app_config(text.replace("  max_depth: 8\n", "").replace("  rr_depth: 4\n", ""))
    assert (c2.tracer.max_depth, c2.tracer.rr_depth) == (0, 0)


def test_scene_features_follow_the_reference_host():
    assert app.scene_features(hpt_scenes.cornell_box()) == 0
    f = app.scene_features(hpt_scenes.displaced_icosphere_lights_medium(level=1))
    assert f == hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT


def _rgbe(rgb):
    m = rgb.max(axis=-1)
    e = np.where(m > 1e-32, np.floor(np.log2(np.maximum(m, 1e-38))) + 1, 0).astype(np.int32)
    scale = np.where(m > 1e-32, np.ldexp(1.0, 8 - e), 0.0)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    out[..., :3] = np.clip(rgb * scale[..., None], 0, 255).astype(np.uint8)
    out[..., 3] = np.where(m > 1e-32, e + 128, 0)
    return out


def test_radiance_hdr_reader_flat_and_rle(tmp_path):
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0, 40, (5, 12, 3))
    px = _rgbe(rgb)
    expect = px[..., :3].astype(np.float64) * np.ldexp(1.0, px[..., 3].astype(np.int32) - 136)[..., None]
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 5 +X 12\n"
    flat = tmp_path / "flat.hdr"
    flat.write_bytes(head + px.tobytes())
    assert np.allclose(app.read_hdr(str(flat)), expect, rtol=1e-6)
    # new-style RLE: each scanline as 4 channel runs (one literal run each)
    body = b""
    for y in range(5):
        body += bytes([2, 2, 0, 12])
        for c in range(4):
            body += bytes([12]) + px[y, :, c].tobytes()
    rle = tmp_path / "rle.hdr"
    rle.write_bytes(head + body)
    assert np.allclose(app.read_hdr(str(rle)), expect, rtol=1e-6)


def test_oracle_renders_the_reference_cornell_asset():
    sc = gltf_scene.load_gltf(os.path.join(GOLDEN, "cornell-box.gltf")).scene
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(32, 24, 8, 4, tonemap=False)
    rad, alb, _ = o.trace_paths(st, 0)
    assert np.isfinite(rad).all() and rad.mean() > 0.01
    # the red wall is on the left of the image, the green one on the right
    assert alb[12, 5, 0] > alb[12, 5, 1] and alb[12, 26, 1] > alb[12, 26, 0]
