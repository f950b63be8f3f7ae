"""glTF 2.0 scenes for the tracer.

The reference host loads its scene with `scene::cpu::HalaScene::new(path)`
(src/bin/hala-pathtracer.rs:97) from the unvendored hala-renderer crate,
which turns glTF nodes, meshes, materials, cameras and KHR_lights_punctual
lights into the std140 arrays of inc/types.glsl.  That conversion is not in
the reference checkout, so every rule below is this module's own choice,
stated where it is made:

* Geometry: triangle primitives (mode 4) of every mesh node of the default
  scene, with the node's world matrix (parent chain of TRS or matrix) as the
  primitive transform.  POSITION is required; a primitive without NORMAL gets
  flat normals (the glTF rule); without TANGENT, the tangent of `onb(N)`
  (common/sampling.glsl:296-300); without TEXCOORD_0, zeros.
* Materials: pbrMetallicRoughness base colour (alpha = opacity), metallic,
  roughness, emissive (times KHR_materials_emissive_strength), ior
  (KHR_materials_ior, default 1.5), transmission, clearcoat, sheen,
  anisotropy from the KHR_materials_* extensions.  The material is Disney
  when metallic > 0 or any of those extensions is present, diffuse
  (Lambert, or Oren-Nayar above zero roughness) otherwise; `extras` may set
  `type` (0 diffuse, 1 Disney), any Material field by name (ax and ay
  together replace the derived values), and `medium` ({type, color,
  density, anisotropy}).  Texture references are kept as
  texture indices (the images are decoded with PIL).
* Lights (KHR_lights_punctual): point and spot keep colour x intensity
  (candela) as their intensity, spot cones as (cos inner, cos outer);
  directional lights travel along the node's -Z.  `extras.type` 1 makes the
  light a rectangle of param0 x param1 in the node's XY plane emitting along
  its -Z, with radiance 4 * intensity / area (the exporter's W/(4 pi) candela
  spread over a Lambertian emitter); `extras.type` 2 a sphere of radius
  param0 with radiance intensity / (pi r^2).
* Cameras: the node's -Z is forward, +Y up, +X right; perspective keeps yfov
  (the image aspect comes from the render size), `extras.focal_distance` and
  `extras.aperture` give the thin lens (default 1 and 0); orthographic keeps
  xmag/ymag.  Up to 8 cameras, in node order.
"""
from __future__ import annotations

import base64
import json
import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

import hpt

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_NORM_DIV = {np.int8: 127.0, np.uint8: 255.0, np.int16: 32767.0, np.uint16: 65535.0}


class GltfError(ValueError):
    pass


@dataclass
class Texture:
    """An RGBA8 image as the material texture table holds it."""
    rgba: np.ndarray  # (H, W, 4) uint8
    name: str = ""


@dataclass
class GltfScene:
    scene: hpt.Scene
    textures: list = field(default_factory=list)   # [Texture]
    names: dict = field(default_factory=dict)      # kind -> [names]


def _read_document(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"glTF":
        version, length = struct.unpack_from("<II", data, 4)
        if version != 2:
            raise GltfError(f"unsupported GLB version {version}")
        off, doc, binchunk = 12, None, None
        while off < length:
            clen, ctype = struct.unpack_from("<II", data, off)
            chunk = data[off + 8: off + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                binchunk = chunk
            off += 8 + clen
        if doc is None:
            raise GltfError("GLB without a JSON chunk")
        return doc, binchunk
    return json.loads(data.decode("utf-8")), None


def _load_buffers(doc, base_dir, binchunk):
    out = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if binchunk is None:
                raise GltfError(f"buffer {i} has no uri and the file has no BIN chunk")
            out.append(binchunk)
        elif uri.startswith("data:"):
            out.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base_dir, uri), "rb") as f:
                out.append(f.read())
        if len(out[-1]) < b["byteLength"]:
            raise GltfError(f"buffer {i} is shorter than its byteLength")
    return out


def _accessor(doc, buffers, index):
    acc = doc["accessors"][index]
    if "sparse" in acc:
        raise GltfError("sparse accessors are not supported")
    dtype = np.dtype(_COMPONENT[acc["componentType"]])
    n = _NCOMP[acc["type"]]
    count = acc["count"]
    if "bufferView" not in acc:
        return np.zeros((count, n), dtype)
    view = doc["bufferViews"][acc["bufferView"]]
    buf = buffers[view["buffer"]]
    start = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = view.get("byteStride", 0) or dtype.itemsize * n
    if stride == dtype.itemsize * n:
        arr = np.frombuffer(buf, dtype, count * n, start).reshape(count, n)
    else:
        rows = np.frombuffer(buf, np.uint8, (count - 1) * stride + dtype.itemsize * n, start)
        arr = np.stack([np.frombuffer(rows[i * stride: i * stride + dtype.itemsize * n].tobytes(), dtype)
                        for i in range(count)])
    if acc.get("normalized") and dtype.type in _NORM_DIV:
        arr = np.maximum(arr.astype(np.float32) / _NORM_DIV[dtype.type], -1.0)
    return arr


def _quat_matrix(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def node_matrix(node):
    """Local 4x4 (row-major numpy) of a node: `matrix` (column-major in the
    file) or T * R * S."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T
    m = np.eye(4)
    s = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    m[:3, :3] = _quat_matrix(node.get("rotation", [0, 0, 0, 1])) * s[None, :]
    m[:3, 3] = node.get("translation", [0, 0, 0])
    return m


def _world_nodes(doc):
    """(node index, world matrix) for every node of the default scene."""
    scenes = doc.get("scenes", [])
    if scenes:
        roots = scenes[doc.get("scene", 0)]["nodes"]
    else:
        children = {c for n in doc.get("nodes", []) for c in n.get("children", [])}
        roots = [i for i in range(len(doc.get("nodes", []))) if i not in children]
    out = []

    def visit(i, parent, depth):
        if depth > 256:
            raise GltfError("node hierarchy too deep (cycle?)")
        world = parent @ node_matrix(doc["nodes"][i])
        out.append((i, world))
        for c in doc["nodes"][i].get("children", []):
            visit(c, world, depth + 1)

    for r in roots:
        visit(r, np.eye(4), 0)
    return out


def _onb_tangent(n):
    """Tangent of common/sampling.glsl:296-300: normalize(cross(up, N)) with
    up = +Z unless N is (anti)parallel to it, then +X."""
    up = np.where((np.abs(n[:, 2]) < 0.9999999)[:, None], [[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]])
    t = np.cross(up, n)
    return t / np.maximum(np.linalg.norm(t, axis=1, keepdims=True), 1e-30)


def _primitive_arrays(doc, buffers, prim):
    if prim.get("mode", 4) != 4:
        return None
    attrs = prim["attributes"]
    if "POSITION" not in attrs:
        raise GltfError("primitive without POSITION")
    pos = _accessor(doc, buffers, attrs["POSITION"]).astype(np.float64)
    if "indices" in prim:
        idx = _accessor(doc, buffers, prim["indices"]).reshape(-1).astype(np.uint32)
    else:
        idx = np.arange(len(pos), dtype=np.uint32)
    if len(idx) % 3:
        raise GltfError("triangle index count is not a multiple of 3")
    if "NORMAL" in attrs:
        nrm = _accessor(doc, buffers, attrs["NORMAL"]).astype(np.float64)
    else:  # flat normals: one vertex per corner
        pos = pos[idx]
        tri = pos.reshape(-1, 3, 3)
        fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
        nrm = np.repeat(fn, 3, axis=0)
        attrs = {k: v for k, v in attrs.items() if k == "POSITION"}
        idx = np.arange(len(pos), dtype=np.uint32)
    if "TANGENT" in attrs:
        tan = _accessor(doc, buffers, attrs["TANGENT"])[:, :3].astype(np.float64)
    else:
        tan = _onb_tangent(nrm)
    uv = (_accessor(doc, buffers, attrs["TEXCOORD_0"]).astype(np.float64) if "TEXCOORD_0" in attrs
          else np.zeros((len(pos), 2)))
    v = np.zeros((len(pos), hpt.VERTEX_FLOATS), np.float32)
    v[:, 0:3], v[:, 3:6], v[:, 6:9], v[:, 9:11] = pos, nrm, tan, uv
    return v, idx.reshape(-1, 3)


def _set_material(m, gm, texture_slot):
    pbr = gm.get("pbrMetallicRoughness", {})
    ext = gm.get("extensions", {})
    extras = gm.get("extras", {}) or {}
    bc = pbr.get("baseColorFactor", [1, 1, 1, 1])
    metallic = float(pbr.get("metallicFactor", 1.0))
    roughness = float(pbr.get("roughnessFactor", 1.0))
    disney_ext = {"KHR_materials_transmission", "KHR_materials_clearcoat", "KHR_materials_sheen",
                  "KHR_materials_ior", "KHR_materials_anisotropy", "KHR_materials_specular"}
    kind = int(extras.get("type", hpt.MATERIAL_DISNEY if (metallic > 0.0 or disney_ext & set(ext))
                          else hpt.MATERIAL_DIFFUSE))
    m["type"] = kind
    m["base_color"] = bc[:3]
    m["opacity"] = bc[3] if len(bc) > 3 else 1.0
    m["metallic"] = metallic
    m["roughness"] = roughness
    strength = float(ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0))
    m["emission"] = [c * strength for c in gm.get("emissiveFactor", [0, 0, 0])]
    m["ior"] = float(ext.get("KHR_materials_ior", {}).get("ior", 1.5))
    m["specular_transmission"] = float(ext.get("KHR_materials_transmission", {}).get("transmissionFactor", 0.0))
    cc = ext.get("KHR_materials_clearcoat", {})
    m["clearcoat"] = float(cc.get("clearcoatFactor", 0.0))
    m["clearcoat_roughness"] = float(cc.get("clearcoatRoughnessFactor", 0.0))
    m["clearcoat_tint"] = (1.0, 1.0, 1.0)
    sh = ext.get("KHR_materials_sheen", {})
    if sh:
        m["sheen"] = float(max(sh.get("sheenColorFactor", [0, 0, 0])))
    m["sheen_tint"] = 0.5
    m["anisotropic"] = float(ext.get("KHR_materials_anisotropy", {}).get("anisotropyStrength", 0.0))
    med = extras.get("medium")
    if med:
        m["medium_type"] = int(med.get("type", hpt.MEDIUM_NONE))
        m["medium_color"] = med.get("color", [1, 1, 1])
        m["medium_density"] = float(med.get("density", 0.0))
        m["medium_anisotropy"] = float(med.get("anisotropy", 0.0))
    for k, val in extras.items():  # any Material field by name
        if k in hpt.MATERIAL_DTYPE.names and k != "type" and not k.endswith("_map_index"):
            m[k] = val
    if "ax" in extras and "ay" in extras:
        pass  # given explicitly
    elif kind == hpt.MATERIAL_DISNEY:
        aspect = math.sqrt(1.0 - 0.9 * float(m["anisotropic"]))
        r2 = float(m["roughness"]) ** 2
        m["ax"] = max(0.001, r2 / aspect)
        m["ay"] = max(0.001, r2 * aspect)
    else:  # Oren-Nayar A, B from sigma = roughness * pi / 2 (oren_nayar.glsl:37-43)
        sigma2 = (float(m["roughness"]) * math.pi / 2) ** 2
        m["ax"] = 1.0 - sigma2 / (2.0 * (sigma2 + 0.33))
        m["ay"] = 0.45 * sigma2 / (sigma2 + 0.09)
    m["base_color_map_index"] = texture_slot(pbr.get("baseColorTexture"))
    m["metallic_roughness_map_index"] = texture_slot(pbr.get("metallicRoughnessTexture"))
    m["normal_map_index"] = texture_slot(gm.get("normalTexture"))
    m["emission_map_index"] = texture_slot(gm.get("emissiveTexture"))


def _light(gl, extras, world):
    L = np.zeros(1, hpt.LIGHT_DTYPE)[0]
    color = np.asarray(gl.get("color", [1, 1, 1]), np.float64)
    intensity = float(gl.get("intensity", 1.0))
    pos = world[:3, 3]
    rot = world[:3, :3]
    fwd = rot @ np.array([0.0, 0.0, -1.0])
    fwd /= np.linalg.norm(fwd)
    kind = int(extras.get("type", -1)) if extras else -1
    if kind == 1:  # rectangle
        w, h = float(extras.get("param0", 1.0)), float(extras.get("param1", 1.0))
        ex = rot @ np.array([1.0, 0.0, 0.0])
        ey = rot @ np.array([0.0, 1.0, 0.0])
        ex /= np.linalg.norm(ex)
        ey /= np.linalg.norm(ey)
        # quad normal = normalize(cross(v, u)) (inc/light/quad.glsl) must be fwd
        u, v = ex * w, ey * h
        if np.dot(np.cross(v, u), fwd) < 0:  # mirrored node: swap the edges
            u, v = ey * h, ex * w
        L["type"] = hpt.LIGHT_QUAD
        L["position"] = pos - 0.5 * u - 0.5 * v
        L["u"], L["v"] = u, v
        L["area"] = w * h
        L["intensity"] = color * (4.0 * intensity / max(w * h, 1e-30))
        return L
    if kind == 2:  # sphere
        r = float(extras.get("param0", 0.1))
        L["type"] = hpt.LIGHT_SPHERE
        L["position"] = pos
        L["radius"] = r
        L["area"] = 4.0 * math.pi * r * r
        L["intensity"] = color * (intensity / (math.pi * r * r))
        return L
    t = gl.get("type", "point")
    L["intensity"] = color * intensity
    L["position"] = pos
    if t == "directional":
        L["type"] = hpt.LIGHT_DIRECTIONAL
        L["u"] = fwd
    elif t == "spot":
        spot = gl.get("spot", {})
        L["type"] = hpt.LIGHT_SPOT
        L["u"] = fwd
        L["v"] = (math.cos(float(spot.get("innerConeAngle", 0.0))),
                  math.cos(float(spot.get("outerConeAngle", math.pi / 4))), 0.0)
    else:
        L["type"] = hpt.LIGHT_POINT
    return L


def _camera(gc, world):
    c = np.zeros(1, hpt.CAMERA_DTYPE)[0]
    rot = world[:3, :3]
    axes = [rot @ a for a in (np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, -1.0]))]
    right, up, fwd = (a / np.linalg.norm(a) for a in axes)
    c["position"] = world[:3, 3]
    c["right"], c["up"], c["forward"] = right, up, fwd
    extras = gc.get("extras", {}) or {}
    if gc.get("type") == "orthographic":
        o = gc["orthographic"]
        c["type"] = hpt.CAMERA_ORTHOGRAPHIC
        c["focal_distance_or_xmag"] = float(o["xmag"])
        c["aperture_or_ymag"] = float(o["ymag"])
    else:
        p = gc.get("perspective", {})
        c["type"] = hpt.CAMERA_PERSPECTIVE
        c["yfov"] = float(p.get("yfov", math.radians(39.0)))
        c["focal_distance_or_xmag"] = float(extras.get("focal_distance", 1.0))
        c["aperture_or_ymag"] = float(extras.get("aperture", 0.0))
    return c


def load_gltf(path: str) -> GltfScene:
    """Reads a .gltf (with external or embedded buffers) or .glb file."""
    doc, binchunk = _read_document(path)
    base_dir = os.path.dirname(os.path.abspath(path))
    if doc.get("asset", {}).get("version", "2.0").split(".")[0] != "2":
        raise GltfError("not a glTF 2.0 asset")
    buffers = _load_buffers(doc, base_dir, binchunk)
    textures, tex_slot_of = [], {}

    def texture_slot(info):
        if not info or "index" not in info:
            return hpt.INVALID_INDEX
        ti = info["index"]
        if ti not in tex_slot_of:
            tex_slot_of[ti] = len(textures)
            textures.append(_decode_texture(doc, buffers, base_dir, ti))
        return tex_slot_of[ti]

    gmats = doc.get("materials", [])
    mats = hpt.new_materials(max(1, len(gmats) + 1))
    for i, gm in enumerate(gmats):
        _set_material(mats[i], gm, texture_slot)
    default_mat = len(gmats)  # glTF's default material: white, metallic 1, roughness 1
    _set_material(mats[default_mat], {}, texture_slot)

    prims, lights, cams, names = [], [], [], {"meshes": [], "lights": [], "cameras": []}
    light_defs = doc.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])
    for ni, world in _world_nodes(doc):
        node = doc["nodes"][ni]
        if "mesh" in node:
            mesh = doc["meshes"][node["mesh"]]
            for p in mesh["primitives"]:
                arrays = _primitive_arrays(doc, buffers, p)
                if arrays is None:
                    continue
                v, ix = arrays
                prims.append(hpt.Primitive(v, ix, int(p.get("material", default_mat)), world.astype(np.float32)))
                names["meshes"].append(mesh.get("name", node.get("name", "")))
        ext = node.get("extensions", {}).get("KHR_lights_punctual")
        if ext is not None:
            gl = light_defs[ext["light"]]
            lights.append(_light(gl, gl.get("extras"), world))
            names["lights"].append(gl.get("name", ""))
        if "camera" in node and len(cams) < hpt.MAX_CAMERAS:
            cams.append(_camera(doc["cameras"][node["camera"]], world))
            names["cameras"].append(node.get("name", ""))
    if not cams:
        raise GltfError("the scene has no camera")
    if len(lights) > hpt.MAX_LIGHTS:
        raise GltfError(f"{len(lights)} lights; the renderer takes at most {hpt.MAX_LIGHTS}")
    scene = hpt.Scene(np.array(cams, hpt.CAMERA_DTYPE), np.array(lights, hpt.LIGHT_DTYPE).reshape(-1),
                      mats, prims, textures=[t.rgba for t in textures])
    return GltfScene(scene, textures, names)


def _decode_texture(doc, buffers, base_dir, ti):
    from PIL import Image
    import io

    tex = doc["textures"][ti]
    img = doc["images"][tex["source"]]
    if "uri" in img:
        uri = img["uri"]
        if uri.startswith("data:"):
            data = base64.b64decode(uri.split(",", 1)[1])
        else:
            with open(os.path.join(base_dir, uri), "rb") as f:
                data = f.read()
    else:
        view = doc["bufferViews"][img["bufferView"]]
        b = buffers[view["buffer"]]
        data = b[view.get("byteOffset", 0): view.get("byteOffset", 0) + view["byteLength"]]
    im = Image.open(io.BytesIO(data)).convert("RGBA")
    return Texture(np.asarray(im, np.uint8).copy(), img.get("name", ""))


# ---------------------------------------------------------------- writing
def write_gltf(path: str, scene: hpt.Scene) -> None:
    """Writes a .gltf + .bin of a Scene: one mesh node per primitive (with its
    transform as the node matrix), one camera node per camera, materials with
    every Material field in `extras`, lights as KHR_lights_punctual nodes with
    the conversions load_gltf undoes."""
    base = os.path.splitext(path)[0]
    binname = os.path.basename(base) + ".bin"
    blob = bytearray()
    doc = {"asset": {"version": "2.0", "generator": "hala-pathtracer_amd"}, "scene": 0, "scenes": [{"nodes": []}],
           "nodes": [], "meshes": [], "materials": [], "accessors": [], "bufferViews": [], "cameras": [],
           "extensionsUsed": ["KHR_lights_punctual"], "extensions": {"KHR_lights_punctual": {"lights": []}}}

    def add_view(arr, ctype, kind, target):
        nonlocal blob
        while len(blob) % 4:
            blob.append(0)
        raw = np.ascontiguousarray(arr).tobytes()
        doc["bufferViews"].append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(raw), "target": target})
        blob += raw
        acc = {"bufferView": len(doc["bufferViews"]) - 1, "componentType": ctype, "count": len(arr), "type": kind}
        if kind == "VEC3" and ctype == 5126:
            acc["min"] = [float(x) for x in arr.min(axis=0)]
            acc["max"] = [float(x) for x in arr.max(axis=0)]
        doc["accessors"].append(acc)
        return len(doc["accessors"]) - 1

    def add_node(node):
        doc["nodes"].append(node)
        doc["scenes"][0]["nodes"].append(len(doc["nodes"]) - 1)

    for m in scene.materials:
        extras = {k: (m[k].tolist() if np.ndim(m[k]) else m[k].item()) for k in hpt.MATERIAL_DTYPE.names
                  if not k.endswith("_map_index")}
        doc["materials"].append({"pbrMetallicRoughness": {"baseColorFactor": list(map(float, m["base_color"])) +
                                                          [float(m["opacity"])]}, "extras": extras})
    for p in scene.primitives:
        v = np.asarray(p.vertices, np.float32).reshape(-1, hpt.VERTEX_FLOATS)
        ix = np.asarray(p.indices, np.uint32).reshape(-1)
        attrs = {"POSITION": add_view(v[:, 0:3], 5126, "VEC3", 34962),
                 "NORMAL": add_view(v[:, 3:6], 5126, "VEC3", 34962),
                 "TANGENT": add_view(np.concatenate([v[:, 6:9], np.ones((len(v), 1), np.float32)], 1), 5126,
                                     "VEC4", 34962),
                 "TEXCOORD_0": add_view(v[:, 9:11], 5126, "VEC2", 34962)}
        prim = {"attributes": attrs, "indices": add_view(ix, 5125, "SCALAR", 34963), "mode": 4}
        if p.material_index != hpt.INVALID_INDEX:
            prim["material"] = int(p.material_index)
        doc["meshes"].append({"primitives": [prim]})
        add_node({"mesh": len(doc["meshes"]) - 1,
                  "matrix": [float(x) for x in np.asarray(p.transform, np.float64).T.reshape(-1)]})
    for c in scene.cameras:
        r, u, f = (np.asarray(c[k], np.float64) for k in ("right", "up", "forward"))
        m = np.eye(4)
        m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, -f, c["position"]
        if c["type"] == hpt.CAMERA_ORTHOGRAPHIC:
            cam = {"type": "orthographic", "orthographic": {"xmag": float(c["focal_distance_or_xmag"]),
                                                              "ymag": float(c["aperture_or_ymag"]),
                                                              "znear": 0.01, "zfar": 1e4}}
        else:
            cam = {"type": "perspective", "perspective": {"yfov": float(c["yfov"]), "znear": 0.01, "zfar": 1e4},
                   "extras": {"focal_distance": float(c["focal_distance_or_xmag"]),
                              "aperture": float(c["aperture_or_ymag"])}}
        doc["cameras"].append(cam)
        add_node({"camera": len(doc["cameras"]) - 1, "matrix": [float(x) for x in m.T.reshape(-1)]})
    for L in scene.lights:
        t = int(L["type"])
        pos = np.asarray(L["position"], np.float64)
        m = np.eye(4)
        gl = {"color": [1.0, 1.0, 1.0], "intensity": 1.0}
        I = np.asarray(L["intensity"], np.float64)
        if t == hpt.LIGHT_QUAD:
            u, v = np.asarray(L["u"], np.float64), np.asarray(L["v"], np.float64)
            w, h = np.linalg.norm(u), np.linalg.norm(v)
            n = np.cross(v, u)
            n /= np.linalg.norm(n)
            m[:3, 0], m[:3, 1], m[:3, 2] = u / w, v / h, -n
            m[:3, 3] = pos + 0.5 * u + 0.5 * v
            area = w * h
            scale = max(float(I.max()), 1e-30)
            gl.update(type="point", color=(I / scale).tolist(), intensity=scale * area / 4.0,
                      extras={"type": 1, "param0": float(w), "param1": float(h)})
        elif t == hpt.LIGHT_SPHERE:
            r = float(L["radius"])
            m[:3, 3] = pos
            scale = max(float(I.max()), 1e-30)
            gl.update(type="point", color=(I / scale).tolist(), intensity=scale * math.pi * r * r,
                      extras={"type": 2, "param0": r})
        else:
            m[:3, 3] = pos
            scale = max(float(I.max()), 1e-30)
            gl.update(color=(I / scale).tolist(), intensity=scale)
            if t in (hpt.LIGHT_DIRECTIONAL, hpt.LIGHT_SPOT):
                f = np.asarray(L["u"], np.float64)
                f /= np.linalg.norm(f)
                a = np.array([0.0, 0.0, 1.0]) if abs(f[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
                x = np.cross(a, f)
                x /= np.linalg.norm(x)
                y = np.cross(-f, x)
                m[:3, 0], m[:3, 1], m[:3, 2] = x, y, -f
            if t == hpt.LIGHT_DIRECTIONAL:
                gl["type"] = "directional"
            elif t == hpt.LIGHT_SPOT:
                gl["type"] = "spot"
                gl["spot"] = {"innerConeAngle": math.acos(min(1.0, float(L["v"][0]))),
                              "outerConeAngle": math.acos(min(1.0, float(L["v"][1])))}
            else:
                gl["type"] = "point"
        doc["extensions"]["KHR_lights_punctual"]["lights"].append(gl)
        add_node({"matrix": [float(x) for x in m.T.reshape(-1)],
                  "extensions": {"KHR_lights_punctual": {"light": len(doc["extensions"]["KHR_lights_punctual"]["lights"]) - 1}}})
    doc["buffers"] = [{"byteLength": len(blob), "uri": binname}]
    with open(base + ".bin", "wb") as f:
        f.write(bytes(blob))
    with open(path, "w") as f:
        json.dump(doc, f)
