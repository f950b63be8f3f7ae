"""Seeded synthetic scenes for the path tracer's workloads and tests.

Every generator is deterministic for a given seed.  Scenes are Z-up, the
convention of the reference's environment lookup
(shaders/src/hala-pathtracer/inc/env_map.glsl:22-26).

Material parameter conversions the reference performs in its (unvendored)
renderer are fixed here as follows:
  * Disney ax, ay: aspect = sqrt(1 - 0.9*anisotropic), ax = max(0.001, r^2/aspect),
    ay = max(0.001, r^2*aspect) (the 0.001 clamp cited at inc/bxdf/disney.glsl:420);
  * diffuse roughness -> Oren-Nayar A, B in ax, ay with sigma = roughness*pi/2
    (the formula kept as a comment at inc/bxdf/oren_nayar.glsl:37-43);
  * "specular 0.5" -> ior 1.5 (F0 = 0.04 through disney.glsl:24-25).
"""
from __future__ import annotations

import math

import numpy as np

import hpt

# ------------------------------------------------------------------ helpers


def look_at_camera(position, target, up=(0.0, 0.0, 1.0), yfov_deg=39.0, aperture=0.0, focal_distance=1.0):
    cam = np.zeros(1, hpt.CAMERA_DTYPE)
    p = np.asarray(position, np.float64)
    f = np.asarray(target, np.float64) - p
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    cam["position"] = p
    cam["forward"] = f
    cam["right"] = r
    cam["up"] = u
    cam["yfov"] = math.radians(yfov_deg)
    cam["focal_distance_or_xmag"] = focal_distance
    cam["aperture_or_ymag"] = aperture
    cam["type"] = hpt.CAMERA_PERSPECTIVE
    return cam


def diffuse_material(m, color, roughness=0.0, emission=(0, 0, 0)):
    m["type"] = hpt.MATERIAL_DIFFUSE
    m["base_color"] = color
    m["roughness"] = roughness
    m["emission"] = emission
    sigma2 = (roughness * math.pi / 2) ** 2
    m["ax"] = 1.0 - sigma2 / (2.0 * (sigma2 + 0.33))
    m["ay"] = 0.45 * sigma2 / (sigma2 + 0.09)


def disney_material(m, color, metallic=0.0, roughness=0.5, anisotropic=0.0, sheen=0.0, sheen_tint=0.5,
                    clearcoat=0.0, clearcoat_roughness=0.1, specular_tint=0.0, subsurface=0.0,
                    specular_transmission=0.0, ior=1.5):
    m["type"] = hpt.MATERIAL_DISNEY
    m["base_color"] = color
    m["metallic"] = metallic
    m["roughness"] = roughness
    m["anisotropic"] = anisotropic
    m["sheen"] = sheen
    m["sheen_tint"] = sheen_tint
    m["clearcoat"] = clearcoat
    m["clearcoat_roughness"] = clearcoat_roughness
    m["specular_tint"] = specular_tint
    m["subsurface"] = subsurface
    m["specular_transmission"] = specular_transmission
    m["ior"] = ior
    aspect = math.sqrt(1.0 - 0.9 * anisotropic)
    m["ax"] = max(0.001, roughness * roughness / aspect)
    m["ay"] = max(0.001, roughness * roughness * aspect)


def quad_mesh(p0, e1, e2):
    """Two triangles over the parallelogram p0 + s*e1 + t*e2; normal e1 x e2."""
    p0, e1, e2 = (np.asarray(x, np.float64) for x in (p0, e1, e2))
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n)
    t = e1 / np.linalg.norm(e1)
    v = np.zeros((4, hpt.VERTEX_FLOATS), np.float32)
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    for i, (s, q) in enumerate(corners):
        v[i, 0:3] = p0 + s * e1 + q * e2
        v[i, 3:6] = n
        v[i, 6:9] = t
        v[i, 9:11] = (s, q)
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def cube_mesh():
    """[-1,1]^3 with outward face normals: 24 vertices, 12 triangles."""
    vs, ids = [], []
    faces = [((-1, -1, -1), (0, 2, 0), (2, 0, 0)), ((-1, -1, 1), (2, 0, 0), (0, 2, 0)),
             ((-1, -1, -1), (2, 0, 0), (0, 0, 2)), ((-1, 1, -1), (0, 0, 2), (2, 0, 0)),
             ((-1, -1, -1), (0, 0, 2), (0, 2, 0)), ((1, -1, -1), (0, 2, 0), (0, 0, 2))]
    for p0, e1, e2 in faces:
        v, ix = quad_mesh(p0, e1, e2)
        ids.append(ix + len(vs) * 4)
        vs.append(v)
    return np.concatenate(vs), np.concatenate(ids)


def trs(translation, rotation_xyzw, scale):
    x, y, z, w = rotation_xyzw
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(scale)
    m[:3, 3] = translation
    return m.astype(np.float32)


def uv_sphere_mesh(center, radius, segments=64, rings=32):
    """UV sphere with fan poles: segments*(rings-2)*2 + 2*segments triangles."""
    cx, cy, cz = center
    verts = []
    for j in range(rings + 1):
        th = math.pi * j / rings
        for i in range(segments + 1):
            ph = 2 * math.pi * i / segments
            n = (math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))
            t = (-math.sin(ph), math.cos(ph), 0.0)
            verts.append((cx + radius * n[0], cy + radius * n[1], cz + radius * n[2], *n, *t,
                          i / segments, j / rings))
    v = np.array(verts, np.float32)
    ids = []
    row = segments + 1
    for j in range(rings):
        for i in range(segments):
            a, b = j * row + i, j * row + i + 1
            c, d = (j + 1) * row + i, (j + 1) * row + i + 1
            if j == 0:
                ids.append((a, c, d))
            elif j == rings - 1:
                ids.append((a, c, b))
            else:
                ids.append((a, c, d))
                ids.append((a, d, b))
    return v, np.array(ids, np.uint32)


# ---------------------------------------------------------------- scenes

def cornell_box(width=512, height=512):
    """Procedural Cornell box: 5 wall quads, short and tall box (34 triangles),
    Lambert white 0.73 / red (0.65,0.05,0.05) / green (0.12,0.45,0.15),
    0.25x0.25 quad light of radiance (17,12,4) just below the ceiling, pinhole
    camera with a 39 degree vertical field of view, black sky."""
    mats = hpt.new_materials(3)
    diffuse_material(mats[0], (0.73, 0.73, 0.73))
    diffuse_material(mats[1], (0.65, 0.05, 0.05))
    diffuse_material(mats[2], (0.12, 0.45, 0.15))
    prims = []
    walls = [((-1, -1, 0), (2, 0, 0), (0, 2, 0), 0),   # floor, normal +z
             ((-1, -1, 2), (0, 2, 0), (2, 0, 0), 0),   # ceiling, normal -z
             ((-1, 1, 0), (0, 0, 2), (2, 0, 0), 0),    # back, normal -y
             ((-1, -1, 0), (0, 2, 0), (0, 0, 2), 1),   # left (red), normal +x
             ((1, -1, 0), (0, 0, 2), (0, 2, 0), 2)]    # right (green), normal -x
    for p0, e1, e2, mi in walls:
        v, ix = quad_mesh(p0, e1, e2)
        prims.append(hpt.Primitive(v, ix, mi))
    cv, ci = cube_mesh()
    # Node transforms of ShortBox / TallBox in the reference's assets/cornell-box.gltf.
    short = trs((0.3286310136318207, -0.37459200620651245, 0.30000001192092896),
                (0.8019842505455017, 0.5973451733589172, 0.0, 5.680963965914998e-08),
                (0.29740551114082336, 0.30219709873199463, 0.30000001192092896))
    tall = trs((-0.335438996553421, 0.2914150059223175, 0.6000000238418579),
               (0.986013650894165, 0.166665181517601, 0.0, 7.563915005448507e-08),
               (0.3036448657512665, 0.29886960983276367, 0.6000000238418579))
    prims.append(hpt.Primitive(cv, ci, 0, short))
    prims.append(hpt.Primitive(cv, ci, 0, tall))
    lights = np.zeros(1, hpt.LIGHT_DTYPE)
    lights["type"] = hpt.LIGHT_QUAD
    lights["position"] = (-0.125, -0.125, 1.999)
    lights["u"] = (0.25, 0.0, 0.0)
    lights["v"] = (0.0, 0.25, 0.0)
    lights["area"] = 0.0625
    lights["intensity"] = (17.0, 12.0, 4.0)
    cam = look_at_camera((0.0, -3.85, 1.0), (0.0, 0.0, 1.0), yfov_deg=39.0)
    return hpt.Scene(cam, lights, mats, prims, ground_color=(0, 0, 0), sky_color=(0, 0, 0))


def checker_texture(n=8, size=64, a=(230, 230, 230), b=(40, 60, 200)):
    """RGBA8 checkerboard of n x n squares."""
    y, x = np.mgrid[0:size, 0:size]
    on = ((x * n // size) + (y * n // size)) % 2 == 0
    t = np.empty((size, size, 4), np.uint8)
    t[..., :3] = np.where(on[..., None], np.array(a, np.uint8), np.array(b, np.uint8))
    t[..., 3] = 255
    return t


def bump_normal_map(size=64, freq=4, amp=0.6):
    """RGBA8 tangent-space normal map of sinusoidal bumps (z stored at 1.0 so
    that the reference's z * 5 before the remap keeps it facing out)."""
    y, x = np.mgrid[0:size, 0:size] / size
    nx = amp * np.cos(2 * math.pi * freq * x)
    ny = amp * np.cos(2 * math.pi * freq * y)
    t = np.empty((size, size, 4), np.uint8)
    t[..., 0] = np.clip((nx * 0.5 + 0.5) * 255 + 0.5, 0, 255)
    t[..., 1] = np.clip((ny * 0.5 + 0.5) * 255 + 0.5, 0, 255)
    t[..., 2] = 255
    t[..., 3] = 255
    return t


def metal_rough_texture(size=32):
    """RGBA8 metallic (blue) / roughness (green) ramp: metallic along u,
    roughness along v."""
    y, x = np.mgrid[0:size, 0:size] / (size - 1)
    t = np.zeros((size, size, 4), np.uint8)
    t[..., 1] = (0.1 + 0.8 * y) * 255
    t[..., 2] = x * 255
    t[..., 3] = 255
    return t


def textured_cornell_box(width=512, height=512):
    """The Cornell box with material maps: a checkerboard base-colour map on
    the floor, a bump normal map on the back wall, an emission map on the
    ceiling, and a Disney tall box whose metallic/roughness come from a map."""
    sc = cornell_box(width, height)
    mats = np.concatenate([sc.materials, hpt.new_materials(4)])
    diffuse_material(mats[3], (1.0, 1.0, 1.0))
    mats[3]["base_color_map_index"] = 0
    diffuse_material(mats[4], (0.73, 0.73, 0.73))
    mats[4]["normal_map_index"] = 1
    diffuse_material(mats[5], (0.73, 0.73, 0.73))
    mats[5]["emission_map_index"] = 0
    disney_material(mats[6], (0.8, 0.6, 0.3), metallic=0.0, roughness=0.5)
    mats[6]["metallic_roughness_map_index"] = 2
    sc.materials = mats
    sc.primitives[0].material_index = 3   # floor
    sc.primitives[2].material_index = 4   # back wall
    sc.primitives[1].material_index = 5   # ceiling
    sc.primitives[6].material_index = 6   # tall box
    sc.textures = [checker_texture(), bump_normal_map(), metal_rough_texture()]
    return sc


def sky_env_map(width=2048, height=1024, seed=0, sun_elevation_deg=35.0, sun_azimuth=0.3,
                sun_radiance=1.0e4, sun_diameter_deg=0.5):
    """Lat-long sky: gradient from horizon white to zenith blue, a dim ground
    below the horizon, lognormal(0, 0.5) texel noise and a sun disc."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi
    cz = np.cos(theta)[:, None]
    t = np.clip(cz, 0.0, 1.0)
    zenith = np.array([0.25, 0.45, 1.0], np.float32)
    horizon = np.array([1.0, 1.0, 1.0], np.float32)
    ground = np.array([0.25, 0.22, 0.18], np.float32)
    sky = horizon[None, None, :] * (1 - t[..., None]) + zenith[None, None, :] * t[..., None]
    img = np.where(cz[..., None] > 0, sky, ground[None, None, :]) * np.ones((height, width, 1), np.float32)
    img = img * rng.lognormal(0.0, 0.5, size=(height, width, 1)).astype(np.float32)
    # Sun: texels whose centre direction lies within half the disc diameter.
    sth = np.sin(theta)[:, None]
    phi = (u[None, :]) * 2 * np.pi
    d = np.stack([-sth * np.cos(phi), sth * np.sin(phi), np.broadcast_to(cz, (height, width))], axis=-1)
    el, az = math.radians(sun_elevation_deg), sun_azimuth * 2 * math.pi
    s = np.array([-math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])
    ang = np.degrees(np.arccos(np.clip(d @ s, -1, 1)))
    img[ang <= sun_diameter_deg / 2] = sun_radiance
    return img.astype(np.float32)


def disney_grid(seed=0, width=1920, height=1080, env_width=2048, env_height=1024, segments=64, rings=32,
                subsurface=False):
    """7x7 Disney UV spheres (64x32 segments) on a Lambert ground plane:
    metallic = col/6, roughness = 0.05 + 0.95*row/6, ior 1.5; sheen,
    clearcoat, specular tint and anisotropy U(0,1) from the seed; sky
    environment map with a sun, sampled by the alias table.  `subsurface`:
    each sphere's subsurface weight is 0.25 + 0.75 * row/6 (the fake
    subsurface lobe of inc/bxdf/disney.glsl:93-112)."""
    rng = np.random.default_rng(seed)
    n = 7
    mats = hpt.new_materials(n * n + 1)
    prims = []
    radius, spacing = 0.4, 1.0
    for row in range(n):
        for col in range(n):
            k = row * n + col
            disney_material(mats[k], tuple(rng.uniform(0.2, 0.9, 3)), metallic=col / 6.0,
                            roughness=0.05 + 0.95 * row / 6.0, anisotropic=rng.uniform(),
                            sheen=rng.uniform(), clearcoat=rng.uniform(), specular_tint=rng.uniform())
            if subsurface:
                mats[k]["subsurface"] = 0.25 + 0.75 * row / 6.0
            c = ((col - 3) * spacing, (row - 3) * spacing, radius)
            v, ix = uv_sphere_mesh(c, radius, segments, rings)
            prims.append(hpt.Primitive(v, ix, k))
    diffuse_material(mats[n * n], (0.5, 0.5, 0.5))
    v, ix = quad_mesh((-20, -20, 0), (40, 0, 0), (0, 40, 0))
    prims.append(hpt.Primitive(v, ix, n * n))
    cam = look_at_camera((0.0, -7.5, 5.0), (0.0, 0.0, 0.3), yfov_deg=45.0)
    env = sky_env_map(env_width, env_height, seed)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, prims, env_rgb=env)


def _hash3(ix, iy, iz, seed):
    """Integer lattice hash -> [0, 1) (one pcg-style round, uint32)."""
    m = np.uint32(0xFFFFFFFF)
    v = (ix.astype(np.uint32) * np.uint32(73856093)) ^ (iy.astype(np.uint32) * np.uint32(19349663)) ^ \
        (iz.astype(np.uint32) * np.uint32(83492791)) ^ np.uint32(seed * 2654435761 & 0xFFFFFFFF)
    v = (v * np.uint32(1664525) + np.uint32(1013904223)) & m
    v ^= v >> np.uint32(16)
    v = (v * np.uint32(2246822519)) & m
    v ^= v >> np.uint32(13)
    return (v >> np.uint32(8)).astype(np.float64) / float(1 << 24)


def value_noise(p, seed=0):
    """Trilinear value noise with smoothstep weights, in [0, 1)."""
    i = np.floor(p)
    f = p - i
    u = f * f * (3 - 2 * f)
    i = i.astype(np.int64)
    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (u[:, 0] if dx else 1 - u[:, 0]) * (u[:, 1] if dy else 1 - u[:, 1]) * (u[:, 2] if dz else 1 - u[:, 2])
                out = out + w * _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def fbm(p, octaves=6, lacunarity=2.0, gain=0.5, seed=0):
    """Fractal sum of value noise in [-1, 1) scaled by sum of amplitudes."""
    amp, freq, total, norm = 1.0, 1.0, 0.0, 0.0
    for o in range(octaves):
        total = total + amp * (2 * value_noise(p * freq, seed + o) - 1)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def icosphere(level):
    """Unit icosphere: 20 * 4^level triangles, vertices deduplicated per level."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e[:, 0] * len(v) + e[:, 1], return_inverse=True)
        a, b = uniq // len(v), uniq % len(v)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + len(v)
        v = np.concatenate([v, mid])
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def smooth_mesh(pos, faces):
    """Vertices (N, 11) with area-weighted smooth normals and tangents
    orthogonal to them, uv from spherical angles."""
    e1 = pos[faces[:, 1]] - pos[faces[:, 0]]
    e2 = pos[faces[:, 2]] - pos[faces[:, 0]]
    fn = np.cross(e1, e2)
    n = np.zeros_like(pos)
    for k in range(3):
        np.add.at(n, faces[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    ref = np.where(np.abs(n[:, 2:3]) < 0.9, np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))
    t = np.cross(ref, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    v = np.zeros((len(pos), hpt.VERTEX_FLOATS), np.float32)
    v[:, 0:3] = pos
    v[:, 3:6] = n
    v[:, 6:9] = t
    v[:, 9] = (np.arctan2(pos[:, 1], pos[:, 0]) / (2 * np.pi)) % 1.0
    v[:, 10] = np.arccos(np.clip(n[:, 2], -1, 1)) / np.pi
    return v, faces.astype(np.uint32)


def displaced_icosphere_lights_medium(level=8, seed=0, amplitude=0.1, width=1920, height=1080,
                                      boundary="transparent"):
    """Large mesh + mixed lights + medium: an icosphere (20*4^level
    triangles; level 8 = 1,310,720) displaced by fBm (6 octaves, lacunarity
    2, gain 0.5, amplitude 0.1) on a ground quad; one point, spot (inner 20,
    outer 30 degrees), directional, rect and sphere (r = 0.2) light; a
    homogeneous scattering sphere (sigma_s 0.5, sigma_a 0.05, HG g 0.3) behind
    an opacity-0 boundary; thin-lens camera (f = 50 mm on a 24 mm sensor,
    aperture 0.05) focused on the mesh.  Needs MEDIUM|VOL_MIS|TRANSPARENT.
    `boundary` "glass" or "rough_glass" makes the medium's boundary a Disney
    dielectric (specular transmission 1, ior 1.5, roughness 0.02 or 0.35)
    instead: paths refract into the medium and the volume-MIS shadow walk
    passes the refractive surface (inc/lighting.glsl:45-47)."""
    mats = hpt.new_materials(4)
    disney_material(mats[0], (0.8, 0.55, 0.35), metallic=0.2, roughness=0.45, clearcoat=0.3, sheen=0.2)
    diffuse_material(mats[1], (0.6, 0.6, 0.6), roughness=0.3)
    if boundary == "transparent":
        # medium boundary: invisible surface carrying the scattering medium
        diffuse_material(mats[2], (1.0, 1.0, 1.0))
        mats[2]["opacity"] = 0.0
    else:
        disney_material(mats[2], (1.0, 1.0, 1.0), roughness={"glass": 0.02, "rough_glass": 0.35}[boundary],
                        specular_transmission=1.0, ior=1.5)
    mats[2]["medium_type"] = hpt.MEDIUM_SCATTER
    mats[2]["medium_density"] = 0.55
    mats[2]["medium_color"] = 0.5 / 0.55
    mats[2]["medium_anisotropy"] = 0.3
    disney_material(mats[3], (0.9, 0.9, 0.9), roughness=0.2)
    v, f = icosphere(level)
    d = 1.0 + amplitude * fbm(v * 2.0, seed=seed)
    pos = v * d[:, None] * 0.8 + np.array([0.0, 0.0, 0.9])
    sv, sf = smooth_mesh(pos, f)
    prims = [hpt.Primitive(sv, sf, 0)]
    q, qi = quad_mesh((-6, -6, 0), (12, 0, 0), (0, 12, 0))
    prims.append(hpt.Primitive(q, qi, 1))
    mv, mf = uv_sphere_mesh((1.6, 0.4, 0.6), 0.6, 32, 16)
    prims.append(hpt.Primitive(mv, mf, 2))
    lights = np.zeros(5, hpt.LIGHT_DTYPE)
    lights[0]["type"] = hpt.LIGHT_POINT
    lights[0]["position"] = (-1.5, -1.5, 2.5)
    lights[0]["intensity"] = (6.0, 5.0, 4.0)
    lights[1]["type"] = hpt.LIGHT_SPOT
    lights[1]["position"] = (1.5, -2.5, 3.0)
    axis = np.array([-1.5, 2.5, -2.2])
    lights[1]["u"] = axis / np.linalg.norm(axis)
    lights[1]["v"] = (math.cos(math.radians(20)), math.cos(math.radians(30)), 0.0)
    lights[1]["intensity"] = (20.0, 20.0, 24.0)
    lights[2]["type"] = hpt.LIGHT_DIRECTIONAL
    dd = np.array([0.3, 0.5, -0.8])
    lights[2]["u"] = dd / np.linalg.norm(dd)
    lights[2]["v"] = (0.0, 0.0, 0.0)
    lights[2]["intensity"] = (1.0, 0.95, 0.9)
    lights[3]["type"] = hpt.LIGHT_QUAD
    lights[3]["position"] = (-0.5, 1.0, 3.0)
    lights[3]["u"] = (1.0, 0.0, 0.0)
    lights[3]["v"] = (0.0, 0.5, 0.0)
    lights[3]["area"] = 0.5
    lights[3]["intensity"] = (8.0, 8.0, 8.0)
    lights[4]["type"] = hpt.LIGHT_SPHERE
    lights[4]["position"] = (-1.6, 0.8, 1.6)
    lights[4]["radius"] = 0.2
    lights[4]["area"] = 4 * math.pi * 0.04
    lights[4]["intensity"] = (10.0, 6.0, 3.0)
    eye, target = np.array([0.0, -4.5, 1.6]), np.array([0.0, 0.0, 0.9])
    yfov = 2 * math.degrees(math.atan(12.0 / 50.0))
    cam = look_at_camera(eye, target, yfov_deg=yfov, aperture=0.05, focal_distance=float(np.linalg.norm(target - eye)))
    return hpt.Scene(cam, lights, mats, prims, ground_color=(0.05, 0.05, 0.06), sky_color=(0.1, 0.12, 0.16))


MEDIUM_SCENE_FEATURES = hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT


def transparent_layers(layers=8):
    """`layers` parallel quads above a Lambert floor (albedo 0.7), alternately
    opacity 0 and 0.5, the odd ones carrying an absorbing medium (density
    0.8) that paths enter on their way up (normals face the floor); a point
    light above them, a directional light, sky.  The camera looks under the
    layers at the floor: a camera ray that hits a layer first ends there
    (its pass-through wraps depth 0, pathtracer.rgen.glsl:199,247), but paths
    that bounce off the floor pass through up to `layers` quads, and with
    TRANSPARENT each pass is a bounce that does not count toward depth
    (pathtracer.rgen.glsl:196-200)."""
    mats = hpt.new_materials(3)
    diffuse_material(mats[0], (0.7, 0.7, 0.7))
    diffuse_material(mats[1], (0.9, 0.4, 0.3))
    mats[1]["opacity"] = 0.0
    diffuse_material(mats[2], (0.3, 0.6, 0.9))
    mats[2]["opacity"] = 0.5
    mats[2]["medium_type"] = hpt.MEDIUM_ABSORB
    mats[2]["medium_density"] = 0.8
    mats[2]["medium_color"] = (0.2, 0.5, 0.7)
    q, qi = quad_mesh((-4, -4, 0), (8, 0, 0), (0, 8, 0))
    prims = [hpt.Primitive(q, qi, 0)]
    for i in range(layers):
        q, qi = quad_mesh((-1.5, -1.5, 0.3 + 0.25 * i), (0, 3, 0), (3, 0, 0))
        prims.append(hpt.Primitive(q, qi, 1 + (i & 1)))
    lights = np.zeros(2, hpt.LIGHT_DTYPE)
    lights[0]["type"] = hpt.LIGHT_POINT
    lights[0]["position"] = (0.6, -0.4, 4.5)
    lights[0]["intensity"] = (9.0, 8.0, 7.0)
    lights[1]["type"] = hpt.LIGHT_DIRECTIONAL
    dd = np.array([0.2, 0.3, -0.93])
    lights[1]["u"] = dd / np.linalg.norm(dd)
    lights[1]["intensity"] = (0.8, 0.8, 0.9)
    cam = look_at_camera((0.0, -5.0, 0.2), (0.0, 0.0, 0.0), yfov_deg=40.0)
    return hpt.Scene(cam, lights, mats, prims, ground_color=(0.1, 0.1, 0.1), sky_color=(0.3, 0.35, 0.45))


def triangle_soup(count, seed=0):
    """`count` triangles: v0 uniform in the unit cube, v1 and v2 at
    v0 + lognormal(-4, 0.5) * random unit vector; Lambert 0.5."""
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(0, 1, (count, 3))
    def offs():
        d = rng.normal(size=(count, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return d * rng.lognormal(-4.0, 0.5, (count, 1))
    v1, v2 = v0 + offs(), v0 + offs()
    p = np.stack([v0, v1, v2], axis=1).reshape(-1, 3)
    e1, e2 = v1 - v0, v2 - v0
    nrm = np.cross(e1, e2)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    tan = e1 / np.maximum(np.linalg.norm(e1, axis=1, keepdims=True), 1e-30)
    v = np.zeros((count * 3, hpt.VERTEX_FLOATS), np.float32)
    v[:, 0:3] = p
    v[:, 3:6] = np.repeat(nrm, 3, axis=0)
    v[:, 6:9] = np.repeat(tan, 3, axis=0)
    ix = np.arange(count * 3, dtype=np.uint32).reshape(-1, 3)
    mats = hpt.new_materials(1)
    diffuse_material(mats[0], (0.5, 0.5, 0.5))
    cam = look_at_camera((0.5, -1.5, 0.5), (0.5, 0.5, 0.5), yfov_deg=50.0)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, [hpt.Primitive(v, ix, 0)])


def batched_job_scenes(seed=0):
    """The 8 resident scenes of the batched and served workloads, as
    (name, scene, shader permutation): the Cornell box, the Disney grid, the
    large mesh with lights and a medium, and five random triangle soups of
    1e5..2e6 triangles (log-uniform, seeded)."""
    rng = np.random.default_rng(seed)
    scenes = [("cornell", cornell_box(), 0),
              ("disney_grid", disney_grid(seed=0), 0),
              ("icosphere_lights_medium", displaced_icosphere_lights_medium(level=8), MEDIUM_SCENE_FEATURES)]
    for i in range(5):
        n = int(round(math.exp(rng.uniform(math.log(1e5), math.log(2e6)))))
        scenes.append((f"soup_{n}", triangle_soup(n, seed=100 + i), 0))
    return scenes


def nested_triangles(count=200, ratio=0.7):
    """`count` right triangles in the x = s plane with legs s along y and z,
    s = ratio^i: each is ratio times the previous and nearer the origin, so
    the binned SAH peels a few off per level and the tree is deep (21 levels
    at the defaults), deeper than the traversal's LDS short stack."""
    s = ratio ** np.arange(count, dtype=np.float64)
    p = np.zeros((count, 3, 3))
    p[:, :, 0] = s[:, None]
    p[:, 1, 1] = s
    p[:, 2, 2] = s
    v = np.zeros((count * 3, hpt.VERTEX_FLOATS), np.float32)
    v[:, 0:3] = p.reshape(-1, 3)
    v[:, 3] = 1.0   # normal +x
    v[:, 7] = 1.0   # tangent +y
    ix = np.arange(count * 3, dtype=np.uint32).reshape(-1, 3)
    mats = hpt.new_materials(1)
    diffuse_material(mats[0], (0.6, 0.6, 0.6))
    cam = look_at_camera((1.5, 0.4, 0.4), (0.0, 0.05, 0.05), yfov_deg=40.0)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, [hpt.Primitive(v, ix, 0)])


def furnace_box(albedo=0.5, emission=1.0, half_size=10.0):
    """Camera inside a closed, flat-shaded Lambert cube that also emits:
    every path gathers emission * sum_k albedo^k exactly (no lights, black
    sky).  Flat faces keep the shading frame orthonormal, so no cosine sample
    falls below the horizon.  The cube is large against the reference's
    absolute ray offset (EPS = 3e-4): a bounce from within EPS of an edge can
    place the next origin outside a small cube and let the path escape."""
    mats = hpt.new_materials(1)
    diffuse_material(mats[0], (albedo, albedo, albedo), emission=(emission, emission, emission))
    v, ix = cube_mesh()
    s = np.diag([half_size, half_size, half_size, 1.0]).astype(np.float32)
    cam = look_at_camera((0.1 * half_size, -0.2 * half_size, 0.05 * half_size),
                         (0.3 * half_size, half_size, 0.2 * half_size), yfov_deg=70.0)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, [hpt.Primitive(v, ix, 0, s)],
                     ground_color=(0, 0, 0), sky_color=(0, 0, 0))


def furnace_sphere(albedo=0.5, emission=1.0, segments=48, rings=24):
    """Camera inside a closed Lambert UV sphere that also emits (smooth
    normals: the interpolated shading frame is not exactly orthonormal)."""
    mats = hpt.new_materials(1)
    diffuse_material(mats[0], (albedo, albedo, albedo), emission=(emission, emission, emission))
    v, ix = uv_sphere_mesh((0, 0, 0), 1.0, segments, rings)
    cam = look_at_camera((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), yfov_deg=60.0)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, [hpt.Primitive(v, ix, 0)],
                     ground_color=(0, 0, 0), sky_color=(0, 0, 0))


# ------------------------------------------------------------------ training scenes
def _random_material(m, rng, textures):
    """Lambert/Oren-Nayar or Disney with random lobes; sometimes a checker
    base-colour map or a bump normal map from `textures`."""
    color = tuple(rng.uniform(0.03, 0.95, 3) * (rng.uniform(0.5, 1.0) if rng.uniform() < 0.5 else 1.0))
    if rng.uniform() < 0.3:
        diffuse_material(m, color, roughness=float(rng.uniform(0.0, 1.0)))
    else:
        metallic = float(rng.choice([0.0, 1.0, rng.uniform()], p=[0.5, 0.25, 0.25]))
        transmission = float(rng.uniform(0.6, 1.0)) if rng.uniform() < 0.12 and metallic < 0.5 else 0.0
        disney_material(m, color, metallic=metallic, roughness=float(rng.uniform(0.02, 1.0)),
                        anisotropic=float(rng.uniform() * (rng.uniform() < 0.3)),
                        sheen=float(rng.uniform() * (rng.uniform() < 0.3)),
                        clearcoat=float(rng.uniform() * (rng.uniform() < 0.3)),
                        specular_tint=float(rng.uniform() * (rng.uniform() < 0.3)),
                        specular_transmission=transmission, ior=float(rng.uniform(1.3, 1.8)))
    if textures is not None and rng.uniform() < 0.2:
        m["base_color"] = (1.0, 1.0, 1.0)
        m["base_color_map_index"] = 0 if rng.uniform() < 0.5 else 2
    if textures is not None and rng.uniform() < 0.1:
        m["normal_map_index"] = 1


def _random_lights(rng, center, extent, count):
    lights = np.zeros(count, hpt.LIGHT_DTYPE)
    for i in range(count):
        kind = int(rng.choice([hpt.LIGHT_POINT, hpt.LIGHT_SPOT, hpt.LIGHT_DIRECTIONAL, hpt.LIGHT_QUAD,
                               hpt.LIGHT_SPHERE], p=[0.2, 0.15, 0.15, 0.3, 0.2]))
        pos = center + rng.uniform(-1, 1, 3) * extent * np.array([1.0, 1.0, 0.5]) + np.array([0, 0, extent])
        tint = rng.uniform(0.6, 1.0, 3)
        power = float(np.exp(rng.uniform(np.log(2.0), np.log(40.0))))
        lights[i]["type"] = kind
        lights[i]["position"] = pos
        if kind == hpt.LIGHT_SPOT:
            axis = center - pos
            lights[i]["u"] = axis / np.linalg.norm(axis)
            inner = rng.uniform(10, 30)
            lights[i]["v"] = (math.cos(math.radians(inner)), math.cos(math.radians(inner + rng.uniform(5, 15))), 0.0)
            lights[i]["intensity"] = tint * power
        elif kind == hpt.LIGHT_DIRECTIONAL:
            d = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), -rng.uniform(0.3, 1.0)])
            lights[i]["u"] = d / np.linalg.norm(d)
            lights[i]["intensity"] = tint * power / 15.0
        elif kind == hpt.LIGHT_QUAD:
            a, b = rng.uniform(0.1, 0.8) * extent, rng.uniform(0.1, 0.8) * extent
            lights[i]["position"] = pos - np.array([a / 2, b / 2, 0])
            lights[i]["u"] = (a, 0.0, 0.0)
            lights[i]["v"] = (0.0, b, 0.0)
            lights[i]["area"] = a * b
            lights[i]["intensity"] = tint * power / max(0.2, a * b)
        elif kind == hpt.LIGHT_SPHERE:
            r = rng.uniform(0.05, 0.3) * extent
            lights[i]["radius"] = r
            lights[i]["area"] = 4 * math.pi * r * r
            lights[i]["intensity"] = tint * power / max(0.5, 4 * math.pi * r * r)
        else:
            lights[i]["intensity"] = tint * power
    return lights


RANDOM_LAYOUTS = (0, 1, 2)  # room, outdoor, displaced icosphere with mixed lights


def random_scene(seed, width=256, height=256, layouts=RANDOM_LAYOUTS):
    """A seeded random scene for training the denoiser: (scene, features).
    One of three layouts -- a walled room with objects and ceiling lights, an
    outdoor ground plane under a sky map with a sun, or a displaced
    icosphere with mixed lights -- filled with random spheres, boxes and
    materials (Lambert, Oren-Nayar, Disney with metal, clearcoat, sheen and
    glass lobes, textured or bump-mapped), sometimes a participating medium
    and a thin-lens camera.  `layouts` restricts the draw (their weights
    renormalised): without layout 2 no displaced icosphere appears, which
    keeps the large-mesh benchmark scene's family out of a training set."""
    rng = np.random.default_rng(seed)
    weights = [0.35, 0.35, 0.3]
    if set(int(k) for k in layouts) != {0, 1, 2}:
        keep = [w if k in layouts else 0.0 for k, w in enumerate(weights)]
        if not any(keep):
            raise ValueError("random_scene: no layout to draw from")
        weights = [w / sum(keep) for w in keep]
    layout = int(rng.choice(3, p=weights))
    textures = [checker_texture(int(rng.integers(4, 16)), 64, tuple(int(c) for c in rng.integers(120, 255, 3)),
                                tuple(int(c) for c in rng.integers(0, 120, 3))),
                bump_normal_map(64, int(rng.integers(2, 8)), float(rng.uniform(0.2, 0.8))),
                checker_texture(int(rng.integers(2, 6)), 64, tuple(int(c) for c in rng.integers(60, 255, 3)),
                                tuple(int(c) for c in rng.integers(60, 255, 3)))]
    n_obj = int(rng.integers(3, 12))
    mats = hpt.new_materials(n_obj + 8)
    for i in range(n_obj + 8):
        _random_material(mats[i], rng, textures)
    prims, features = [], 0
    env, sky, ground = None, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
    if layout == 0:      # room
        s = float(rng.uniform(1.5, 3.0))
        for k, (p0, e1, e2) in enumerate([((-s, -s, 0), (2 * s, 0, 0), (0, 2 * s, 0)),
                                          ((-s, -s, 2 * s), (0, 2 * s, 0), (2 * s, 0, 0)),
                                          ((-s, s, 0), (0, 0, 2 * s), (2 * s, 0, 0)),
                                          ((-s, -s, 0), (0, 2 * s, 0), (0, 0, 2 * s)),
                                          ((s, -s, 0), (0, 0, 2 * s), (0, 2 * s, 0))]):
            diffuse_material(mats[n_obj + k], tuple(rng.uniform(0.1, 0.9, 3)), roughness=float(rng.uniform(0, 0.5)))
            if rng.uniform() < 0.2:
                mats[n_obj + k]["base_color"] = (1.0, 1.0, 1.0)
                mats[n_obj + k]["base_color_map_index"] = 0
            v, ix = quad_mesh(p0, e1, e2)
            prims.append(hpt.Primitive(v, ix, n_obj + k))
        lights = _random_lights(rng, np.array([0.0, 0.0, 0.0]), s, int(rng.integers(1, 4)))
        for L in lights:  # keep lights inside the room
            L["position"] = np.clip(L["position"], -0.9 * s, 0.9 * s) + np.array([0, 0, s])
        eye = np.array([rng.uniform(-0.5, 0.5) * s, -0.95 * s, rng.uniform(0.4, 1.4) * s])
        target = np.array([rng.uniform(-0.3, 0.3) * s, 0.0, rng.uniform(0.2, 0.8) * s])
        spread, zmin = 0.8 * s, 0.0
    elif layout == 1:    # outdoor
        diffuse_material(mats[n_obj], tuple(rng.uniform(0.2, 0.7, 3)), roughness=float(rng.uniform(0, 1)))
        if rng.uniform() < 0.4:
            mats[n_obj]["base_color"] = (1.0, 1.0, 1.0)
            mats[n_obj]["base_color_map_index"] = 2
        v, ix = quad_mesh((-30, -30, 0), (60, 0, 0), (0, 60, 0))
        prims.append(hpt.Primitive(v, ix, n_obj))
        env = sky_env_map(512, 256, seed, sun_elevation_deg=float(rng.uniform(5, 80)),
                          sun_azimuth=float(rng.uniform()), sun_radiance=float(rng.uniform(2e3, 2e4)),
                          sun_diameter_deg=float(rng.uniform(0.5, 3.0)))
        lights = _random_lights(rng, np.zeros(3), 2.0, int(rng.integers(0, 2)))
        d = rng.uniform(3.0, 7.0)
        a = rng.uniform(0, 2 * math.pi)
        eye = np.array([d * math.cos(a), d * math.sin(a), rng.uniform(0.5, 4.0)])
        target = np.array([0.0, 0.0, rng.uniform(0.2, 0.8)])
        spread, zmin = 2.0, 0.0
    else:                # displaced icosphere with mixed lights
        v, f = icosphere(int(rng.integers(4, 8)))
        d = 1.0 + rng.uniform(0.03, 0.15) * fbm(v * rng.uniform(1.0, 3.0), seed=seed)
        r0 = rng.uniform(0.5, 0.9)
        sv, sf = smooth_mesh(v * d[:, None] * r0 + np.array([0.0, 0.0, r0 + 0.1]), f)
        prims.append(hpt.Primitive(sv, sf, 0))
        diffuse_material(mats[n_obj], tuple(rng.uniform(0.3, 0.7, 3)), roughness=float(rng.uniform(0, 0.5)))
        v, ix = quad_mesh((-6, -6, 0), (12, 0, 0), (0, 12, 0))
        prims.append(hpt.Primitive(v, ix, n_obj))
        lights = _random_lights(rng, np.array([0.0, 0.0, 0.5]), 2.0, int(rng.integers(2, 6)))
        sky = tuple(rng.uniform(0.0, 0.2, 3))
        ground = tuple(rng.uniform(0.0, 0.08, 3))
        eye = np.array([rng.uniform(-2, 2), -rng.uniform(3.0, 5.0), rng.uniform(0.8, 2.5)])
        target = np.array([0.0, 0.0, r0])
        spread, zmin = 2.0, 0.0
    # objects
    for i in range(1 if layout == 2 else 0, n_obj):
        c = np.array([rng.uniform(-spread, spread), rng.uniform(-spread, spread), zmin])
        size = float(rng.uniform(0.15, 0.5)) * (spread / 2.0 if layout == 0 else 1.0)
        if rng.uniform() < 0.55:
            c[2] += size
            v, ix = uv_sphere_mesh(c, size, 32, 16)
            prims.append(hpt.Primitive(v, ix, i))
        else:
            cv, ci = cube_mesh()
            ang = rng.uniform(0, math.pi)
            prims.append(hpt.Primitive(cv, ci, i, trs((c[0], c[1], size), (0.0, 0.0, math.sin(ang / 2), math.cos(ang / 2)),
                                                   (size, size * rng.uniform(0.5, 1.5), size))))
    if rng.uniform() < 0.15:  # an emissive object
        k = int(rng.integers(0, n_obj))
        mats[k]["emission"] = tuple(rng.uniform(0.5, 1.0, 3) * rng.uniform(1.0, 8.0))
    if rng.uniform() < (0.5 if layout == 2 else 0.2):  # a participating medium behind an invisible boundary
        mi = n_obj + 6
        diffuse_material(mats[mi], (1.0, 1.0, 1.0))
        mats[mi]["opacity"] = 0.0
        mats[mi]["medium_type"] = hpt.MEDIUM_SCATTER if rng.uniform() < 0.7 else hpt.MEDIUM_ABSORB
        mats[mi]["medium_density"] = float(rng.uniform(0.2, 2.0))
        mats[mi]["medium_color"] = tuple(rng.uniform(0.3, 1.0, 3))
        mats[mi]["medium_anisotropy"] = float(rng.uniform(-0.3, 0.6))
        c = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0]) * spread * 0.6
        r = float(rng.uniform(0.3, 0.8)) * (spread / 2.0)
        c[2] = r + 0.05
        v, ix = uv_sphere_mesh(c, r, 32, 16)
        prims.append(hpt.Primitive(v, ix, mi))
        features = MEDIUM_SCENE_FEATURES
    if len(lights) == 0 and env is None:
        lights = _random_lights(rng, np.zeros(3), 2.0, 1)
    # A narrow field of view in part of the scenes gives the pixel scale of
    # full-HD frames on these small training images.
    yfov = float(rng.uniform(30, 60)) if rng.uniform() < 0.6 else float(rng.uniform(5, 15))
    aperture = float(rng.uniform(0.01, 0.06)) if rng.uniform() < (0.5 if layout == 2 else 0.2) else 0.0
    cam = look_at_camera(eye, target, yfov_deg=yfov, aperture=aperture,
                         focal_distance=float(np.linalg.norm(target - eye)))
    sc = hpt.Scene(cam, lights, mats, prims, env_rgb=env, ground_color=ground, sky_color=sky, textures=textures)
    return sc, features
