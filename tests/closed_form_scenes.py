"""Scenes whose expected radiance has a closed form, and those forms.

Shared by the oracle pins (tests/test_oracle_media_nee.py) and the GPU
checks of the same scenes (tests/test_gpu_closed_forms.py).  Each closed
form follows from the geometry and the physics the reference's shaders
implement, not from the restated code:

* card in a sphere -- the camera sees the back of a tiny Lambert card at the
  centre of a transparent sphere of radius R.  The card's material carries a
  medium, so a bounce off its back enters it (pathtracer.rgen.glsl:233-243),
  and every bounce direction meets the sphere after the same distance R.
  The medium's ABSORB, EMISSIVE and SCATTER branches
  (pathtracer.rgen.glsl:157-191) then have exact expectations under a
  uniform sky of radiance Ls.
* slab under a point light -- a Lambert floor lit through a wide slab whose
  opacity-0 boundary carries a medium: the volume-MIS shadow walk
  (inc/lighting.glsl:7-61) gives Beer-Lambert transmittance over the chord.
* env-lit plane -- a Lambert plane under a constant environment map: every
  point sees the whole upper hemisphere, so radiance = albedo * L for any
  correct combination of light sampling and BSDF sampling
  (inc/lighting.glsl:75-150 with the MIS weight of env_map_eval.rcall.glsl:23).
* sphere lights over a plane -- the irradiance of a uniform sphere fully above
  the horizon is pi * Le * sin^2(alpha) * cos(theta_c), whatever the
  sampling (inc/light/sphere.glsl, lights.rint.glsl, lighting.glsl:154-240);
  without volume MIS the reference's offset shadow rays darken it by a
  computable amount (sphere_lights_expected).
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate

import hpt
import hpt_scenes

EPS = 3e-4  # common/common.glsl:14, the reference's ray offset


# ------------------------------------------------------------ card in a sphere

def card_in_sphere(medium_type, color, density, anisotropy=0.0, radius=1.0, card_albedo=0.6, sky=0.8,
                   segments=256, rings=128):
    """A 2 mm Lambert card at the origin facing +y, seen from (0, -0.5, 0)
    through a narrow pinhole; it carries the medium.  Around it a sphere of
    radius `radius` whose opacity-0 surface carries none.  Uniform sky of
    radiance `sky`.  Needs the TRANSPARENT permutation (and MEDIUM)."""
    mats = hpt.new_materials(2)
    hpt_scenes.diffuse_material(mats[0], (card_albedo,) * 3)
    mats[0]["medium_type"] = medium_type
    mats[0]["medium_color"] = color
    mats[0]["medium_density"] = density
    mats[0]["medium_anisotropy"] = anisotropy
    hpt_scenes.diffuse_material(mats[1], (1.0, 1.0, 1.0))
    mats[1]["opacity"] = 0.0
    h = 1e-3
    cv, ci = hpt_scenes.quad_mesh((-h, 0.0, -h), (0.0, 0.0, 2 * h), (2 * h, 0.0, 0.0))  # normal +y
    sv, si = hpt_scenes.uv_sphere_mesh((0.0, 0.0, 0.0), radius, segments, rings)
    cam = hpt_scenes.look_at_camera((0.0, -0.5, 0.0), (0.0, 0.0, 0.0), yfov_deg=0.02)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, [hpt.Primitive(cv, ci, 0), hpt.Primitive(sv, si, 1)],
                     ground_color=(sky,) * 3, sky_color=(sky,) * 3)


def card_absorb(card_albedo, sky, color, density, radius):
    """Per path: the bounce enters the absorbing medium, crosses R - EPS of it
    (the continuation starts EPS past the card) and escapes to the sky:
    rho * Ls * exp(-(1 - c) * density * (R - EPS))."""
    k = (1.0 - np.asarray(color, np.float64)) * density
    return card_albedo * sky * np.exp(-k * (radius - EPS))


def card_emissive(card_albedo, sky, color, density, radius):
    """Per path: rho * (c * density * (R - EPS) + Ls)."""
    return card_albedo * (np.asarray(color, np.float64) * density * (radius - EPS) + sky)


def hg_density(mu, g):
    """Henyey-Greenstein density of the cosine to the propagation direction,
    normalised over mu in [-1, 1]."""
    return 0.5 * (1.0 - g * g) / (1.0 + g * g - 2.0 * g * mu) ** 1.5


def card_scatter(card_albedo, sky, albedo, sigma, g, radius, max_depth):
    """Expected radiance with a scattering medium (sigma_t = sigma, single-
    scattering albedo `albedo`, HG g) around the card.
    max_depth 1: only paths that cross the sphere unscattered reach the sky:
        rho * Ls * exp(-sigma R).
    max_depth 2: plus paths scattered once at distance r (density
    sigma e^{-sigma r}) into a direction of cosine mu to the outward radius
    (HG), which then escape unscattered over d(r, mu) = -r mu +
    sqrt(R^2 - r^2 (1 - mu^2)); a second scattering ends the path:
        rho * Ls * (exp(-sigma R) + albedo * P1)."""
    R = radius - EPS
    p0 = math.exp(-sigma * R)
    if max_depth == 1:
        return card_albedo * sky * p0
    assert max_depth == 2

    def inner(mu, r):
        d = -r * mu + math.sqrt(max(R * R - r * r * (1.0 - mu * mu), 0.0))
        return hg_density(mu, g) * math.exp(-sigma * d)

    p1, _ = integrate.dblquad(lambda mu, r: sigma * math.exp(-sigma * r) * inner(mu, r), 0.0, R, -1.0, 1.0,
                              epsabs=1e-10, epsrel=1e-9)
    return card_albedo * sky * (p0 + albedo * p1)


# ------------------------------------------------------- slab under a point light

SLAB_Z0, SLAB_Z1 = 2.0, 2.5
SLAB_LIGHT = (0.3, -0.2, 4.0)
SLAB_INTENSITY = 5.0
FLOOR_ALBEDO = 0.5


def slab_point_light(medium_type, color, density):
    """Lambert floor (z = 0) lit by a point light at z = 4 through a slab
    z in [2, 2.5] (80 x 80 wide) whose opacity-0 boundary carries the
    medium; the camera at z = 1 looks straight down, under the slab."""
    mats = hpt.new_materials(2)
    hpt_scenes.diffuse_material(mats[0], (FLOOR_ALBEDO,) * 3)
    hpt_scenes.diffuse_material(mats[1], (1.0, 1.0, 1.0))
    mats[1]["opacity"] = 0.0
    mats[1]["medium_type"] = medium_type
    mats[1]["medium_color"] = color
    mats[1]["medium_density"] = density
    fv, fi = hpt_scenes.quad_mesh((-50, -50, 0), (100, 0, 0), (0, 100, 0))
    bv, bi = hpt_scenes.cube_mesh()
    half_h = 0.5 * (SLAB_Z1 - SLAB_Z0)
    s = np.diag([40.0, 40.0, half_h, 1.0]).astype(np.float32)
    s[2, 3] = 0.5 * (SLAB_Z0 + SLAB_Z1)
    lights = np.zeros(1, hpt.LIGHT_DTYPE)
    lights[0]["type"] = hpt.LIGHT_POINT
    lights[0]["position"] = SLAB_LIGHT
    lights[0]["intensity"] = (SLAB_INTENSITY,) * 3
    cam = hpt_scenes.look_at_camera((0.0, 0.0, 1.0), (0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), yfov_deg=60.0)
    return hpt.Scene(cam, lights, mats, [hpt.Primitive(fv, fi, 0), hpt.Primitive(bv, bi, 1, s)],
                     ground_color=(0, 0, 0), sky_color=(0, 0, 0))


def floor_points(rays):
    """Where the camera rays meet the floor z = 0 (float64)."""
    o = rays["origin"].astype(np.float64)
    d = rays["direction"].astype(np.float64)
    t = -o[:, 2] / d[:, 2]
    return o + d * t[:, None]


def slab_direct(p, k):
    """Point light through the slab at floor points p: Lambert
    rho/pi * cos * I / dist^2 times exp(-k (l - EPS)), l = the slab thickness
    over the cosine (the walk re-enters EPS past the entry face)."""
    v = np.asarray(SLAB_LIGHT, np.float64)[None, :] - p
    dist = np.linalg.norm(v, axis=1)
    cos = v[:, 2] / dist
    chord = (SLAB_Z1 - SLAB_Z0) / cos
    tr = np.exp(-np.asarray(k, np.float64)[None, :] * (chord - EPS)[:, None])
    return (FLOOR_ALBEDO / math.pi * cos * SLAB_INTENSITY / dist ** 2)[:, None] * tr


# ----------------------------------------------------------------- env-lit plane

def env_lit_plane(env_radiance=(0.9, 0.7, 0.4), albedo=0.5, env_size=(32, 16)):
    """A Lambert plane (z = 0) under a constant environment map, seen from
    above: radiance = albedo * L (NEE of the map + BSDF samples that escape)."""
    mats = hpt.new_materials(1)
    hpt_scenes.diffuse_material(mats[0], (albedo,) * 3)
    fv, fi = hpt_scenes.quad_mesh((-50, -50, 0), (100, 0, 0), (0, 100, 0))
    w, h = env_size
    env = np.ones((h, w, 3), np.float32) * np.asarray(env_radiance, np.float32)
    cam = hpt_scenes.look_at_camera((0.0, -1.0, 1.5), (0.0, 0.5, 0.0), yfov_deg=50.0)
    return hpt.Scene(cam, np.zeros(0, hpt.LIGHT_DTYPE), mats, [hpt.Primitive(fv, fi, 0)], env_rgb=env)


# --------------------------------------------------------- sphere lights over a plane

SPHERE_LIGHTS = (  # centre, radius, radiance
    ((0.4, 0.3, 1.4), 0.25, (6.0, 4.0, 2.0)),
    ((-0.5, -0.2, 1.1), 0.15, (2.0, 5.0, 9.0)),
)


def sphere_lights_plane(albedo=0.5):
    """Lambert floor lit by two sphere lights fully above its horizon, black
    sky; the camera looks down at the floor between them."""
    mats = hpt.new_materials(1)
    hpt_scenes.diffuse_material(mats[0], (albedo,) * 3)
    fv, fi = hpt_scenes.quad_mesh((-50, -50, 0), (100, 0, 0), (0, 100, 0))
    lights = np.zeros(len(SPHERE_LIGHTS), hpt.LIGHT_DTYPE)
    for i, (c, r, le) in enumerate(SPHERE_LIGHTS):
        lights[i]["type"] = hpt.LIGHT_SPHERE
        lights[i]["position"] = c
        lights[i]["radius"] = r
        lights[i]["area"] = 4.0 * math.pi * r * r
        lights[i]["intensity"] = le
    cam = hpt_scenes.look_at_camera((0.0, 0.0, 0.9), (0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), yfov_deg=70.0)
    return hpt.Scene(cam, lights, mats, [hpt.Primitive(fv, fi, 0)], ground_color=(0, 0, 0), sky_color=(0, 0, 0))


def sphere_lights_radiance(p, albedo=0.5):
    """rho/pi * sum_i pi Le_i sin^2(alpha_i) cos(theta_i) at floor points p
    (each sphere entirely above the floor's horizon: checked).  This is the
    ideal value; sphere_lights_expected adds the reference's shadow-ray
    offsets."""
    out = np.zeros((len(p), 3))
    for c, r, le in SPHERE_LIGHTS:
        v = np.asarray(c, np.float64)[None, :] - p
        d = np.linalg.norm(v, axis=1)
        cos = v[:, 2] / d
        assert (np.arcsin(cos) > np.arcsin(r / d)).all(), "sphere crosses the horizon"
        out += (albedo * (r / d) ** 2 * cos)[:, None] * np.asarray(le, np.float64)[None, :]
    return out


def _ray_sphere_t1(o, u, c, r):
    """First intersection distance of rays o + t u with a sphere (inf: none)."""
    op = c[None, ...] - o if o.ndim == c.ndim + 1 else c - o
    b = np.sum(op * u, axis=-1)
    det = b * b - np.sum(op * op, axis=-1) + r * r
    return np.where(det >= 0, b - np.sqrt(np.maximum(det, 0.0)), np.inf)


def sphere_lights_expected(p, albedo=0.5, n_radial=192, n_phi=96, offsets=True):
    """Expected radiance at floor points p (depth 1, black sky) of the
    reference's estimator: light sampling and BSDF sampling with power-
    heuristic weights (inc/lighting.glsl:154-240), integrated numerically over
    each sphere's visible disc (midpoint rule in the cone angle and azimuth).

    With `offsets`, a light sample counts only if its shadow ray -- from
    p + n * EPS (lighting.glsl:72) along the direction from p, any-hit up to
    dist - 2 EPS (lighting.glsl:215) -- does not meet the light's own sphere
    first.  Near the disc's lower rim the offset origin sees the sample point
    behind the sphere's own surface (cos_q < |m.n| / 2 for the sphere normal
    m there), a region independent of EPS's size: the reference's sphere
    lights come out a few percent dark.  Without `offsets` the integral is
    the ideal irradiance of sphere_lights_radiance (a check of the
    quadrature)."""
    n = np.array([0.0, 0.0, 1.0])
    out = np.zeros((len(p), 3))
    for c, r, le in SPHERE_LIGHTS:
        c = np.asarray(c, np.float64)
        area = 4.0 * math.pi * r * r
        v = c[None, :] - p
        dc = np.linalg.norm(v, axis=1)
        w = v / dc[:, None]
        alpha = np.arcsin(r / dc)
        # orthonormal frame around the centre direction
        a = np.where(np.abs(w[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
        t = np.cross(w, a)
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        bvec = np.cross(w, t)
        th = (np.arange(n_radial) + 0.5) / n_radial  # fraction of alpha
        ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
        total = np.zeros(len(p))
        for k in range(n_radial):
            ang = th[k] * alpha  # (P,)
            dom = np.sin(ang) * alpha / n_radial * (2 * math.pi / n_phi)  # solid angle per cell
            u = (np.cos(ang)[:, None, None] * w[:, None, :]
                 + np.sin(ang)[:, None, None] * (np.cos(ph)[None, :, None] * t[:, None, :]
                                                 + np.sin(ph)[None, :, None] * bvec[:, None, :]))
            dist = _ray_sphere_t1(p[:, None, :], u, c, r)  # distance to the sample point q
            q = p[:, None, :] + u * dist[..., None]
            m = (q - c) / r
            cos_q = -np.sum(m * u, axis=-1)
            cos_p = u[..., 2]
            p_light = dist ** 2 / (area * 0.5 * cos_q)
            p_bsdf = cos_p / math.pi
            w_light = p_light ** 2 / (p_light ** 2 + p_bsdf ** 2)
            vis = np.ones_like(dist)
            if offsets:
                o = p[:, None, :] + n * EPS
                t1 = _ray_sphere_t1(o, u, c, r)
                vis = (t1 > dist - 2 * EPS).astype(np.float64)
            f = albedo / math.pi * cos_p
            total += np.sum(f * (w_light * vis + (1.0 - w_light)), axis=1) * dom
        out += total[:, None] * np.asarray(le, np.float64)[None, :]
    return out
