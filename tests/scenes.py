"""Small hand-built scenes with closed-form answers, shared by the CPU
(oracle) and GPU tests. Built through the product host API (HostScene)."""
import numpy as np

from marmel_amd import HostScene, abi


def material(**kw):
    m = np.zeros(1, abi.MATERIAL_DTYPE)[0]
    m["ior"] = 1.0
    m["texture_scale"] = 1.0
    m["texture_color_a"] = (1.0, 1.0, 1.0)
    for k, v in kw.items():
        m[k] = v
    return m


def prim(kind, mat, p0=(0, 0, 0), p1=(0, 0, 0), p2=(0, 0, 0), r0=0.0, r1=0.0, flags=0):
    p = np.zeros(1, abi.PRIM_DTYPE)[0]
    p["kind"], p["material"], p["flags"] = kind, mat, flags
    p["p0"], p["p1"], p["p2"] = p0, p1, p2
    p["r0"], p["r1"] = r0, r1
    return p


def sky(horizon=(0, 0, 0), zenith=(0, 0, 0), sun_color=(0, 0, 0), sun_dir=(0.0, 1.0, 0.0),
        glow=0.0, clouds=False, sun_radius=0.0):
    s = abi.Sky()
    s.sun_dir = abi.vec3(sun_dir)
    s.sun_color = abi.vec3(sun_color)
    s.horizon_color = abi.vec3(horizon)
    s.zenith_color = abi.vec3(zenith)
    s.gradient_gamma = 0.6
    s.sun_glow_exponent = 350.0
    s.sun_glow_strength = glow
    s.cloud_height = 120.0
    s.cloud_scale = 0.0025
    # coverage above every density value switches the cloud layer off
    s.cloud_coverage = 0.5 if clouds else 5.0
    s.cloud_softness = 0.12
    s.cloud_sharpness = 1.5
    s.cloud_octaves = 5
    s.seed = 1337
    s.sun_radius = sun_radius
    return s


def build(prims, mats, sk, water=(0.0, -1), attrs=None):
    hs = HostScene()
    hs.add_materials(np.array(mats, dtype=abi.MATERIAL_DTYPE))
    hs.add_prims(np.array(prims, dtype=abi.PRIM_DTYPE),
                 None if attrs is None else np.array(attrs, dtype=abi.ATTR_DTYPE))
    hs.set_sky(sk)
    hs.set_water(*water)
    hs.build_bvh()
    return hs


def white_furnace():
    """One white Lambert sphere under a uniform sky of radiance 1, no sun: every
    path that leaves the convex sphere returns exactly 1."""
    return build([prim(abi.PRIM_SPHERE, 0, p0=(0, 0, 0), r0=1.0)],
                 [material(albedo=(1, 1, 1))], sky(horizon=(1, 1, 1), zenith=(1, 1, 1)))


def sun_plane(albedo=0.6, sun=(0.3, 0.8, -0.52)):
    """A Lambert plane under a point sun in a black sky: the primary-hit
    radiance is albedo * sun_color * (N . L) exactly (Blinn-Phong NEE has no
    1/pi, the point sun has solid angle 1)."""
    s = np.array(sun, float)
    s /= np.linalg.norm(s)
    return build([prim(abi.PRIM_PLANE, 0, p0=(0, 0, 0), p1=(0, 1, 0))],
                 [material(albedo=(albedo, albedo, albedo))],
                 sky(sun_color=(1.0, 0.9, 0.8), sun_dir=tuple(s))), s


def lamp_over_plane(h=4.0, R=0.5, le=10.0, albedo=0.5):
    """A Lambert plane and one emissive PBR sphere of radiance le straight
    above at height h, black sky: the NEE estimate at the foot point has mean
    albedo * le * pi * R^2 / h^2 (cap of half-angle asin(R/h) about the
    normal; Blinn-Phong NEE without 1/pi)."""
    mats = [material(albedo=(albedo, albedo, albedo)),
            material(pbr=1, albedo=(1, 1, 1), emissive=(le, le, le), roughness=1.0)]
    prims = [prim(abi.PRIM_SPHERE, 1, p0=(0, h, 0), r0=R),
             prim(abi.PRIM_PLANE, 0, p0=(0, 0, 0), p1=(0, 1, 0))]
    return build(prims, mats, sky())


def down_ray(x=0.0, z=0.0, y=3.0, tilt=(0.0, 0.0)):
    r = np.zeros(1, abi.RAY_DTYPE)
    d = np.array([tilt[0], -1.0, tilt[1]])
    d /= np.linalg.norm(d)
    r["ox"], r["oy"], r["oz"] = x, y, z
    r["dx"], r["dy"], r["dz"] = d
    return r


def fog(density, color, height=0.0, falloff=0.0, inscatter=0.5, g=0.7, noise_scale=0.0,
        noise_amount=0.0, steps=0):
    f = abi.Fog()
    f.density, f.color, f.height, f.height_falloff = density, abi.vec3(color), height, falloff
    f.inscatter_strength, f.sun_anisotropy = inscatter, g
    f.noise_scale, f.noise_amount, f.shadow_steps = noise_scale, noise_amount, steps
    return f


# The fog blocks of the reference's own scene files (scenes/ashenvale_forest,
# nordic_trees, cyberpunk_anamorphic, ocean_reef_abyss .scene), densities x4 so
# a small test scene shows them, plus the step clamps (3 -> 4, 200 -> 128 / 16).
FOG_SETS = {
    "ashenvale": fog(0.064, (0.36, 0.15, 0.40), 5.2, 0.18, 0.45, 0.82, 0.16, 0.25, 10),
    "nordic": fog(0.024, (0.84, 0.88, 0.94), 5.5, 0.22, 0.35, 0.74, 0.16, 0.35, 0),
    "cyberpunk": fog(0.060, (0.02, 0.025, 0.05), 10.0),
    "reef": fog(0.012, (0.65, 0.78, 0.90), 1.8, 0.50, 0.60, 0.75, 0.20, 0.10, 16),
    "steps3": fog(0.05, (0.5, 0.55, 0.6), 2.0, 0.3, 0.8, -0.99, 0.3, 0.9, 3),
    "steps200": fog(0.03, (0.5, 0.55, 0.6), 0.0, 0.0, 0.6, 0.99, 0.0, 0.0, 200),
}


def fog_rays():
    """Rays for the fog checks on the forest-and-pond workload: from the air
    (some hit the pond and refract in), from under the water (the path starts
    in the water medium) and from above the fog layer's top."""
    from checkers import seeded_rays
    a = seeded_rays(400, 21, [-10, 0.2, -10], [10, 6, 10])
    b = seeded_rays(150, 22, [-3.5, -0.9, -2.5], [3.5, -0.1, 2.5])
    c = seeded_rays(50, 23, [-10, 26, -10], [10, 30, 10])
    return np.concatenate([a, b, c])


# ---- thin lens, anamorphic squeeze and albedo textures ---------------------

def _uv_wall(mat, x0, x1, y0, y1, z, nx, ny, smooth_rows):
    """A vertical grid of nx x ny quads at depth z, two uv triangles each
    (uv = position in the grid, spanning [0, 1.5] x [0, 1]); the quads of the
    rows in smooth_rows also carry tilted vertex normals."""
    prims, attrs = [], []
    xs, ys = np.linspace(x0, x1, nx + 1), np.linspace(y0, y1, ny + 1)
    for j in range(ny):
        for i in range(nx):
            c = [(xs[i], ys[j]), (xs[i + 1], ys[j]), (xs[i + 1], ys[j + 1]), (xs[i], ys[j + 1])]
            for tri in ((0, 1, 2), (0, 2, 3)):
                v = [(c[k][0], c[k][1], z) for k in tri]
                uv = [(1.5 * (c[k][0] - x0) / (x1 - x0), (c[k][1] - y0) / (y1 - y0), 0.0) for k in tri]
                a = np.zeros(1, abi.ATTR_DTYPE)[0]
                flags = abi.PRIM_UV
                if j in smooth_rows:
                    flags |= abi.PRIM_SMOOTH
                    for k, key in enumerate(("n0", "n1", "n2")):
                        nv = np.array([0.25 * np.sin(3.0 * v[k][0]), 0.15 * np.cos(2.0 * v[k][1]), 1.0])
                        a[key] = tuple(nv / np.linalg.norm(nv))
                for k, key in enumerate(("uv0", "uv1", "uv2")):
                    a[key] = uv[k]
                prims.append(prim(abi.PRIM_TRIANGLE, mat, v[0], v[1], v[2], flags=flags))
                attrs.append(a)
    return prims, attrs


def lens_texture_scene():
    """Every albedo texture kind of the GPU path on every primitive kind that
    carries surface coordinates: a CHECKER ground plane and box, a PBR sphere
    with a 3D checker, UV_CHECKER on a sphere (spherical uv), a tapered
    cylinder, a uv-mapped triangle wall (half of it with smooth normals) and a
    quad whose scale falls back to 1; plus a mirror, a glass sphere and a
    small emissive lamp, under a sun with a disk, glow and clouds."""
    T = abi
    mats = [
        material(albedo=(0.85, 0.85, 0.85), texture_kind=T.TEXTURE_CHECKER, texture_scale=1.0,
                 texture_color_a=(0.9, 0.85, 0.8), texture_color_b=(0.1, 0.15, 0.25)),
        material(pbr=1, albedo=(1, 1, 1), roughness=0.35, metallic=0.3,
                 texture_kind=T.TEXTURE_CHECKER, texture_scale=0.35,
                 texture_color_a=(0.9, 0.3, 0.2), texture_color_b=(0.2, 0.3, 0.9)),
        material(albedo=(0.8, 0.7, 0.6), texture_kind=T.TEXTURE_UV_CHECKER, texture_scale=8.0,
                 texture_color_a=(1.0, 1.0, 1.0), texture_color_b=(0.05, 0.05, 0.05)),
        material(albedo=(0.9, 0.9, 0.9), texture_kind=T.TEXTURE_UV_CHECKER, texture_scale=4.0,
                 texture_color_a=(0.9, 0.2, 0.2), texture_color_b=(0.9, 0.9, 0.2)),
        material(albedo=(0.7, 0.7, 0.7), specular=(0.3, 0.3, 0.3), shininess=40.0,
                 texture_kind=T.TEXTURE_CHECKER, texture_scale=0.25,
                 texture_color_a=(0.2, 0.8, 0.3), texture_color_b=(0.9, 0.9, 0.9)),
        material(pbr=1, albedo=(0.95, 0.9, 0.85), roughness=0.6, metallic=0.0,
                 texture_kind=T.TEXTURE_UV_CHECKER, texture_scale=6.0,
                 texture_color_a=(0.3, 0.3, 0.9), texture_color_b=(0.9, 0.6, 0.1)),
        material(albedo=(0.9, 0.9, 0.9), reflectivity=1.0),
        material(albedo=(1, 1, 1), transparency=1.0, ior=1.5),
        material(pbr=1, albedo=(1, 1, 1), emissive=(8.0, 7.0, 6.0), roughness=1.0),
        material(albedo=(0.8, 0.8, 0.8), texture_kind=T.TEXTURE_UV_CHECKER, texture_scale=-2.0,
                 texture_color_a=(0.1, 0.9, 0.9), texture_color_b=(0.9, 0.1, 0.9)),
    ]
    prims = [
        prim(T.PRIM_PLANE, 0, p0=(0, 0, 0), p1=(0, 1, 0)),
        prim(T.PRIM_SPHERE, 8, p0=(-0.5, 3.2, -1.5), r0=0.3),
        prim(T.PRIM_SPHERE, 1, p0=(-2.2, 1.0, 0.0), r0=1.0),
        prim(T.PRIM_SPHERE, 2, p0=(0.3, 0.8, 1.2), r0=0.8),
        prim(T.PRIM_SPHERE, 6, p0=(2.4, 0.7, -0.5), r0=0.7),
        prim(T.PRIM_SPHERE, 7, p0=(1.2, 0.5, 2.6), r0=0.5),
        prim(T.PRIM_BOX, 4, p0=(-0.6, 0.5, -2.0), p1=(0.5, 0.5, 0.5)),
        prim(T.PRIM_CYLINDER, 5, p0=(3.2, 0.0, 1.8), p1=(3.2, 1.6, 1.8), r0=0.45, r1=0.3),
    ]
    attrs = [np.zeros(1, abi.ATTR_DTYPE)[0] for _ in prims]
    wp, wa = _uv_wall(3, -4.5, 4.5, 0.0, 3.0, -4.0, 6, 4, smooth_rows=(2, 3))
    qp, qa = _uv_wall(9, 1.8, 3.0, 1.6, 2.6, -2.5, 1, 1, smooth_rows=())
    sk = sky(horizon=(0.45, 0.5, 0.6), zenith=(0.15, 0.25, 0.5), sun_color=(1.4, 1.3, 1.2),
             sun_dir=tuple(np.array([0.4, 0.7, 0.3]) / np.linalg.norm([0.4, 0.7, 0.3])),
             glow=0.5, clouds=True, sun_radius=0.5)
    return build(prims + wp + qp, mats, sk, attrs=attrs + wa + qa)


# Thin-lens cameras of the lens-and-texture scene: (aperture, focus distance,
# anamorphic squeeze). A squeeze of 1 is the round lens; above 1 the lens
# disc is compressed horizontally, below 1 widened.
LENS_EYE, LENS_AT, LENS_VFOV = (0.5, 2.4, 8.5), (0.0, 0.9, 0.0), 45.0
LENS_CAMS = {
    "round": (0.15, 8.4, 1.0),
    "anamorphic": (0.25, 6.0, 1.8),
    "wide": (0.2, 10.0, 0.55),
}


def lens_camera(hs, name, aspect):
    ap, fd, sq = LENS_CAMS[name]
    cam = hs.camera(aspect, eye=list(LENS_EYE), at=list(LENS_AT), up=[0, 1, 0], vfov=LENS_VFOV)
    cam.aperture, cam.focus_distance, cam.anamorphic_squeeze = ap, fd, sq
    return cam
