"""Seeded scene generators for the benchmark workloads, built through the
tinypt-style object API of this package.

* ``cornell_spheres`` — the analytic two-spheres-in-a-box scene: six radius-1e5
  wall spheres (two tinted, four white diffuse), a mirror and a glass sphere
  (r = 16.5, IOR 1.5) and an emissive r = 600 light sphere (emission 12), laid
  out like smallpt / tinypt's ``make_debug2`` (src/examples/scene_factory.cpp:225-291).
* ``triangle_room`` — an axis-aligned box room plus 200 instanced meshes
  (icospheres and subdivided boxes, random rotation, scale U(0.2, 2)), baked
  into ONE world-space TriangleMesh of ~265k triangles; materials 60 %
  Lambertian / 15 % metal / 10 % dielectric / 15 % GGX; four quad area lights.

The reference has no API to build meshes from arrays (only ``from_obj``);
``TriangleMesh.from_arrays`` stands in for an OBJ with one shape.
"""
import math

import numpy as np

from . import _C


def rgb(v):
    return _C.RGBTexture([float(x) for x in v])


def material(surface, emission=None):
    m = _C.Material(surface)
    if emission is not None:
        m.emission_texture = rgb(emission)
    return m


def lookat_rotation(forward, up=(0.0, 0.0, 1.0)):
    """Camera rotation with columns (u, v, w), w = -forward (cpu/camera.cpp:17)."""
    f = np.asarray(forward, dtype=np.float64)
    f = f / np.linalg.norm(f)
    u = np.cross(f, np.asarray(up, dtype=np.float64))
    u = u / np.linalg.norm(u)
    w = -f
    v = np.cross(w, u)
    return np.stack([u, v, w], axis=1).astype(np.float32)


# ---------------------------------------------------------------- cornell spheres
def cornell_spheres(width=512, height=384):
    def sphere(center, radius, surface, emission=None):
        s = _C.Sphere()
        s.location = [float(c) for c in center]
        s.radius = float(radius)
        s.material = material(surface, emission)
        return s

    white = _C.Lambertian(rgb((0.75, 0.75, 0.75)))
    objs = [
        sphere((1e5 + 1, 40.8, 81.6), 1e5, _C.Lambertian(rgb((0.75, 0.25, 0.25)))),   # left
        sphere((-1e5 + 99, 40.8, 81.6), 1e5, _C.Lambertian(rgb((0.25, 0.25, 0.75)))),  # right
        sphere((50, 40.8, 1e5), 1e5, white),                                           # back
        sphere((50, 40.8, -1e5 + 170), 1e5, white),                                    # front
        sphere((50, 1e5, 81.6), 1e5, white),                                           # bottom
        sphere((50, -1e5 + 81.6, 81.6), 1e5, white),                                   # top
        sphere((27, 16.5, 47), 16.5, _C.Metal(rgb((0.999, 0.999, 0.999)))),
        sphere((73, 16.5, 78), 16.5, _C.Dielectric(rgb((0.999, 0.999, 0.999)), 1.5)),
        sphere((50, 681.6 - 0.27, 81.6), 600, _C.Lambertian(rgb((0, 0, 0))), emission=(12, 12, 12)),
    ]
    cam = _C.Camera()
    # smallpt's camera at z = 295.6 starts its rays 140 units forward (inside the
    # front wall); tinypt's pinhole has no near plane, so the eye sits at z = 160.
    cam.location = [50.0, 52.0, 160.0]
    w = np.array([0.0, 0.042612, 1.0])
    w /= np.linalg.norm(w)
    u = np.array([1.0, 0.0, 0.0])
    v = np.cross(w, u)
    cam.rotation = np.stack([u, v, w], axis=1).astype(np.float32)
    cam.width, cam.height = width, height
    cam.fov = math.radians(55.0)
    return _C.Scene(cam, objs, [], None)


# ---------------------------------------------------------------- meshes
def icosphere(level):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    return V, np.array(faces, dtype=np.int64), V.copy()


def subdivided_box(n):
    verts, faces = [], []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            base = len(verts)
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            for i in range(n + 1):
                for j in range(n + 1):
                    p = [0.0, 0.0, 0.0]
                    p[axis] = sign
                    p[a1] = -1 + 2 * i / n
                    p[a2] = -1 + 2 * j / n
                    verts.append(p)
            for i in range(n):
                for j in range(n):
                    q00 = base + i * (n + 1) + j
                    q10, q01, q11 = q00 + n + 1, q00 + 1, q00 + n + 2
                    if sign > 0:
                        faces += [(q00, q10, q11), (q00, q11, q01)]
                    else:
                        faces += [(q00, q11, q10), (q00, q01, q11)]
    return np.array(verts, dtype=np.float64), np.array(faces, dtype=np.int64)


def random_rotation(rng):
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def room_materials(rng, n, lambert_texture=None):
    """Per-instance materials: 60% Lambertian U(0.1,0.9), 15% metal, 10% dielectric IOR U(1.3,1.7),
    15% GGX roughness U(0.05,0.6) (alpha = roughness, shininess = 2/alpha^2 - 2).
    `lambert_texture` replaces the Lambertian albedo (the draws stay the same)."""
    mats = []
    for _ in range(n):
        r = rng.random()
        if r < 0.60:
            a = rgb(rng.uniform(0.1, 0.9, 3))
            mats.append(material(_C.Lambertian(a if lambert_texture is None else lambert_texture)))
        elif r < 0.75:
            mats.append(material(_C.Metal(rgb(rng.uniform(0.5, 0.95, 3)))))
        elif r < 0.85:
            mats.append(material(_C.Dielectric(rgb((0.999, 0.999, 0.999)), float(rng.uniform(1.3, 1.7)))))
        else:
            rough = float(rng.uniform(0.05, 0.6))
            shin = 2.0 / (rough * rough) - 2.0
            mats.append(material(_C.Glossy(shin, rgb(rng.uniform(0.5, 0.95, 3)), _C.Glossy.MICRO_TROWBRIDGE_REITZ)))
    return mats


def _local_uv(lv, is_ico):
    """Texture coordinates of an instance's local vertices: spherical for the
    icospheres, per-face planar for the subdivided boxes."""
    if is_ico:
        return np.stack([0.5 + np.arctan2(lv[:, 1], lv[:, 0]) / (2 * np.pi),
                         0.5 + np.arcsin(np.clip(lv[:, 2], -1.0, 1.0)) / np.pi], 1)
    ax = np.argmax(np.abs(lv), axis=1)
    a1, a2 = (ax + 1) % 3, (ax + 2) % 3
    idx = np.arange(len(lv))
    return np.stack([(lv[idx, a1] + 1) * 0.5, (lv[idx, a2] + 1) * 0.5], 1)


def triangle_room_arrays(seed=1, instances=200, ico_level=3, box_n=11, n_ico=150, uv=False):
    """World-space (vertices, normals-or-None, faces9, face_material, materials-spec) for the room;
    with `uv`, also per-vertex texture coordinates of the instances (the faces then
    reference them; the room box has none)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = 40.0, 40.0, 20.0
    # room box (inward facing), no vertex normals
    rv = np.array([[0, 0, 0], [X, 0, 0], [X, Y, 0], [0, Y, 0], [0, 0, Z], [X, 0, Z], [X, Y, Z], [0, Y, Z]], dtype=np.float64)
    rf = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                   [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.int64)
    ico_v, ico_f, _ = icosphere(ico_level)
    box_v, box_f = subdivided_box(box_n)
    kinds = np.array([0] * n_ico + [1] * (instances - n_ico))
    rng.shuffle(kinds)
    V = [rv]
    N = [np.zeros_like(rv)]
    UV = [np.zeros((len(rv), 2))]
    F9 = []
    FM = []
    # room faces: material 0, no normals
    for f in rf:
        F9.append([f[0], -1, -1, f[1], -1, -1, f[2], -1, -1])
        FM.append(0)
    nv = len(rv)
    nn = len(rv)
    for i in range(instances):
        R = random_rotation(rng)
        s = float(rng.uniform(0.2, 2.0))
        pos = np.array([rng.uniform(3, X - 3), rng.uniform(6, Y - 3), rng.uniform(0.5, 12.0)])
        if kinds[i] == 0:
            lv, lf, ln = ico_v, ico_f, ico_v
        else:
            lv, lf, ln = box_v, box_f, None
        wv = (lv * s) @ R.T + pos
        V.append(wv)
        if uv:
            UV.append(_local_uv(lv, kinds[i] == 0))
        if ln is not None:
            N.append(ln @ R.T)
        t = nv if uv else None
        for f in lf:
            ta, tb, tc = (f[0] + t, f[1] + t, f[2] + t) if uv else (-1, -1, -1)
            if ln is not None:
                F9.append([f[0] + nv, f[0] + nn, ta, f[1] + nv, f[1] + nn, tb, f[2] + nv, f[2] + nn, tc])
            else:
                F9.append([f[0] + nv, -1, ta, f[1] + nv, -1, tb, f[2] + nv, -1, tc])
            FM.append(1 + i)
        nv += len(lv)
        if ln is not None:
            nn += len(ln)
    verts = np.concatenate(V).astype(np.float32)
    normals = np.concatenate(N).astype(np.float32)
    faces9 = np.array(F9, dtype=np.int32)
    fm = np.array(FM, dtype=np.int32)
    if uv:
        return verts, normals, faces9, fm, rng, np.concatenate(UV).astype(np.float32)
    return verts, normals, faces9, fm, rng


# (icosphere level, box subdivision) pairs of similar face counts: 20 * 4^l vs 12 * n^2
_ROOM_SIZES = ((0, 1), (1, 3), (2, 5), (3, 10), (4, 21))


def room_size_for(tris):
    """(instances, ico_level, box_n, n_ico) for a room of about `tris` triangles:
    75% icospheres, the instance size whose count lands closest to 200."""
    cands = []
    for lvl, bn in _ROOM_SIZES:
        per = 0.75 * 20 * 4 ** lvl + 0.25 * 12 * bn * bn
        inst = max(4, int(round((tris - 12) / per)))
        cands.append((abs(math.log(inst / 200.0)), inst, lvl, bn))
    _, inst, lvl, bn = min(cands)
    return inst, lvl, bn, int(round(0.75 * inst))


def triangle_room(width=1024, height=768, seed=1, tris=None):
    """The seeded triangle room; `tris` picks instance count and tessellation for
    about that many triangles (None: 200 instances, 264,612 triangles)."""
    if tris is None:
        inst, lvl, bn, n_ico = 200, 3, 11, 150
    else:
        inst, lvl, bn, n_ico = room_size_for(tris)
    verts, normals, faces9, fm, rng = triangle_room_arrays(seed, inst, lvl, bn, n_ico)
    mats = [material(_C.Lambertian(rgb((0.75, 0.75, 0.75))))] + room_materials(rng, inst)
    mesh = _C.TriangleMesh.from_arrays(verts, faces9, mats, fm, normals, None)
    lights = []
    for cx, cy in ((10, 10), (30, 10), (10, 30), (30, 30)):
        r = _C.Rectangle()
        r.location = [float(cx), float(cy), 19.9]
        r.dimension = [6.0, 6.0]
        e = float(rng.uniform(5, 20))
        r.material = material(_C.Lambertian(rgb((0, 0, 0))), emission=(e, e, e))
        lights.append(r)
    cam = _C.Camera()
    cam.location = [20.0, 1.0, 9.0]
    cam.rotation = lookat_rotation((0.0, 1.0, -0.15))
    cam.width, cam.height = width, height
    cam.fov = math.radians(65.0)
    return _C.Scene(cam, [mesh] + lights, [], None)


def value_noise(rng, size, cells):
    """size x size bilinear value noise over a (cells + 1)^2 random lattice, in [0, 1]."""
    g = rng.random((cells + 1, cells + 1))
    x = np.arange(size) * (cells / size)
    i = x.astype(np.int64)
    f = x - i
    a = g[i][:, i] * (1 - f)[None, :] + g[i][:, i + 1] * f[None, :]
    b = g[i + 1][:, i] * (1 - f)[None, :] + g[i + 1][:, i + 1] * f[None, :]
    return a * (1 - f)[:, None] + b * f[:, None]


def procedural_textures(seed=3, albedo_size=1024, height_size=512):
    """(RGBA8 albedo: checker x value noise, alpha 0 where the noise is low;
    FP32 height map: value noise), both seeded."""
    rng = np.random.default_rng(seed)
    n = 0.5 * value_noise(rng, albedo_size, 8) + 0.3 * value_noise(rng, albedo_size, 32) + \
        0.2 * value_noise(rng, albedo_size, 128)
    y, x = np.mgrid[0:albedo_size, 0:albedo_size]
    checker = ((x // 64 + y // 64) % 2).astype(np.float64)
    c0, c1 = np.array([0.85, 0.75, 0.55]), np.array([0.25, 0.35, 0.6])
    rgb = (c0 * (1 - checker)[..., None] + c1 * checker[..., None]) * (0.5 + n)[..., None]
    rgba = np.empty((albedo_size, albedo_size, 4), np.uint8)
    rgba[..., :3] = np.clip(rgb * 255.0, 0, 255).astype(np.uint8)
    rgba[..., 3] = np.where(n < 0.38, 0, 255).astype(np.uint8)
    height = (0.6 * value_noise(rng, height_size, 16) + 0.4 * value_noise(rng, height_size, 64)).astype(np.float32)
    return rgba, height


def textured_room(width=1920, height=1080, seed=1, texture_seed=3, cutout=0.2, bump=0.5):
    """The 264,612-triangle room with per-vertex texture coordinates, a 1024^2
    RGBA8 checker / value-noise albedo on its Lambertian and glossy instances,
    alpha cut-out (the texture's alpha channel) on `cutout` of the instances and a
    512^2 FP32 height map as bump map on `bump` of them."""
    verts, normals, faces9, fm, rng, uvs = triangle_room_arrays(seed, uv=True)
    rgba, hmap = procedural_textures(texture_seed)
    albedo_img = _C.Image.from_linear(rgba)
    albedo = _C.RGBTexture(albedo_img)
    alpha = _C.AlphaTexture(albedo_img, 3)
    bumpt = _C.BumpTexture(_C.Image.from_linear(hmap), 4.0)
    mats = [material(_C.Lambertian(rgb((0.75, 0.75, 0.75))))]
    trng = np.random.default_rng(seed * 7919 + texture_seed)
    for m in room_materials(rng, 200, albedo):
        if trng.random() < cutout:
            m.alpha_texture = alpha
        if trng.random() < bump:
            m.bump_texture = bumpt
        mats.append(m)
    mesh = _C.TriangleMesh.from_arrays(verts, faces9, mats, fm, normals, uvs)
    lights = []
    for cx, cy in ((10, 10), (30, 10), (10, 30), (30, 30)):
        r = _C.Rectangle()
        r.location = [float(cx), float(cy), 19.9]
        r.dimension = [6.0, 6.0]
        e = float(rng.uniform(5, 20))
        r.material = material(_C.Lambertian(rgb((0, 0, 0))), emission=(e, e, e))
        lights.append(r)
    cam = _C.Camera()
    cam.location = [20.0, 1.0, 9.0]
    cam.rotation = lookat_rotation((0.0, 1.0, -0.15))
    cam.width, cam.height = width, height
    cam.fov = math.radians(65.0)
    return _C.Scene(cam, [mesh] + lights, [], None)


def procedural_sky(width=2048, height=1024, seed=11):
    """RGBA32F equirectangular HDR sky (row 0 = bottom = nadir, as the texture
    convention): a horizon-to-zenith gradient with value-noise clouds and a bright
    sun disk; the last channel is 1."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height            # 0 = nadir, 1 = zenith
    u = (np.arange(width) + 0.5) / width
    elev = (v - 0.5) * np.pi                          # -pi/2 .. pi/2
    up = np.clip(np.sin(elev), 0.0, 1.0)[:, None]
    horizon = np.array([0.9, 0.85, 0.8]) * 1.5
    zenith = np.array([0.25, 0.45, 0.95]) * 1.2
    ground = np.array([0.15, 0.13, 0.12])
    sky = horizon * (1 - up[..., None]) + zenith * up[..., None]
    sky = np.where((elev < 0)[:, None, None], ground, sky) * np.ones((height, width, 1))
    clouds = value_noise(rng, max(width, height), 24)[:height, :width]
    sky = sky * (1.0 + 0.6 * np.clip(clouds - 0.5, 0, 1)[..., None] * (up[..., None] > 0))
    # sun disk (elevation 35 deg, azimuth at u = 0.3)
    su, se = 0.3, math.radians(35.0)
    d = np.sqrt(((u[None, :] - su) * 2 * np.pi * np.cos(elev[:, None])) ** 2 + (elev[:, None] - se) ** 2)
    sky = sky + (d < 0.02)[..., None] * np.array([2000.0, 1800.0, 1500.0])
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = sky.astype(np.float32)
    return out


def city_arrays(seed=21, target_tris=6_000_000, grid=512, extent=2000.0):
    """World-space (vertices, faces (F, 3), face material, material count) of a
    seeded city: a grid x grid displaced heightfield (2 (grid - 1)^2 triangles)
    plus extruded box buildings (10 triangles each, no bottom) with log-normal
    heights of mean 30, filling the rest of `target_tris`."""
    rng = np.random.default_rng(seed)
    # heightfield
    h = 25.0 * value_noise(rng, grid, 8) + 5.0 * value_noise(rng, grid, 48)
    xs = np.linspace(0.0, extent, grid)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    tv = np.stack([X.ravel(), Y.ravel(), h.ravel()], 1)
    i, j = np.meshgrid(np.arange(grid - 1), np.arange(grid - 1), indexing="xy")
    a = (j * grid + i).ravel()
    b, c, d = a + 1, a + grid + 1, a + grid
    tf = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    # buildings on a jittered grid of lots
    nb = max(0, (target_tris - len(tf)) // 10)
    side = int(math.ceil(math.sqrt(nb)))
    lot = extent / side
    k = np.arange(nb)
    cx = (k % side + 0.5) * lot + rng.uniform(-0.15, 0.15, nb) * lot
    cy = (k // side + 0.5) * lot + rng.uniform(-0.15, 0.15, nb) * lot
    hw = rng.uniform(0.25, 0.42, (nb, 2)) * lot
    sigma = 0.6
    height = rng.lognormal(math.log(30.0) - sigma * sigma / 2, sigma, nb)
    gi = np.clip((cx / extent * (grid - 1)).astype(np.int64), 0, grid - 1)
    gj = np.clip((cy / extent * (grid - 1)).astype(np.int64), 0, grid - 1)
    base = h[gj, gi] - 1.0
    corners = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float64)
    bx = cx[:, None] + corners[None, :, 0] * hw[:, 0:1]
    by = cy[:, None] + corners[None, :, 1] * hw[:, 1:2]
    bot = np.stack([bx, by, np.repeat(base[:, None], 4, 1)], 2)
    top = np.stack([bx, by, np.repeat((base + height)[:, None], 4, 1)], 2)
    bv = np.concatenate([bot, top], 1).reshape(-1, 3)  # 8 vertices per building
    box = np.array([[4, 5, 6], [4, 6, 7],                      # roof
                    [0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5],  # walls
                    [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int64)
    bf = (box[None, :, :] + (len(tv) + 8 * k)[:, None, None]).reshape(-1, 3)
    verts = np.concatenate([tv, bv]).astype(np.float32)
    faces = np.concatenate([tf, bf]).astype(np.int32)
    n_mat = 64
    fm = np.concatenate([np.zeros(len(tf), np.int32),
                         np.repeat(rng.integers(1, n_mat, nb).astype(np.int32), 10)])
    return verts, faces, fm, n_mat, rng


def city(width=1920, height=1080, seed=21, target_tris=6_000_000, point_lights=1000, importance=True):
    """The configs' city: ~6M triangles, a 2048x1024 HDR sky sampled through its
    CDF (`importance`), one sun (distant light) and `point_lights` point lights;
    the terrain is diffuse, the buildings use a palette of GGX and dielectric
    materials."""
    verts, faces, fm, n_mat, rng = city_arrays(seed, target_tris)
    mats = [material(_C.Lambertian(rgb((0.35, 0.33, 0.3))))]
    for _ in range(n_mat - 1):
        if rng.random() < 0.7:
            rough = float(rng.uniform(0.1, 0.6))
            shin = 2.0 / (rough * rough) - 2.0
            mats.append(material(_C.Glossy(shin, rgb(rng.uniform(0.3, 0.9, 3)), _C.Glossy.MICRO_TROWBRIDGE_REITZ)))
        else:
            mats.append(material(_C.Dielectric(rgb((0.95, 0.97, 0.99)), float(rng.uniform(1.4, 1.6)))))
    mesh = _C.TriangleMesh.from_arrays(verts, faces, mats, fm)
    lights = []
    sun = _C.DistantLight()
    sun.rotation = _C.rotation_from_direction([math.cos(math.radians(35)) * math.cos(0.3 * 2 * math.pi),
                                               math.cos(math.radians(35)) * math.sin(0.3 * 2 * math.pi),
                                               math.sin(math.radians(35))])
    sun.power = 3.0
    sun.color = [1.0, 0.95, 0.85]
    lights.append(sun)
    ext = float(verts[:, 0].max())
    for _ in range(point_lights):
        pl = _C.PointLight()
        pl.location = [float(rng.uniform(0, ext)), float(rng.uniform(0, ext)), float(rng.uniform(8.0, 20.0))]
        pl.color = [1.0, 0.8, 0.55]
        pl.power = float(rng.uniform(0.5, 2.0))
        lights.append(pl)
    bg = _C.EnvTexture(_C.RGBTexture(_C.Image.from_linear(procedural_sky())), importance)
    cam = _C.Camera()
    cam.location = [ext * 0.1, ext * 0.1, 180.0]
    cam.rotation = lookat_rotation((1.0, 1.0, -0.35))
    cam.width, cam.height = width, height
    cam.fov = math.radians(60.0)
    return _C.Scene(cam, [mesh], lights, bg)


# Workload presets.  Only the two-spheres box asks for Russian roulette (after
# depth 5); the others name a maximum depth and no roulette (rr_depth -1), and the
# city names no depth, so it takes the triangle room's 8.
CONFIG0 = dict(width=512, height=384, spp=64, max_depth=10, rr_depth=5)
CONFIG1 = dict(width=1024, height=768, spp=256, max_depth=8, rr_depth=-1)
CONFIG2 = dict(count=64, tris=(2000, 500000), side=(128, 1024), spp=(16, 256), max_depth=8, rr_depth=-1)
CONFIG3 = dict(width=1920, height=1080, max_depth=12, rr_depth=-1)
CONFIG4 = dict(width=1920, height=1080, spp=1024, max_depth=8, rr_depth=-1)


def batch_specs(count=64, seed=2024, tris=(2000, 500000), side=(128, 1024), spp=(16, 256)):
    """Seeded specs of a batch of independent triangle rooms: triangle counts
    log-uniform in `tris`, square resolutions log-uniform in `side` (multiples of
    8), samples per pixel uniform in `spp`."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        t = int(round(math.exp(rng.uniform(math.log(tris[0]), math.log(tris[1])))))
        s = int(round(2.0 ** rng.uniform(math.log2(side[0]), math.log2(side[1])) / 8.0)) * 8
        n = int(rng.integers(spp[0], spp[1] + 1))
        specs.append(dict(seed=seed * 1000 + i, tris=t, width=s, height=s, spp=n))
    return specs


def batch_scenes(specs):
    """Host scenes of `batch_specs` entries."""
    return [triangle_room(sp["width"], sp["height"], sp["seed"], sp["tris"]) for sp in specs]
