"""Scenes built through tinypt_amd.wrapper the way tinypt's Python example
builds its Cornell-sphere scene (li-plus/tinypt python/examples/main.py:34-96:
five wall rectangles, a mirror and a glass sphere, a disc light, a camera
looking along +y), and the same scene built directly on the `_C` objects.

`extras` adds what the wrapper offers beyond the reference's Python package: a
GGX sphere with a '#rrggbb' colour, an alpha-cut, bump-mapped glossy
rectangle, a point light, a sun (DistantLight) and a '#rrggbb' sky."""
import math

import numpy as np

import tinypt_amd.wrapper as tinypt
from tinypt_amd import _C


def rot_x(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float32)


def rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float32)


# box corners, as in the example
X1, Y1, Z1 = 1.0, -170.0, 0.0
X2, Y2, Z2 = 99.0, 0.0, 81.6
XM, YM, ZM = (X1 + X2) / 2, (Y1 + Y2) / 2, (Z1 + Z2) / 2
XS, YS, ZS = X2 - X1, Y2 - Y1, Z2 - Z1
WALLS = [  # location, rotation, dimension, albedo
    ((X1, YM, ZM), rot_y(90), (ZS, YS), (0.75, 0.25, 0.25)),
    ((X2, YM, ZM), rot_y(90), (ZS, YS), (0.25, 0.25, 0.75)),
    ((XM, Y2, ZM), rot_x(90), (XS, ZS), (0.75, 0.75, 0.75)),
    ((XM, YM, Z1), None, (XS, YS), (0.75, 0.75, 0.75)),
    ((XM, YM, Z2), None, (XS, YS), (0.75, 0.75, 0.75)),
]
MIRROR = (16.5, (27.0, -47.0, 16.5))
GLASS = (16.5, (73.0, -78.0, 16.5))
DISC = (18.0, (50.0, -81.6, 81.6 - 0.01), (12.0, 12.0, 12.0))


def camera_rotation():
    w = np.array([0, -0.999093, 0.042573], dtype=np.float32)
    u = np.array([1, 0, 0], dtype=np.float32)
    return np.vstack((u, np.cross(w, u), w)).T


def height_ramp(n=32):
    return np.tile(np.linspace(0.0, 1.0, n, dtype=np.float32), (n, 1))


def alpha_stripes(n=32):
    a = np.ones((n, n), dtype=np.float32)
    a[:, ::4] = 0.0
    return a


def cornell_wrapper(width, height, extras=False):
    objects = [tinypt.Rectangle(location=loc, rotation=rot, dimension=dim, surface=tinypt.Lambertian(texture=alb))
               for loc, rot, dim, alb in WALLS]
    objects += [
        tinypt.Sphere(radius=MIRROR[0], location=MIRROR[1], surface=tinypt.Metal(texture=(1, 1, 1))),
        tinypt.Sphere(location=GLASS[1], radius=GLASS[0], surface=tinypt.Dielectric(texture=(1, 1, 1), ior=1.5)),
        tinypt.Circle(radius=DISC[0], location=DISC[1], emission=DISC[2]),
    ]
    lights, background = [], None
    if extras:
        objects += [
            tinypt.Sphere(radius=10, location=(50, -40, 10), surface=tinypt.Glossy(200, "#C08040", model="ggx")),
            tinypt.Rectangle(location=(50, -20, 30), rotation=rot_x(90), dimension=(30, 30),
                             surface=tinypt.Mix(tinypt.Glossy(50, (0.8, 0.8, 0.8)), tinypt.Lambertian("#4080C0"), 0.3),
                             alpha=tinypt.AlphaTexture(image=alpha_stripes()),
                             bump=tinypt.BumpTexture(height_ramp(), scale=8.0)),
        ]
        lights = [tinypt.PointLight(location=(30, -100, 60), color=(1, 0.9, 0.8), power=2000, radius=1.0),
                  tinypt.DistantLight(rotation=rot_x(-40), color=(1, 1, 1), power=2, angle=math.radians(5))]
        background = tinypt.EnvTexture(texture="#5689BE")
    camera = tinypt.Camera(location=(50, -295.6, 52), rotation=camera_rotation(), resolution=(width, height),
                           fov=math.radians(39.32))
    return tinypt.Scene(camera=camera, objects=objects, lights=lights, background=background)


def cornell_native(width, height):
    """cornell_wrapper(extras=False) spelled out on `_C` with every default the
    reference wrapper fills in made explicit."""
    def rgb(c):
        return _C.RGBTexture([float(x) for x in c])

    def material(surface, emission=None):
        m = _C.Material(surface)
        if emission is not None:
            m.emission_texture = rgb(emission)
        return m

    objects = []
    for loc, rot, dim, alb in WALLS:
        r = _C.Rectangle()
        r.location = list(loc)
        if rot is not None:
            r.rotation = rot
        r.dimension = list(dim)
        r.material = material(_C.Lambertian(rgb(alb)))
        objects.append(r)
    for (radius, loc), surface in ((MIRROR, _C.Metal(rgb((1, 1, 1)))), (GLASS, _C.Dielectric(rgb((1, 1, 1)), 1.5))):
        s = _C.Sphere()
        s.radius, s.location, s.material = radius, list(loc), material(surface)
        objects.append(s)
    c = _C.Circle()
    c.radius, c.location = DISC[0], list(DISC[1])
    c.material = material(_C.Lambertian(rgb((0, 0, 0))), DISC[2])
    objects.append(c)
    cam = _C.Camera()
    cam.location = [50.0, -295.6, 52.0]
    cam.rotation = camera_rotation()
    cam.width, cam.height = width, height
    cam.fov = math.radians(39.32)
    return _C.Scene(cam, objects, [], _C.EnvTexture(rgb((0, 0, 0))))
