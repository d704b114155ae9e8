"""Object-level front end with the class names, constructor arguments, defaults
and errors of tinypt's Python package (li-plus/tinypt
python/tinypt/wrapper.py:51-270 and __init__.py), over this package's `_C`.

A script written against ``import tinypt`` runs unchanged with::

    import tinypt_amd.wrapper as tinypt
    scene = make_scene(...).to("hip")
    image = tinypt.PathTracer("hip").render(scene, spp).numpy()

Reference behaviour kept:
- colours are RGB triples or ``"#rrggbb"`` strings, converted to linear RGB
  by ``(c / 255) ** 2.2`` (wrapper.py:9-16); any other string raises
  RuntimeError, and so does a vector of the wrong shape (wrapper.py:19-30);
- a texture argument is an RGBTexture, an H x W x 3 display-space array (an
  image map) or a colour (wrapper.py:33-43);
- a Material without a surface is a black Lambertian, a Scene without a
  background has a black EnvTexture, a Circle without a location sits at the
  origin (wrapper.py:107-108, 218-219, 253-254);
- ``PathTracer(device)`` defaults to ``"cpu"``, the reference's CPU renderer,
  which this package does not ship: it raises RuntimeError naming the devices
  it has.

Added beyond the reference's Python package (its ``_C`` binds some of them):
``Glossy`` (Blinn-Phong or Trowbridge-Reitz microfacets), ``Add`` / ``Mix``
combinations of two surfaces, ``PointLight``, ``ObjectLight``,
``AlphaTexture`` and ``BumpTexture`` with their maps, the thin lens of
``Camera``, the importance flag of ``EnvTexture`` and render options passed
through ``PathTracer.render``.
"""
import warnings

import numpy as np

from . import _C

__all__ = [
    "PathTracer",
    "Scene",
    "Camera",
    "Sphere",
    "Rectangle",
    "Circle",
    "TriangleMesh",
    "Lambertian",
    "Metal",
    "Dielectric",
    "Glossy",
    "Add",
    "Mix",
    "DistantLight",
    "PointLight",
    "ObjectLight",
    "EnvLight",
    "Material",
    "RGBTexture",
    "AlphaTexture",
    "BumpTexture",
    "EnvTexture",
]

_GAMMA = 2.2


def _hex_rgb(text):
    """'#rgb' / '#rrggbb' (and their alpha forms) as integer channels, the
    forms PIL.ImageColor.getrgb accepts after a '#'."""
    digits = text[1:]
    if len(digits) in (3, 4):
        digits = "".join(c * 2 for c in digits)
    if len(digits) not in (6, 8):
        raise ValueError(f"unknown color specifier: {text!r}")
    try:
        return tuple(int(digits[i:i + 2], 16) for i in range(0, len(digits), 2))
    except ValueError:
        raise ValueError(f"unknown color specifier: {text!r}") from None


def _parse_color(color):
    if isinstance(color, str):
        if not color.startswith("#"):
            raise RuntimeError("not implemented")
        color = tuple((c / 255.0) ** _GAMMA for c in _hex_rgb(color))
    return _parse_vec3f(color)


def _parse_vec2f(v):
    arr = np.asarray(v, dtype=np.float32)
    if arr.shape != (2,):
        raise RuntimeError(f"Invalid Vec2f {v}")
    return arr


def _parse_vec3f(v):
    arr = np.array(v, dtype=np.float32)
    if arr.shape != (3,):
        raise RuntimeError(f"Invalid Vec3f {v}")
    return arr


def _rotation(r):
    return np.asarray(r, dtype=np.float32)


def _parse_rgb_texture(texture):
    if texture is None:
        return RGBTexture()
    if isinstance(texture, RGBTexture):
        return texture
    if isinstance(texture, np.ndarray) and texture.ndim == 3:
        return RGBTexture(image=texture)
    return RGBTexture(color=_parse_color(texture))


def _native(x):
    return x._obj if hasattr(x, "_obj") else x


# ---------------------------------------------------------------- integrator and scene
class PathTracer(object):
    def __init__(self, device="cpu") -> None:
        self._obj = _C.PathTracer(device)

    def render(self, scene, num_samples, **options):
        """The rendered image (``_C.Image``: ``numpy()`` for display RGB).
        `options` (max_depth, rr_depth, seed, hdr, fp64, ...) are this
        package's additions; without them the reference's defaults hold."""
        return self._obj.render(scene._obj, num_samples, **options)


class Scene(object):
    def __init__(self, camera, objects=(), lights=None, background=None, obj=None) -> None:
        if obj is not None:
            self._obj = obj
            return
        if background is None:
            background = EnvTexture(texture=(0, 0, 0))
        self._obj = _C.Scene(camera._obj, [_native(o) for o in objects], [_native(l) for l in (lights or [])],
                             _native(background))

    def to(self, device):
        return Scene(camera=None, obj=self._obj.to(device))

    @property
    def device(self):
        return self._obj.device


class Camera(object):
    def __init__(self, location, rotation, resolution, fov, aperture=None, focus_dist=None) -> None:
        self._obj = _C.Camera()
        self._obj.location = _parse_vec3f(location)
        self._obj.rotation = _rotation(rotation)
        self._obj.width, self._obj.height = resolution
        self._obj.fov = fov
        if aperture is not None:
            self._obj.aperture = float(aperture)
        if focus_dist is not None:
            self._obj.focus_dist = float(focus_dist)

    @property
    def location(self):
        return self._obj.location

    @property
    def rotation(self):
        return self._obj.rotation

    @property
    def width(self):
        return self._obj.width

    @property
    def height(self):
        return self._obj.height


# ---------------------------------------------------------------- objects
class Object3d(object):
    pass


class Sphere(Object3d):
    def __init__(self, radius, location, surface=None, emission=None, alpha=None, bump=None) -> None:
        self._obj = _C.Sphere()
        self._obj.radius = float(radius)
        self._obj.location = _parse_vec3f(location)
        self._obj.material = Material(surface, emission, alpha, bump)._obj

    @property
    def radius(self):
        return self._obj.radius

    @property
    def location(self):
        return self._obj.location


class Rectangle(Object3d):
    def __init__(self, location=None, rotation=None, dimension=None, surface=None, emission=None, alpha=None,
                 bump=None) -> None:
        self._obj = _C.Rectangle()
        if location is not None:
            self._obj.location = _parse_vec3f(location)
        if rotation is not None:
            self._obj.rotation = _rotation(rotation)
        if dimension is not None:
            self._obj.dimension = _parse_vec2f(dimension)
        self._obj.material = Material(surface, emission, alpha, bump)._obj

    @property
    def location(self):
        return self._obj.location

    @property
    def rotation(self):
        return self._obj.rotation

    @property
    def dimension(self):
        return self._obj.dimension

    @property
    def material(self):
        return self._obj.material


class Circle(Object3d):
    def __init__(self, radius=1, location=None, rotation=None, surface=None, emission=None, alpha=None,
                 bump=None) -> None:
        self._obj = _C.Circle()
        self._obj.radius = radius
        self._obj.location = _parse_vec3f((0, 0, 0) if location is None else location)
        if rotation is not None:
            self._obj.rotation = _rotation(rotation)
        self._obj.material = Material(surface, emission, alpha, bump)._obj


class TriangleMesh(Object3d):
    @staticmethod
    def from_obj(path):
        """Meshes and materials of a Wavefront OBJ / MTL file (the MTL decides
        each face's material)."""
        mesh = TriangleMesh()
        mesh._obj = _C.TriangleMesh.from_obj(path, None)
        return mesh


# ---------------------------------------------------------------- surfaces
class BXDF(object):
    pass


class Lambertian(BXDF):
    def __init__(self, texture=None) -> None:
        self._obj = _C.Lambertian(_parse_rgb_texture(texture)._obj)


class Metal(BXDF):
    def __init__(self, texture=None) -> None:
        self._obj = _C.Metal(_parse_rgb_texture(texture)._obj)


class Dielectric(BXDF):
    def __init__(self, texture=None, ior=1.5) -> None:
        self._obj = _C.Dielectric(_parse_rgb_texture(texture)._obj, ior)


class Glossy(BXDF):
    """Microfacet reflection of the given shininess: ``model`` "blinn_phong"
    (the reference's default) or "ggx" (Trowbridge-Reitz)."""

    _MODELS = {"blinn_phong": _C.Glossy.MICRO_BLINN_PHONG, "ggx": _C.Glossy.MICRO_TROWBRIDGE_REITZ}

    def __init__(self, shininess, texture=None, model="blinn_phong") -> None:
        if model not in self._MODELS:
            raise RuntimeError(f"Unknown microfacet model {model!r}: expected one of {sorted(self._MODELS)}")
        self._obj = _C.Glossy(float(shininess), _parse_rgb_texture(texture)._obj, self._MODELS[model])


class Add(BXDF):
    """Sum of two surfaces (the reference's AddBXDF)."""

    def __init__(self, first, second) -> None:
        self._obj = _C.AddBXDF(first._obj, second._obj)


class Mix(BXDF):
    """Weighted choice of two surfaces (the reference's MixBXDF)."""

    def __init__(self, first, second, weight=0.5) -> None:
        self._obj = _C.MixBXDF(first._obj, second._obj, float(weight))


# ---------------------------------------------------------------- textures and materials
class RGBTexture(object):
    def __init__(self, color=None, image=None) -> None:
        self._obj = _C.RGBTexture()
        if image is not None:
            if color is not None:
                warnings.warn("both color and image are specified, using image as texture")
            self._obj.map = _C.Image(image)
        elif color is not None:
            self._obj.value = _parse_color(color)


class AlphaTexture(object):
    """Opacity: a constant `value`, or channel `channel` of `image` (an
    H x W or H x W x C array of opacities in [0, 1], row 0 at the top)."""

    def __init__(self, value=None, image=None, channel=0) -> None:
        if image is not None:
            self._obj = _C.AlphaTexture(_C.Image.from_linear(np.asarray(image, dtype=np.float32), True), channel)
        elif value is not None:
            self._obj = _C.AlphaTexture(float(value))
        else:
            self._obj = _C.AlphaTexture()


class BumpTexture(object):
    """Bump map from a height map (H x W float array, row 0 at the top),
    its gradient scaled by `scale`."""

    def __init__(self, height_map=None, scale=64.0) -> None:
        if height_map is None:
            self._obj = _C.BumpTexture()
        else:
            hm = np.asarray(height_map, dtype=np.float32)
            if hm.ndim != 2:
                raise RuntimeError(f"Invalid height map of shape {hm.shape}")
            self._obj = _C.BumpTexture(_C.Image.from_linear(hm[:, :, None], True), float(scale))


class EnvTexture(object):
    def __init__(self, texture, rotation=None, importance=None) -> None:
        texture = _parse_rgb_texture(texture)
        self._obj = _C.EnvTexture(texture._obj)
        if rotation is not None:
            self._obj.rotation = _rotation(rotation)
        if importance is not None:
            self._obj.importance = bool(importance)


class Material(object):
    def __init__(self, surface=None, emission=None, alpha=None, bump=None) -> None:
        if surface is None:
            surface = Lambertian(texture=(0, 0, 0))
        self._obj = _C.Material(surface._obj)
        if emission is not None:
            self._obj.emission_texture = _parse_rgb_texture(emission)._obj
        if alpha is not None:
            self._obj.alpha_texture = alpha._obj
        if bump is not None:
            self._obj.bump_texture = bump._obj


# ---------------------------------------------------------------- lights
class DistantLight(object):
    def __init__(self, rotation=None, color=(1, 1, 1), power=1, angle=0) -> None:
        self._obj = _C.DistantLight()
        if rotation is not None:
            self._obj.rotation = _rotation(rotation)
        self._obj.color = _parse_vec3f(color)
        self._obj.power = power
        self._obj.angle = angle


class PointLight(object):
    def __init__(self, location, color=(1, 1, 1), power=1, radius=0) -> None:
        self._obj = _C.PointLight()
        self._obj.location = _parse_vec3f(location)
        self._obj.color = _parse_vec3f(color)
        self._obj.power = power
        self._obj.radius = radius


class ObjectLight(object):
    """An emissive object added to the scene's light mixture (it must also be
    among the scene's objects to be seen)."""

    def __init__(self, obj) -> None:
        self._obj = _C.ObjectLight(_native(obj))


class EnvLight(object):
    def __init__(self, env=None) -> None:
        self._obj = _C.EnvLight(_native(env)) if env is not None else _C.EnvLight()
