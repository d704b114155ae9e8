"""tinypt_amd.wrapper, the reference's Python front end (li-plus/tinypt
python/tinypt/wrapper.py:51-270) over this package: colour and vector parsing,
the defaults it fills in, its errors, and scenes that flatten exactly like the
same scene built on `_C` (checked by rendering both with the oracle)."""
import numpy as np
import pytest
from PIL import ImageColor

import oracle_py as O
import tinypt_amd.wrapper as tinypt
import wrapper_scenes as W


@pytest.mark.parametrize("spec", ["#5689BE", "#ffffff", "#000", "#0a0B0c", "#12345678"])
def test_hex_channels_match_pil(spec):
    # the reference converts '#...' through PIL.ImageColor.getrgb (wrapper.py:11)
    assert tinypt._hex_rgb(spec) == ImageColor.getrgb(spec)


def test_colors_are_linearised_with_gamma_22():
    c = tinypt._parse_color("#5689BE")
    want = np.array([(0x56 / 255) ** 2.2, (0x89 / 255) ** 2.2, (0xBE / 255) ** 2.2], dtype=np.float32)
    assert c.dtype == np.float32 and np.array_equal(c, want)
    assert np.array_equal(tinypt._parse_color((0.5, 1, 2)), np.array([0.5, 1, 2], dtype=np.float32))
    with pytest.raises(RuntimeError, match="not implemented"):
        tinypt._parse_color("red")
    with pytest.raises(RuntimeError, match="Invalid Vec3f"):
        tinypt._parse_color("#12345678")  # four channels
    with pytest.raises(RuntimeError, match="Invalid Vec3f"):
        tinypt.Sphere(radius=1, location=(0, 0))
    with pytest.raises(RuntimeError, match="Invalid Vec2f"):
        tinypt.Rectangle(dimension=(1, 2, 3))


def test_defaults_filled_in_like_the_reference():
    m = tinypt.Material()
    assert list(m._obj.surface.texture.value) == [0, 0, 0]  # black Lambertian
    assert list(tinypt.Circle()._obj.location) == [0, 0, 0] and tinypt.Circle()._obj.radius == 1
    tex = tinypt.Lambertian()._obj.texture  # no texture: the default RGBTexture
    assert list(tex.value) == list(tinypt.RGBTexture()._obj.value)
    img = np.random.default_rng(0).random((4, 6, 3), dtype=np.float32)
    mapped = tinypt.Lambertian(texture=img)._obj.texture
    assert mapped.map is not None and mapped.map.width == 6 and mapped.map.height == 4
    with pytest.warns(UserWarning, match="using image as texture"):
        t = tinypt.RGBTexture(color=(1, 0, 0), image=img)
    assert t._obj.map.width == 6
    d = tinypt.DistantLight()
    assert list(d._obj.color) == [1, 1, 1] and d._obj.power == 1
    with pytest.raises(RuntimeError, match="microfacet model"):
        tinypt.Glossy(10, model="phong")


def test_devices():
    with pytest.raises(RuntimeError, match="device hip"):
        tinypt.PathTracer()  # the reference's default device is its CPU renderer
    with pytest.raises(ValueError, match="Unknown device"):
        tinypt.PathTracer("tpu")
    s = W.cornell_wrapper(16, 12)
    assert s.device == "cpu"


def test_wrapper_scene_flattens_like_native_scene():
    # the wrapper's scene and the same scene spelled out on _C render identically
    # in the oracle (same flattened description: objects, materials, defaults)
    a = O.OracleScene(W.cornell_wrapper(32, 24)._obj)
    b = O.OracleScene(W.cornell_native(32, 24))
    p = O.params(4, seed=5, max_depth=5)
    ia, ib = a.render(24, p), b.render(24, p)
    assert ia.mean() > 0.01 and np.array_equal(ia, ib)


def test_extras_build_and_render_in_the_oracle():
    s = W.cornell_wrapper(32, 24, extras=True)
    img = O.OracleScene(s._obj).render(24, O.params(2, seed=1, max_depth=4, hdr=True))
    assert np.all(np.isfinite(img)) and img.mean() > 0.0


def test_mesh_from_obj_through_the_wrapper(tmp_path):
    # TriangleMesh.from_obj through the wrapper (wrapper.py:200-205) loads the file as _C does
    from tinypt_amd import _C

    (tmp_path / "quad.mtl").write_text("newmtl red\nKd 0.8 0.1 0.1\n")
    (tmp_path / "quad.obj").write_text("mtllib quad.mtl\nv 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nusemtl red\nf 1 2 3 4\n")
    path = str(tmp_path / "quad.obj")
    mesh = tinypt.TriangleMesh.from_obj(path)
    assert mesh._obj.num_faces == 2  # the quad split in two
    cam = tinypt.Camera(location=(0.5, 3, 0.5), rotation=W.rot_x(90) @ W.rot_y(180), resolution=(8, 6), fov=0.8)
    light = dict(radius=0.4, location=(0.5, 1.5, 1.2), emission=(5, 5, 5))
    a = tinypt.Scene(camera=cam, objects=[mesh, tinypt.Sphere(**light)])
    b = _C.Scene(cam._obj, [_C.TriangleMesh.from_obj(path, None), tinypt.Sphere(**light)._obj], [],
                 _C.EnvTexture(_C.RGBTexture([0.0, 0.0, 0.0])))
    p = O.params(4, seed=2, max_depth=3, hdr=True)
    ia, ib = O.OracleScene(a._obj).render(6, p), O.OracleScene(b).render(6, p)
    assert ia.max() > 0 and np.array_equal(ia, ib)
