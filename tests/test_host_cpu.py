"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol include/tinypt_hip.h declares, the Device / Scene / PathTracer facade
raises the reference's exception types for the reference's error conditions
(src/tinypt/defs.h:68-78, scene.h:24-37, path_tracer.h:14-44), and the scene
flattening feeds the oracle the reference's object-light order."""
import ctypes
import os
import re

import numpy as np
import pytest

import tinypt_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "tinypt_hip.h")).read()
    return sorted(set(re.findall(r"\b(tpt_hip_\w+)\s*\(", src)))


def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(T.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(T.C_ABI_SYMBOLS)
    assert T.lib().tpt_hip_abi_version() == T.ABI_VERSION == 3


def test_device_strings():
    with pytest.raises(ValueError, match="Unknown device"):
        T.PathTracer("tpu")
    with pytest.raises(RuntimeError, match="CUDA is not enabled"):
        T.PathTracer("cuda")
    with pytest.raises(RuntimeError, match="reference build"):
        T.PathTracer("cpu")


def small_scene():
    from tinypt_amd import scenes

    return scenes.cornell_spheres(16, 12)


def test_scene_to_errors():
    s = small_scene()
    assert s.device == "cpu"
    assert s.to("cpu").device == "cpu"
    with pytest.raises(RuntimeError, match="CUDA is not enabled"):
        s.to("cuda")
    with pytest.raises(ValueError):
        s.to("metal")


@pytest.mark.skipif(T.hip_device_count() > 0, reason="checks the no-device path")
def test_hip_unavailable_is_loud():
    with pytest.raises(RuntimeError, match="hip is not available"):
        small_scene().to("hip")
    # the raw C ABI reports the same failure without throwing
    h = ctypes.c_void_p()
    keep = small_scene()  # the description lives as long as its Scene
    rc = T.lib().tpt_hip_scene_create(keep._desc_ptr(), ctypes.byref(h))
    assert rc != 0 and not h.value
    assert "hip is not available" in T.last_error()


def test_c_abi_rejects_invalid_input():
    L = T.lib()
    assert L.tpt_hip_scene_create(None, None) == -1
    assert L.tpt_hip_render(None, None, None, None) == -1
    assert "null" in T.last_error()
    assert L.tpt_hip_render_batch(None, None, 0, None, None, None, None) == -1
    assert "null" in T.last_error()


def test_material_and_light_validation():
    lam = T.Lambertian(T.RGBTexture([1, 1, 1]))
    with pytest.raises(ValueError, match="MixBXDF"):
        T.MixBXDF(lam, lam, 1.5)
    dl = T.DistantLight()
    with pytest.raises(ValueError, match="angle"):
        dl.angle = -0.1
    with pytest.raises(ValueError):
        T.TriangleMesh.from_arrays(np.zeros((3, 2), np.float32), np.zeros((1, 3), np.int32))


def test_object_light_order_matches_reference_dfs():
    import oracle_py as O
    from tinypt_amd import scenes

    host = scenes.triangle_room(32, 24)
    lights = O.OracleScene(host).object_lights()
    # four emissive rectangles after the mesh: the DFS stack pops them last-first (cpu/scene.cpp:9-19)
    n = len(lights)
    assert n == 4 and [k for k, _ in lights] == [0] * 4
    ids = [i for _, i in lights]
    assert ids == sorted(ids, reverse=True)


def test_image_conventions():
    # Image(array) applies pow(2.2) and flips rows (pybind.cpp:77-90); numpy() undoes both
    a = np.random.default_rng(0).random((5, 7, 3)).astype(np.float32)
    im = T.Image(a)
    raw = im.raw()
    assert np.allclose(raw[::-1], a ** 2.2, rtol=1e-6)
    assert np.allclose(im.numpy(), a, rtol=1e-5, atol=1e-6)
    lin = T.Image.from_linear(np.arange(24, dtype=np.uint8).reshape(2, 3, 4))
    assert lin.channels == 4 and np.allclose(lin.raw()[1, 2], np.arange(20, 24) / 255.0)


def test_mesh_transform_is_baked():
    import oracle_py as O

    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    m = T.TriangleMesh.from_arrays(v, np.array([[0, 1, 2]], np.int32))
    m.location = [0, 0, 5]
    cam = T.Camera()
    cam.width = cam.height = 4
    o = O.OracleScene(T.Scene(cam, [m], [], None))
    rays = np.array([[0.2, 0.2, 10, 1e-3, 0, 0, -1, 1e20]], dtype=np.float32)
    p, t = o.trace_closest(rays)
    assert p[0] == 0 and t[0] == np.float32(5.0)


def test_object_and_env_lights_join_the_mixture_in_reference_order():
    """cpu/scene.cpp:9-30: emissive objects in DFS order, then the non-delta lights
    passed to Scene() in their order (delta lights go to their own group), then the
    background's EnvLight when it is a map."""
    import oracle_py as O
    import zoo

    scene = zoo.feature_zoo(16, 12, importance=True, object_lights=True)
    base = zoo.feature_zoo(16, 12, importance=True)
    got = O.OracleScene(scene).object_lights()
    dfs = O.OracleScene(base).object_lights()
    assert dfs[-1] == (1, -1) and all(k == 0 for k, _ in dfs[:-1])
    # ObjectLight(sphere) -> its primitive; ObjectLight(mesh) -> kind 2 with the object index
    # (objects: floor, back, 5 spheres, icosphere mesh 7, quad mesh 8, circle 9); EnvLight(bg)
    sphere_gid = 2
    assert got == dfs[:-1] + [(0, sphere_gid), (2, 7), (1, -1), (2, 8), (1, -1)]


def test_object_and_env_light_errors():
    import zoo

    s = T.Sphere()
    outside = T.Sphere()
    cam = T.Camera()
    with pytest.raises(ValueError, match="ObjectLight"):
        T.Scene(cam, [s], [T.ObjectLight(outside)], None)
    env = T.EnvTexture(T.RGBTexture([1.0, 1.0, 1.0]))
    with pytest.raises(ValueError, match="EnvLight"):
        T.Scene(cam, [s], [T.EnvLight(env)], None)
    with pytest.raises(ValueError, match="EnvLight"):
        T.Scene(cam, [s], [T.EnvLight()], env)
    T.Scene(cam, [s], [T.ObjectLight(s), T.EnvLight(env)], env)  # both accepted


def test_mesh_object_light_pdf_is_its_sampling_density():
    """ObjectLight of a two-shape mesh (a 2 x 3 rectangle split into two triangles,
    one per shape): shape-uniform then face-uniform sampling, pdf = mean over
    shapes of the mean face pdf (cpu/object.cpp:352-365).  If the pdf is the
    sampler's density, E[1 / pdf] over its samples is the solid angle the mesh
    subtends: 4 atan(ab / (2h sqrt(4h^2 + a^2 + b^2))) for a centred a x b
    rectangle at distance h."""
    import math

    import oracle_py as O

    a, b, h = 2.0, 3.0, 1.5
    v = np.array([[-a / 2, -b / 2, h], [a / 2, -b / 2, h], [a / 2, b / 2, h], [-a / 2, b / 2, h]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    faces9 = np.repeat(f, 3, axis=1).astype(np.int32)
    faces9[:, [1, 2, 4, 5, 7, 8]] = -1
    m = T.Material(T.Lambertian(T.RGBTexture([0.5, 0.5, 0.5])))
    mesh = T.TriangleMesh.from_arrays(v, faces9, [m], np.zeros(2, np.int32), None, None, [0, 1, 2])
    cam = T.Camera()
    o = O.OracleScene(T.Scene(cam, [mesh], [T.ObjectLight(mesh)], None))
    assert o.object_lights() == [(2, 0)]
    pos = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    p = ctypes.c_float()
    inv = []
    for seed in range(20000):
        assert O.lib().oracle_light_sample(o.h, 0, pos.ctypes.data, seed, d.ctypes.data, ctypes.byref(p)) == 0
        assert p.value > 0
        inv.append(1.0 / p.value)
    want = 4 * math.atan(a * b / (2 * h * math.sqrt(4 * h * h + a * a + b * b)))
    assert abs(np.mean(inv) / want - 1) < 0.02, (np.mean(inv), want)
