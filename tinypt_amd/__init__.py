"""tinypt_amd — the MI355X (gfx950) backend of tinypt, `device="hip"`.

The public classes mirror tinypt's pybind module `_C` (li-plus/tinypt
src/tinypt/pybind.cpp:67-181): ``PathTracer(device).render(scene, spp)``,
``Scene(camera, objects, lights, background).to(device)``, ``Camera``,
``Sphere``/``Rectangle``/``Circle``/``TriangleMesh``, ``Lambertian``/``Metal``/
``Dielectric``/``Glossy``/``AddBXDF``/``MixBXDF``, ``Material``, the textures and
the delta lights.  Rendering runs through ``lib/libtinypt_hip.so`` (the C ABI in
``include/tinypt_hip.h``); there is no CPU fallback: importing fails loudly when
the native module is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtinypt_hip.so")

try:
    from . import _C  # noqa: F401
except ImportError as e:  # pragma: no cover - exercised only on a broken install
    raise ImportError(
        "tinypt_amd native module is not built; run __graft_entry__.build() "
        f"(make -C tinypt_amd/csrc && make -C tinypt_amd/host): {e}"
    ) from e

from ._C import (  # noqa: E402,F401
    AddBXDF,
    AlphaTexture,
    BumpTexture,
    BXDF,
    Camera,
    Circle,
    Dielectric,
    DistantLight,
    EnvTexture,
    Glossy,
    Image,
    Lambertian,
    Light,
    Material,
    Metal,
    MixBXDF,
    Object3d,
    PathTracer,
    PointLight,
    Rectangle,
    RGBTexture,
    Scene,
    Sphere,
    TriangleMesh,
    hip_device_count,
    rotation_from_direction,
    rotation_from_euler,
    sobel_gradient,
)

_lib = None


class _RenderStats(ctypes.Structure):
    _fields_ = [
        ("render_ms", ctypes.c_double),
        ("extend_ms", ctypes.c_double),
        ("shade_ms", ctypes.c_double),
        ("shadow_ms", ctypes.c_double),
        ("other_ms", ctypes.c_double),
        ("paths", ctypes.c_uint64),
        ("rays_closest", ctypes.c_uint64),
        ("rays_shadow", ctypes.c_uint64),
        ("extend_launches", ctypes.c_uint64),
        ("nodes_visited", ctypes.c_uint64),
        ("tris_tested", ctypes.c_uint64),
    ]


# every symbol include/tinypt_hip.h declares
C_ABI_SYMBOLS = (
    "tpt_hip_abi_version",
    "tpt_hip_device_count",
    "tpt_hip_last_error",
    "tpt_hip_scene_create",
    "tpt_hip_scene_destroy",
    "tpt_hip_scene_info",
    "tpt_hip_render",
    "tpt_hip_trace_closest",
    "tpt_hip_trace_occluded",
    "tpt_hip_last_trace_counters",
)


def lib():
    """ctypes handle on libtinypt_hip.so with the C ABI prototypes set."""
    global _lib
    if _lib is not None:
        return _lib
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_uint32
    L.tpt_hip_abi_version.restype = ctypes.c_int
    L.tpt_hip_device_count.restype = ctypes.c_int
    L.tpt_hip_last_error.restype = ctypes.c_char_p
    L.tpt_hip_scene_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.tpt_hip_scene_destroy.argtypes = [vp]
    L.tpt_hip_scene_info.argtypes = [vp] + [ctypes.POINTER(u64)] * 5
    L.tpt_hip_render.argtypes = [vp, vp, vp, ctypes.POINTER(_RenderStats)]
    L.tpt_hip_trace_closest.argtypes = [vp, vp, u64, vp, vp, u32, vp]
    L.tpt_hip_trace_occluded.argtypes = [vp, vp, u64, vp, u32, vp]
    L.tpt_hip_last_trace_counters.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    for name in C_ABI_SYMBOLS:
        if name not in ("tpt_hip_abi_version", "tpt_hip_device_count", "tpt_hip_last_error", "tpt_hip_scene_destroy"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def last_error():
    return lib().tpt_hip_last_error().decode()


def _check(rc):
    if rc != 0:
        msg = last_error()
        if rc == -1:
            raise ValueError(msg)
        raise RuntimeError(msg)


def trace_closest(scene_hip, rays, count_traffic=False):
    """Closest hit through the C ABI for an (N, 8) float32 host array of rays
    (org.xyz, tmin, dir.xyz, tmax).  Returns (prim_id int32[N], t float32[N])."""
    import numpy as np

    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    prim = np.empty(n, dtype=np.int32)
    t = np.empty(n, dtype=np.float32)
    flags = 2 if count_traffic else 0
    _check(lib().tpt_hip_trace_closest(scene_hip._hip_ptr(), rays.ctypes.data, n, prim.ctypes.data, t.ctypes.data, flags, None))
    return prim, t


def trace_occluded(scene_hip, rays):
    import numpy as np

    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    occ = np.empty(n, dtype=np.uint8)
    _check(lib().tpt_hip_trace_occluded(scene_hip._hip_ptr(), rays.ctypes.data, n, occ.ctypes.data, 0, None))
    return occ


def last_trace_counters(scene_hip):
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    _check(lib().tpt_hip_last_trace_counters(scene_hip._hip_ptr(), ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def scene_info(scene_hip):
    v = [ctypes.c_uint64() for _ in range(5)]
    _check(lib().tpt_hip_scene_info(scene_hip._hip_ptr(), *[ctypes.byref(x) for x in v]))
    return dict(arena_bytes=v[0].value, nodes=v[1].value, triangles=v[2].value, analytic=v[3].value,
                node_bytes=v[4].value)
