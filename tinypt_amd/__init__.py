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
    EnvLight,
    EnvTexture,
    Glossy,
    Image,
    Lambertian,
    Light,
    Material,
    Metal,
    MixBXDF,
    Object3d,
    ObjectLight,
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


class _RenderParams(ctypes.Structure):
    _fields_ = [
        ("spp", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("rr_depth", ctypes.c_int32),
        ("shadow_bounded", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("output", ctypes.c_int32),
        ("accum_fp64", ctypes.c_int32),
        ("queue_capacity", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("count_traffic", ctypes.c_int32),
    ]


class _CacheParams(ctypes.Structure):
    _fields_ = [
        ("frames", ctypes.c_int32),
        ("cache_depth", ctypes.c_int32),
        ("train_records", ctypes.c_int32),
        ("reset", ctypes.c_int32),
        ("learning_rate", ctypes.c_float),
        ("train_steps", ctypes.c_int32),
        ("weight_seed", ctypes.c_uint64),
        ("suffix_len", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
    ]


class _CacheStats(ctypes.Structure):
    _fields_ = [
        ("render_ms", ctypes.c_double),
        ("infer_ms", ctypes.c_double),
        ("train_ms", ctypes.c_double),
        ("last_loss", ctypes.c_double),
        ("queries", ctypes.c_uint64),
        ("records", ctypes.c_uint64),
        ("steps", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
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
    "tpt_hip_render_noalloc",
    "tpt_hip_trace_closest",
    "tpt_hip_trace_occluded",
    "tpt_hip_last_trace_counters",
    "tpt_hip_render_batch",
    "tpt_hip_render_cached",
    "tpt_hip_cache_weights",
    "tpt_hip_cache_eval",
    "tpt_hip_cache_train",
    "tpt_hip_progress_create",
    "tpt_hip_progress_add",
    "tpt_hip_progress_wait",
    "tpt_hip_progress_read",
    "tpt_hip_progress_destroy",
)


ABI_VERSION = 3  # TPT_ABI_VERSION of include/tinypt_hip.h (the ctypes structures above follow it)


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
    L.tpt_hip_render_noalloc.argtypes = [vp, ctypes.POINTER(_RenderParams), vp, vp]
    L.tpt_hip_trace_closest.argtypes = [vp, vp, u64, vp, vp, u32, vp]
    L.tpt_hip_trace_occluded.argtypes = [vp, vp, u64, vp, u32, vp]
    L.tpt_hip_last_trace_counters.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.tpt_hip_render_batch.argtypes = [ctypes.POINTER(vp), vp, i32, vp, ctypes.POINTER(vp), vp, ctypes.POINTER(ctypes.c_double)]
    L.tpt_hip_render_cached.argtypes = [vp, ctypes.POINTER(_RenderParams), ctypes.POINTER(_CacheParams), vp,
                                        ctypes.POINTER(_CacheStats)]
    L.tpt_hip_cache_weights.argtypes = [vp, vp, vp]
    L.tpt_hip_cache_eval.argtypes = [vp, vp, i32, vp]
    L.tpt_hip_cache_train.argtypes = [vp, vp, vp, i32, ctypes.c_float, vp, vp]
    L.tpt_hip_progress_create.argtypes = [vp, ctypes.POINTER(_RenderParams), ctypes.POINTER(_CacheParams), ctypes.POINTER(vp)]
    L.tpt_hip_progress_add.argtypes = [vp, i32]
    L.tpt_hip_progress_wait.argtypes = [vp]
    L.tpt_hip_progress_read.argtypes = [vp, vp, ctypes.POINTER(i32)]
    L.tpt_hip_progress_destroy.argtypes = [vp]
    L.tpt_hip_progress_destroy.restype = None
    for name in C_ABI_SYMBOLS:
        if name not in ("tpt_hip_abi_version", "tpt_hip_device_count", "tpt_hip_last_error", "tpt_hip_scene_destroy",
                        "tpt_hip_progress_destroy"):
            getattr(L, name).restype = ctypes.c_int
    if L.tpt_hip_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI version {L.tpt_hip_abi_version()}, this package expects "
                           f"{ABI_VERSION}: rebuild the library (make -C tinypt_amd/csrc)")
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


def render_noalloc(scene_hip, out_ptr, spp, max_depth=5, rr_depth=-1, shadow_bounded=False, seed=0, hdr=False,
                   fp64=False, queue_capacity=0, block_size=0, stream=0):
    """tpt_hip_render_noalloc: enqueue a render on `stream` (a hipStream_t as an
    int, 0 = the null stream) into the device buffer at `out_ptr` (H x W x 3
    float32, row 0 at the bottom); allocates nothing and does not synchronise."""
    p = _RenderParams(spp=spp, max_depth=max_depth, rr_depth=rr_depth, shadow_bounded=int(shadow_bounded), seed=seed,
                      output=1 if hdr else 0, accum_fp64=int(fp64), queue_capacity=queue_capacity,
                      block_size=block_size)
    _check(lib().tpt_hip_render_noalloc(scene_hip._hip_ptr(), ctypes.byref(p), ctypes.c_void_p(out_ptr),
                                        ctypes.c_void_p(stream)))


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


# ---------------------------------------------------------------- learned radiance cache
def render_cached(scene_hip, frames, cache_depth=2, train_records=1 << 16, learning_rate=1e-3, reset=False,
                  weight_seed=0, max_depth=5, rr_depth=-1, shadow_bounded=False, seed=0, fp64=False, train_steps=8,
                  suffix_len=1):
    """Progressive render with the learned radiance cache (cache_depth < 0: plain
    progressive frames).  suffix_len > 0 ends training paths in the cache too,
    suffix_len vertices past their cache vertex (self-training); 0 traces them
    to max_depth.  train_steps: Adam steps per frame, on consecutive
    mini-batches of the frame's train_records records (textured room 1920x1080,
    256 frames: 4 / 8 / 16 steps give equal-error speedups 2.30 / 2.42 / 2.25).  Returns (H x W x 3 float32 HDR
    mean, row 0 at the bottom, stats dict)."""
    import numpy as np

    cam = scene_hip.camera()
    out = np.empty((cam.height, cam.width, 3), dtype=np.float32)
    p = _RenderParams(spp=1, max_depth=max_depth, rr_depth=rr_depth, shadow_bounded=int(shadow_bounded), seed=seed,
                      output=1, accum_fp64=int(fp64))
    c = _CacheParams(frames=frames, cache_depth=cache_depth, train_records=train_records, reset=int(reset),
                     learning_rate=learning_rate, train_steps=train_steps, weight_seed=weight_seed,
                     suffix_len=suffix_len)
    st = _CacheStats()
    _check(lib().tpt_hip_render_cached(scene_hip._hip_ptr(), ctypes.byref(p), ctypes.byref(c), out.ctypes.data,
                                       ctypes.byref(st)))
    return out, {f: getattr(st, f) for f, _ in _CacheStats._fields_ if f != "pad_"}


class ProgressiveRender:
    """Progressive render session (tpt_hip_progress_*): frames of one sample per
    pixel accumulate on the device across calls; add() enqueues and returns at
    once, read() waits and returns (H x W x 3 float32 HDR mean, frames so far).
    cache_depth >= 0 enables the learned radiance cache (arguments as in
    render_cached).  Keep the scene alive while the session is."""

    def __init__(self, scene_hip, max_depth=5, rr_depth=-1, shadow_bounded=False, seed=0, fp64=False,
                 cache_depth=-1, train_records=1 << 16, learning_rate=1e-3, reset=False, weight_seed=0,
                 train_steps=8, suffix_len=1):
        self._scene = scene_hip
        cam = scene_hip.camera()
        self._shape = (cam.height, cam.width, 3)
        p = _RenderParams(spp=1, max_depth=max_depth, rr_depth=rr_depth, shadow_bounded=int(shadow_bounded),
                          seed=seed, output=1, accum_fp64=int(fp64))
        c = None
        if cache_depth >= 0:
            c = _CacheParams(frames=1, cache_depth=cache_depth, train_records=train_records, reset=int(reset),
                             learning_rate=learning_rate, train_steps=train_steps, weight_seed=weight_seed,
                             suffix_len=suffix_len)
        h = ctypes.c_void_p()
        _check(lib().tpt_hip_progress_create(scene_hip._hip_ptr(), ctypes.byref(p),
                                             None if c is None else ctypes.byref(c), ctypes.byref(h)))
        self._h = h

    def add(self, frames):
        _check(lib().tpt_hip_progress_add(self._h, int(frames)))
        return self

    def wait(self):
        _check(lib().tpt_hip_progress_wait(self._h))

    def read(self):
        import numpy as np

        out = np.empty(self._shape, dtype=np.float32)
        n = ctypes.c_int32()
        _check(lib().tpt_hip_progress_read(self._h, out.ctypes.data, ctypes.byref(n)))
        return out, n.value

    def close(self):
        if self._h:
            lib().tpt_hip_progress_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def cache_weights(scene_hip, set=None):
    """Returns the cache network's FP32 weights (5, 64, 64) [layer][in][out];
    `set` replaces them first (and resets the Adam state)."""
    import numpy as np

    get = np.empty((5, 64, 64), dtype=np.float32)
    src = None
    if set is not None:
        src = np.ascontiguousarray(set, dtype=np.float32)
        assert src.shape == (5, 64, 64)
    _check(lib().tpt_hip_cache_weights(scene_hip._hip_ptr(), None if src is None else src.ctypes.data, get.ctypes.data))
    return get


def cache_eval(scene_hip, inputs_bf16):
    """Network output (n, 3) for (n, 64) inputs given as BF16 bit patterns (uint16)."""
    import numpy as np

    x = np.ascontiguousarray(inputs_bf16, dtype=np.uint16)
    out = np.empty((x.shape[0], 3), dtype=np.float32)
    _check(lib().tpt_hip_cache_eval(scene_hip._hip_ptr(), x.ctypes.data, x.shape[0], out.ctypes.data))
    return out


def cache_train(scene_hip, inputs_bf16, targets, lr):
    """One training step; returns (loss before the step, gradient (5, 64, 64))."""
    import numpy as np

    x = np.ascontiguousarray(inputs_bf16, dtype=np.uint16)
    t = np.ascontiguousarray(targets, dtype=np.float32)
    loss = ctypes.c_float()
    grad = np.empty((5, 64, 64), dtype=np.float32)
    _check(lib().tpt_hip_cache_train(scene_hip._hip_ptr(), x.ctypes.data, t.ctypes.data, x.shape[0], lr,
                                     ctypes.byref(loss), grad.ctypes.data))
    return loss.value, grad
