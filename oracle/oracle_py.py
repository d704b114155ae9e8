"""TEST INFRASTRUCTURE ONLY — ctypes binding of the CPU parity oracle
(oracle/_build/libtpt_oracle.so).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtpt_oracle.so")
_lib = None


class RenderParams(ctypes.Structure):
    """tpt_render_params (include/tinypt_hip.h)."""

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


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_rng_run.argtypes = [u64, vp, i32, vp]
        L.oracle_rng_state.argtypes = [u64, vp]
        L.oracle_stream_seed.argtypes = [u64, u64, u64]
        L.oracle_stream_seed.restype = u64
        L.oracle_scene_create.argtypes = [vp]
        L.oracle_scene_create.restype = vp
        L.oracle_scene_destroy.argtypes = [vp]
        L.oracle_object_lights.argtypes = [vp, vp, vp, i32]
        L.oracle_trace_closest.argtypes = [vp, vp, u64, vp, vp, ctypes.c_int, ctypes.c_int]
        L.oracle_trace_occluded.argtypes = [vp, vp, u64, vp, ctypes.c_int]
        L.oracle_intersect_surface.argtypes = [vp, vp, u64, i32, vp]
        L.oracle_intersect_surface.restype = i32
        L.oracle_intersect_surface_batch.argtypes = [vp, vp, u64, vp, vp, ctypes.c_int]
        L.oracle_render.argtypes = [vp, ctypes.POINTER(RenderParams), vp, i32, i32, ctypes.c_int]
        L.oracle_radiance_sample.argtypes = [vp, ctypes.POINTER(RenderParams), i32, i32, i32, vp]
        L.oracle_bvh_octant_golden.argtypes = [ctypes.c_int, vp, vp, vp]
        L.oracle_distribution1d.argtypes = [vp, i32, ctypes.c_float, vp, vp]
        L.oracle_distribution2d.argtypes = [vp, i32, i32, ctypes.c_float, ctypes.c_float, vp, vp]
        L.oracle_env_uv2dir.argtypes = [ctypes.c_float, ctypes.c_float, vp]
        L.oracle_env_dir2uv.argtypes = [vp, vp]
        L.oracle_sobel.argtypes = [vp, i32, i32, ctypes.c_float, vp]
        L.oracle_rotation_from_direction.argtypes = [vp, vp]
        L.oracle_rotation_from_euler.argtypes = [vp, vp, i32, vp]
        L.oracle_delta_light_sample.argtypes = [vp, i32, vp, u64, vp]
        L.oracle_light_sample.argtypes = [vp, i32, vp, u64, vp, vp]
        L.oracle_bxdf_eval.argtypes = [vp, i32, vp, vp, i32, vp, vp, vp]
        L.oracle_bxdf_sample.argtypes = [vp, i32, vp, vp, i32, u64, vp]
        L.oracle_micro_terms.argtypes = [i32, ctypes.c_float, vp, vp, vp, vp, vp, vp]
        L.oracle_micro_sample_h.argtypes = [i32, ctypes.c_float, vp, u64, i32, vp]
        L.oracle_camera_ray.argtypes = [vp, ctypes.c_float, ctypes.c_float, u64, vp, vp]
        _lib = L
    return _lib


def params(spp, max_depth=5, rr_depth=-1, shadow_bounded=False, seed=0, hdr=False, fp64=False):
    return RenderParams(spp, max_depth, rr_depth, 1 if shadow_bounded else 0, seed, 1 if hdr else 0, 1 if fp64 else 0, 0, 0, 0)


class OracleScene:
    """Oracle view of a host scene (anything with ``_desc_ptr()``)."""

    def __init__(self, scene):
        self._keep = scene
        self.h = lib().oracle_scene_create(scene._desc_ptr())
        if not self.h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.width = scene_width(scene)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    def object_lights(self):
        kinds = np.zeros(65536, dtype=np.int32)
        ids = np.zeros(65536, dtype=np.int32)
        n = lib().oracle_object_lights(self.h, kinds.ctypes.data, ids.ctypes.data, 65536)
        return list(zip(kinds[:n].tolist(), ids[:n].tolist()))

    def trace_closest(self, rays, mode=0, threads=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        prim = np.empty(n, dtype=np.int32)
        t = np.empty(n, dtype=np.float32)
        lib().oracle_trace_closest(self.h, rays.ctypes.data, n, prim.ctypes.data, t.ctypes.data, mode, threads)
        return prim, t

    def trace_occluded(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.empty(rays.shape[0], dtype=np.uint8)
        lib().oracle_trace_occluded(self.h, rays.ctypes.data, rays.shape[0], occ.ctypes.data, threads)
        return occ

    def intersect_surface(self, ray8, key=0, depth=0):
        r = np.ascontiguousarray(ray8, dtype=np.float32)
        out = np.zeros(10, dtype=np.float32)
        prim = lib().oracle_intersect_surface(self.h, r.ctypes.data, key, depth, out.ctypes.data)
        return prim, out

    def intersect_surface_batch(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        prim = np.empty(n, dtype=np.int32)
        out = np.empty((n, 10), dtype=np.float32)
        lib().oracle_intersect_surface_batch(self.h, rays.ctypes.data, n, prim.ctypes.data, out.ctypes.data, threads)
        return prim, out

    def render(self, height, p, y0=0, y1=None, threads=0):
        y1 = height if y1 is None else y1
        out = np.zeros((height, self.width, 3), dtype=np.float32)
        lib().oracle_render(self.h, ctypes.byref(p), out.ctypes.data, y0, y1, threads)
        return out

    def radiance_sample(self, p, x, y, s):
        out = np.zeros(3, dtype=np.float32)
        lib().oracle_radiance_sample(self.h, ctypes.byref(p), x, y, s, out.ctypes.data)
        return out


def scene_width(scene):
    # tpt_scene_desc begins with tpt_camera: center[3], rot[9], width, height
    buf = (ctypes.c_int32 * 14).from_address(scene._desc_ptr())
    return int(buf[12])


def scene_height(scene):
    buf = (ctypes.c_int32 * 14).from_address(scene._desc_ptr())
    return int(buf[13])


BLINN, GGX = 3, 4  # TPT_BXDF_GLOSSY_BLINN, TPT_BXDF_GLOSSY_GGX


def micro_terms(kind, shininess, n, l, v, h):
    """(D(n, h), G(l, n, v)) of a Glossy lobe's microfacet model."""
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (n, l, v, h)]
    D, G = ctypes.c_float(), ctypes.c_float()
    rc = lib().oracle_micro_terms(kind, shininess, *[x.ctypes.data for x in a], ctypes.byref(D), ctypes.byref(G))
    assert rc == 0
    return D.value, G.value


def micro_sample_h(kind, shininess, n, seed, count):
    """(count, 3) microfacet normals from one RandEngine(seed) stream."""
    n = np.ascontiguousarray(n, dtype=np.float32)
    out = np.empty((count, 3), dtype=np.float32)
    assert lib().oracle_micro_sample_h(kind, shininess, n.ctypes.data, seed, count, out.ctypes.data) == 0
    return out


def camera_ray(oracle_scene, x, y, seed=0):
    """(origin, direction) of Camera::shoot_ray(x, y) with RandEngine(seed) for the lens."""
    o = np.empty(3, dtype=np.float32)
    d = np.empty(3, dtype=np.float32)
    lib().oracle_camera_ray(oracle_scene.h, x, y, seed, o.ctypes.data, d.ctypes.data)
    return o, d
