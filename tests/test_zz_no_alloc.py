"""tpt_hip_render_noalloc (the render entry point that allocates nothing: its path
pool lives in a fixed device array of the library) against tpt_hip_render: the
same image bit for bit, on consecutive calls that reuse the array (the second
with a smaller scene, then with several pipeline waves and FP64 sums), and its
size limits rejected with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs to own device memory and a stream."""

    def __init__(self):
        L = ctypes.CDLL("libamdhip64.so")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipStreamDestroy.argtypes = [vp]
        self.L = L

    def malloc(self, n):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), n) == 0
        return p.value

    def to_host(self, ptr, shape):
        out = np.empty(shape, dtype=np.float32)
        assert self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out


def test_noalloc_render_equals_render(hip):
    from tinypt_amd import scenes

    rt = _Hip()
    stream = ctypes.c_void_p()
    assert rt.L.hipStreamCreate(ctypes.byref(stream)) == 0
    buf = rt.malloc(96 * 64 * 3 * 4)
    try:
        cases = [
            # (scene, spp, render options, extra noalloc options)
            (scenes.triangle_room(96, 64, seed=3, tris=20000), 16, dict(max_depth=8, rr_depth=-1, hdr=True), {}),
            (scenes.cornell_spheres(48, 40), 8, dict(max_depth=10, rr_depth=5), {}),  # smaller, LDR
            # several pipeline waves of 3 samples per pixel, FP64 sums, b = 2 sub-pixels
            (scenes.triangle_room(64, 48, seed=4, tris=5000), 64, dict(max_depth=6, rr_depth=3, fp64=True),
             dict(queue_capacity=64 * 48 * 3)),
        ]
        pt = hip.PathTracer("hip")
        for i, (host, spp, kw, extra) in enumerate(cases):
            dev = host.to("hip")
            ref = pt.render(dev, spp, seed=11 + i, **kw).raw()
            hip.render_noalloc(dev, buf, spp, seed=11 + i, stream=stream.value, **kw, **extra)
            assert rt.L.hipStreamSynchronize(stream) == 0
            got = rt.to_host(buf, ref.shape)
            print(f"case {i}: {ref.shape[1]}x{ref.shape[0]} {spp} spp, mean {got.mean():.5f}")
            assert np.all(np.isfinite(got))
            assert np.array_equal(got, ref), f"case {i}: noalloc image differs from tpt_hip_render"
        # limits: more pixels than the array holds, a larger queue, counting
        big = scenes.cornell_spheres(1024, 768).to("hip")
        with pytest.raises(ValueError, match="exceed its limit"):
            hip.render_noalloc(big, buf, 1, stream=stream.value)
        small = scenes.cornell_spheres(16, 16).to("hip")
        with pytest.raises(ValueError, match="queue_capacity"):
            hip.render_noalloc(small, buf, 1, queue_capacity=(1 << 19) + 1, stream=stream.value)
    finally:
        rt.L.hipStreamSynchronize(stream)
        rt.L.hipFree(buf)
        rt.L.hipStreamDestroy(stream)
