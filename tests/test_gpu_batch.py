"""Batch mode: many independent scenes through one captured pipeline wave,
replayed on up to three path pools (one per stream of the first scene).  Paths
run scene-major and sample-major, so every pixel's samples are accumulated in
sample order and each scene's image must equal its single-scene render bit for
bit, whatever the queue capacity (pipeline waves that split a scene, or hold
several scenes), the workgroup size and the number of pools."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(hip):
    import zoo
    from tinypt_amd import scenes

    hosts = [
        scenes.cornell_spheres(64, 48),
        scenes.triangle_room(48, 40, seed=3, tris=5000),
        zoo.feature_zoo(40, 32, importance=True),  # textures, alpha, bump, delta + env lights
        scenes.triangle_room(32, 32, seed=4, tris=2000),
        scenes.triangle_room(24, 56, seed=5, tris=12000),
    ]
    devs = [h.to("hip") for h in hosts]
    return hosts, devs, [16, 4, 8, 3, 5], [1, 2, 3, 4, 5]


def singles(hip, devs, spps, seeds, **kw):
    pt = hip.PathTracer("hip")
    return [pt.render(d, n, seed=s, **kw).raw() for d, n, s in zip(devs, spps, seeds)]


@pytest.mark.parametrize("cap,block", [(4096, 64), (9000, 128), (1 << 18, 256)])
def test_batch_equals_single_renders(hip, batch, cap, block):
    hosts, devs, spps, seeds = batch
    kw = dict(max_depth=6, rr_depth=3, hdr=True)
    ref = singles(hip, devs, spps, seeds, **kw)
    pt = hip.PathTracer("hip")
    imgs = pt.render_batch(devs, spps, seeds, queue_capacity=cap, block_size=block, **kw)
    st = pt.last_batch_stats()
    print(f"cap {cap} block {block}: {st['waves']} waves, {st['paths']} paths, {st['render_ms']:.1f} ms")
    assert st["block_size"] == block
    assert st["waves"] == -(-st["paths"] // st["queue_capacity"])
    assert len(st["scene_ms"]) == len(devs) and all(t > 0 for t in st["scene_ms"])
    for i, (a, b) in enumerate(zip(imgs, ref)):
        assert a.raw().shape == b.shape
        assert np.array_equal(a.raw(), b), f"scene {i} differs"


def test_batch_ldr_fp64_and_autotune(hip, batch):
    hosts, devs, spps, seeds = batch
    kw = dict(max_depth=5, fp64=True)
    ref = singles(hip, devs, spps, seeds, **kw)
    pt = hip.PathTracer("hip")
    imgs = pt.render_batch(devs, spps, seeds, **kw)  # capacity and block size autotuned
    st = pt.last_batch_stats()
    print(f"tuned: cap {st['queue_capacity']} block {st['block_size']} in {st['tune_ms']:.1f} ms")
    assert st["block_size"] in (64, 128, 256, 512)
    assert 0 < st["queue_capacity"] <= 1 << 22
    for i, (a, b) in enumerate(zip(imgs, ref)):
        assert np.array_equal(a.raw(), b), f"scene {i} differs"


def test_batch_rejects_bad_input(hip, batch):
    hosts, devs, spps, seeds = batch
    pt = hip.PathTracer("hip")
    with pytest.raises(ValueError):
        pt.render_batch(devs, spps[:-1], seeds)
    with pytest.raises(ValueError):
        pt.render_batch(devs[:1], [0], [1])
    with pytest.raises(RuntimeError, match="Expect scene on hip"):
        pt.render_batch([hosts[0]], [4], [1])


@pytest.mark.parametrize("streams", ["1", "2"])
def test_batch_pools_equal_single_renders(hip, batch, monkeypatch, streams):
    # the batch takes its pool count from the first scene's streams (TPT_STREAMS at
    # scene creation): one and two pools give the single renders' images too
    hosts, _, spps, seeds = batch
    monkeypatch.setenv("TPT_STREAMS", streams)
    devs = [h.to("hip") for h in hosts]
    kw = dict(max_depth=6, rr_depth=3, hdr=True)
    ref = singles(hip, devs, spps, seeds, **kw)
    imgs = hip.PathTracer("hip").render_batch(devs, spps, seeds, queue_capacity=4096, block_size=64, **kw)
    for i, (a, b) in enumerate(zip(imgs, ref)):
        assert np.array_equal(a.raw(), b), f"scene {i} differs"


def test_batch_pools_reused_across_calls(hip, batch):
    # the path pools stay on the first (fewest-paths) scene's handle: a smaller batch
    # after a larger one, then the larger again, reuse them and still give the
    # single renders' images
    hosts, devs, spps, seeds = batch
    kw = dict(max_depth=5, rr_depth=3, hdr=True)
    ref = singles(hip, devs, spps, seeds, **kw)
    pt = hip.PathTracer("hip")
    order = [(list(range(5)), 1 << 16), ([3, 0], 1 << 12), (list(range(5)), 1 << 14)]
    for idx, cap in order:
        imgs = pt.render_batch([devs[i] for i in idx], [spps[i] for i in idx], [seeds[i] for i in idx],
                               queue_capacity=cap, block_size=128, **kw)
        for i, img in zip(idx, imgs):
            assert np.array_equal(img.raw(), ref[i]), f"scene {i} differs (batch {idx}, capacity {cap})"


@pytest.mark.parametrize("large_log2,streams", [(13, "3"), (13, "1"), (10, "3"), (13, "2")])
def test_batch_large_scenes_as_plain_waves(hip, batch, monkeypatch, large_log2, streams):
    # scenes of at least 2^TPT_BATCH_LARGE paths run as plain pipeline waves of their
    # own after the batched ones (2^13: the spheres box and the feature zoo; 2^10:
    # every scene, no batched waves); the images still equal the single renders
    hosts, _, spps, seeds = batch
    monkeypatch.setenv("TPT_STREAMS", streams)
    devs = [h.to("hip") for h in hosts]
    monkeypatch.setenv("TPT_BATCH_LARGE", str(large_log2))
    kw = dict(max_depth=6, rr_depth=3, hdr=True)
    ref = singles(hip, devs, spps, seeds, **kw)
    pt = hip.PathTracer("hip")
    for cap, block in ((4096, 64), (0, 0)):
        imgs = pt.render_batch(devs, spps, seeds, queue_capacity=cap, block_size=block, **kw)
        st = pt.last_batch_stats()
        assert len(st["scene_ms"]) == len(devs) and all(t > 0 for t in st["scene_ms"])
        for i, (a, b) in enumerate(zip(imgs, ref)):
            assert np.array_equal(a.raw(), b), f"scene {i} differs (2^{large_log2}, capacity {cap})"


@pytest.mark.parametrize("graph", ["1", "0"])
def test_batch_one_graph_or_wave_replays(hip, batch, monkeypatch, graph):
    # TPT_BATCH_GRAPH=1 captures the whole batch (batched waves, large scenes'
    # plain waves, finalize) into one hipGraph launched once; 0 (the default)
    # replays one captured wave per pool from the host; same images either way
    hosts, devs, spps, seeds = batch
    monkeypatch.setenv("TPT_BATCH_GRAPH", graph)
    monkeypatch.setenv("TPT_BATCH_LARGE", "13")
    kw = dict(max_depth=6, rr_depth=3, hdr=True)
    ref = singles(hip, devs, spps, seeds, **kw)
    pt = hip.PathTracer("hip")
    for _ in range(2):  # twice: nothing of one call's graph outlives it
        imgs = pt.render_batch(devs, spps, seeds, queue_capacity=4096, block_size=128, **kw)
        st = pt.last_batch_stats()
        assert st["graphs"] == int(graph)
        times = st["scene_ms"]
        assert len(times) == len(devs) and all(t > 0 for t in times)
        for i, (a, b) in enumerate(zip(imgs, ref)):
            assert np.array_equal(a.raw(), b), f"scene {i} differs (one graph {graph})"
