"""Renders from several host threads at once (render releases the GIL): each
thread with its own PathTracer and scene handle gets exactly the image of the
same call made alone, for plain, batched and cached renders."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_concurrent_renders_equal_sequential(hip):
    from tinypt_amd import scenes

    room = scenes.triangle_room(96, 64).to("hip")
    spheres = scenes.cornell_spheres(80, 60).to("hip")
    cached = scenes.triangle_room(64, 48, seed=3, tris=20000).to("hip")
    rooms = [scenes.triangle_room(48, 32, seed=s, tris=5000).to("hip") for s in (5, 6)]
    kw = dict(max_depth=6, hdr=True)

    def plain(dev, seed):
        return hip.PathTracer("hip").render(dev, 16, seed=seed, **kw).raw()

    def batch(seed):
        return [im.raw() for im in hip.PathTracer("hip").render_batch(rooms, [8, 12], [seed, seed + 1], **kw)]

    def cache(seed):  # learning rate 0: fixed weights, a deterministic image
        img, _ = hip.render_cached(cached, 8, cache_depth=2, learning_rate=0.0, reset=True, weight_seed=5, seed=seed,
                                   max_depth=6)
        return img

    jobs = {"room": lambda k: plain(room, 1 + k), "spheres": lambda k: plain(spheres, 11 + k),
            "batch": lambda k: batch(21 + k), "cache": lambda k: cache(31 + k)}
    reps = 3
    alone = {(name, k): fn(k) for name, fn in jobs.items() for k in range(reps)}
    together, errors = {}, []

    def run(name, fn):
        try:
            for k in range(reps):
                together[(name, k)] = fn(k)
        except Exception as e:  # reported below
            errors.append(f"{name}: {e!r}")

    threads = [threading.Thread(target=run, args=(n, f)) for n, f in jobs.items()]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for key, img in alone.items():
        got = together[key]
        if isinstance(img, list):
            assert all(np.array_equal(a, b) for a, b in zip(img, got)), key
        else:
            assert np.array_equal(img, got), key
