"""Image parity of the HIP renderer against the CPU oracle on the workloads the
benchmark runs, at sizes the oracle renders in seconds:

- the sampler's sub-pixel blocks b = 3 and b = 4 and the integer-division drop
  of samples (cpu/path_tracer.cpp:10-21, 29-43: spp 144 -> b = 3, 16 samples per
  sub-pixel; spp 150 -> b = 3, 150 / 9 = 16 samples per sub-pixel, 144 in all;
  spp 256 -> b = 4) on the triangle room, LDR output (per-sub-pixel clipping);
- the textured room (RGBA8 albedo, alpha cut-out, FP32 bump maps), depth 12;
- the 6M-triangle city (HDR sky sampled through its CDF, sun, 1000 point lights,
  GGX and dielectric buildings), depth 8;
- a batch of independent rooms drawn from the 64-scene batch mix (its triangle
  counts, seeds and samples per pixel; resolutions divided by 8), rendered as one
  batch and compared with the oracle and with single renders.

Bars as in test_gpu_render.py: relMSE(gpu, oracle same seed) <= 0.1 x
relMSE(oracle seed A, oracle seed B), and a stated fraction of pixel channels
within 1e-4.  At many samples per pixel the per-channel fraction says little (a
pixel differs as soon as one of its paths diverged through a libm ulp), so it is
only required at a few samples per pixel; the relMSE bar is what catches a wrong
sampler, since other sub-pixel jitter or another sample count gives an error of
the size of the noise floor.
"""
import numpy as np
import pytest

import oracle_py as O
from test_gpu_render import check, render_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spp,block,eff", [(144, 3, 144), (150, 3, 144), (256, 4, 256)])
def test_sampler_blocks_match_oracle(hip, spp, block, eff):
    from tinypt_amd import scenes

    host = scenes.triangle_room(64, 48)
    g, o, o2 = render_both(hip, host, spp, seed=31, max_depth=8, rr_depth=-1)
    # the effective sample count (b x b sub-pixels of spp / b^2 samples each)
    dev = host.to("hip")
    pt = hip.PathTracer("hip")
    pt.render(dev, spp, seed=31, max_depth=8, rr_depth=-1)
    assert pt.last_stats()["paths"] == 64 * 48 * eff
    assert eff == (spp // (block * block)) * block * block
    check(g, o, o2, 0.0)


def test_textured_room_matches_oracle(hip):
    from tinypt_amd import scenes

    g, o, o2 = render_both(hip, scenes.textured_room(160, 90), 4, seed=41, max_depth=12, rr_depth=-1, hdr=True)
    assert g.mean() > 0.01
    check(g, o, o2, 0.75)


def test_city_matches_oracle(hip):
    from tinypt_amd import scenes

    host = scenes.city(96, 54)  # ~6M triangles, sky CDF, sun, 1000 point lights
    g, o, o2 = render_both(hip, host, 2, seed=43, max_depth=8, rr_depth=-1, hdr=True)
    assert g.mean() > 0.01
    check(g, o, o2, 0.75)


def test_batch_mix_matches_oracle_and_single_renders(hip):
    from tinypt_amd import scenes

    specs = [s for s in scenes.batch_specs(64) if s["tris"] <= 50000][:8]
    hosts = [scenes.triangle_room(max(16, s["width"] // 8), max(16, s["height"] // 8), s["seed"], s["tris"])
             for s in specs]
    spps = [s["spp"] for s in specs]
    seeds = [s["seed"] for s in specs]
    devs = [h.to("hip") for h in hosts]
    kw = dict(max_depth=8, rr_depth=-1)
    pt = hip.PathTracer("hip")
    imgs = pt.render_batch(devs, spps, seeds, **kw)  # capacity and block size autotuned
    st = pt.last_batch_stats()
    print(f"batch: {st['paths']} paths, capacity {st['queue_capacity']}, block {st['block_size']}")
    for i, (h, d, n, sd) in enumerate(zip(hosts, devs, spps, seeds)):
        single = pt.render(d, n, seed=sd, **kw).raw()
        g = imgs[i].raw()
        assert np.array_equal(g, single), f"scene {i}: batch differs from its single render"
        oracle = O.OracleScene(h)
        hh = g.shape[0]
        o = oracle.render(hh, O.params(n, max_depth=8, rr_depth=-1, seed=sd))
        o2 = oracle.render(hh, O.params(n, max_depth=8, rr_depth=-1, seed=sd + 1000))
        print(f"scene {i}: {g.shape[1]}x{g.shape[0]} {n} spp {specs[i]['tris']} tris", end=": ")
        check(g, o, o2, 0.0)
