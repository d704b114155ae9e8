"""Benchmark: the plain HIP wavefront path tracer on the seeded ~265k-triangle
room (1024x768, 256 spp, max depth 8, no Russian roulette), and the learned
radiance cache against the plain tracer at equal relMSE on the textured room
(1920x1080, depth 12) against a 16,384-spp FP64 render.

One step = one full render of the room (201,326,592 paths) through
PathTracer("hip").render; the scene is resident in HBM before the timed region.
Prints ONE JSON line with Mpaths/s, the extend kernel's roofline (VALU issue:
the kernel's bound; its HBM bytes from counters and its algorithmic bytes beside
it), the CPU oracle timed on a bounded sample, and the radiance-cache rows.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--spp S] [--no-cpu] [--no-cache]

Other workloads (one JSON line each, same fields):
  --workload batch    64 seeded rooms (2k-500k triangles, 128^2-1024^2 pixels,
                      16-256 spp) through PathTracer.render_batch: up to three
                      path pools on as many streams, small scenes in shared
                      captured pipeline waves, large ones in plain waves,
                      autotuned queue capacity and workgroup size; also the
                      same batch as one hipGraph launch (TPT_BATCH_GRAPH=1);
                      per-scene completion times and the same scenes rendered
                      one by one
  --workload spheres  the analytic two-spheres box, 512x384, 64 spp, depth 10, RR 5
  --workload city     the ~6M-triangle city (HDR sky CDF, sun, 1000 point lights),
                      1920x1080 at 1024 spp: plain tracer, batched pipeline
                      and radiance cache, against a 16,384-spp FP64 render
  --workload cache    the radiance-cache comparison alone
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + equal-time relMSE vs 16k-spp oracle, plain vs radiance-cache"
HBM_PEAK_GBS = 8000.0
# VALU issue peak of an MI355X: 256 CUs x 4 SIMDs of 16 lanes, one wave64 VALU
# instruction per SIMD every 4 cycles, 2.4 GHz peak engine clock
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
# extend algorithmic bytes: per ray 32 (ray) + 4 (queue index) + 16 (hit), per triangle record 48,
# per node visit the node record's size (64 B for the quantized 4-wide layout; from scene_info)
RAY_BYTES, TRI_BYTES = 52, 48
# Hardware counters of the extend kernel (k_extend_chunk), averaged per dispatch over one render
# with dispatches serialised, from rocprofv3 --pmc passes:
#   valu_insts     SQ_INSTS_VALU (wave64 VALU instructions issued);
#   thread_cycles / active_inst  SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU (their ratio: active
#                  lanes per VALU instruction);
#   fetch_kib / write_kib  FETCH_SIZE / WRITE_SIZE (KiB), scaled by what the counters read for a
#                  1 GiB device copy (`bench.py --copy-calibration`: FETCH_SIZE reads half the copied
#                  bytes -> x2.0, WRITE_SIZE all of them -> x1.0).
# Counters cannot be read inside this process, so these are the profiled figures, reported only for
# the workload they were taken on.
# Room 1024x768, 256 spp, depth 8: `bench.py --steps 1 --warmup 0 --no-cpu --no-count --no-cache`
# (117 dispatches).
EXTEND_COUNTERS = {"valu_insts": 1122033576, "thread_cycles": 37290147763, "active_inst": 1127273312,
                   "fetch_kib": 1632454.4, "write_kib": 306188.9}
# City 1920x1080, depth 8: one 16-spp render (18 dispatches, waves of 8 samples per pixel).
CITY_EXTEND_COUNTERS = {"valu_insts": 1413430719, "thread_cycles": 49919942694, "active_inst": 1417850773,
                        "fetch_kib": 3085105.1, "write_kib": 262906.4}
# Two-spheres box 512x384, 64 spp, depth 10, RR 5: the whole-path kernel's counters for one
# render (one dispatch of all 12.6M paths, lanes refilled; 99% of the render's kernel time).
SPHERES_PATH_COUNTERS = {"valu_insts": 2447618600, "thread_cycles": 72325034179, "active_inst": 2517330930,
                         "fetch_kib": 476.7, "write_kib": 436596.3, "dispatches": 1}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
DEFAULT_ROOM = (1024, 768, 256)


def copy_calibration(mib=1024):
    """A plain device-to-device copy of known size through torch (for scaling the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters of the extend kernel)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 3}))


def host_cpu():
    """(model name, threads this process may use) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    return model, threads


def cpu_baseline(spp=64, rows=256, width=1024, height=768):
    """The CPU oracle (C++/OpenMP restatement of the reference CPU renderer, built
    -O3 -march=x86-64-v3 -ffp-contract=off) on a bounded strip of the same
    workload: `rows` rows of the 1024x768 image, depth 8, no roulette."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    from tinypt_amd import scenes

    host = scenes.triangle_room(width, height)
    oracle = O.OracleScene(host)
    model, threads = host_cpu()
    y0 = height // 2 - rows // 2
    c = scenes.CONFIG1
    p = O.params(spp, max_depth=c["max_depth"], rr_depth=c["rr_depth"], seed=1)
    t0 = time.perf_counter()
    oracle.render(height, p, y0, y0 + rows, threads=threads)
    dt = time.perf_counter() - t0
    paths = width * rows * spp
    # the extend baseline: the oracle's closest hits (the reference structure: linear top
    # level, median-split BVHTree per shape) over the 1M-ray set of the parity tests
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import raysets

    cam = host.camera()
    rays = raysets.extend_set(oracle, (cam.location, cam.rotation, width, height, cam.fov), 1_000_000, seed=0x5EED)
    t0 = time.perf_counter()
    oracle.trace_closest(rays, mode=0, threads=threads)
    dt_rays = time.perf_counter() - t0
    return {
        "value": paths / dt / 1e6,
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"triangle room rows {y0}-{y0 + rows} of 1024x768 at {spp} spp, depth 8 ({paths} paths, {dt:.1f} s)",
        "cpu": model,
        "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp",
        "extend": {"value": round(len(rays) / dt_rays / 1e6, 2), "unit": "Mrays/s",
                   "sample": f"closest hits of the 1M-ray set (500k camera + 500k cosine secondary rays), {dt_rays:.2f} s"},
    }


def oracle_sample(host, height, spp, max_depth, rr_depth, y0, y1, what):
    """The CPU oracle on rows [y0, y1) of a host scene (a bounded sample of a workload)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    model, threads = host_cpu()
    oracle = O.OracleScene(host)
    p = O.params(spp, max_depth=max_depth, rr_depth=rr_depth, seed=1)
    t0 = time.perf_counter()
    oracle.render(height, p, y0, y1, threads=threads)
    dt = time.perf_counter() - t0
    width = host.camera().width
    paths = width * (y1 - y0) * spp
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{what}: rows {y0}-{y1} of {width}x{height} at {spp} spp ({paths} paths, {dt:.1f} s)",
            "cpu": model, "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--no-cache", action="store_true", help="room: skip the radiance-cache rows")
    ap.add_argument("--render-only", action="store_true",
                    help="city: one render at --spp (and a counting render unless --no-count), for counter passes")
    ap.add_argument("--block-size", type=int, default=0, help="room: threads per workgroup (0 = the library's default)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--workload", choices=["room", "batch", "spheres", "cache", "city"], default="room")
    ap.add_argument("--ref-spp", type=int, default=16384, help="cache, city: samples per pixel of the reference")
    ap.add_argument("--cache-suffix", type=int, default=1,
                    help="cache: training paths end in the cache this many vertices past their cache vertex "
                         "(0: they run to the maximum depth)")
    ap.add_argument("--cache-frames", type=str, default="16,32,64,128,256", help="cache: frame counts of the rows")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--no-seq", action="store_true", help="batch: skip the one-by-one comparison")
    ap.add_argument("--batch-mix", choices=["mixed", "small"], default="mixed",
                    help="batch: mixed = 2k-500k triangles, 128-1024 px, 16-256 spp; "
                         "small = 2k-50k triangles, 64-256 px, 4-32 spp")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by sample index with an RCCL reduce of the accumulation buffer)",
              file=sys.stderr)
        sys.exit(2)

    import tinypt_amd

    if tinypt_amd.hip_device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "spheres":
        bench_spheres(args)
        return
    if args.workload == "city":
        bench_city(args)
        return
    if args.workload == "cache":
        bench_cache(args)
        return
    bench_room(args)


def extend_roofline(pt, host, info, kw, spp, ext_ms, count_stats, counters, streams=3):
    """Roofline of the extend kernel.  Its bound is VALU issue (counters: ~87% of
    the SIMD cycles issue VALU instructions while it runs alone), so `frac` is
    the issue rate of the kernel alone (a one-stream scene handle, no other
    kernel beside it) against the VALU issue peak; its HBM bytes (counters) and
    the survey's algorithmic bytes are reported beside it.  `counters`: the
    profiled per-dispatch counters of this workload, or None."""
    c = count_stats
    rays = c["rays_closest"]
    launches = max(1, c["extend_launches"])
    alg_bytes = rays * RAY_BYTES + c["nodes_visited"] * info["node_bytes"] + c["tris_tested"] * TRI_BYTES
    # the same kernel without the other streams' co-running waves (a scene handle
    # created with TPT_STREAMS=1, one untimed warm-up render, then one measured)
    prev = os.environ.get("TPT_STREAMS")
    os.environ["TPT_STREAMS"] = "1"
    try:
        dev1 = host.to("hip")
    finally:
        if prev is None:
            os.environ.pop("TPT_STREAMS")
        else:
            os.environ["TPT_STREAMS"] = prev
    pt.render(dev1, spp, **kw)
    pt.render(dev1, spp, **kw)
    iso = pt.last_stats()
    del dev1
    iso_s = iso["extend_ms"] / launches / 1e3
    co_s = (sum(ext_ms) / len(ext_ms)) / launches / 1e3
    insts = traffic = lanes = None
    if counters is not None:
        k = counters
        insts = k["valu_insts"]
        traffic = (k["fetch_kib"] * FETCH_SCALE + k["write_kib"] * WRITE_SCALE) * 1024
        lanes = k["thread_cycles"] / k["active_inst"]
    valu = insts / iso_s / 1e9 if insts else None
    alg = alg_bytes / launches
    roofline = {
        "bound": "valu",
        "achieved": round(valu, 1) if valu else None,
        "peak": VALU_PEAK_GINST,
        "unit": "Gwaveinst/s",
        "frac": round(valu / VALU_PEAK_GINST, 4) if valu else None,
        "traffic": round(traffic) if traffic else None,
        "kernel": "k_extend_chunk",
        "avg_launch_ms": round(iso_s * 1e3, 4),
        "valu_insts_per_launch": round(insts) if insts else None,
        "active_lanes_per_inst": round(lanes, 1) if lanes else None,
        "hbm": {"achieved": round(traffic / iso_s / 1e9, 1) if traffic else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(traffic / iso_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "source": "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch / isolated launch time"},
        "algorithmic": {"bytes_per_launch": round(alg), "achieved": round(alg / iso_s / 1e9, 1), "unit": "GB/s",
                        "per_ray": "52 + 64 N_node + 48 N_tri (node and triangle fetches, mostly L2 / MALL hits)"},
        # the timed renders run pipeline waves at once on three streams, so each extend
        # launch shares the CUs with the other waves' shading
        "co_scheduled": {"streams": streams, "avg_launch_ms": round(co_s * 1e3, 4),
                         "achieved": round(insts / co_s / 1e9, 1) if insts else None},
    }
    extra = {
        "extend_mrays_per_s": round(rays / (iso["extend_ms"] / 1e3) / 1e6, 1),
        "nodes_per_ray": round(c["nodes_visited"] / max(1, rays), 2),
        "tris_per_ray": round(c["tris_tested"] / max(1, rays), 2),
        "extend_share": round(iso["extend_ms"] / iso["render_ms"], 3),
    }
    return roofline, extra


def bench_room(args):
    """The triangle room: one step = one render of 1024x768 at 256 spp, depth 8."""
    import tinypt_amd
    from tinypt_amd import scenes

    t0 = time.perf_counter()
    host = scenes.triangle_room(args.width, args.height)
    dev = host.to("hip")
    setup_s = time.perf_counter() - t0
    info = tinypt_amd.scene_info(dev)
    pt = tinypt_amd.PathTracer("hip")
    c1 = scenes.CONFIG1
    kw = dict(max_depth=c1["max_depth"], rr_depth=c1["rr_depth"], seed=1)
    if args.block_size:
        kw["block_size"] = args.block_size
    print(f"# scene ready in {setup_s:.1f} s: {info}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        pt.render(dev, args.spp, **kw)
        print(f"# warmup {i} done: {pt.last_stats()['render_ms']:.1f} ms", file=sys.stderr, flush=True)
    ms = []
    ext_ms = []
    stats = None
    for i in range(args.steps):
        t = time.perf_counter()
        pt.render(dev, args.spp, **kw)  # synchronous: returns after the device finished and the image is on the host
        ms.append((time.perf_counter() - t) * 1e3)
        stats = pt.last_stats()
        ext_ms.append(stats["extend_ms"])
        print(f"# step {i}: {ms[-1]:.1f} ms (extend {stats['extend_ms']:.1f} ms)", file=sys.stderr, flush=True)
    ms_per_step = sum(ms) / len(ms)
    paths = stats["paths"]
    value = paths / (ms_per_step / 1e3) / 1e6

    roofline = None
    extra = {}
    if not args.no_count:
        # counting render: identical paths (same seed), extend counts node / triangle fetches
        pt.render(dev, args.spp, count_traffic=True, **kw)
        c = pt.last_stats()
        profiled = EXTEND_COUNTERS if (args.width, args.height, args.spp) == DEFAULT_ROOM else None
        roofline, extra = extend_roofline(pt, host, info, kw, args.spp, ext_ms, c, profiled)
        extra = {"mrays_per_s": round((c["rays_closest"] + c["rays_shadow"]) / (ms_per_step / 1e3) / 1e6, 1), **extra}
        print(f"# roofline: {roofline}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()
    if cpu:
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)
    cache = None
    if not args.no_cache:
        cache = cache_comparison(args, "16,64,256")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"triangle_room_{info['triangles']}tri_{args.width}x{args.height}_{args.spp}spp_depth8",
            "tracer": "plain HIP wavefront",
            "paths_per_step": paths,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    if cache:
        out["radiance_cache_textured_room"] = cache
    print(json.dumps(out))


def base_line(value, ms_per_step, args, config, dtype="f32"):
    return {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
    }


def bench_batch(args):
    """Independent seeded rooms through one captured pipeline wave."""
    import tinypt_amd
    from tinypt_amd import scenes

    ranges = {"mixed": dict(tris=(2000, 500000), side=(128, 1024), spp=(16, 256)),
              "small": dict(tris=(2000, 50000), side=(64, 256), spp=(4, 32))}[args.batch_mix]
    specs = scenes.batch_specs(count=args.batch_count, **ranges)
    t0 = time.perf_counter()
    hosts = scenes.batch_scenes(specs)
    devs = [h.to("hip") for h in hosts]
    print(f"# {len(devs)} scenes ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    spps = [sp["spp"] for sp in specs]
    seeds = [sp["seed"] for sp in specs]
    c2 = scenes.CONFIG2
    kw = dict(max_depth=c2["max_depth"], rr_depth=c2["rr_depth"])
    pt = tinypt_amd.PathTracer("hip")
    t = time.perf_counter()
    pt.render_batch(devs, spps, seeds, **kw)  # autotunes queue capacity and workgroup size
    st = pt.last_batch_stats()
    cap, block = st["queue_capacity"], st["block_size"]
    print(f"# tuned in {st['tune_ms']:.0f} ms: capacity {cap}, block {block}; first batch {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    for i in range(max(0, args.warmup - 1)):
        pt.render_batch(devs, spps, seeds, queue_capacity=cap, block_size=block, **kw)
    ms = []
    for i in range(args.steps):
        t = time.perf_counter()
        pt.render_batch(devs, spps, seeds, queue_capacity=cap, block_size=block, **kw)
        ms.append((time.perf_counter() - t) * 1e3)
        st = pt.last_batch_stats()
        print(f"# step {i}: {ms[-1]:.1f} ms ({st['waves']} waves)", file=sys.stderr, flush=True)
    ms_per_step = sum(ms) / len(ms)
    paths = st["paths"]
    done = sorted(st["scene_ms"])
    out = base_line(paths / (ms_per_step / 1e3) / 1e6, ms_per_step, args, {
        "workload": (f"batch_{len(devs)}_rooms_{ranges['tris'][0] // 1000}k-{ranges['tris'][1] // 1000}k_tris_"
                     f"{ranges['side'][0]}-{ranges['side'][1]}px_{ranges['spp'][0]}-{ranges['spp'][1]}spp_depth8"),
        "tracer": ("plain HIP wavefront, three path pools on three streams; scenes under 2^23 paths share batched "
                   "pipeline waves (one captured wave per pool, replayed), larger ones run plain pipeline waves of "
                   "their own" + ("; the whole batch captured in one hipGraph and launched once" if st.get("graphs")
                                  else "")),
        "paths_per_step": paths,
        "queue_capacity": cap,
        "block_size": block,
        "waves": st["waves"],
        "tune_ms": round(st["tune_ms"], 1),
        "triangles": [min(sp["tris"] for sp in specs), max(sp["tris"] for sp in specs)],
    })
    # per-kernel times of the batch would need a profiled batch run (kernel trace or
    # counters), which this bench does not make; the batch runs the room's kernels,
    # whose roofline is the default line's
    out["roofline"] = None
    out["roofline_note"] = ("no per-kernel time: the batched waves run as captured graphs replayed from the host; "
                            "the extend roofline of the same kernels is the default room line's")
    out["cpu_baseline"] = None if args.no_cpu else oracle_sample(
        hosts[0], specs[0]["height"], specs[0]["spp"], c2["max_depth"], c2["rr_depth"], 0, specs[0]["height"],
        f"the batch's first scene ({specs[0]['tris']} triangles), whole image")
    out["scene_completion_ms"] = {"min": round(done[0], 1), "median": round(done[len(done) // 2], 1),
                                  "p90": round(done[int(0.9 * (len(done) - 1))], 1), "max": round(done[-1], 1)}
    # the same batch with the capacity tuned up to 2^24 slots (above the configuration's
    # 2^18..2^22): larger pipeline waves amortise each wave's fixed launches
    os.environ["TPT_BATCH_MAX_CAP"] = "24"
    try:
        pt.render_batch(devs, spps, seeds, **kw)
        st24 = pt.last_batch_stats()
        t = time.perf_counter()
        pt.render_batch(devs, spps, seeds, queue_capacity=st24["queue_capacity"], block_size=st24["block_size"], **kw)
        ms24 = (time.perf_counter() - t) * 1e3
    finally:
        os.environ.pop("TPT_BATCH_MAX_CAP")
    out["tuned_up_to_2^24"] = {"mpaths_per_s": round(paths / (ms24 / 1e3) / 1e6, 2), "ms": round(ms24, 1),
                               "queue_capacity": st24["queue_capacity"], "block_size": st24["block_size"]}
    # the same batch captured into one hipGraph and launched once
    os.environ["TPT_BATCH_GRAPH"] = "1"
    try:
        t = time.perf_counter()
        pt.render_batch(devs, spps, seeds, queue_capacity=cap, block_size=block, **kw)
        ms1 = (time.perf_counter() - t) * 1e3
        g1 = pt.last_batch_stats()["graphs"]
    finally:
        os.environ.pop("TPT_BATCH_GRAPH")
    out["one_graph_launch"] = {"mpaths_per_s": round(paths / (ms1 / 1e3) / 1e6, 2), "ms": round(ms1, 1), "graphs": g1}
    if not args.no_seq:
        t = time.perf_counter()
        for d, n, sd in zip(devs, spps, seeds):
            pt.render(d, n, seed=sd, **kw)
        seq_s = time.perf_counter() - t
        out["sequential_mpaths_per_s"] = round(paths / seq_s / 1e6, 2)
        out["batch_speedup_vs_sequential"] = round(seq_s * 1e3 / ms_per_step, 3)
    print(json.dumps(out))


def cache_comparison(args, frame_counts, width=1920, height=1080):
    """The textured room (RGBA8 albedo + alpha cut-out + FP32 bump maps), 1920x1080,
    depth 12: the plain tracer and the learned radiance cache, both against a plain
    FP64 reference of --ref-spp samples per pixel.  Plain relMSE falls as c / time
    (unbiased Monte Carlo), so the plain tracer's time to reach each cached
    render's relMSE is c / relMSE; the ratio to the cached render's time is the
    equal-error speedup.  Every frame-count row is reported."""
    import numpy as np

    import tinypt_amd
    from tinypt_amd import scenes

    c3 = scenes.CONFIG3
    W, H = width, height
    host = scenes.textured_room(W, H)
    dev = host.to("hip")
    pt = tinypt_amd.PathTracer("hip")
    kw = dict(max_depth=c3["max_depth"], rr_depth=c3["rr_depth"])
    t = time.perf_counter()
    ref = pt.render(dev, args.ref_spp, seed=777, hdr=True, fp64=True, **kw).raw()
    ref_s = time.perf_counter() - t
    print(f"# textured room reference {args.ref_spp} spp in {ref_s:.1f} s", file=sys.stderr, flush=True)

    def rel(x):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    pt.render(dev, 4, seed=1, hdr=True, **kw)  # warm-up
    plain = []
    for spp in (16, 32, 64, 128, 256):
        t = time.perf_counter()
        img = pt.render(dev, spp, seed=1, hdr=True, **kw).raw()
        dt = time.perf_counter() - t
        plain.append((spp, dt, rel(img)))
        print(f"# plain {spp:4d} spp: {dt * 1e3:8.1f} ms relMSE {plain[-1][2]:.5f}", file=sys.stderr, flush=True)
    c = sum(e * dt for _, dt, e in plain) / len(plain)  # relMSE * time of the plain tracer
    cached = []
    tinypt_amd.render_cached(dev, 2, cache_depth=2, reset=True, seed=3, suffix_len=args.cache_suffix, **kw)  # warm-up
    for frames in [int(x) for x in frame_counts.split(",")]:
        t = time.perf_counter()
        img, st = tinypt_amd.render_cached(dev, frames, cache_depth=2, train_records=1 << 16, learning_rate=1e-3,
                                           reset=True, weight_seed=5, seed=1, suffix_len=args.cache_suffix, **kw)
        dt = time.perf_counter() - t
        e = rel(img)
        cached.append(dict(frames=frames, ms=round(dt * 1e3, 1), relmse=round(e, 5),
                           plain_ms_same_relmse=round(c / e * 1e3, 1), speedup=round(c / e / dt, 3),
                           infer_ms=round(st["infer_ms"], 1), train_ms=round(st["train_ms"], 1),
                           queries=int(st["queries"]), records=int(st["records"]), steps=int(st["steps"]),
                           mean_bias=round(float(img.mean() / ref.mean() - 1.0), 4)))
        print(f"# cached {frames:4d} frames: {dt * 1e3:8.1f} ms relMSE {e:.5f} -> equal-error speedup "
              f"{cached[-1]['speedup']:.2f} (bias {cached[-1]['mean_bias']:+.4f})", file=sys.stderr, flush=True)
    mlp = mlp_roofline(host, kw, args.cache_suffix)
    print(f"# MLP kernels alone: {mlp}", file=sys.stderr, flush=True)
    return {
        "workload": f"textured_room_264612tri_{W}x{H}_depth12_rgba8_alpha_bump",
        "roofline": mlp,
        "reference": f"plain HIP tracer, {args.ref_spp} spp, FP64 accumulation ({ref_s:.1f} s)",
        "cache": ("5x64 BF16 MFMA MLP, Adam 1e-3, 2^16 suffixes/frame, 8 steps/frame, cache after 2 bounces, "
                  f"training suffixes end in the cache after {args.cache_suffix} more vertices"),
        "plain_relmse_times_s": round(c, 5),
        "plain": [dict(spp=s, ms=round(dt * 1e3, 1), relmse=round(e, 5)) for s, dt, e in plain],
        "cached": cached,
        # the row of the most frames: the longest, least noisy comparison
        "equal_relmse_speedup": cached[-1]["speedup"],
        "equal_relmse_speedup_frames": cached[-1]["frames"],
    }


NRC_FLOP_PER_QUERY = 4 * 64 * 64 * 2 + 64 * 3 * 2  # four 64x64 hidden layers + the 3-output layer
NRC_BYTES_PER_QUERY = 128 + 4 + 16  # 64 BF16 inputs, the record's row index, the float4 prediction
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def mlp_roofline(host, kw, suffix, frames=16):
    """Roofline of the cache's inference kernel (k_nrc_infer) alone: a cached render on a
    one-lane scene handle, where the library's HIP events around each frame's inference
    and training see no other kernel beside them."""
    import tinypt_amd

    prev = os.environ.get("TPT_STREAMS")
    os.environ["TPT_STREAMS"] = "1"
    try:
        dev1 = host.to("hip")
    finally:
        if prev is None:
            os.environ.pop("TPT_STREAMS")
        else:
            os.environ["TPT_STREAMS"] = prev
    tinypt_amd.render_cached(dev1, 2, cache_depth=2, reset=True, seed=3, suffix_len=suffix, **kw)  # warm-up
    _, st = tinypt_amd.render_cached(dev1, frames, cache_depth=2, train_records=1 << 16, learning_rate=1e-3,
                                     reset=True, weight_seed=5, seed=1, suffix_len=suffix, **kw)
    del dev1
    infer_s = st["infer_ms"] / frames / 1e3
    q = st["queries"]
    tflops = q * NRC_FLOP_PER_QUERY / infer_s / 1e12
    gbs = q * NRC_BYTES_PER_QUERY / infer_s / 1e9
    steps = 8 * frames
    return {
        "bound": "mfma", "achieved": round(tflops, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None, "kernel": "k_nrc_infer",
        "avg_launch_ms": round(infer_s * 1e3, 4), "queries_per_launch": q, "flop_per_query": NRC_FLOP_PER_QUERY,
        "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "bytes_per_query": NRC_BYTES_PER_QUERY},
        "train_step_ms": round(st["train_ms"] / steps, 4),
        "source": "one-lane cached render, 16 frames: HIP events around each frame's inference and training",
    }


def bench_cache(args):
    """The radiance-cache comparison alone (every frame-count row); `value` is the
    cached tracer's Mpaths/s at the row of the most frames."""
    W, H = 1920, 1080
    if (args.width, args.height) != (1024, 768):
        W, H = args.width, args.height
    res = cache_comparison(args, args.cache_frames, W, H)
    last = res["cached"][-1]
    paths = W * H * last["frames"]
    out = base_line(paths / (last["ms"] / 1e3) / 1e6, last["ms"], args, {
        "workload": res["workload"],
        "tracer": f"radiance cache ({res['cache']})",
        "paths_per_step": paths,
        "reference": res["reference"],
    })
    out["roofline"] = res["roofline"]
    out["cpu_baseline"] = None
    out["equal_relmse_speedup_vs_plain"] = res["equal_relmse_speedup"]
    out["plain_relmse_times_s"] = res["plain_relmse_times_s"]
    out["plain"] = res["plain"]
    out["cached"] = res["cached"]
    print(json.dumps(out))


def bench_city(args):
    """The ~6M-triangle city, 1920x1080 at 1024 spp, depth 8: the plain tracer at the
    configuration's 1024 spp (timed: `value`), the batched pipeline (one captured
    pipeline wave per pool, replayed) on the same render, and the learned radiance
    cache, each measured by relMSE against a plain FP64 render of --ref-spp samples
    per pixel (16,384 by default)."""
    import numpy as np

    import tinypt_amd
    from tinypt_amd import scenes

    c = scenes.CONFIG4
    t = time.perf_counter()
    host = scenes.city(c["width"], c["height"])
    dev = host.to("hip")
    info = tinypt_amd.scene_info(dev)
    print(f"# city ready in {time.perf_counter() - t:.1f} s: {info}", file=sys.stderr, flush=True)
    pt = tinypt_amd.PathTracer("hip")
    kw = dict(max_depth=c["max_depth"], rr_depth=c["rr_depth"])
    if args.render_only:
        pt.render(dev, args.spp, seed=1, hdr=True, **kw)
        print(f"# render {args.spp} spp: {pt.last_stats()}", file=sys.stderr, flush=True)
        if not args.no_count:
            pt.render(dev, args.spp, seed=1, hdr=True, count_traffic=True, **kw)
            print(f"# counting render: {pt.last_stats()}", file=sys.stderr, flush=True)
        return
    pt.render(dev, 4, seed=1, hdr=True, **kw)  # warm-up
    t = time.perf_counter()
    ref = pt.render(dev, args.ref_spp, seed=777, hdr=True, fp64=True, **kw).raw()
    ref_s = time.perf_counter() - t
    print(f"# reference {args.ref_spp} spp: {ref_s:.1f} s", file=sys.stderr, flush=True)

    def rel(x):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    t = time.perf_counter()
    img = pt.render(dev, c["spp"], seed=1, hdr=True, **kw).raw()
    t_plain = time.perf_counter() - t
    st = pt.last_stats()
    rows = {"plain_1024spp": dict(ms=round(t_plain * 1e3, 1), relmse=round(rel(img), 6))}
    print(f"# plain {c['spp']} spp: {t_plain:.2f} s", file=sys.stderr, flush=True)
    pt.render_batch([dev], [16], [1], hdr=True, **kw)  # autotunes queue capacity and workgroup size
    bst = pt.last_batch_stats()
    cap, blk = bst["queue_capacity"], bst["block_size"]
    t = time.perf_counter()
    imgs = pt.render_batch([dev], [c["spp"]], [1], hdr=True, queue_capacity=cap, block_size=blk, **kw)
    rows["batch_1024spp"] = dict(ms=round((time.perf_counter() - t) * 1e3, 1),
                                 equals_plain=bool(np.array_equal(imgs[0].raw(), img)),
                                 queue_capacity=cap, block_size=blk, tune_ms=round(bst["tune_ms"], 1))
    for spp in (64, 256):
        t = time.perf_counter()
        p_img = pt.render(dev, spp, seed=2, hdr=True, **kw).raw()
        rows[f"plain_{spp}spp"] = dict(ms=round((time.perf_counter() - t) * 1e3, 1), relmse=round(rel(p_img), 5))
        t = time.perf_counter()
        cimg, cst = tinypt_amd.render_cached(dev, spp, cache_depth=2, train_records=1 << 16, reset=True, seed=2,
                                             weight_seed=5, suffix_len=args.cache_suffix, **kw)
        rows[f"cached_{spp}frames"] = dict(ms=round((time.perf_counter() - t) * 1e3, 1), relmse=round(rel(cimg), 5),
                                           mean_bias=round(float(cimg.mean() / ref.mean() - 1.0), 4))
    for k, v in rows.items():
        print(f"# {k}: {v}", file=sys.stderr, flush=True)
    # extend roofline on 16-spp renders (two full waves of 8 samples per pixel, as each wave of
    # the 1024-spp render): co-scheduled timing, counting render, one-stream handle
    pt.render(dev, 16, seed=1, hdr=True, **kw)
    ext16 = pt.last_stats()["extend_ms"]
    pt.render(dev, 16, seed=1, hdr=True, count_traffic=True, **kw)
    cnt = pt.last_stats()
    roofline, extra = extend_roofline(pt, host, info, dict(seed=1, hdr=True, **kw), 16, [ext16], cnt,
                                      CITY_EXTEND_COUNTERS if (c["width"], c["height"]) == (1920, 1080) else None)
    print(f"# roofline: {roofline}", file=sys.stderr, flush=True)
    out = base_line(st["paths"] / t_plain / 1e6, t_plain * 1e3, args, {
        "workload": f"city_{info['triangles']}tri_{c['width']}x{c['height']}_{c['spp']}spp_depth8_"
                    "hdr_sky_cdf_sun_1000_point_lights",
        "tracer": "plain HIP wavefront",
        "paths_per_step": st["paths"],
        "reference": f"plain HIP tracer, {args.ref_spp} spp, FP64 accumulation ({ref_s:.1f} s)",
    })
    out["roofline"] = roofline
    out.update(extra)
    out["cpu_baseline"] = None if args.no_cpu else oracle_sample(
        host, c["height"], 4, c["max_depth"], c["rr_depth"], c["height"] // 2 - 16, c["height"] // 2 + 16,
        "the city at 4 spp")
    out["vs_reference"] = rows
    print(json.dumps(out))


def path_roofline(pt, make_host, kw, spp, counters):
    """Roofline of k_megapath, the analytic scene's one path kernel (its time is
    the render's shading time): its VALU issue rate alone (a one-stream scene
    handle; counters: the workload's profiled per-render sums) against the
    nominal issue peak of 4 cycles per wave64 instruction, with its HBM bytes
    beside it.  The path state stays in registers, so HBM carries only the
    colours written for accumulation."""
    prev = os.environ.get("TPT_STREAMS")
    os.environ["TPT_STREAMS"] = "1"
    try:
        dev1 = make_host().to("hip")
    finally:
        if prev is None:
            os.environ.pop("TPT_STREAMS")
        else:
            os.environ["TPT_STREAMS"] = prev
    pt.render(dev1, spp, **kw)
    pt.render(dev1, spp, **kw)
    iso = pt.last_stats()
    del dev1
    k = counters
    shade_s = iso["shade_ms"] / 1e3
    valu = k["valu_insts"] / shade_s / 1e9
    traffic = (k["fetch_kib"] * FETCH_SCALE + k["write_kib"] * WRITE_SCALE) * 1024
    return {
        "bound": "valu",
        "achieved": round(valu, 1),
        "peak": VALU_PEAK_GINST,
        "unit": "Gwaveinst/s",
        "frac": round(valu / VALU_PEAK_GINST, 4),
        "traffic": round(traffic / k["dispatches"]),
        "kernel": "k_megapath",
        "avg_launch_ms": round(iso["shade_ms"] / k["dispatches"], 4),
        "valu_insts_per_launch": round(k["valu_insts"] / k["dispatches"]),
        "active_lanes_per_inst": round(k["thread_cycles"] / k["active_inst"], 1),
        "kernel_share": round(iso["shade_ms"] / iso["render_ms"], 3),
        "note": ("frac above 1: this instruction mix issues faster than the nominal 4 cycles per wave64 instruction"
                 if valu > VALU_PEAK_GINST else None),
        "hbm": {"achieved": round(traffic / shade_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(traffic / shade_s / 1e9 / HBM_PEAK_GBS, 4),
                "source": "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per render / isolated kernel time"},
    }


def bench_spheres(args):
    """The analytic two-spheres box (9 spheres), 512x384 at 64 spp, depth 10, RR after 5."""
    import tinypt_amd
    from tinypt_amd import scenes

    c = scenes.CONFIG0
    dev = scenes.cornell_spheres(c["width"], c["height"]).to("hip")
    pt = tinypt_amd.PathTracer("hip")
    kw = dict(max_depth=c["max_depth"], rr_depth=c["rr_depth"], seed=1)
    for _ in range(args.warmup):
        pt.render(dev, c["spp"], **kw)
    ms = []
    for _ in range(args.steps):
        t = time.perf_counter()
        pt.render(dev, c["spp"], **kw)
        ms.append((time.perf_counter() - t) * 1e3)
    st = pt.last_stats()
    ms_per_step = sum(ms) / len(ms)
    out = base_line(st["paths"] / (ms_per_step / 1e3) / 1e6, ms_per_step, args, {
        "workload": f"cornell_spheres_9_{c['width']}x{c['height']}_{c['spp']}spp_depth10_rr5",
        "tracer": "plain HIP, whole paths per thread (no BVH)",
        "paths_per_step": st["paths"],
    })
    out["roofline"] = path_roofline(pt, lambda: scenes.cornell_spheres(c["width"], c["height"]), kw, c["spp"],
                                    SPHERES_PATH_COUNTERS)
    print(f"# roofline: {out['roofline']}", file=sys.stderr, flush=True)
    out["cpu_baseline"] = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O

        oracle = O.OracleScene(scenes.cornell_spheres(c["width"], c["height"]))
        threads = len(os.sched_getaffinity(0))
        p = O.params(c["spp"], max_depth=c["max_depth"], rr_depth=c["rr_depth"], seed=1)
        t = time.perf_counter()
        oracle.render(c["height"], p, 0, c["height"], threads=threads)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(st["paths"] / dt / 1e6, 3), "unit": "Mpaths/s", "cores": threads,
                               "kind": "port", "sample": f"the full frame ({st['paths']} paths, {dt:.1f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
