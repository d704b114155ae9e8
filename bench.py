"""Benchmark: the plain HIP wavefront path tracer on the seeded ~265k-triangle
room (1024x768, 256 spp, max depth 8, Russian roulette after depth 5).

One step = one full render of that workload (201,326,592 paths) through
PathTracer("hip").render; the scene is resident in HBM before the timed region.
Prints ONE JSON line with Mpaths/s, the extend kernel's roofline figure
(algorithmic bytes from a counting render with the same seed / HIP-event time
of the extend launches) and the CPU oracle timed on a bounded sample.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--spp S] [--no-cpu]
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
# extend algorithmic bytes: per ray 32 (ray) + 4 (queue index) + 16 (hit), per triangle record 48,
# per node visit the node record's size (64 B for the quantized 4-wide layout; from scene_info)
RAY_BYTES, TRI_BYTES = 52, 48


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


def cpu_baseline(spp=64, rows=256, width=1024, height=768):
    """The CPU oracle (C++/OpenMP restatement of the reference CPU renderer) on a
    bounded strip of the same workload: `rows` rows of the 1024x768 image."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    from tinypt_amd import scenes

    host = scenes.triangle_room(width, height)
    oracle = O.OracleScene(host)
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    y0 = height // 2 - rows // 2
    p = O.params(spp, max_depth=8, rr_depth=5, seed=1)
    t0 = time.perf_counter()
    oracle.render(height, p, y0, y0 + rows, threads=threads)
    dt = time.perf_counter() - t0
    paths = width * rows * spp
    return {
        "value": paths / dt / 1e6,
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"triangle room rows {y0}-{y0 + rows} of 1024x768 at {spp} spp ({paths} paths, {dt:.1f} s)",
    }


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
    ap.add_argument("--copy-calibration", action="store_true")
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
    from tinypt_amd import scenes

    if tinypt_amd.hip_device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    t0 = time.perf_counter()
    host = scenes.triangle_room(args.width, args.height)
    dev = host.to("hip")
    setup_s = time.perf_counter() - t0
    info = tinypt_amd.scene_info(dev)
    pt = tinypt_amd.PathTracer("hip")
    kw = dict(max_depth=8, rr_depth=5, seed=1)
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
        rays = c["rays_closest"]
        alg_bytes = rays * RAY_BYTES + c["nodes_visited"] * info["node_bytes"] + c["tris_tested"] * TRI_BYTES
        ext_time_s = (sum(ext_ms) / len(ext_ms)) / 1e3
        achieved = alg_bytes / ext_time_s / 1e9
        roofline = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_extend",
            "node_bytes": info["node_bytes"],
            "bytes_per_launch": alg_bytes / max(1, c["extend_launches"]),
            "avg_launch_ms": (sum(ext_ms) / len(ext_ms)) / max(1, c["extend_launches"]),
        }
        extra = {
            "mrays_per_s": round((rays + c["rays_shadow"]) / (ms_per_step / 1e3) / 1e6, 1),
            "extend_mrays_per_s": round(rays / ext_time_s / 1e6, 1),
            "nodes_per_ray": round(c["nodes_visited"] / max(1, rays), 2),
            "tris_per_ray": round(c["tris_tested"] / max(1, rays), 2),
            "extend_share": round((sum(ext_ms) / len(ext_ms)) / ms_per_step, 3),
        }
    cpu = None if args.no_cpu else cpu_baseline()
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
            "workload": f"triangle_room_{info['triangles']}tri_{args.width}x{args.height}_{args.spp}spp_depth8_rr5",
            "tracer": "plain HIP wavefront",
            "paths_per_step": paths,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
