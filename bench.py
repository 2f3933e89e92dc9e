"""Headline benchmark: M paths/s on the 1920x1080, 16 spp, depth-8 terrain frame.

Workload: seeded fBm heightfield (1024^2 vertices, 2,093,058 triangles) plus
64 subdivision-4 icospheres (327,680 triangles) with 40% diffuse / 30% GGX
conductor / 30% rough dielectric materials, 16 blackbody quad emitters and a
constant environment; hero-4-wavelength path tracing, Sobol (fastowen) seed 1.
A step is one full frame of 1920*1080*16 = 33,177,600 paths with the scene
resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per closest-hit query: queue index (4) + ray (32) + hit (24)
# + 64 per BVH node line + 48 per triangle record.
RAY_BYTES = 4 + 32 + 24
NODE_BYTES = 64
TRI_BYTES = 48


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="mega,wave,sorted")
    ap.add_argument("--workload", default="terrain", choices=["terrain", "cornell", "volume"])
    ap.add_argument("--profile", action="store_true",
                    help="profiling run: frames only (no roofline replay, no CPU leg), then a 1 GiB device "
                         "copy used to calibrate FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


WORKLOADS = {
    # name: (synthetic kind, width, height, spp, options, workload label)
    "terrain": (1, 1920, 1080, 16, {}, "terrain-1080p-16spp-depth8"),
    "cornell": (0, 256, 256, 64, {}, "cornell-256-64spp-depth5"),
    "volume": (3, 512, 512, 32, {"hg_g": 0.3}, "volume-512-32spp-depth32-g0.3"),
}


def cpu_baseline(synth, width, height, budget):
    """Oracle (scalar C++ restatement) on a bounded band of pixels, all host threads."""
    import rypath_amd as rp
    from tests import oracle_lib

    threads = os.cpu_count() or 1
    orc = oracle_lib.Oracle()
    t0 = time.time()
    oscene = orc.scene(synth.desc)
    build = time.time() - t0
    row0 = height // 2
    count = width * 4
    req = rp.request(seed=1, pixel_begin=row0 * width, pixel_count=count, sample_count=1)
    t0 = time.time()
    oscene.sample(req, threads)
    dt = time.time() - t0
    rows = 4
    # scale the band so the timed sample takes about `budget` seconds
    if dt < budget / 3:
        rows = max(4, min(height // 2, int(4 * budget / max(dt, 1e-3))))
        req = rp.request(seed=1, pixel_begin=(height // 2 - rows // 2) * width, pixel_count=width * rows,
                         sample_count=1)
        t0 = time.time()
        oscene.sample(req, threads)
        dt = time.time() - t0
    paths = width * rows
    return {"value": paths / dt / 1e6, "unit": "M paths/s", "cores": threads, "kind": "port",
            "sample": f"{rows} full-width rows x 1 spp ({paths} paths) of the same frame; oracle BVH build "
                      f"{build:.1f}s excluded"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (paths shard by pixel or sample "
              f"range; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    import rypath_amd as rp

    torch.cuda.init()
    kind, w, h, spp, opts, label = WORKLOADS[args.workload]
    if args.workload != "terrain":
        args.width, args.height, args.spp = w, h, spp
    synth = rp.Synthetic(kind, seed=1, width=args.width, height=args.height, **opts)
    t0 = time.time()
    scene = rp.Scene(synth.desc)
    info = scene.info()
    film = torch.empty((args.height, args.width, 3), dtype=torch.float32, device="cuda")
    paths = args.width * args.height * args.spp
    modes = {"mega": rp.MODE_MEGAKERNEL, "wave": rp.MODE_WAVEFRONT, "sorted": rp.MODE_WAVEFRONT_SORTED}
    results = {}
    for name in args.modes.split(","):
        mode = modes[name]
        cfg = rp.frame_config(args.spp, mode=mode, seed=1)
        for _ in range(args.warmup):
            scene.render_frame_device(cfg, film.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = []
        for _ in range(args.steps):
            stats.append(scene.render_frame_device(cfg, film.data_ptr()))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        results[name] = {"ms_per_step": dt * 1e3, "mpaths": paths / dt / 1e6, "stats": stats[-1]}
        print(f"[bench] {name}: {dt * 1e3:.1f} ms/frame, {paths / dt / 1e6:.1f} M paths/s, "
              f"rays {stats[-1]['extension_rays'] + stats[-1]['shadow_rays']}", file=sys.stderr, flush=True)
    if args.profile:
        src = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 1 GiB read + 1 GiB written", file=sys.stderr, flush=True)
        return
    best = max(results, key=lambda k: results[k]["mpaths"])
    headline = "sorted" if "sorted" in results else best
    r = results[headline]
    roofline = None
    if headline in ("wave", "sorted"):
        cfg = rp.frame_config(args.spp, mode=modes[headline], seed=1, flags=rp.FRAME_COUNT_TRAFFIC)
        cnt = scene.render_frame_device(cfg, film.data_ptr())
        st = r["stats"]
        alg = RAY_BYTES * cnt["extension_rays"] + NODE_BYTES * cnt["node_fetches"] + TRI_BYTES * cnt["triangle_tests"]
        achieved = alg / st["trace_seconds"] / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_wf_extend (closest hit)",
                    "per_ray_bytes": round(alg / cnt["extension_rays"], 1),
                    "nodes_per_ray": round(cnt["node_fetches"] / cnt["extension_rays"], 2),
                    "tris_per_ray": round(cnt["triangle_tests"] / cnt["extension_rays"], 2),
                    "avg_launch_ms": round(st["trace_seconds"] / max(st["extend_launches"], 1) * 1e3, 3)}
    cpu = None if args.no_cpu else cpu_baseline(synth, args.width, args.height, args.cpu_seconds)
    st = r["stats"]
    line = {
        "metric": "M paths/s @1080p 16spp" if args.workload == "terrain" else f"M paths/s ({label})", "value": round(r["mpaths"], 2), "unit": "M paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": label, "width": args.width, "height": args.height,
                   "spp": args.spp, "max_depth": synth.d.max_depth, "triangles": synth.d.triangle_count,
                   "mode": headline},
        "modes": {k: {"M_paths_s": round(v["mpaths"], 2), "ms_per_frame": round(v["ms_per_step"], 2),
                      "G_rays_s": round((v["stats"]["extension_rays"] + v["stats"]["shadow_rays"]) /
                                        (v["ms_per_step"] * 1e-3) / 1e9, 3)} for k, v in results.items()},
        "bvh_build_s": round(info["build_seconds"], 2), "upload_s": round(info["upload_seconds"], 2),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
