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
# Algorithmic bytes per any-hit (shadow) query: the 64-B shadow request
# (origin/tmax, direction/MIS weight, contribution, metadata) + 64 per node
# line + 48 per triangle record (the radiance update of a visible ray is not
# counted).
SHADOW_RAY_BYTES = 64


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
    ap.add_argument("--modes", default="mega,wave,sorted,pool")
    ap.add_argument("--workload", default="terrain", choices=["terrain", "cornell", "volume", "batch", "inverse"])
    ap.add_argument("--fd-trials", type=int, default=4, help="inverse: independent sample sets of the gradient check")
    ap.add_argument("--fd-rough-trials", type=int, default=0,
                    help="inverse: extra sample sets for roughness at --fd-rough-step (its 1e-3 quotient is noisy)")
    ap.add_argument("--fd-rough-step", type=float, default=0.01)
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="terrain: skip the 64-scene batch key")
    ap.add_argument("--pool-config", default=None,
                    help="pool mode: 'capacity,group_size,sort_bits' instead of running the autotuner")
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


PMC_SUMMARY = ROOT / "profiles" / "terrain_1080p_hbm_traffic.md"


TRAVERSAL_KERNELS = ("void rpd::k_wf_extend<false", "void rpd::k_wf_trace<false", "void rpd::k_wf_shadow<false")


def pmc_traffic(frame_mode):
    """HBM bytes per traversal launch (closest hit, fused closest hit + shadow,
    shadow alone) of the 1080p terrain frame, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by its copy calibration).  None
    when the summary is absent or was recorded for another frame mode than
    the timed one."""
    launches, mib = 0, 0.0
    try:
        text = PMC_SUMMARY.read_text()
        if f"--modes {frame_mode} " not in text:
            return None
        for line in text.splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cells and cells[0].startswith(TRAVERSAL_KERNELS):
                launches += int(cells[1])
                mib += float(cells[4]) + float(cells[5])
    except (OSError, ValueError, IndexError):
        return None
    return round(mib * 2 ** 20 / launches) if launches else None


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_cornell_full(threads):
    """The whole 256x256, 64 spp, depth-5 Cornell frame (4,194,304 paths) on
    the oracle, all host threads: M paths/s and seconds."""
    import rypath_amd as rp
    from tests import oracle_lib

    orc = oracle_lib.Oracle()
    synth = rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=256, height=256)
    oscene = orc.scene(synth.desc)
    t0 = time.time()
    oscene.render_frame(64, seed=1, threads=threads)
    dt = time.time() - t0
    return {"workload": "cornell-256-64spp-depth5", "M_paths_s": round(256 * 256 * 64 / dt / 1e6, 3),
            "seconds": round(dt, 2)}


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
            "cpu_model": cpu_model(),
            "sample": f"{rows} full-width rows x 1 spp ({paths} paths) of the same frame; oracle BVH build "
                      f"{build:.1f}s excluded",
            "full_frame_cornell": cpu_cornell_full(threads)}


def inverse_line(args, dt, W, H, spp, nparams):
    return {
        "metric": f"adjoint gradients/s ({W}x{H}, {spp} spp forward + {spp} spp adjoint, {nparams} params)",
        "value": round(1 / dt, 2), "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "inverse-cornell-fog-256-16spp", "width": W, "height": H, "spp": spp,
                   "params": nparams, "loss": "mean squared XYZ error vs 1024 spp target"},
        "M_paths_s": round(2 * W * H * spp / dt / 1e6, 2),
    }


class InverseScene:
    """The inverse-rendering workload: the 256x256 Cornell box with a fog box
    and 58 parameters (48 wall reflectance bins, 2 GGX roughnesses, 8 fog
    sigma_s bins), the L2 loss against a 1024 spp target rendered from
    parameters perturbed by N(0, 0.1) in logit/log space."""

    W = H = 256
    spp = 16

    def __init__(self):
        import numpy as np
        import torch

        import rypath_amd as rp

        self.rp = rp
        synth = rp.Synthetic(rp.SYNTH_INVERSE, seed=2, width=self.W, height=self.H)
        self.synth = synth
        self.scene = rp.Scene(synth.desc)
        params, values = rp.inverse_params(synth)
        self.params, self.values = params, values.astype(np.float64)
        rng = np.random.default_rng(7)
        z = rng.normal(0, 0.1, len(values))
        target_values = self.values.copy()
        refl = slice(0, 48)
        logit = np.log(self.values[refl] / (1 - self.values[refl]))
        target_values[refl] = 1 / (1 + np.exp(-(logit + z[refl])))
        target_values[48:] = self.values[48:] * np.exp(z[48:])
        self.img = torch.empty((self.H, self.W, 3), dtype=torch.float32, device="cuda")
        self.target = torch.empty_like(self.img)
        self.scene.set_params(params, target_values.astype(np.float32))
        self.scene.render_frame_device(rp.frame_config(1024, mode=rp.MODE_WAVEFRONT, seed=999),
                                       self.target.data_ptr())
        self.scene.set_params(params, self.values.astype(np.float32))
        torch.cuda.synchronize()

    def loss_and_grad(self, cfg, want_grad=True):
        import numpy as np

        self.scene.render_frame_device(cfg, self.img.data_ptr())
        diff = self.img.double() - self.target.double()
        loss = float((diff * diff).mean())
        if not want_grad:
            return loss, None
        dl = (2.0 * diff / diff.numel()).float().contiguous()
        g, _ = self.scene.backward_device(cfg, self.params, dl.data_ptr())
        return loss, g.astype(np.float64)

    def step_seconds(self, steps, warmup):
        """Mean wall time of one step: 16 spp forward (frame wavefront), the loss
        and its image gradient (PyTorch), 16 spp path-replay adjoint."""
        import torch

        rp = self.rp
        for k in range(warmup):
            self.loss_and_grad(rp.frame_config(self.spp, mode=rp.MODE_WAVEFRONT, seed=10 + k))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(steps):
            self.loss, _ = self.loss_and_grad(rp.frame_config(self.spp, mode=rp.MODE_WAVEFRONT, seed=100 + k))
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps


def inverse_workload(args):
    """Inverse rendering step (InverseScene), then the gradient check:
    directional derivatives per parameter group against central differences
    with step 1e-3 and 4096 spp common random numbers, over independent
    sample sets."""
    import numpy as np
    import torch

    import rypath_amd as rp

    inv = InverseScene()
    W, H, spp = inv.W, inv.H, inv.spp
    fd_spp, fd_h = 4096, 1e-3
    scene, params, values, img, target = inv.scene, inv.params, inv.values, inv.img, inv.target
    loss_and_grad = inv.loss_and_grad
    dt = inv.step_seconds(args.steps, args.warmup)
    print(f"[bench] inverse step: {dt * 1e3:.2f} ms ({1 / dt:.1f} gradients/s), loss {inv.loss:.4e}",
          file=sys.stderr, flush=True)

    # Gradient check.  Roulette decisions flip with the parameters and make
    # central differences at a 1e-3 step far noisier than 2 %; the check runs
    # forward and adjoint with roulette off (RP_FRAME_NO_ROULETTE), which
    # changes no derivative term.  The loss is linearised at an independent
    # pilot render (fixed weights w = dLoss/dXYZ of the pilot): with weights
    # taken from the same samples as the derivative, the two estimators would
    # differ by their different covariance with the image noise, not by the
    # derivative.  Each group is probed along the pilot adjoint gradient's own
    # direction, so the directional derivative is as large as the group allows.
    groups = {"walls": slice(0, 48), "roughness": slice(48, 50), "sigma_s": slice(50, 58)}
    if args.fd_trials < 2:
        print(json.dumps(dict(inverse_line(args, dt, W, H, spp, len(params)), gradient_check=None)), flush=True)
        return
    flags = rp.FRAME_NO_ROULETTE
    pilot = rp.frame_config(fd_spp, mode=rp.MODE_WAVEFRONT, seed=4999, flags=flags)
    scene.render_frame_device(pilot, img.data_ptr())
    wts = (2.0 * (img.double() - target.double()) / img.numel())
    wts_np = wts.float().cpu().numpy()
    g0, _ = scene.backward(pilot, params, wts_np)
    g0 = g0.astype(np.float64)

    def functional(cfg):
        scene.render_frame_device(cfg, img.data_ptr())
        return float((img.double() * wts).sum())

    dirs = {}
    for name, sl in groups.items():
        v = np.zeros(len(values))
        v[sl] = g0[sl] * np.abs(values[sl])
        dirs[name] = v / np.linalg.norm(v)
    fd = {n: [] for n in groups}
    ad = {n: [] for n in groups}
    t = time.perf_counter()
    for trial in range(args.fd_trials):
        cfg = rp.frame_config(fd_spp, mode=rp.MODE_WAVEFRONT, seed=5000 + trial, flags=flags)
        g, _ = scene.backward(cfg, params, wts_np)
        g = g.astype(np.float64)
        for name, v in dirs.items():
            ad[name].append(float(g @ v))
            ls = []
            for sgn in (+1, -1):
                scene.set_params(params, (values + sgn * fd_h * v).astype(np.float32))
                ls.append(functional(cfg))
            scene.set_params(params, values.astype(np.float32))
            fd[name].append((ls[0] - ls[1]) / (2 * fd_h))
        print(f"[bench] gradient check trial {trial}: " +
              ", ".join(f"{n} adjoint {ad[n][-1]:.4e} fd {fd[n][-1]:.4e}" for n in groups), file=sys.stderr, flush=True)
    # Roughness: VNDF sampling moves directions with the roughness, and at a
    # 1e-3 step the difference quotient's discontinuity noise (variance ~
    # 1 / (spp * step)) is far above 2 %.  A larger step on more sample sets,
    # compared with the adjoint mean of the trials above.
    rough_fd = []
    for trial in range(args.fd_rough_trials):
        cfg = rp.frame_config(fd_spp, mode=rp.MODE_WAVEFRONT, seed=7000 + trial, flags=flags)
        v = dirs["roughness"]
        ls = []
        for sgn in (+1, -1):
            scene.set_params(params, (values + sgn * args.fd_rough_step * v).astype(np.float32))
            ls.append(functional(cfg))
        scene.set_params(params, values.astype(np.float32))
        rough_fd.append((ls[0] - ls[1]) / (2 * args.fd_rough_step))
    check_s = time.perf_counter() - t
    check = {}
    for name in groups:
        a = np.array(ad[name])
        f = np.array(fd[name])
        n = len(a)
        se = lambda x: float(x.std(ddof=1) / np.sqrt(n)) if n > 1 else None
        check[name] = {"adjoint": float(a.mean()), "adjoint_se": se(a), "central_difference": float(f.mean()),
                       "central_difference_se": se(f),
                       "rel_err": float(abs(f.mean() - a.mean()) / abs(a.mean())),
                       # same sample set per trial: the paired difference has the smaller spread
                       "paired_rel_err_se": (se(f - a) / abs(a.mean())) if n > 1 else None}
    if rough_fd:
        f = np.array(rough_fd)
        a = float(np.mean(ad["roughness"]))
        check["roughness_step_%g" % args.fd_rough_step] = {
            "step": args.fd_rough_step, "trials": len(f), "adjoint": a, "central_difference": float(f.mean()),
            "central_difference_se": float(f.std(ddof=1) / np.sqrt(len(f))) if len(f) > 1 else None,
            "rel_err": float(abs(f.mean() - a) / abs(a))}
    worst = max(c["rel_err"] for n, c in check.items() if n in groups)
    line = dict(inverse_line(args, dt, W, H, spp, len(params)),
                gradient_check={"step": fd_h, "spp": fd_spp, "trials": args.fd_trials, "roulette": False,
                                "direction": "adjoint gradient per group (independent pilot)",
                                "functional": "loss linearised at an independent pilot render", "groups": check,
                                "max_rel_err": round(worst, 5), "seconds": round(check_s, 1)})
    print(json.dumps(line), flush=True)


def batch_results(scenes, steps, warmup):
    """`scenes` seeded uneven scenes (rp.batch_mix) in one pool against the
    same scenes rendered one after another in the plain (megakernel) and the
    wavefront form: aggregate rates, and a per-scene table of the sequential
    frame times."""
    import torch

    import rypath_amd as rp

    t0 = time.time()
    mix = rp.batch_mix(scenes)
    items = [(rp.Scene(s.desc), spp, seed) for s, spp, seed in mix]
    build = time.time() - t0
    total = sum(sc.width * sc.height * spp for sc, spp, _ in items)
    tris = sum(s.d.triangle_count for s, _, _ in mix)
    films = {sc: torch.empty((sc.height, sc.width, 3), dtype=torch.float32, device="cuda") for sc, _, _ in items}
    out = {}
    per_scene = [{"scene": i, "width": sc.width, "spp": spp, "triangles": mix[i][0].d.triangle_count,
                  "paths": sc.width * sc.height * spp} for i, (sc, spp, _) in enumerate(items)]

    def sequential(mode, key):
        for _ in range(warmup):
            for sc, spp, seed in items:
                sc.render_frame_device(rp.frame_config(spp, mode=mode, seed=seed), films[sc].data_ptr())
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            for i, (sc, spp, seed) in enumerate(items):
                t1 = time.perf_counter()
                sc.render_frame_device(rp.frame_config(spp, mode=mode, seed=seed), films[sc].data_ptr())
                per_scene[i][key] = round((time.perf_counter() - t1) * 1e3, 3)  # synchronous call
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    for name, mode, key in (("sequential_plain", rp.MODE_MEGAKERNEL, "plain_ms"),
                            ("sequential_wavefront", rp.MODE_WAVEFRONT, "wavefront_ms"),
                            ("sequential_sorted", rp.MODE_WAVEFRONT_SORTED, "sorted_ms")):
        dt = sequential(mode, key)
        out[name] = {"ms": round(dt * 1e3, 2), "M_paths_s": round(total / dt / 1e6, 2)}
        print(f"[bench] batch {name}: {dt * 1e3:.1f} ms, {total / dt / 1e6:.1f} M paths/s", file=sys.stderr,
              flush=True)
    batch = rp.Batch(items)
    t = time.time()
    cfg, _ = batch.autotune()
    tune = time.time() - t
    for name, c in (("batched_persistent", rp.Batch.config(scheduler=rp.SCHED_PERSISTENT, group_size=128)),
                    ("batched_autotuned", cfg)):
        for _ in range(warmup):
            batch.render(c, films=False)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            _, st = batch.render(c, films=False)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        out[name] = {"ms": round(dt * 1e3, 2), "M_paths_s": round(total / dt / 1e6, 2), "config": rp.config_dict(c)}
        print(f"[bench] batch {name}: {dt * 1e3:.1f} ms, {total / dt / 1e6:.1f} M paths/s", file=sys.stderr,
              flush=True)
    best = max(("batched_persistent", "batched_autotuned"), key=lambda k: out[k]["M_paths_s"])
    return {"workload": "batch-64-uneven", "scenes": len(items), "paths": total, "triangles": tris, "max_depth": 8,
            "steps": steps, "warmup": warmup, "best": best, "M_paths_s": out[best]["M_paths_s"],
            "ms": out[best]["ms"], "modes": out,
            "batched_vs_sequential_plain": round(out[best]["M_paths_s"] / out["sequential_plain"]["M_paths_s"], 3),
            "batched_vs_sequential_wavefront": round(out[best]["M_paths_s"] /
                                                     out["sequential_wavefront"]["M_paths_s"], 3),
            # against the faster of the two single-scene frame paths
            "batched_vs_sequential_best_frame": round(out[best]["M_paths_s"] / max(
                out["sequential_wavefront"]["M_paths_s"], out["sequential_sorted"]["M_paths_s"]), 3),
            "bvh_build_s": round(build, 2), "autotune_s": round(tune, 2), "per_scene": per_scene}


def batch_workload(args):
    """64 seeded uneven scenes in one pool vs the same scenes rendered one after another."""
    b = batch_results(args.batch_scenes, args.steps, args.warmup)
    line = {
        "metric": "M paths/s (64 uneven scenes, one pool)", "value": b["M_paths_s"], "unit": "M paths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": b["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": b["workload"], "scenes": b["scenes"], "paths": b["paths"], "triangles": b["triangles"],
                   "max_depth": 8},
        "modes": b["modes"], "batched_vs_sequential_plain": b["batched_vs_sequential_plain"],
        "batched_vs_sequential_wavefront": b["batched_vs_sequential_wavefront"],
        "batched_vs_sequential_best_frame": b["batched_vs_sequential_best_frame"],
        "bvh_build_s": b["bvh_build_s"], "autotune_s": b["autotune_s"], "per_scene": b["per_scene"],
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: this build renders on one GPU "
              f"(paths would shard by pixel tiles or sample ranges, one film reduction per frame)", file=sys.stderr)
        sys.exit(2)
    import torch

    import rypath_amd as rp

    torch.cuda.init()
    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "inverse":
        return inverse_workload(args)
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
        if name == "pool":
            batch = rp.Batch([(scene, args.spp, 1)])
            if args.pool_config:
                cap, gs, bits = (int(x) for x in args.pool_config.split(","))
                bcfg = rp.Batch.config(scheduler=rp.SCHED_WAVEFRONT, capacity=cap, group_size=gs, sort_bits=bits,
                                       use_graph=1)
            else:
                bcfg, _ = batch.autotune()
            for _ in range(args.warmup):
                batch.render(bcfg, films=False)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                _, bst = batch.render(bcfg, films=False)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            results[name] = {"ms_per_step": dt * 1e3, "mpaths": paths / dt / 1e6,
                             "stats": dict(bst, trace_seconds=0.0, extend_launches=0),
                             "config": rp.config_dict(bcfg)}
            print(f"[bench] pool {rp.config_dict(bcfg)}: {dt * 1e3:.1f} ms/frame, {paths / dt / 1e6:.1f} M paths/s",
                  file=sys.stderr, flush=True)
            continue
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
    headline = best
    frame_mode = max((k for k in results if k in ("wave", "sorted")), key=lambda k: results[k]["mpaths"], default=None)
    r = results[headline]
    roofline = None
    if frame_mode is not None:
        cfg = rp.frame_config(args.spp, mode=modes[frame_mode], seed=1, flags=rp.FRAME_COUNT_TRAFFIC)
        cnt = scene.render_frame_device(cfg, film.data_ptr())
        st = results[frame_mode]["stats"]
        # The traversal launches of the timed frame: the first bounce's closest
        # hits (k_wf_extend), every later bounce's closest hits together with
        # the previous bounce's shadow rays (k_wf_trace), and shadow rays alone
        # before the latency tail (k_wf_shadow).  The instrumented replay (no
        # tail: every query in a bounce kernel) gives the algorithmic bytes per
        # closest-hit and per any-hit query; the timed frame's launches traced
        # extend_rays and shadow_kernel_rays of them.
        ext_per_ray = (RAY_BYTES * cnt["extension_rays"] + NODE_BYTES * cnt["node_fetches"] +
                       TRI_BYTES * cnt["triangle_tests"]) / cnt["extension_rays"]
        sh_rays = max(cnt["shadow_rays"], 1)
        sh_per_ray = (SHADOW_RAY_BYTES * cnt["shadow_rays"] + NODE_BYTES * cnt["shadow_node_fetches"] +
                      TRI_BYTES * cnt["shadow_triangle_tests"]) / sh_rays
        st = results[frame_mode]["stats"]
        alg = ext_per_ray * st["extend_rays"] + sh_per_ray * st["shadow_kernel_rays"]
        t_trav = st["trace_seconds"] + st["shadow_seconds"]
        launches_ = max(st["extend_launches"] + st["shadow_launches"], 1)
        achieved = alg / t_trav / 1e9
        avg_launch = t_trav / launches_
        traffic = pmc_traffic(frame_mode) if (args.workload == "terrain" and (args.width, args.height, args.spp) ==
                                              (1920, 1080, 16)) else None
        # measured read peaks on this device (rp_probe_read_bandwidth): an
        # 8 GiB buffer streams from HBM; 24 MiB read 400 times stays in the
        # XCDs' L2 (3 MiB per XCD, above the 32 KiB L1 share of each CU)
        hbm_meas = rp.probe_read_bandwidth(8 << 30, 2)
        l2_meas = rp.probe_read_bandwidth(24 << 20, 400)
        print(f"[bench] measured read bandwidth: HBM {hbm_meas:.0f} GB/s, L2 {l2_meas:.0f} GB/s",
              file=sys.stderr, flush=True)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    # the kernels move far fewer HBM bytes than they fetch through
                    # the caches: they are limited by latency and lane divergence,
                    # not by the HBM roof that `frac` prices them against
                    "limiter": "latency/divergence",
                    "frac_by_measured_traffic": (round(traffic / avg_launch / 1e9 / HBM_PEAK_GBS, 4)
                                                 if traffic else None),
                    "measured_read_peak_hbm": round(hbm_meas, 1), "frac_of_measured_hbm": round(achieved / hbm_meas, 4),
                    "measured_read_peak_l2": round(l2_meas, 1), "frac_of_measured_l2": round(achieved / l2_meas, 4),
                    "traffic": traffic,
                    "traffic_source": PMC_SUMMARY.relative_to(ROOT).as_posix() if traffic else None,
                    "algorithmic_bytes_per_launch": round(alg / launches_),
                    "kernel": f"BVH traversal ({frame_mode} frame): k_wf_extend (first bounce, closest hit), "
                              "k_wf_trace (closest hit + the previous bounce's shadow rays), k_wf_shadow "
                              "(shadow rays alone)",
                    "launches": {"k_wf_extend": st["extend_launches"] - st["fused_launches"],
                                 "k_wf_trace": st["fused_launches"], "k_wf_shadow": st["shadow_launches"]},
                    "avg_launch_ms": round(avg_launch * 1e3, 3),
                    "closest_hit": {"per_ray_bytes": round(ext_per_ray, 1), "rays": st["extend_rays"],
                                    "nodes_per_ray": round(cnt["node_fetches"] / cnt["extension_rays"], 2),
                                    "tris_per_ray": round(cnt["triangle_tests"] / cnt["extension_rays"], 2),
                                    "tail_extension_rays": st["extension_rays"] - st["extend_rays"]},
                    "any_hit": {"per_ray_bytes": round(sh_per_ray, 1), "rays": st["shadow_kernel_rays"],
                                "fused_rays": st["fused_shadow_rays"],
                                "nodes_per_ray": round(cnt["shadow_node_fetches"] / sh_rays, 2),
                                "tris_per_ray": round(cnt["shadow_triangle_tests"] / sh_rays, 2)},
                    "ns_per_query": round(t_trav / max(st["extend_rays"] + st["shadow_kernel_rays"], 1) * 1e9, 4),
                    "stack_spills_per_ray": round(cnt["stack_spills"] / max(cnt["extension_rays"] +
                                                                            cnt["shadow_rays"], 1), 5)}
    # the second headline figure: inverse-rendering gradients per second
    adjoint = None
    if args.workload == "terrain":
        inv = InverseScene()
        dt_inv = inv.step_seconds(10, 2)
        adjoint = dict(inverse_line(args, dt_inv, inv.W, inv.H, inv.spp, len(inv.params)), steps=10, warmup=2,
                       gradient_check={
                           "step": 1e-3, "target_rel_err": 0.02, "groups": ["walls", "roughness", "sigma_s"],
                           "test": "tests/test_gpu_adjoint.py::"
                                   "test_adjoint_within_2_percent_of_central_differences_at_1e_3",
                           "method": "GPU adjoint against central differences of the forward render with "
                                     "every sampling decision frozen at theta (common paths)",
                           "moving_paths": "bench.py --workload inverse --fd-trials 6 --fd-rough-trials 32"})
        for k in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
            adjoint.pop(k, None)
        print(f"[bench] inverse step: {dt_inv * 1e3:.2f} ms ({1 / dt_inv:.1f} gradients/s)", file=sys.stderr,
              flush=True)
        del inv
    # the uneven 64-scene batch beside the single frame (few steps)
    batch64 = batch_results(args.batch_scenes, 2, 1) if args.workload == "terrain" and not args.no_batch else None
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
                                        (v["ms_per_step"] * 1e-3) / 1e9, 3),
                      **({"autotuned": v["config"]} if "config" in v else {})} for k, v in results.items()},
        "batched_vs_plain": (round(results["pool"]["mpaths"] / results["mega"]["mpaths"], 3)
                             if "pool" in results and "mega" in results else None),
        "bvh_build_s": round(info["build_seconds"], 2), "upload_s": round(info["upload_seconds"], 2),
        "roofline": roofline, "cpu_baseline": cpu, "adjoint": adjoint, "batch64": batch64,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
