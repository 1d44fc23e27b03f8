"""Benchmark of the epcq searching + sampling path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--workload batched|box|sphere|large] [--mode compat|knn]

Default workload (batched): the 1M-point 20-box scene seen by 64 views of
mixed resolution ({200x150, 400x300, 640x512, 800x800, 1600x1200}, f = 1111 *
W / 800, eyes on a sphere of radius 4), r = 0.005, K = 8, fp32.  A step is one
pass over all 64 views: per view the hash-table build (projection, per-pixel
point lists, hash table) and the sampling of every pixel's ray (samples,
neighbour indices and, in knn mode, neighbour depths and point-to-ray
distances).  The batched form runs a step as one replay of a captured HIP
graph of persistent kernels with an autotuned launch shape; `value` is its
rays/s.  Beside it the same 64 views go through the plain per-view drop-in
calls (pts2plane + render_image_hashtable_topk, view after view, as the
reference's renderer does), and through per-view compact sampling; the
ratios are reported.  Inputs are resident in HBM before timing starts.

Workloads box / sphere / large time the per-view path on one or sixteen
views (the single-view scene configurations and the 20M-point one).
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hashpoint-code_amd"))

HBM_PEAK_GBS = 8000.0
# VALU issue: 256 CUs x 2.4 GHz x one wave instruction per CU per cycle
VALU_PEAK_PER_S = 256 * 2.4e9
# committed rocprofv3 summaries of the batched workload (compat mode): HBM
# bytes per launch (FETCH_SIZE / WRITE_SIZE, copy-calibrated) and VALU wave
# instructions per launch of the sampling kernel
PROFILES = {
    ("box20-1M-64views-mixed", "compat"): ("profiles/batched_box20-1M-64views_hbm_traffic.txt",
                                           "profiles/batch_sample_kernel_box20-1M-64views_sq_counters.txt"),
}
METRIC = "rays/s search+sample, batched/served vs plain per-view; p50/p99 latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "serving", "box", "sphere", "large", "aggregator", "render", "rays"])
    ap.add_argument("--mode", default=None, choices=["compat", "knn"],
                    help="neighbour mode (default: compat, knn for the large workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="batched: default launch shape, no autotuning")
    ap.add_argument("--no-serve", action="store_true", help="batched: skip the serving sweep")
    ap.add_argument("--tuning", default=None, help="launch shape tile_w,tile_h,tile_cap,flags[,heavy] (no autotuning)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (counter calibration) and exit")
    ap.add_argument("--plain-passes", type=int, default=3, help="batched: timed passes of the per-view forms")
    ap.add_argument("--heavy", type=int, default=None,
                    help="per-view workloads: candidate count above which a ray is sampled by a whole wave "
                         "(skips the autotuning of the launch shape on the first view)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    return ap.parse_args()


def profiled(workload, mode):
    """(HBM bytes per launch, VALU wave instructions per launch, sources) of
    the sampling kernel from the committed profiles, or Nones."""
    names = PROFILES.get((workload, mode))
    if not names:
        return None, None, None
    traffic = valu = None
    try:
        with open(os.path.join(ROOT, names[0])) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    traffic = float(line.rsplit(":", 1)[1])
        with open(os.path.join(ROOT, names[1])) as f:
            for line in f:
                if line.startswith("SQ_INSTS_VALU"):
                    valu = float(line.split()[1])
    except OSError:
        return None, None, None
    return traffic, valu, list(names)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_workload(name):
    from epcq import synth
    if name == "sphere":
        pts, cams, kw = synth.workload_sphere_small()
        return dict(name="sphere-100k-400x400", pts=pts, cams=cams, kw=kw, radius=None, F=0, feat_dtype=None)
    if name == "box":
        pts, cams, kw = synth.workload_box_scene()
        return dict(name="box20-1M-800x800", pts=pts, cams=cams, kw=kw, radius=None, F=16, feat_dtype="f32")
    if name == "batched":
        pts, cams, kw = synth.workload_batched()
        return dict(name="box20-1M-64views-mixed", pts=pts, cams=cams, kw=kw, radius=None, F=0, feat_dtype=None)
    pts, radius, cams, kw = synth.workload_large()
    return dict(name="box200-20M-16x1920x1080", pts=pts, cams=cams, kw=kw, radius=radius, F=32, feat_dtype="bf16")


def camera_specs(cams, dev):
    import numpy as np
    import torch
    from epcq.epcq import CameraSpec
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return [CameraSpec(c2w=t(c.c2w), w2c=t(c.w2c), K=t(c.K), fx=c.fx, fy=c.fy, cx=c.cx, cy=c.cy,
                       width=c.width, height=c.height) for c in cams]


def percentiles(ms):
    import numpy as np
    a = np.asarray(ms, dtype=np.float64)
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)), "mean_ms": float(a.mean())}


def out_bytes_per_ray(s_max, K, knn, F=0, fe=4):
    """Bytes the sampler writes per ray: count, s_max depths, s_max x K
    indices (+ depths and distances in knn mode, + weights and aggregated
    features with features)."""
    b = 4 + s_max * (4 + 4 * K)
    if knn:
        b += s_max * 8 * K
    if F:
        b += s_max * (4 * K + 4 * F)
    return b


# ----------------------------------------------------------------------------
def bench_batched(args, wl):
    import numpy as np
    import torch

    import epcq.csrc as C
    from epcq import autotune
    from epcq.batch import MultiViewSampler
    from epcq.epcq import NeuralPoints, RenderOptions

    dev = torch.device("cuda:0")
    opt = RenderOptions(**wl["kw"])
    xyz = torch.from_numpy(wl["pts"]).to(dev)
    n = xyz.shape[0]
    cams = wl["cams"]
    pairs = [(xyz, c) for c in cams]
    V = len(pairs)
    rays_per_step = sum(c.width * c.height for c in cams)
    K = opt.num_point_cloud_per_sp
    knn = args.mode == "knn"

    # ---- autotuned launch shape for the batch ---------------------------------
    probe = MultiViewSampler(opt, mode=args.mode, device=dev, max_views=V)
    nbytes = probe.required_bytes(pairs)
    s_max = probe.s_max
    probe.close()
    del probe
    arena = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    tune_log = {}
    if args.tuning:
        vals = [int(x) for x in args.tuning.split(",")]
        best = dict(autotune.DEFAULT)
        best.update(zip(("tile_w", "tile_h", "tile_cap", "flags", "heavy"), vals))
    elif args.no_tune:
        best = dict(autotune.DEFAULT)
    else:
        cands = {}

        def run(t):
            key = (t["tile_w"], t["tile_h"], t["tile_cap"], t["flags"], t["heavy"])
            if key not in cands:
                torch.cuda.synchronize()  # the previous candidate is idle before its arena is rewritten
                s = MultiViewSampler(opt, mode=args.mode, device=dev, max_views=V, tuning=t)
                s.adopt_arena(arena)
                s.set_views(pairs)
                cands.clear()  # one live candidate at a time (they share the arena)
                cands[key] = s
            cands[key].run()

        t0 = time.perf_counter()
        best, tune_log = autotune.tune(f"batch/{wl['name']}/{args.mode}", run, reps=2, log=log)
        for s in cands.values():
            s.close()
        cands.clear()
        log(f"autotune: best {best} in {time.perf_counter() - t0:.1f} s")

    sampler = MultiViewSampler(opt, mode=args.mode, device=dev, max_views=V, tuning=best)
    sampler.adopt_arena(arena)
    sampler.set_views(pairs)
    assert sampler.rays == rays_per_step
    sampler.capture()
    stream = sampler.stream

    # ---- batched: graph replays -----------------------------------------------
    for _ in range(args.warmup):
        sampler.replay()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        sampler.replay()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [a.elapsed_time(b) for a, b in evs]
    value = rays_per_step * args.steps / wall
    log(f"batched: {wall * 1e3 / args.steps:.3f} ms/step, {value / 1e9:.3f} G rays/s")

    # stage split of the same pass (eager, events between stages on its stream)
    prof = [sampler.profile() for _ in range(max(3, min(args.steps, 10)))]
    torch.cuda.synchronize()
    stage = {k: float(np.median([p[k] for p in prof])) for k in ("bin", "place", "build", "sample")}
    n_in = 0
    n_samples = 0
    for i in range(V):
        o = sampler.view_outputs(i)
        n_in += int(o["hash"][:, 1].sum().item())
        n_samples += int(o["n_samples"].sum().item())
    build_bytes = V * 16 * n + 8 * rays_per_step
    sample_bytes = 8 * rays_per_step + 16 * n_in + out_bytes_per_ray(s_max, K, knn) * rays_per_step
    sample_gbs = sample_bytes / (stage["sample"] * 1e-3) / 1e9
    build_gbs = build_bytes / (stage["build"] * 1e-3) / 1e9

    # ---- plain per-view drop-in and per-view compact forms ---------------------
    npts = NeuralPoints(pc_type="pc", points=xyz)
    specs = camera_specs(cams, dev)
    ws = torch.empty(max(C.build_hashtable_workspace_size(n, c.height, c.width) for c in cams),
                     dtype=torch.uint8, device=dev)

    def plain_pass():
        for cs in specs:
            idx, depth, table = npts.pts2plane(cs, opt, workspace=ws)
            out = npts.render_image_hashtable_topk(cs, opt, idx, depth, table, mode=args.mode)
            del out, idx, depth, table

    def compact_pass():
        for cs in specs:
            idx, depth, table = npts.pts2plane(cs, opt, workspace=ws)
            out = npts.knn_sampling(cs, opt, idx, table, mode=args.mode, s_max=s_max)
            del out, idx, depth, table

    def timed(fn, passes):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(passes):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / passes

    C.set_tuning(**best)
    if args.plain_passes > 0:
        plain_s = timed(plain_pass, args.plain_passes)
        compact_s = timed(compact_pass, args.plain_passes)
        log(f"plain per-view drop-in: {plain_s * 1e3:.2f} ms/pass; per-view compact: {compact_s * 1e3:.2f} ms/pass")
    else:
        plain_s = compact_s = None
    plain = rays_per_step / plain_s if plain_s else None
    compact = rays_per_step / compact_s if compact_s else None
    ratio = lambda x: value / x if x else None
    ms = lambda x: x * 1e3 if x else None
    traffic, valu, prof_src = profiled(wl["name"], args.mode)

    record = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["name"], "points": n, "views": V, "rays_per_step": rays_per_step,
                   "resolutions": "200x150,400x300,640x512,800x800,1600x1200 (uniform draw)",
                   "radius": float(opt.radius_threshold), "K": K, "mode": args.mode, "s_max": s_max,
                   "form": "batched: one captured HIP graph per step", "tuning": best},
        "latency_per_step": percentiles(step_ms),
        "plain_per_view": {"value": plain, "unit": "rays/s", "ms_per_pass": ms(plain_s),
                           "calls": "pts2plane + render_image_hashtable_topk (dense outputs) per view",
                           "batched_speedup": ratio(plain)},
        "per_view_compact": {"value": compact, "unit": "rays/s", "ms_per_pass": ms(compact_s),
                             "calls": "pts2plane + knn_sampling (compact outputs) per view",
                             "batched_speedup": ratio(compact)},
        "stages_ms": stage,
        "samples_per_step": n_samples,
        "roofline": {"bound": "hbm", "kernel": "batch_sample_kernel",
                     "achieved": sample_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sample_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "avg_ms": stage["sample"], "bytes_per_launch": sample_bytes,
                     "bytes_model": "8 B/ray hash + 16 B per in-view point + written outputs per ray",
                     "traffic_source": prof_src[0] if prof_src else None},
        "issue_roofline": ({"bound": "valu", "valu_wave_instructions": valu,
                            "achieved_per_s": valu / (stage["sample"] * 1e-3), "peak_per_s": VALU_PEAK_PER_S,
                            "frac": valu / (stage["sample"] * 1e-3) / VALU_PEAK_PER_S, "source": prof_src[1]}
                           if valu else None),
        "build": {"avg_ms": stage["build"], "achieved_GBs": build_gbs, "frac": build_gbs / HBM_PEAK_GBS,
                  "bytes_per_step": build_bytes, "bytes_model": "16 B/point/view + 8 B/pixel"},
    }
    if tune_log:
        record["autotune_ms"] = tune_log
    sampler.close()
    del sampler, arena
    if not args.no_serve:
        record["serving"] = bench_serving(args, tuning=best)
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline_views(wl, opt, args.cpu_seconds, args.mode)
    return record


# ----------------------------------------------------------------------------
def serving_clouds(dev):
    """The four resident clouds: 0.5M sphere, 1M / 20 boxes, 2M / 40 boxes,
    5M / 80 boxes (each box scene with its ground plane)."""
    import torch
    from epcq import synth
    t = lambda a: torch.from_numpy(a).to(dev)
    return {
        "sphere-0.5M": t(synth.sphere_cloud(500_000, seed=10)),
        "box20-1M": t(synth.box_scene(1_000_000, n_boxes=20, seed=11)),
        "box40-2M": t(synth.box_scene(2_000_000, n_boxes=40, seed=12)),
        "box80-5M": t(synth.box_scene(5_000_000, n_boxes=80, seed=13)),
    }


def bench_serving(args, fractions=(0.3, 0.6, 0.9), seconds=4.0, tuning=None, clouds=None):
    """Poisson request streams against the continuous-batching server.  The
    capacity (requests/s with every request queued at t = 0) sets the swept
    rates: `fractions` of it, each for about `seconds` of arrivals (at least
    200 requests).  Latency = completion seen by the host - scheduled
    arrival."""
    import numpy as np
    import torch

    from epcq.epcq import RenderOptions
    from epcq.serve import PointServer, poisson_trace, run_trace

    dev = torch.device("cuda:0")
    clouds = clouds or serving_clouds(dev)
    opt = RenderOptions(radius_threshold=0.005, num_point_cloud_per_sp=8, topK=8)
    server = PointServer(clouds, opt, mode=args.mode, slots=4, slot_bytes=3 << 30, max_views=32, tuning=tuning)

    def summarize(reqs, wall):
        lat = np.array([q.latency for q in reqs]) * 1e3
        rays = sum(q.rays for q in reqs)
        return {"requests": len(reqs), "rays_per_s": rays / wall, "req_per_s": len(reqs) / wall,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean())}

    # warm: capture every slot's graph
    warm = poisson_trace(1e9, 4 * len(server.slots), clouds, seed=19)
    run_trace(server, warm, keep_results=False)
    # capacity: everything queued at once
    burst = poisson_trace(1e9, 400, clouds, seed=21)
    burst = [(0.0, c, cam) for _, c, cam in burst]
    t0 = time.perf_counter()
    reqs = run_trace(server, burst, keep_results=False)
    wall = time.perf_counter() - t0
    cap = summarize(reqs, wall)
    log(f"serving capacity: {cap['req_per_s']:.0f} req/s, {cap['rays_per_s'] / 1e9:.3f} G rays/s")
    rates = []
    for i, f in enumerate(fractions):
        rate = f * cap["req_per_s"]
        count = int(max(200, min(4000, rate * seconds)))
        trace = poisson_trace(rate, count, clouds, seed=20 + 10 * i)
        t0 = time.perf_counter()
        reqs = run_trace(server, trace, keep_results=False)
        wall = time.perf_counter() - t0
        r = summarize(reqs, wall)
        r["rate_req_per_s"] = rate
        rates.append(r)
        log(f"serving at {rate:.0f} req/s: p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, "
            f"{r['rays_per_s'] / 1e9:.3f} G rays/s")
    out = {"clouds": {k: int(v.shape[0]) for k, v in clouds.items()}, "slots": len(server.slots),
           "slot_bytes": server.slots[0].nbytes, "max_views_per_batch": server.max_views,
           "max_inflight_slots": server.max_inflight, "capacity": cap, "rates": rates,
           "resolutions": "200x150,400x300,640x512,800x800,1600x1200 (uniform draw)", "mode": args.mode}
    server.close()
    return out


def serving_record(args):
    s = bench_serving(args, fractions=(0.2, 0.4, 0.6, 0.8, 0.95), seconds=6.0)
    top = max(s["rates"], key=lambda r: r["rate_req_per_s"])
    return {"metric": METRIC, "value": top["rays_per_s"], "unit": "rays/s", "n_gpus": 1,
            "steps": top["requests"], "warmup": 4 * s["slots"], "ms_per_step": top["mean_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "serving-4clouds-poisson", "form": "served, continuous batching",
                       "radius": 0.005, "K": 8, "mode": args.mode},
            "serving": s}


# ----------------------------------------------------------------------------
def bench_per_view(args, wl):
    import numpy as np
    import torch

    import epcq.csrc as C
    from epcq import autotune
    from epcq.epcq import NeuralPoints, RenderOptions

    dev = torch.device("cuda:0")
    pts = wl["pts"]
    n = len(pts)
    opt = RenderOptions(**wl["kw"])
    xyz = torch.from_numpy(pts).to(dev)
    radius = torch.from_numpy(wl["radius"]).to(dev) if wl["radius"] is not None else None
    feat = None
    if wl["F"]:
        g = torch.Generator(device=dev).manual_seed(1)
        feat = torch.randn(n, wl["F"], device=dev, generator=g)
        if wl["feat_dtype"] == "bf16":
            feat = feat.to(torch.bfloat16)
    npts = NeuralPoints(pc_type="pc", points=xyz)
    views = camera_specs(wl["cams"], dev)
    rays_per_step = sum(c.width * c.height for c in wl["cams"])
    r_max = float(wl["radius"].max()) if wl["radius"] is not None else float(opt.radius_threshold)
    s_max = max(1, C.max_live_samples(opt, r_max))
    ws = torch.empty(max(C.build_hashtable_workspace_size(n, c.height, c.width) for c in wl["cams"]),
                     dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    ev = []

    # launch shape of the per-view sampler: given, or autotuned on the first view
    if args.tuning or args.heavy is not None:
        best = dict(autotune.DEFAULT)
        if args.tuning:
            best.update(zip(("tile_w", "tile_h", "tile_cap", "flags", "heavy"),
                            (int(x) for x in args.tuning.split(","))))
        if args.heavy is not None:
            best["heavy"] = args.heavy
    else:
        cs = views[0]
        idx, depth, table = npts.pts2plane(cs, opt, workspace=ws)

        def one(t):
            autotune.apply(t)
            npts.knn_sampling(cs, opt, idx, table, radius=radius, mode=args.mode, feat=feat, s_max=s_max)

        t0 = time.perf_counter()
        best, _ = autotune.tune(f"view/{wl['name']}/{args.mode}", one, reps=2, heavy=(C.HEAVY_DEFAULT, 512, 2048, 0),
                                log=log)
        log(f"autotune: best {best} in {time.perf_counter() - t0:.1f} s")
        del idx, depth, table
    autotune.apply(best)

    def step(record):
        for cs in views:
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)] if record else None
            if record:
                e[0].record(stream)
            idx, depth, table = npts.pts2plane(cs, opt, workspace=ws)
            if record:
                e[1].record(stream)
            out = npts.knn_sampling(cs, opt, idx, table, radius=radius, mode=args.mode, feat=feat, s_max=s_max)
            if record:
                e[2].record(stream)
                ev.append(e)
            del out, idx, depth, table

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    build_ms = np.array([a.elapsed_time(b) for a, b, _ in ev])
    sample_ms = np.array([b.elapsed_time(c) for _, b, c in ev])
    value = rays_per_step * args.steps / wall

    V = len(views)
    K = opt.num_point_cloud_per_sp
    F = wl["F"]
    fe = 2 if wl["feat_dtype"] == "bf16" else 4
    # hash table in use by the last view: in-view points
    idx, depth, table = npts.pts2plane(views[0], opt, workspace=ws)
    n_in = int(table.view(-1, 2)[:, 1].sum().item())
    P0 = views[0].width * views[0].height
    build_bytes = 20 * n + 8 * P0  # pts2plane: xyz in; sorted idx, depth, hash out
    sample_bytes = (8 * P0 + n_in * (16 + (4 if radius is not None else 0) + (F * fe if F else 0))
                    + out_bytes_per_ray(s_max, K, args.mode == "knn", F) * P0)
    sample_gbs = sample_bytes / (float(sample_ms.mean()) * 1e-3) / 1e9
    build_gbs = build_bytes / (float(build_ms.mean()) * 1e-3) / 1e9
    record = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 features, f32 geometry" if wl["feat_dtype"] == "bf16" else "f32",
        "data": "synthetic",
        "config": {"workload": wl["name"], "points": n, "views": V, "rays_per_step": rays_per_step,
                   "radius": r_max, "per_point_radius": radius is not None, "K": K, "features": F,
                   "mode": args.mode, "s_max": s_max, "form": "per-view", "tuning": C.get_tuning()},
        "latency_per_view": percentiles(build_ms + sample_ms),
        "roofline": {"bound": "hbm", "kernel": "sample_tile_kernel (knn_sampling)",
                     "achieved": sample_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sample_gbs / HBM_PEAK_GBS, "traffic": None, "avg_ms": float(sample_ms.mean()),
                     "bytes_per_launch": sample_bytes},
        "build": {"avg_ms": float(build_ms.mean()), "achieved_GBs": build_gbs, "frac": build_gbs / HBM_PEAK_GBS,
                  "bytes_per_view": build_bytes},
    }
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline_views(wl, opt, args.cpu_seconds, args.mode)
    return record


# ----------------------------------------------------------------------------
def cpu_model():
    """The host CPU's model name (as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_views(wl, opt, target_s, mode, reps=5):
    """The oracle (a C++ restatement of the reference path, OpenMP) on the
    host cores, over a bounded, stratified sample: one view of each
    resolution class of the workload; per view the full hash build and the
    sampling of a fixed set of uniformly drawn random rays, each timed as one
    warm-up and then the median of `reps` runs.  The ray count per view is
    sized so the whole sample takes about `target_s`.  rays/s of the whole
    workload = total rays / sum over classes of (views x build time + rays x
    sampling time per ray)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.set_threads(threads)
    pts = wl["pts"]
    oopt = oracle.options_from(opt)
    m = oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT
    cams = wl["cams"]
    classes = {}
    for c in cams:
        classes.setdefault((c.width, c.height), []).append(c)
    rng = np.random.default_rng(12345)
    per_class = []
    budget = target_s / (len(classes) * (reps + 1))  # seconds per sampling run

    def median_time(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    for (w, h), group in sorted(classes.items()):
        cam = group[0]
        t_build = median_time(lambda: oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width))
        oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
        ocam = oracle.camera_from(cam)
        P = w * h
        feat = np.zeros((len(pts), wl["F"]), np.float32) if wl["F"] else None

        def sample(rays):
            res = oracle.quick_sampling(pts, ocam, oopt, oi, oh, radius=wl["radius"], mode=m, rays=rays, with_nb=True)
            if feat is not None:
                live = res[3].reshape(-1)
                oracle.idw(pts, feat, res[0].reshape(-1, 3)[live], res[1].reshape(-1, res[1].shape[-1])[live])

        # size the ray set from a probe run, then time it
        probe = rng.integers(0, P, size=min(P, 2000)).astype(np.int32)
        t0 = time.perf_counter()
        sample(probe)
        per_ray = max((time.perf_counter() - t0) / probe.size, 1e-9)
        count = int(min(P, max(2000, budget / per_ray)))
        rays = rng.integers(0, P, size=count).astype(np.int32)
        t_samp = median_time(lambda: sample(rays))
        per_class.append({"resolution": f"{w}x{h}", "views": len(group), "build_s": t_build,
                          "rays_timed": count, "sample_s": t_samp, "rays_per_view": P})
    total = sum(c.width * c.height for c in cams)
    denom = sum(pc["views"] * (pc["build_s"] + pc["rays_per_view"] * pc["sample_s"] / pc["rays_timed"])
                for pc in per_class)
    rate = total / denom
    desc = "; ".join(f'{pc["resolution"]}: build {pc["build_s"] * 1e3:.1f} ms, {pc["rays_timed"]} random rays '
                     f'in {pc["sample_s"] * 1e3:.1f} ms' for pc in per_class)
    return {"value": rate, "unit": "rays/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "method": f"one view per resolution class, 1 warm-up + median of {reps} runs per timing",
            "sample": f"{desc}; {mode} mode, scaled to {total} rays over {len(cams)} views; oracle with "
                      f"{threads} OpenMP threads",
            "per_class": per_class}


# algorithmic FLOPs of the Aggregator: per valid (sample, neighbour) pair the
# four point layers and the alpha head, per sample the colour branch
AGG_FLOP_PAIR = 2 * (284 * 256 + 256 * 256 + 263 * 256 + 256 * 256 + 256)
AGG_FLOP_SAMPLE = 2 * (280 * 128 + 128 * 128 + 128 * 128 + 128 * 3)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def bench_aggregator(args):
    """Aggregator forward (the reference renderer's MLP after sampling) on
    the live samples of an 800x800 view with two samples per ray: 1,280,000
    samples, K = 8 neighbours (90 % valid), 32-dim features; bf16 matrix
    cores with fp32 accumulation.  Beside it the fp32 torch formulation of
    the reference module on the same GPU (rocBLAS GEMMs) and on the host."""
    import numpy as np
    import torch

    import epcq.csrc as C
    from epcq.epcq import Aggregator

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from aggregator_ref import aggregator_reference, random_inputs

    dev = torch.device("cuda:0")
    M, K = 1_280_000, 8
    inputs = random_inputs(M, K, seed=1, p_valid=0.9, empty_every=0)
    dinp = [t.to(dev) for t in inputs]
    pairs = int(inputs[5].sum())
    torch.manual_seed(0)
    mod = Aggregator().to(dev)
    stream = torch.cuda.current_stream(dev)
    with torch.no_grad():
        pk = mod.packed_params()
        for _ in range(args.warmup):
            C.aggregator_forward(pk, *dinp)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(stream)
            C.aggregator_forward(pk, *dinp)
            e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = [e0.elapsed_time(e1) for e0, e1 in evs]
        # the fp32 torch formulation of the reference module on this GPU
        ref_ms = []
        for i in range(3):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            aggregator_reference(mod, *dinp)
            torch.cuda.synchronize()
            ref_ms.append((time.perf_counter() - t1) * 1e3)
        # and on the host, on a sample of the reference's chunk size
        ncpu = 20000
        cpu_inp = [t[:ncpu] for t in inputs]
        mod_cpu = Aggregator()
        mod_cpu.load_state_dict({k: v.cpu() for k, v in mod.state_dict().items()})
        aggregator_reference(mod_cpu, *cpu_inp)
        cpu_t = []
        for _ in range(3):
            t1 = time.perf_counter()
            aggregator_reference(mod_cpu, *cpu_inp)
            cpu_t.append(time.perf_counter() - t1)
    avg = float(np.mean(ms))
    flops = pairs * AGG_FLOP_PAIR + M * AGG_FLOP_SAMPLE
    ach = flops / (avg * 1e-3) / 1e12
    log(f"aggregator: {avg:.3f} ms per forward, {M / avg * 1e3 / 1e6:.1f} M samples/s, {ach:.1f} TFLOP/s")
    return {
        "metric": "samples/s aggregator forward (point + colour MLPs, IDW aggregation)",
        "value": M / (avg * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": avg, "wall_s": wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "aggregator-1.28M-samples-K8", "samples": M, "K": K, "valid_pairs": pairs,
                   "features": 32, "layers": "284-256-256 | 263-256-256 | 256-1 | 280-128-128-128-3"},
        "roofline": {"bound": "mfma", "kernel": "agg_pair_kernel + agg_color_kernel",
                     "achieved": ach, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "flops_model": "valid pairs x 2(284x256 + 256x256 + 263x256 + "
                     "256x256 + 256) + samples x 2(280x128 + 2x128x128 + 128x3)"},
        "torch_fp32_gpu": {"ms_per_step": float(np.median(ref_ms)), "value": M / (float(np.median(ref_ms)) * 1e-3),
                           "unit": "samples/s", "speedup": float(np.median(ref_ms)) / avg},
        "cpu_baseline": {"value": ncpu / float(np.median(cpu_t)), "unit": "samples/s",
                         "cores": torch.get_num_threads(), "kind": "port", "cpu_model": cpu_model(),
                         "sample": f"{ncpu} samples (the reference's chunk size), fp32 torch formulation of the "
                                   f"reference module, median of 3 after a warm-up"},
    }


def bench_render(args):
    """Volume-render epilogue (clamp_t fused into render, and render's
    backward) on the rays of an 800x800 image with S = 16 samples each:
    memory-bound per-ray scans.  Algorithmic bytes per ray: forward reads t,
    sigma (4S each) and rgb (12S) and writes the colour (12); backward also
    reads the colour gradient (12) and writes the sigma and rgb gradients
    (4S + 12S)."""
    import numpy as np
    import torch

    import epcq.csrc as C
    from oracle import oracle

    dev = torch.device("cuda:0")
    R, S = 800 * 800, 16
    rng = np.random.default_rng(5)
    gaps = rng.uniform(0.0, 0.012, size=(R, S))
    t = (rng.uniform(2.0, 6.0, size=(R, 1)) + np.cumsum(gaps, axis=1)).astype(np.float32)
    sigma = rng.exponential(20.0, size=(R, S)).astype(np.float32)
    rgb = rng.uniform(0, 1, size=(R, S, 3)).astype(np.float32)
    g = rng.normal(size=(R, 3)).astype(np.float32)
    td, sd, cd, gd = (torch.from_numpy(x).to(dev) for x in (t, sigma, rgb, g))
    stream = torch.cuda.current_stream(dev)

    def timed(fn, n):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
        for e0, e1 in evs:
            e0.record(stream)
            fn()
            e1.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))

    fwd = timed(lambda: C.render_rays(sd, cd, white_bg=True, t_val=td), args.steps)
    bwd = timed(lambda: C.render_rays_backward(sd, cd, gd, white_bg=True, t_val=td), args.steps)
    fb = R * (20 * S + 12)
    bb = R * (20 * S + 12 + 16 * S)
    # the reference's torch formulation (epcq.py:508-530) on this GPU
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_oracle_render import torch_reference

    ref = timed(lambda: torch_reference(td, sd, cd, True), 3)
    n_cpu = 100_000
    oracle.render(sigma[:n_cpu], rgb[:n_cpu], True, t=t[:n_cpu])
    cpu = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.render(sigma[:n_cpu], rgb[:n_cpu], True, t=t[:n_cpu])
        cpu.append(time.perf_counter() - t0)
    ach = fb / (fwd * 1e-3) / 1e9
    log(f"render: forward {fwd:.3f} ms ({ach:.0f} GB/s), backward {bwd:.3f} ms "
        f"({bb / (bwd * 1e-3) / 1e9:.0f} GB/s), torch fp32 {ref:.3f} ms")
    return {
        "metric": "rays/s volume-render epilogue (clamp_t + render forward)", "value": R / (fwd * 1e-3),
        "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": fwd,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "render-800x800-S16", "rays": R, "samples_per_ray": S, "white_bg": True},
        "backward": {"ms": bwd, "rays_per_s": R / (bwd * 1e-3), "achieved_GBs": bb / (bwd * 1e-3) / 1e9,
                     "bytes_per_launch": bb},
        "roofline": {"bound": "hbm", "kernel": "render_vec_kernel", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": fb,
                     "bytes_model": "per ray 4S (t) + 4S (sigma) + 12S (rgb) + 12 (colour)"},
        "torch_fp32_gpu": {"ms_per_step": ref, "speedup": ref / fwd},
        "cpu_baseline": {"value": n_cpu / float(np.median(cpu)), "unit": "rays/s", "cores": 1, "kind": "port",
                         "cpu_model": cpu_model(), "sample": f"{n_cpu} rays through the oracle's render, "
                                                              "median of 3 after a warm-up"},
    }


def bench_rays(args):
    """Training ray subset (opt/render_image.py:82-88): a batch of 64 x 64
    random rays of an 800x800 view of the 1M-point box scene sampled in the
    dense drop-in layout (S = 16, Kc = 8), by render_image_hashtable_topk
    with training_ray_index, against the reference's way: sample the whole
    image, then index the outputs with the batch.  Each step includes the
    hash-table build of the view (pts2plane), as a training step does."""
    import numpy as np
    import torch

    from epcq import synth
    from epcq.epcq import NeuralPoints, RenderOptions
    from oracle import oracle

    dev = torch.device("cuda:0")
    pts, cams, kw = synth.workload_box_scene()
    cam = cams[0]
    opt = RenderOptions(**dict(kw, num_point_cloud_per_sp=8))
    npts = NeuralPoints(pc_type="pc", points=torch.from_numpy(pts).to(dev))
    cs = camera_specs([cam], dev)[0]
    rng = np.random.default_rng(6)
    B = 64 * 64
    uv = np.stack([rng.integers(0, cam.width, B), rng.integers(0, cam.height, B)], axis=1).astype(np.int64)
    tri = torch.from_numpy(uv).to(dev)
    stream = torch.cuda.current_stream(dev)

    def subset():
        idx, depth, table = npts.pts2plane(cs, opt)
        return npts.render_image_hashtable_topk(cs, opt, idx, depth, table, training_ray_index=tri)

    def full_then_index():
        idx, depth, table = npts.pts2plane(cs, opt)
        sp, pc, st, m = npts.render_image_hashtable_topk(cs, opt, idx, depth, table)
        v, u = tri[:, 1], tri[:, 0]
        return sp[v, u], pc[v, u], st[v, u], m[v, u]

    def timed(fn, n):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
        for e0, e1 in evs:
            e0.record(stream)
            fn()
            e1.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))

    a = subset()
    b = full_then_index()
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y), "subset sampling differs from full-image sampling + indexing"
    ms_sub = timed(subset, args.steps)
    ms_full = timed(full_then_index, args.steps)
    # CPU: the oracle's build and its sampling of the same ray list
    oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    rays = (uv[:, 1] * cam.width + uv[:, 0]).astype(np.int32)
    ocam, oopt = oracle.camera_from(cam), oracle.options_from(opt)
    cpu = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
        oracle.quick_sampling(pts, ocam, oopt, oi, oh, mode=oracle.MODE_COMPAT, rays=rays)
        cpu.append(time.perf_counter() - t0)
    log(f"training rays: {ms_sub:.3f} ms per batch of {B} (build + subset sampling), "
        f"full image + indexing {ms_full:.3f} ms")
    return {
        "metric": "rays/s training ray batch (build + sampling of the batch's rays)", "value": B / (ms_sub * 1e-3),
        "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_sub,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "box20-1M-800x800-batch4096", "points": len(pts), "rays": B, "S": opt.num_sample_points_per_ray,
                   "Kc": opt.num_point_cloud_per_sp, "layout": "dense drop-in"},
        "full_image_then_index": {"ms_per_step": ms_full, "speedup": ms_full / ms_sub},
        "roofline": None,
        "cpu_baseline": {"value": B / float(np.median(cpu)), "unit": "rays/s", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                         "kind": "port", "cpu_model": cpu_model(),
                         "sample": "the oracle's build of the view and its sampling of the same 4096 rays, "
                                   "median of 3"},
    }


def calibrate_copy(mib):
    """A plain device-to-device copy of `mib` MiB, five times: the reference
    point that scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE)."""
    import torch
    n = mib << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 5}), flush=True)


def main():
    args = parse()
    if args.mode is None:
        args.mode = "knn" if args.workload == "large" else "compat"
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(views would be sharded round-robin across GPUs, one process per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serving":
        print(json.dumps(serving_record(args)), flush=True)
        return
    if args.workload == "aggregator":
        print(json.dumps(bench_aggregator(args)), flush=True)
        return
    if args.workload == "render":
        print(json.dumps(bench_render(args)), flush=True)
        return
    if args.workload == "rays":
        print(json.dumps(bench_rays(args)), flush=True)
        return
    wl = load_workload(args.workload)
    if args.workload == "batched":
        record = bench_batched(args, wl)
    else:
        record = bench_per_view(args, wl)
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
