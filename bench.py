"""Benchmark of the epcq searching + sampling path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload box|sphere|batched|large]

A step is one pass of the path over the workload's views: per view the hash
table build (fused pts2plane) and the K-nearest sampling (samples, neighbour
indices, ray depths, point-to-ray distances and, where the workload has
features, inverse-distance weights and weighted features).  Inputs are
resident in HBM before timing starts.  Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hashpoint-code_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="box", choices=["box", "sphere", "batched", "large"])
    ap.add_argument("--mode", default="knn", choices=["knn", "compat"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline sample time")
    return ap.parse_args()


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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(views would be sharded round-robin across GPUs, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import epcq.csrc as C
    from epcq.epcq import CameraSpec, NeuralPoints, RenderOptions

    dev = torch.device("cuda:0")
    wl = load_workload(args.workload)
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
    views = []
    for c in wl["cams"]:
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        views.append(CameraSpec(c2w=t(c.c2w), w2c=t(c.w2c), K=t(c.K), fx=c.fx, fy=c.fy, cx=c.cx, cy=c.cy,
                                width=c.width, height=c.height))
    rays_per_step = sum(c.width * c.height for c in wl["cams"])
    r_max = float(wl["radius"].max()) if wl["radius"] is not None else float(opt.radius_threshold)
    s_max = max(1, C.max_live_samples(opt, r_max))
    ws = torch.empty(max(C.build_hashtable_workspace_size(n, c.height, c.width) for c in wl["cams"]),
                     dtype=torch.uint8, device=dev)

    stream = torch.cuda.current_stream(dev)
    ev = []  # (build_start, build_end, sample_end) per view per timed step

    def step(record):
        for cs in views:
            e0 = torch.cuda.Event(enable_timing=True) if record else None
            e1 = torch.cuda.Event(enable_timing=True) if record else None
            e2 = torch.cuda.Event(enable_timing=True) if record else None
            if record:
                e0.record(stream)
            idx, depth, table = npts.pts2plane(cs, opt, workspace=ws)
            if record:
                e1.record(stream)
            out = npts.knn_sampling(cs, opt, idx, table, radius=radius, mode=args.mode, feat=feat, s_max=s_max)
            if record:
                e2.record(stream)
                ev.append((e0, e1, e2, cs))
            del out, idx, depth, table

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    build_ms = np.array([a.elapsed_time(b) for a, b, _, _ in ev])
    sample_ms = np.array([b.elapsed_time(c) for _, b, c, _ in ev])
    ms_per_step = wall * 1e3 / args.steps
    value = rays_per_step * args.steps / wall

    # algorithmic bytes per launch (one view), as in the design notes
    P_total = rays_per_step
    views_per_step = len(views)
    K = opt.num_point_cloud_per_sp
    F = wl["F"]
    fe = 2 if wl["feat_dtype"] == "bf16" else 4
    b_out = 4 + s_max * (4 + 12 * K) + ((s_max * (4 * K + 4 * F)) if F else 0)
    build_bytes = views_per_step * (12 * n + 8 * n) + 8 * P_total
    # sample kernel reads: hash 8 B/pixel, each in-image point's index + xyz
    # once (16 B; N_in <= N), features once per point; writes b_out per ray
    sample_bytes = 8 * P_total + views_per_step * (16 * n + (F * fe * n if F else 0)) + b_out * P_total
    sample_avg_ms = float(sample_ms.mean())
    build_avg_ms = float(build_ms.mean())
    sample_gbs = sample_bytes / views_per_step / (sample_avg_ms * 1e-3) / 1e9
    build_gbs = build_bytes / views_per_step / (build_avg_ms * 1e-3) / 1e9

    record = {
        "metric": "rays/s search+sample, batched/served vs plain per-view; p50/p99 latency",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["name"], "points": n, "views": views_per_step, "rays_per_step": rays_per_step,
                   "radius": r_max if wl["radius"] is not None else float(opt.radius_threshold), "K": K,
                   "features": F, "mode": args.mode, "form": "plain per-view"},
        "roofline": {"bound": "hbm", "kernel": "sample_kernel (knn_sampling)", "achieved": sample_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sample_gbs / HBM_PEAK_GBS, "traffic": None,
                     "avg_ms": sample_avg_ms, "bytes_per_launch": sample_bytes / views_per_step},
        "build": {"avg_ms": build_avg_ms, "achieved_GBs": build_gbs, "frac": build_gbs / HBM_PEAK_GBS,
                  "bytes_per_view": build_bytes / views_per_step},
    }
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline(wl, opt, args.cpu_seconds, s_max, args.mode)
    print(json.dumps(record), flush=True)


def cpu_baseline(wl, opt, target_s, s_max, mode):
    """The oracle (a C++ restatement of the reference path, OpenMP) on the
    host cores: full hash build of the first view + sampling of a row-block
    sample of its rays; rays/s = P / (t_build + t_sample * P / sampled)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.set_threads(threads)
    cam = wl["cams"][0]
    pts = wl["pts"]
    P = cam.width * cam.height
    t0 = time.perf_counter()
    oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    t_build = time.perf_counter() - t0
    ocam = oracle.camera_from(cam)
    oopt = oracle.options_from(opt)
    m = oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT
    # grow the sample until it takes ~target_s
    rows = max(1, cam.height // 64)
    sampled, t_sample = 0, 0.0
    while True:
        rays = np.arange(0, rows * cam.width, dtype=np.int32) + (cam.height // 2 - rows // 2) * cam.width
        rays = rays[(rays >= 0) & (rays < P)]
        t1 = time.perf_counter()
        res = oracle.quick_sampling(pts, ocam, oopt, oi, oh, radius=wl["radius"], mode=m, rays=rays, with_nb=True)
        if wl["F"]:
            live = res[3].reshape(-1)
            feat = np.zeros((len(pts), wl["F"]), np.float32)
            oracle.idw(pts, feat, res[0].reshape(-1, 3)[live], res[1].reshape(-1, res[1].shape[-1])[live])
        t_sample = time.perf_counter() - t1
        sampled = rays.size
        if t_sample >= target_s / 4 or rows >= cam.height:
            break
        rows = min(cam.height, rows * 4)
    rate = P / (t_build + t_sample * P / sampled)
    return {"value": rate, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"view 0 of the workload: full hash build ({t_build:.3f} s) + sampling of {sampled} "
                      f"central rays ({t_sample:.3f} s), oracle with {threads} OpenMP threads"}


if __name__ == "__main__":
    main()
