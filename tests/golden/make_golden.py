"""Writes tests/golden/golden.json: SHA-256 digests of the synthetic inputs
(sphere and box scenes, the 64 mixed views, per-point radii) and of the
oracle's outputs on the 100k-point sphere seen by one 400x400 camera
(f = 555.5, r = 0.01, K = 8): the hash table build and the compat and knn
sampling.  The digests pin the input generator and the oracle across hosts;
the GPU tests compare the HIP outputs with the same digests.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hashpoint-code_amd"))

import numpy as np  # noqa: E402

from epcq import synth  # noqa: E402


def camera_arrays(cams):
    return [np.concatenate([c.c2w.reshape(-1), c.w2c.reshape(-1), c.K.reshape(-1),
                            np.array([c.fx, c.fy, c.cx, c.cy], np.float32),
                            np.array([c.width, c.height], np.float32)]) for c in cams]


def sphere_case():
    """Oracle outputs on the small sphere configuration."""
    from oracle import oracle
    pts, cams, kw = synth.workload_sphere_small()
    cam = cams[0]
    idx, depth, table = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width)
    opt = oracle.options(**kw)
    ocam = oracle.camera_from(cam)
    sp_xyz, pc_idx, sp_t, mask = oracle.quick_sampling(pts, ocam, opt, idx, table, mode=oracle.MODE_COMPAT)
    _, kidx, _, _, nb_t, nb_d = oracle.quick_sampling(pts, ocam, opt, idx, table, mode=oracle.MODE_KNN,
                                                      with_nb=True)
    return {
        "build": synth.sha256(idx, depth, table),
        "compat": synth.sha256(sp_xyz, pc_idx, sp_t, mask.astype(np.uint8)),
        "knn": synth.sha256(kidx, nb_t, nb_d),
        "live_samples": int(mask.sum()),
        "in_image_points": int(table[:, 1][table[:, 1] > 0].sum()),
    }


def inputs():
    pts_s, cams_s, _ = synth.workload_sphere_small()
    pts_b = synth.box_scene(1_000_000, n_boxes=20, seed=0, jitter=0.002)
    cams_b = synth.random_views(64, seed_eye=2, seed_res=3)
    return {
        "sphere_100k_seed0": synth.sha256(pts_s),
        "sphere_camera": synth.sha256(*camera_arrays(cams_s)),
        "box20_1M_seed0": synth.sha256(pts_b),
        "views64_seed2_3": synth.sha256(*camera_arrays(cams_b)),
        "radius_100k_seed31": synth.sha256(synth.log_uniform_radius(100_000, seed=31)),
    }


def main():
    out = {"inputs": inputs(), "sphere_400x400": sphere_case()}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
