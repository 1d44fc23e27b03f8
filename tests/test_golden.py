"""The committed golden digests (tests/golden/golden.json, written by
tests/golden/make_golden.py): synthetic inputs and oracle outputs reproduce
them on this host, and the HIP path reproduces the oracle's."""
import json
import os

import numpy as np
import pytest

from epcq import synth

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def test_inputs_match_golden():
    import make_golden
    assert make_golden.inputs() == GOLDEN["inputs"]


def test_oracle_sphere_case_matches_golden():
    import make_golden
    assert make_golden.sphere_case() == GOLDEN["sphere_400x400"]


@pytest.mark.gpu
def test_hip_sphere_case_matches_golden(dev):
    import torch

    import epcq.csrc as C
    from epcq.epcq import NeuralPoints, RenderOptions
    from scenes import camera_spec, to_dev
    pts, cams, kw = synth.workload_sphere_small()
    cam = cams[0]
    g = GOLDEN["sphere_400x400"]
    npts = NeuralPoints(pc_type="pc", points=to_dev(pts, dev))
    cs = camera_spec(cam, dev)
    opt = RenderOptions(**kw)
    idx, depth, table = npts.pts2plane(cs, opt)
    host = lambda t: t.cpu().numpy()
    assert synth.sha256(host(idx), host(depth), host(table)) == g["build"]
    sp_xyz, pc_idx, sp_t, mask = npts.render_image_hashtable_topk(cs, opt, idx, depth, table, mode="compat")
    P, S, Kc = cam.width * cam.height, opt.num_sample_points_per_ray, opt.num_point_cloud_per_sp
    assert int(mask.sum().item()) == g["live_samples"]
    assert synth.sha256(host(sp_xyz).reshape(P, S, 3), host(pc_idx).reshape(P, S, Kc), host(sp_t).reshape(P, S),
                        host(mask).reshape(P, S).astype(np.uint8)) == g["compat"]
    out = npts.knn_sampling(cs, opt, idx, table, mode="knn")
    s = out["sample_t"].shape[1]
    dense = [np.full((P, S, Kc), -1, np.int32), np.zeros((P, S, Kc), np.float32), np.zeros((P, S, Kc), np.float32)]
    for a, k in zip(dense, ("nb_idx", "nb_t", "nb_dist")):
        a[:, :s] = host(out[k])
    assert synth.sha256(*dense) == g["knn"]
    torch.cuda.synchronize()
