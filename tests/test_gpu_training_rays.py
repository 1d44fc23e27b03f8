"""The training form of the drop-in: render_image_hashtable_topk with the
data loader's training_ray_index samples only the listed rays, and its
outputs equal the reference renderer's full-image sampling followed by
indexing with [v, u] (opt/render_image.py:82-88), and the oracle's rows for
the same rays."""
import numpy as np
import pytest
import torch

from epcq import synth
from oracle import oracle
from scenes import bits, to_dev
from test_gpu_sampling import View

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_training_ray_index(dev, mode):
    pts = synth.box_scene(400_000, seed=7)
    cam = synth.look_at_camera([-2.0, -2.5, 2.0], 400, 300, 555.5)
    v = View(pts, cam, dev, radius_threshold=0.005, num_point_cloud_per_sp=8)
    # (u, v, image) rows as the loader's split_ray builds them, 64x64 rays
    g = torch.Generator().manual_seed(3)
    perm = torch.randperm(cam.width * cam.height, generator=g)[:4096]
    tri = torch.stack([perm % cam.width, perm // cam.width, torch.zeros_like(perm)], dim=-1).to(torch.int64)
    print(f"training_ray_index {mode}: {len(tri)} rays", flush=True)
    full = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d, mode=mode)
    part = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d, mode=mode,
                                              training_ray_index=tri.to(dev))
    torch.cuda.synchronize()
    iv, iu = tri[:, 1].to(dev), tri[:, 0].to(dev)
    for f, p_ in zip(full, part):
        ref = f[iv, iu]
        assert p_.shape == ref.shape
        assert np.array_equal(bits(p_.cpu().numpy()), bits(ref.cpu().numpy()))
    rays = (tri[:, 1] * cam.width + tri[:, 0]).to(torch.int32).numpy()
    o_xyz, o_idx, o_t, o_mask = oracle.quick_sampling(v.pts, v.ocam, v.oopt, v.oi, v.oh, rays=rays,
                                                      mode=oracle.MODE_KNN if mode == "knn" else oracle.MODE_COMPAT)
    assert np.array_equal(part[3].cpu().numpy(), o_mask)
    assert np.array_equal(part[1].cpu().numpy(), o_idx)
    assert np.array_equal(bits(part[2].cpu().numpy()), bits(o_t))
    assert np.array_equal(bits(part[0].cpu().numpy()), bits(o_xyz))
    assert o_mask.sum() > 500
    # an empty index list samples the whole image, as the reference does
    again = v.npts.render_image_hashtable_topk(v.cs, v.opt, v.idx_d, None, v.hash_d, mode=mode,
                                               training_ray_index=torch.zeros((0, 3), dtype=torch.int64))
    assert again[0].shape == full[0].shape
