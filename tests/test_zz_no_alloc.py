"""epcq_build_hashtable_noalloc (the build whose workspace is the library's
own device array) through the C ABI: its outputs equal epcq_build_hashtable's
bit for bit, on two calls in a row (the second smaller, so stale scratch from
the first must not leak into it), with and without the depth sort; sizes
beyond its scratch are refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq import synth
from oracle import oracle
from scenes import to_dev

pytestmark = pytest.mark.gpu


def run(lib_fn, xyz, cam, sort, dev, workspace=None):
    n = xyz.shape[0]
    H, W = cam.height, cam.width
    w2c = to_dev(np.ascontiguousarray(cam.w2c[:3, :4]), dev)
    K = to_dev(np.ascontiguousarray(cam.K), dev)
    idx = torch.full((n,), 7, dtype=torch.int32, device=dev)
    depth = torch.full((n,), 7.0, dtype=torch.float32, device=dev)
    table = torch.full((H * W, 2), 7, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    if workspace is None:
        rc = lib_fn(p(xyz), n, p(w2c), p(K), H, W, int(sort), p(idx), p(depth), p(table), stream)
    else:
        rc = lib_fn(p(xyz), n, p(w2c), p(K), H, W, int(sort), p(idx), p(depth), p(table), p(workspace),
                    workspace.numel(), stream)
    assert rc == 0, C._lib.epcq_last_error()
    torch.cuda.synchronize()
    return idx.cpu().numpy(), depth.cpu().numpy(), table.cpu().numpy()


def same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.int32) if x.dtype == np.float32 else x,
                              y.view(np.int32) if y.dtype == np.float32 else y)


@pytest.mark.parametrize("sort", [0, 1])
def test_noalloc_equals_build(dev, sort):
    lib = C._lib
    big = synth.box_scene(1_000_000, seed=3)
    small = synth.sphere_cloud(50_000, seed=4)
    cams = [synth.look_at_camera([2.3, -2.3, 2.3], 800, 800, synth.nerf_focal(800)),
            synth.look_at_camera([0.0, -4.0, 0.0], 160, 120, 222.0)]
    for pts, cam in ((big, cams[0]), (small, cams[1])):  # second call smaller
        print(f"noalloc build: {len(pts)} points, {cam.width}x{cam.height}, sort={sort}", flush=True)
        xyz = to_dev(pts, dev)
        need = lib.epcq_build_hashtable_workspace_size(len(pts), cam.height, cam.width)
        ws = torch.empty(need, dtype=torch.uint8, device=dev)
        ref = run(lib.epcq_build_hashtable, xyz, cam, sort, dev, workspace=ws)
        got = run(lib.epcq_build_hashtable_noalloc, xyz, cam, sort, dev)
        same(got, ref)
        oi, od, oh = oracle.pts2plane(pts, cam.w2c, cam.K, cam.height, cam.width, sort=bool(sort))
        same(got, (oi, od, oh))


def test_noalloc_refuses_oversize():
    lib = C._lib
    n = 40_000_000
    assert lib.epcq_build_hashtable_workspace_size(n, 1080, 1920) > (128 << 20)
    rc = lib.epcq_build_hashtable_noalloc(None, n, None, None, 1080, 1920, 0, None, None, None, None)
    assert rc == 3
    assert "bytes of scratch" in C._lib.epcq_last_error().decode()
