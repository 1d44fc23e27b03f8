"""Scene helpers shared by the tests: synthetic clouds and cameras as device
tensors / spec objects, plus the oracle's view of the same inputs."""
import numpy as np
import torch

from epcq import synth


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def camera_spec(cam, dev):
    from epcq.epcq import CameraSpec
    return CameraSpec(c2w=to_dev(cam.c2w, dev), w2c=to_dev(cam.w2c, dev), K=to_dev(cam.K, dev), fx=cam.fx,
                      fy=cam.fy, cx=cam.cx, cy=cam.cy, width=cam.width, height=cam.height)


def sphere_view(n=20_000, W=96, H=80, f=140.0, seed=0):
    pts = synth.sphere_cloud(n, seed=seed)
    cam = synth.look_at_camera([0.0, -4.0, 0.0], W, H, f)
    return pts, cam


def box_view(n=200_000, W=160, H=120, seed=0, eye=(2.3, -2.3, 2.3)):
    pts = synth.box_scene(n, n_boxes=20, seed=seed)
    cam = synth.look_at_camera(np.asarray(eye), W, H, synth.nerf_focal(W))
    return pts, cam


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a
