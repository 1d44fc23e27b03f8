"""Loaders of the reference's data path without its optional dependencies:
PLY point clouds (the arrays open3d.io.read_point_cloud returns: float64
points, colours scaled to [0, 1]), NeRF-synthetic cameras converted from
Blender to OpenCV axes (checked by projecting points of known placement),
and the Point-NeRF checkpoint branch of NeuralPoints (tensors only)."""
import json
import math

import numpy as np
import pytest
import torch

from epcq.io import nerf_synthetic_cameras, read_ply


def write_ply(path, pts, cols=None, fmt="ascii", xyz_type="float", col_type="uchar", faces=True):
    n = len(pts)
    head = ["ply", f"format {fmt} 1.0", "comment written by the test", f"element vertex {n}"]
    head += [f"property {xyz_type} {c}" for c in "xyz"]
    if cols is not None:
        head += [f"property {col_type} {c}" for c in ("red", "green", "blue")]
    if faces:
        head += ["element face 2", "property list uchar int vertex_indices"]
    head += ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for i in range(n):
                row = [repr(float(v)) for v in pts[i]]
                if cols is not None:
                    row += [str(int(c)) if col_type == "uchar" else repr(float(c)) for c in cols[i]]
                f.write((" ".join(row) + "\n").encode())
            if faces:
                f.write(b"3 0 1 2\n3 1 2 3\n")
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            xt = {"float": "f4", "double": "f8"}[xyz_type]
            fields = [(c, e + xt) for c in "xyz"]
            if cols is not None:
                fields += [(c, e + {"uchar": "u1", "float": "f4"}[col_type]) for c in ("red", "green", "blue")]
            rec = np.zeros(n, dtype=np.dtype(fields))
            for j, c in enumerate("xyz"):
                rec[c] = pts[:, j]
            if cols is not None:
                for j, c in enumerate(("red", "green", "blue")):
                    rec[c] = cols[:, j]
            f.write(rec.tobytes())
            if faces:
                for tri in ((0, 1, 2), (1, 2, 3)):
                    f.write(np.array([3], "u1").tobytes() + np.array(tri, e + "i4").tobytes())


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("xyz_type", ["float", "double"])
def test_read_ply(tmp_path, fmt, xyz_type):
    rng = np.random.default_rng(0)
    pts = rng.normal(size=(50, 3))
    if xyz_type == "float":
        pts = pts.astype(np.float32).astype(np.float64)
    cols = rng.integers(0, 256, size=(50, 3))
    p = tmp_path / "c.ply"
    write_ply(p, pts, cols, fmt=fmt, xyz_type=xyz_type)
    got, col = read_ply(p)
    assert got.dtype == np.float64 and got.shape == (50, 3)
    assert np.array_equal(got, pts)
    assert np.array_equal(col, cols / 255.0)


def test_read_ply_float_colours_and_none(tmp_path):
    pts = np.arange(12, dtype=np.float64).reshape(4, 3) / 7
    pts = pts.astype(np.float32).astype(np.float64)
    cols = np.array([[0.5, 0.25, 1.0]] * 4, np.float32)
    write_ply(tmp_path / "a.ply", pts, cols, fmt="binary_little_endian", col_type="float", faces=False)
    got, col = read_ply(tmp_path / "a.ply")
    assert np.array_equal(got, pts) and np.array_equal(col, cols.astype(np.float64))
    write_ply(tmp_path / "b.ply", pts, None, fmt="ascii", faces=False)
    got, col = read_ply(tmp_path / "b.ply")
    assert np.array_equal(got, pts) and col.shape == (0, 3)
    (tmp_path / "c.ply").write_bytes(b"not a ply\n")
    with pytest.raises(ValueError):
        read_ply(tmp_path / "c.ply")


def test_neural_points_from_ply(tmp_path):
    from epcq.epcq import NeuralPoints
    pts = np.random.default_rng(1).uniform(-1, 1, size=(30, 3)).astype(np.float32).astype(np.float64)
    write_ply(tmp_path / "pc.ply", pts, np.full((30, 3), 255), fmt="binary_little_endian")
    npts = NeuralPoints(pc_type="gt_pc", gt_pc_path=str(tmp_path / "pc.ply"), device="cpu")
    assert npts.xyz_data.dtype == torch.float32 and tuple(npts.xyz_data.shape) == (30, 3)
    assert np.array_equal(npts.xyz_data.numpy(), pts.astype(np.float32))
    assert torch.all(npts.rgb_data == 1.0) and tuple(npts.feat_data.shape) == (30, 32)
    assert not npts.xyz_data.requires_grad and npts.rgb_data.requires_grad


def test_neural_points_from_pointnerf_checkpoint(tmp_path):
    from epcq.epcq import NeuralPoints
    n = 40
    g = torch.Generator().manual_seed(2)
    ckpt = {
        "neural_points.xyz": torch.rand(n, 3, generator=g),
        "neural_points.points_color": torch.rand(1, n, 3, generator=g),
        "neural_points.points_dir": torch.rand(1, n, 3, generator=g),
        "neural_points.points_conf": torch.rand(1, n, 1, generator=g),
        "neural_points.points_embeding": torch.rand(1, n, 32, generator=g),
    }
    path = tmp_path / "scene.pth"
    torch.save(ckpt, path)
    npts = NeuralPoints(pc_type="pointnerf", pointnerf_path=str(path), device="cpu")
    assert torch.equal(npts.xyz_data, ckpt["neural_points.xyz"]) and not npts.xyz_data.requires_grad
    assert torch.equal(npts.rgb_data, ckpt["neural_points.points_color"][0]) and npts.rgb_data.requires_grad
    assert tuple(npts.density_data.shape) == (n, 1) and tuple(npts.feat_data.shape) == (n, 32)
    with pytest.raises(ValueError, match="Unsupported pointnerf_path"):
        NeuralPoints(pc_type="pointnerf", pointnerf_path=str(tmp_path / "scene.npz"), device="cpu")


def project(cam, p):
    pc = cam.w2c[:, :3].astype(np.float64) @ p + cam.w2c[:, 3]
    uvw = cam.K.astype(np.float64) @ pc
    return uvw[:2] / uvw[2], pc[2]


def test_nerf_synthetic_cameras_blender_to_opencv(tmp_path):
    # a Blender camera at (0, -4, 0) looking at the origin, up = +z: its
    # columns are the camera axes x = (1, 0, 0), y (up) = (0, 0, 1), z (back) =
    # (0, -1, 0) in world coordinates
    pose = [[1, 0, 0, 0], [0, 0, -1, -4], [0, 1, 0, 0], [0, 0, 0, 1]]
    angle = 2 * math.atan(0.5)  # focal = W
    tr = {"camera_angle_x": angle, "frames": [{"file_path": "./train/r_0", "transform_matrix": pose}]}
    path = tmp_path / "transforms_train.json"
    path.write_text(json.dumps(tr))
    (cam,) = nerf_synthetic_cameras(str(path), 800, 600)
    assert cam.name == "train-r_0"
    assert cam.fx == pytest.approx(800.0) and cam.cx == 400.0 and cam.cy == 300.0
    uv, z = project(cam, np.zeros(3))
    np.testing.assert_allclose(uv, [400.0, 300.0], atol=1e-4)
    assert z == pytest.approx(4.0)
    uv, _ = project(cam, np.array([0.1, 0.0, 0.0]))   # to the camera's right: u grows
    assert uv[0] == pytest.approx(400.0 + 800.0 * 0.1 / 4.0, rel=1e-5)
    uv, _ = project(cam, np.array([0.0, 0.0, 0.1]))   # up: OpenCV v decreases
    assert uv[1] == pytest.approx(300.0 - 800.0 * 0.1 / 4.0, rel=1e-5)
    np.testing.assert_allclose(cam.c2w[:3, 3], [0.0, -4.0, 0.0])
    np.testing.assert_allclose(cam.c2w @ np.append(cam.w2c, [[0, 0, 0, 1]], axis=0), np.eye(4), atol=1e-6)
