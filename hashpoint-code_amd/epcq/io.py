"""Point-cloud and camera loaders of the reference's data path, without its
optional dependencies.

read_ply             the PLY files NeuralPoints(pc_type='gt_pc') reads with
                     open3d.io.read_point_cloud (epcq/epcq.py:178-207): vertex
                     positions as float64 [N,3] and colours scaled to [0, 1]
                     as float64 [N,3] (empty [0,3] when the file has none),
                     from ASCII or binary (either byte order) files.
nerf_synthetic_cameras
                     the NeRF-synthetic transforms_*.json cameras converted
                     from Blender to OpenCV axes as the reference's loader
                     does (dataset/nerfsynthesis.py:96-107): c2w = pose *
                     diag(1, -1, -1, 1), w2c its inverse, K from
                     camera_angle_x with the principal point at the centre.
"""
import json
import math
import os
from dataclasses import dataclass

import numpy as np

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
    "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _parse_header(f):
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []  # [name, count, [(prop, dtype) | (prop, ('list', count_t, item_t))]]
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header without end_header")
        words = line.decode("ascii", "replace").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "format":
            fmt = words[1]
        elif words[0] == "element":
            elements.append([words[1], int(words[2]), []])
        elif words[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if words[1] == "list":
                elements[-1][2].append((words[4], ("list", _PLY_TYPES[words[2]], _PLY_TYPES[words[3]])))
            else:
                elements[-1][2].append((words[2], _PLY_TYPES[words[1]]))
        elif words[0] == "end_header":
            break
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, elements


def _read_element_binary(f, count, props, endian):
    if all(not isinstance(t, tuple) for _, t in props):
        dt = np.dtype([(name, endian + t) for name, t in props])
        data = np.frombuffer(f.read(dt.itemsize * count), dtype=dt, count=count)
        return data
    # rows with list properties: read one row at a time
    rows = []
    for _ in range(count):
        row = {}
        for name, t in props:
            if isinstance(t, tuple):
                n = int(np.frombuffer(f.read(np.dtype(t[1]).itemsize), endian + t[1])[0])
                row[name] = np.frombuffer(f.read(np.dtype(t[2]).itemsize * n), endian + t[2])
            else:
                row[name] = np.frombuffer(f.read(np.dtype(t).itemsize), endian + t)[0]
        rows.append(row)
    return rows


def read_ply(path):
    """(points float64 [N,3], colours float64 [N,3] in [0, 1] or [0,3])."""
    with open(path, "rb") as f:
        fmt, elements = _parse_header(f)
        vertex = None
        if fmt == "ascii":
            body = f.read().decode("ascii", "replace").split("\n")
            at = 0
            for name, count, props in elements:
                rows = []
                for _ in range(count):
                    while at < len(body) and not body[at].strip():
                        at += 1
                    rows.append(body[at].split())
                    at += 1
                if name == "vertex":
                    cols = {}
                    simple = [p for p in props if not isinstance(p[1], tuple)]
                    if len(simple) != len(props):
                        raise ValueError("PLY vertex element with list properties")
                    arr = np.array(rows, dtype=np.float64).reshape(count, len(props))
                    for j, (pname, t) in enumerate(props):
                        cols[pname] = arr[:, j].astype(t)
                    vertex = cols
                    break
        else:
            endian = "<" if fmt == "binary_little_endian" else ">"
            for name, count, props in elements:
                data = _read_element_binary(f, count, props, endian)
                if name == "vertex":
                    vertex = {pname: np.asarray(data[pname]) for pname, _ in props}
                    break
    if vertex is None:
        raise ValueError("PLY file without a vertex element")
    pts = np.stack([vertex[c].astype(np.float64) for c in ("x", "y", "z")], axis=1)
    if all(c in vertex for c in ("red", "green", "blue")):
        col = np.stack([vertex[c] for c in ("red", "green", "blue")], axis=1)
        if np.issubdtype(col.dtype, np.integer):
            col = col.astype(np.float64) / float(np.iinfo(col.dtype).max)
        else:
            col = col.astype(np.float64)
    else:
        col = np.zeros((0, 3), np.float64)
    return pts, col


@dataclass
class NerfCamera:
    """One camera of a NeRF-synthetic split in OpenCV axes."""
    name: str
    c2w: np.ndarray  # float32 [4,4]
    w2c: np.ndarray  # float32 [3,4]
    K: np.ndarray    # float32 [3,3]
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int


def nerf_synthetic_cameras(transforms, width, height):
    """Cameras of a transforms_*.json (a path or the parsed dict) for images
    of width x height pixels (dataset/nerfsynthesis.py:96-107)."""
    if isinstance(transforms, (str, os.PathLike)):
        with open(transforms) as f:
            transforms = json.load(f)
    angle = float(transforms["camera_angle_x"])
    focal = 0.5 * width / math.tan(0.5 * angle)
    K = np.asarray([[focal, 0, width / 2], [0, focal, height / 2], [0, 0, 1]], np.float32)
    flip = np.diag(np.array([1.0, -1.0, -1.0, 1.0], dtype=np.float32))
    cams = []
    for frame in transforms["frames"]:
        c2w = np.array(frame["transform_matrix"], dtype=np.float32) @ flip  # Blender to OpenCV
        w2c = np.linalg.inv(c2w)[:3, :4].astype(np.float32)
        name = "-".join(str(frame.get("file_path", "")).split("/")[1:])
        cams.append(NerfCamera(name, c2w, w2c, K.copy(), float(K[0, 0]), float(K[1, 1]), float(K[0, 2]),
                               float(K[1, 2]), int(width), int(height)))
    return cams
