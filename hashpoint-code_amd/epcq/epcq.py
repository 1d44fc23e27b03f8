"""Python API of the epcq searching + sampling path (mirrors epcq/epcq.py of
the reference: CameraSpec :15-63, RenderOptions :81-114, NeuralPointsGrads
:133-144, NeuralPoints :147-327, FeatureQuery :329-387, calculate_weight
:434-437).

NeuralPoints binds the ops of `epcq.csrc` exactly as the reference does
(`scatter`, `scatter_hashtable`, `quick_sampling`); `pts2plane` runs the fused
gfx950 build (projection, pixel keys, stable sort and hash table in one
pipeline) instead of torch matmuls + torch.sort + scatter_hashtable, with
pixel lists in the deterministic (pixel, index) order.  The volume-rendering
epilogue of the reference module (clamp_t :508-515, render :517-530) runs on
gfx950 kernels too, with a backward pass for training.  The Aggregator
(:406-507) between them keeps the reference's layers and parameter names and
runs its inference forward on the gfx950 matrix cores (bf16 inputs and
weights, fp32 accumulation); positional_encoding (:389-404) is the torch
form the reference uses.
"""
from dataclasses import dataclass
from typing import Optional, Union

import torch
import torch.nn as nn

from . import utils

_C = utils._get_c_extension()


@dataclass
class CameraSpec:
    c2w: torch.Tensor
    w2c: torch.Tensor
    K: torch.Tensor
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int

    def _to_cpp(self):
        spec = _C.CameraSpec()
        for name in ("c2w", "w2c", "fx", "fy", "cx", "cy", "width", "height"):
            setattr(spec, name, getattr(self, name))
        return spec

    @staticmethod
    def decomposeK(K: torch.Tensor):
        """(fx, fy, cx, cy) of an intrinsics matrix."""
        return K[0, 0], K[1, 1], K[0, 2], K[1, 2]

    def pixelidx2hwidx(self, pixelidx):
        return pixelidx // self.width, pixelidx % self.width

    def hwidx2pixelidx(self, h_idx, w_idx):
        return h_idx * self.width + w_idx


@dataclass
class RenderOptions:
    kernel_size: int = 5
    t_intvl: float = 0.01
    radius_threshold: float = 0.016
    sigma_threshold: float = 1e-4
    background_brightness: float = 1.0
    stop_threshold: float = 0.001
    num_sample_points_per_ray: int = 16
    num_point_cloud_per_sp: int = 4
    num_point_per_ray: int = 256
    sdf2weight_gaussian_alpha: float = 0.2
    sdf2weight_gaussian_gamma: float = 0.9
    sparsity_loss_w: float = 5e-4
    topK: int = 8

    def _to_cpp(self):
        opt = _C.RenderOptions()
        for name in _C.RenderOptions._FIELDS:
            setattr(opt, name, getattr(self, name))
        return opt


@dataclass
class NeuralPointsGrads:
    grad_sigma: torch.Tensor
    grad_sh: torch.Tensor
    mask_out: torch.Tensor

    def _to_cpp(self):
        grads = _C.NeuralPointsGrads()
        grads.grad_sigma_out = self.grad_sigma
        grads.grad_sh_out = self.grad_sh
        grads.mask_out = self.mask_out
        return grads


class NeuralPoints(nn.Module):
    """A point cloud with its per-point attributes, and the search ops.

    pc_type 'pc' wraps the given `points` [N,3]; 'pointnerf' loads a Point-NeRF
    checkpoint (tensors only, torch.load(weights_only=True)); 'gt_pc' reads a
    PLY point cloud (with open3d when that package is installed, else with
    epcq.io.read_ply, which returns the same arrays)."""

    def __init__(self,
                 basis_dim: int = 9,
                 device: Union[torch.device, str] = "cuda:0",
                 pc_type: str = "pointnerf",
                 gt_pc_path: Optional[str] = None,
                 pointnerf_path: str = "pointnerf_ckpt/nerf_synthesis/lego.pth",
                 alpha: float = 5.0,
                 points=None):
        super().__init__()
        if pc_type not in ("pointnerf", "gt_pc", "pc"):
            raise AssertionError(f"unknown pc_type {pc_type!r}")
        self.pc_type = pc_type
        if pc_type == "pointnerf":
            if not str(pointnerf_path).endswith(".pth"):
                raise ValueError("Unsupported pointnerf_path: {}".format(pointnerf_path))
            ckpt = torch.load(pointnerf_path, map_location="cpu", weights_only=True)
            xyz = ckpt["neural_points.xyz"]
            attrs = {
                "rgb_data": ckpt["neural_points.points_color"].squeeze(0),
                "dir_data": ckpt["neural_points.points_dir"].squeeze(0),
                "density_data": ckpt["neural_points.points_conf"].squeeze(0),
                "feat_data": ckpt["neural_points.points_embeding"].squeeze(0),
            }
            self.xyz_data = nn.Parameter(xyz, requires_grad=False)
            for name, value in attrs.items():
                setattr(self, name, nn.Parameter(value, requires_grad=True))
        elif pc_type == "gt_pc":
            if gt_pc_path is None:
                raise ValueError("gt_pc_path is required")
            import numpy as np
            try:
                import open3d as o3d
            except ImportError:
                o3d = None
            if o3d is not None:
                pcd = o3d.io.read_point_cloud(gt_pc_path)
                pts_np, cols_np = np.asarray(pcd.points), np.asarray(pcd.colors)
            else:  # the same arrays from the package's own PLY reader
                from .io import read_ply
                pts_np, cols_np = read_ply(gt_pc_path)
            pts = torch.tensor(pts_np, dtype=torch.float32, device=device)
            cols = torch.tensor(cols_np, dtype=torch.float32, device=device)
            self.basis_dim = basis_dim
            self.xyz_data = nn.Parameter(pts, requires_grad=False)
            self.rgb_data = nn.Parameter(cols, requires_grad=True)
            self.dir_data = nn.Parameter(torch.zeros_like(pts), requires_grad=True)
            self.density_data = nn.Parameter(torch.zeros(pts.shape[0], 1, device=device), requires_grad=True)
            self.feat_data = nn.Parameter(torch.zeros(pts.shape[0], 32, device=device), requires_grad=True)
        else:
            if points is None:
                raise AssertionError("pc_type='pc' needs points")
            self.xyz_data = nn.Parameter(points, requires_grad=False)
            self.rgb_data = nn.Parameter(torch.zeros_like(points), requires_grad=False)
            self.dir_data = nn.Parameter(torch.zeros_like(points), requires_grad=False)
            self.density_data = nn.Parameter(torch.zeros((points.shape[0], 1)), requires_grad=False)
            self.feat_data = nn.Parameter(torch.zeros_like(points), requires_grad=False)
        self.capacity = self.xyz_data.shape[0]
        self.multi_bins_rasterize = _C.__dict__["scatter"]
        self.rasterize_hashtable = _C.__dict__["scatter_hashtable"]
        self.quick_sampling = _C.__dict__["quick_sampling"]

    def _to_cpp(self):
        spec = _C.NeuralPointsSpec()
        spec.xyz_data = self.xyz_data
        return spec

    # -- torch forms of the projection (epcq.py:251-277); the fused build in
    #    pts2plane evaluates the same formulas in a fixed operation order
    def _world2cam(self, cam: CameraSpec, pc_xyz: torch.Tensor = None):
        xyz = self.xyz_data if pc_xyz is None else pc_xyz
        pts_c = (cam.w2c[:3, :3] @ xyz.t() + cam.w2c[:3, 3:4]).t()
        return pts_c, pts_c[:, 2:3]

    def _cam2image(self, pts_c: torch.Tensor, cam: CameraSpec):
        scr = cam.K @ pts_c.t()
        scr = scr / scr[2:3, :]
        return scr[0], scr[1], scr[0].to(torch.int32), scr[1].to(torch.int32)

    def _convetpixel(self, screen_x, screen_y, hw):
        key = screen_x + screen_y * hw[1]
        inside = (screen_x >= 0) & (screen_x < hw[1]) & (screen_y >= 0) & (screen_y < hw[0])
        return torch.where(inside, key, torch.full_like(key, hw[0] * hw[1]))

    def _uv2ray_index(self, uv, width):
        return (uv[:, 1] * width + uv[:, 0]).to(torch.int32)

    def pts2plane(self, cam: CameraSpec, opt: RenderOptions, pc_xyz: torch.Tensor = None, sort: bool = False,
                  workspace: Optional[torch.Tensor] = None):
        """Per-pixel point lists of the view: (sorted_point_idx_list int32
        [N], sorted_depth_list float32 [N], hash_table int32 [H*W, 2])."""
        xyz = self.xyz_data if pc_xyz is None else pc_xyz
        dev = xyz.device
        w2c = cam.w2c.to(device=dev, dtype=torch.float32).contiguous()
        K = cam.K.to(device=dev, dtype=torch.float32).contiguous()
        return _C.build_hashtable(xyz.detach(), w2c, K, [int(cam.height), int(cam.width)], bool(sort),
                                  workspace=workspace)

    def render_image_hashtable_topk(self, cam: CameraSpec,
                                    opt: RenderOptions,
                                    sorted_point_idx_list: torch.Tensor,
                                    sorted_depth_list: torch.Tensor,
                                    hashtable: torch.Tensor,
                                    save_variables_for_debug: bool = False,
                                    mode: str = "compat", cull: bool = True,
                                    training_ray_index: Optional[torch.Tensor] = None):
        """(output_sp_xyz, output_pc_idx, output_sp_t, mask_out) of every
        pixel; `save_variables_for_debug` is accepted and ignored.

        With a non-empty `training_ray_index` ([R, >=2] integer (u, v, ...)
        rows, as the reference's data loader builds them) only those rays are
        sampled, and the outputs are [R,S,3], [R,S,Kc], [R,S], [R,S]: what the
        reference's renderer obtains by indexing the full-image outputs with
        [v, u] (opt/render_image.py:82-88), without sampling the rest."""
        rays = None
        if training_ray_index is not None and len(training_ray_index) != 0:
            tri = training_ray_index.to(device=hashtable.device)
            rays = (tri[:, 1].to(torch.int64) * int(cam.width) + tri[:, 0].to(torch.int64)).to(torch.int32)
            rays = rays.contiguous()
        return self.quick_sampling(self._to_cpp(), cam._to_cpp(), opt._to_cpp(), sorted_point_idx_list,
                                   hashtable, mode=mode, cull=cull, rays=rays)

    def knn_sampling(self, cam: CameraSpec, opt: RenderOptions, sorted_point_idx_list: torch.Tensor,
                     hashtable: torch.Tensor, rays: Optional[torch.Tensor] = None,
                     radius: Optional[torch.Tensor] = None, mode: str = "knn", feat: Optional[torch.Tensor] = None,
                     s_max: Optional[int] = None, cull: bool = True, r_max: Optional[float] = None):
        """Compact K-nearest sampling (see epcq.csrc.knn_sampling)."""
        return _C.knn_sampling(self._to_cpp(), cam._to_cpp(), opt._to_cpp(), sorted_point_idx_list, hashtable,
                               rays=rays, radius=radius, mode=mode, s_max=s_max, feat=feat, cull=cull, r_max=r_max)


class FeatureQuery(nn.Module):
    """Gathers the attributes of each live sample's neighbours
    (epcq.py:329-387)."""

    def generate_rays(self, cam: CameraSpec):
        """Unnormalised pixel-corner ray directions [H,W,3] and origins."""
        dev = cam.c2w.device
        W, H = int(cam.width), int(cam.height)
        x, y = torch.meshgrid(torch.arange(W, dtype=torch.float32, device=dev),
                              torch.arange(H, dtype=torch.float32, device=dev), indexing="xy")
        cx = (x - cam.cx) / cam.fx
        cy = (y - cam.cy) / cam.fy
        R = cam.c2w[:3, :3]
        rays_d = torch.stack([R[i, 0] * cx + R[i, 1] * cy + R[i, 2] for i in range(3)], dim=-1)
        rays_o = cam.c2w[:3, 3].expand_as(rays_d)
        return rays_o, rays_d

    def forward(self, neuralpoints: NeuralPoints, output_sp_xyz: torch.Tensor, output_pc_idx: torch.Tensor,
                mask_out: torch.Tensor, cam: CameraSpec, ray_index_uv: torch.Tensor = None):
        S, Kc = output_pc_idx.shape[-2:]
        live = mask_out.reshape(-1)
        sp = output_sp_xyz.reshape(-1, 3)[live]
        idx = output_pc_idx.reshape(-1, Kc)[live].long()
        valid = idx != -1
        safe = idx.clamp_min(0)

        def gather(attr):
            out = attr[safe]
            return torch.where(valid.unsqueeze(-1), out, torch.zeros((), dtype=out.dtype, device=out.device))

        pc_xyz = gather(neuralpoints.xyz_data)
        pc_rgb = gather(neuralpoints.rgb_data)
        pc_dir = gather(neuralpoints.dir_data)
        pc_feat = gather(neuralpoints.feat_data)
        _, rays_d = self.generate_rays(cam)
        if ray_index_uv is not None and len(ray_index_uv) != 0:
            rays_d = rays_d[ray_index_uv[:, 1], ray_index_uv[:, 0]]
            rays_d = rays_d[:, None, :].expand(-1, S, -1).reshape(-1, 3)
        else:
            rays_d = rays_d[:, :, None, :].expand(-1, -1, S, -1).reshape(-1, 3)
        return pc_xyz, pc_rgb, pc_dir, pc_feat, sp, valid, rays_d[live]


def calculate_weight(dist: torch.Tensor, mask: torch.Tensor = None):
    """Inverse-distance weights (Aggregator.calculate_weight, epcq.py:434-437)."""
    inv = 1.0 / (dist + 1e-6)
    if mask is not None:
        inv = inv * mask.float()
    return inv / (inv.sum(dim=-1, keepdim=True) + 1e-6)


def positional_encoding(positions: torch.Tensor, freqs: int, ori: bool = False):
    """[sin(x_i 2^j), cos(x_i 2^j)] pairs over inputs i and frequencies
    j < freqs, i major (epcq.py:389-404); with ori, [x, sin(all), cos(all)]."""
    bands = 2.0 ** torch.arange(freqs, dtype=torch.float32, device=positions.device)
    D = positions.shape[-1]
    pts = (positions[..., None] * bands).reshape(*positions.shape[:-1], freqs * D)
    if ori:
        return torch.cat([positions, torch.sin(pts), torch.cos(pts)], dim=-1)
    return torch.stack([torch.sin(pts), torch.cos(pts)], dim=-1).reshape(*pts.shape[:-1], 2 * freqs * D)


class Aggregator(nn.Module):
    """Point aggregator of the reference renderer (epcq.py:406-507) with the
    same layers, initialisation and parameter names, so its state dicts load
    as they are.  forward takes the FeatureQuery outputs of the live samples
    (pc_xyz / pc_rgb / pc_dir [M,K,3], pc_feat [M,K,32], output_sp_xyz [M,3],
    valid_sp_mask_out [M,K], rays_d [M,3]) and returns alphas and colours
    scattered to mask_out's shape: [H*W, S] and [H*W, S, 3] for an image mask
    [H,W,S], [R, S] and [R, S, 3] for a ray mask [R,S].

    The forward runs on gfx950 (epcq_aggregator_forward): the point MLP of
    every (sample, neighbour) pair and the colour MLP on the bf16 matrix
    cores with fp32 accumulation, the inverse-distance aggregation in fp32;
    results agree with the fp32 module within bf16 rounding.  It is an
    inference forward: with gradients enabled on the parameters it raises
    (training is not part of this package)."""

    def __init__(self) -> None:
        super().__init__()
        self.block1 = nn.Sequential(nn.Linear(284, 256), nn.LeakyReLU(inplace=True),
                                    nn.Linear(256, 256), nn.LeakyReLU(inplace=True))
        self.block3 = nn.Sequential(nn.Linear(263, 256), nn.LeakyReLU(inplace=True),
                                    nn.Linear(256, 256), nn.LeakyReLU(inplace=True))
        self.alpha_branch = nn.Sequential(nn.Linear(256, 1))
        self.alpha_act = nn.Softplus()
        self.color_branch = nn.Sequential(nn.Linear(280, 128), nn.LeakyReLU(inplace=True),
                                          nn.Linear(128, 128), nn.LeakyReLU(inplace=True),
                                          nn.Linear(128, 128), nn.LeakyReLU(inplace=True),
                                          nn.Linear(128, 3))
        self.color_act = nn.Sigmoid()
        for mod in self.modules():
            if isinstance(mod, nn.Linear):
                nn.init.xavier_uniform_(mod.weight)
                nn.init.constant_(mod.bias, 0.0)
        self.chunksize = 20000  # kept for compatibility; the kernels need no chunking
        self._packed = None
        self._packed_key = None

    def calculate_weight(self, dist, mask=None):
        return calculate_weight(dist, mask)

    def _layers(self):
        lin = lambda m: (m.weight, m.bias)  # noqa: E731
        return {"w1": lin(self.block1[0]), "w2": lin(self.block1[2]), "w3": lin(self.block3[0]),
                "w4": lin(self.block3[2]), "wa": lin(self.alpha_branch[0]), "c1": lin(self.color_branch[0]),
                "c2": lin(self.color_branch[2]), "c3": lin(self.color_branch[4]), "c4": lin(self.color_branch[6])}

    def packed_params(self):
        """bf16 / f32 device copies of the parameters, rebuilt when a
        parameter changes (its storage or version)."""
        key = tuple((p.data_ptr(), p._version, p.device) for p in self.parameters())
        if self._packed is None or key != self._packed_key:
            self._packed = _C.pack_aggregator(self._layers())
            self._packed_key = key
        return self._packed

    def forward(self, pc_xyz, pc_rgb, pc_dir, pc_feat, output_sp_xyz, valid_sp_mask_out, mask_out, rays_d):
        if torch.is_grad_enabled() and any(p.requires_grad for p in self.parameters()):
            raise RuntimeError("Aggregator: the gfx950 kernels implement the inference forward; "
                               "call it under torch.no_grad()")
        if mask_out.dim() not in (2, 3):
            raise ValueError("mask_out shape error")
        alpha, rgb = _C.aggregator_forward(self.packed_params(), pc_xyz.contiguous(), pc_rgb.contiguous(),
                                           pc_dir.contiguous(), pc_feat.contiguous(), output_sp_xyz.contiguous(),
                                           valid_sp_mask_out.contiguous(), rays_d.contiguous())
        live = mask_out.reshape(-1)
        rows = mask_out.shape[0] * mask_out.shape[1] if mask_out.dim() == 3 else mask_out.shape[0]
        S = mask_out.shape[-1]
        fa = torch.zeros(live.numel(), dtype=torch.float32, device=alpha.device).masked_scatter_(live, alpha)
        fc = torch.zeros((live.numel(), 3), dtype=torch.float32, device=rgb.device)
        fc.masked_scatter_(live[:, None].expand(-1, 3), rgb)
        return fa.view(rows, S), fc.view(rows, S, 3)


def clamp_t(t_val: torch.Tensor, vsize: float = 0.004):
    """Sample intervals of sample depths t_val [R,S] (epcq.py:508-515): the
    difference to the next sample, vsize for the last, and vsize wherever the
    difference is below 1e-8 or above 2 * vsize."""
    return _C.clamp_t(t_val.contiguous(), vsize)


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, deltas, sigmas, rgbs, white_bg):
        deltas = deltas.contiguous()
        sigmas = sigmas.contiguous()
        rgbs = rgbs.contiguous()
        out = _C.render_rays(sigmas, rgbs, white_bg=white_bg, deltas=deltas)
        ctx.save_for_backward(deltas, sigmas, rgbs)
        ctx.white_bg = white_bg
        return out

    @staticmethod
    def backward(ctx, grad_out):
        deltas, sigmas, rgbs = ctx.saved_tensors
        gs, gc = _C.render_rays_backward(sigmas, rgbs, grad_out.contiguous(), white_bg=ctx.white_bg, deltas=deltas)
        return None, gs, gc, None


def render(deltas: torch.Tensor, sigmas: torch.Tensor, rgbs: torch.Tensor, white_bg: bool = False):
    """Ray colours [R,3] by alpha compositing (epcq.py:517-530): opacity
    1 - exp(-sigma * delta), transmittance the exclusive running product of
    (1 - opacity + 1e-10), background weighted by the transmittance before
    the last sample.  Differentiable in sigmas [R,S] and rgbs [R,S,3] (the
    intervals deltas [R,S] come from the samples and take no gradient).
    Intervals of one sample per ray are empty ([R,0], see clamp_t): no
    sample contributes and the colour is the background, as in the
    reference."""
    if deltas.dim() == 2 and deltas.shape[1] == 0:
        base = sigmas.new_ones((sigmas.shape[0], 3)) if white_bg else sigmas.new_zeros((sigmas.shape[0], 3))
        return base + 0.0 * (sigmas.sum() + rgbs.sum())  # keeps the graph: zero gradients
    return _Render.apply(deltas, sigmas, rgbs, bool(white_bg))
