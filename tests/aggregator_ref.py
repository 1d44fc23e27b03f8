"""fp32 torch restatement of the reference's Aggregator.forward
(epcq/epcq.py:439-507) on the compacted samples, without its chunk loop (rows
are independent), and random inputs of the FeatureQuery shapes.  Test
infrastructure: the checker of the gfx950 aggregator kernels."""
import torch

from epcq.epcq import calculate_weight, positional_encoding


def aggregator_reference(mod, pc_xyz, pc_rgb, pc_dir, pc_feat, sp_xyz, valid, rays_d):
    """(alpha [M], rgb [M,3]) of module `mod` in fp32, as the reference
    computes them: the point MLPs on the valid pairs only, zeros elsewhere,
    inverse-distance sums per sample, then the colour branch."""
    M, K = valid.shape
    F = pc_feat.shape[-1]
    dist = pc_xyz - sp_xyz[:, None, :]
    distpow = pc_xyz.pow(2) - sp_xyz[:, None, :].pow(2)
    w = calculate_weight(torch.norm(dist, dim=-1), valid)
    v = valid.reshape(-1)
    df = torch.cat([dist, distpow], dim=-1).reshape(-1, 6)[v]
    f = pc_feat.reshape(-1, F)[v]
    h = mod.block1(torch.cat([f, positional_encoding(f, 3), positional_encoding(df, 5)], dim=-1))
    rd = rays_d[:, None, :].expand(-1, K, -1).reshape(-1, 3)[v]
    pd = pc_dir.reshape(-1, 3)[v]
    pr = pc_rgb.reshape(-1, 3)[v]
    h = mod.block3(torch.cat([h, pr, pd - rd, torch.sum(pd * rd, dim=-1, keepdim=True)], dim=-1))
    a = mod.alpha_act(mod.alpha_branch(h) - 1)
    ap = torch.zeros((M * K, 1), dtype=torch.float32, device=a.device)
    ap[v] = a
    alpha = torch.sum(ap.view(M, K) * w, dim=-1)
    fc = torch.zeros((M * K, h.shape[1]), dtype=torch.float32, device=h.device)
    fc[v] = h
    feat = torch.sum(fc.view(M, K, -1) * w.unsqueeze(-1), dim=-2)
    x = torch.cat([feat, positional_encoding(rays_d, 4)], dim=-1)
    rgb = mod.color_act(mod.color_branch(x)) * (1 + 2 * 0.001) - 0.001
    return alpha, rgb


def random_inputs(M, K, seed=0, p_valid=0.8, empty_every=13, F=32):
    """Inputs of the FeatureQuery shapes: neighbours within ~0.01 of their
    sample, unit ray directions, features ~N(0, 1); a fraction of invalid
    neighbours (zero-filled, as the reference's gathers leave them) and every
    empty_every-th sample without any."""
    g = torch.Generator().manual_seed(seed)
    sp = torch.rand((M, 3), generator=g) * 2 - 1
    pc = sp[:, None, :] + torch.randn((M, K, 3), generator=g) * 0.01
    valid = torch.rand((M, K), generator=g) < p_valid
    if empty_every:
        valid[::empty_every] = False
    rgb = torch.rand((M, K, 3), generator=g)
    d = torch.randn((M, K, 3), generator=g)
    d = d / d.norm(dim=-1, keepdim=True)
    feat = torch.randn((M, K, F), generator=g)
    rays = torch.randn((M, 3), generator=g)
    rays = rays / rays.norm(dim=-1, keepdim=True)
    z = ~valid[..., None]
    pc = pc.masked_fill(z, 0.0)
    rgb = rgb.masked_fill(z, 0.0)
    d = d.masked_fill(z, 0.0)
    feat = feat.masked_fill(z, 0.0)
    return pc, rgb, d, feat, sp, valid, rays
