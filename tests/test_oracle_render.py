"""The oracle's volume-rendering epilogue (clamp_t, render and render's
backward, epcq/epcq.py:508-530) pinned by what the mathematics fixes: closed
forms for constant densities, the clamp rule's cases by hand, and the
backward against automatic differentiation of the forward formulas in
float64 (an independent derivation of the same derivative)."""
import numpy as np
import torch

from oracle import oracle


def test_clamp_t_cases():
    t = np.array([[1.0, 1.002, 1.004, 1.004, 1.02, 0.5, 0.503]], np.float32)
    d = oracle.clamp_t(t, vsize=0.004)
    exp = np.array([[np.float32(1.002) - np.float32(1.0),   # inside (0, 0.008]
                     np.float32(1.004) - np.float32(1.002),
                     0.004,                                   # zero gap -> vsize
                     0.004,                                   # 0.016 > 2 vsize -> vsize
                     0.004,                                   # negative -> vsize
                     np.float32(0.503) - np.float32(0.5),
                     0.004]], np.float32)                     # last sample -> vsize
    assert np.array_equal(d, exp)
    # S = 1: the reference slices its appended vsize from an empty interval
    # list, so there is no interval at all, and no sample contributes
    assert oracle.clamp_t(np.array([[3.0], [4.0]], np.float32)).shape == (2, 0)
    out = oracle.render(np.full((2, 1), 50.0, np.float32), np.full((2, 1, 3), 0.5, np.float32), True,
                        t=np.array([[3.0], [4.0]], np.float32))
    assert np.array_equal(out, np.ones((2, 3), np.float32))
    ref = torch_reference(torch.tensor([[3.0]]), torch.tensor([[50.0]]), torch.full((1, 1, 3), 0.5), True)
    assert np.array_equal(ref.numpy(), np.ones((1, 3), np.float32))


def test_render_constant_density_closed_form():
    R, S = 7, 16
    rng = np.random.default_rng(0)
    sig = rng.uniform(0.5, 80.0, size=R).astype(np.float32)
    delta = np.full((R, S), 0.01, np.float32)
    c = rng.uniform(0, 1, size=(R, 3)).astype(np.float32)
    rgb = np.repeat(c[:, None, :], S, axis=1)
    sigma = np.repeat(sig[:, None], S, axis=1)
    for white in (False, True):
        out = oracle.render(sigma, rgb, white_bg=white, delta=delta)
        a = np.exp(-sig.astype(np.float64) * 0.01) + 1e-10
        # sum_k (1 - e) a^k c + a^(S-1) bg, e = exp(-sigma delta)
        e = np.exp(-sig.astype(np.float64) * 0.01)
        weights = (1 - e)[:, None] * a[:, None] ** np.arange(S)[None, :]
        expect = weights.sum(1)[:, None] * c + (a ** (S - 1))[:, None] * (1.0 if white else 0.0)
        np.testing.assert_allclose(out, expect, rtol=2e-5, atol=2e-6)


def test_render_transparent_and_opaque():
    S = 5
    rgb = np.arange(3 * S, dtype=np.float32).reshape(1, S, 3) / 15
    delta = np.full((1, S), 0.004, np.float32)
    # no density: the background only
    assert np.array_equal(oracle.render(np.zeros((1, S), np.float32), rgb, True, delta=delta),
                          np.ones((1, 3), np.float32))
    # an opaque first sample: its colour, transmittance 1e-10 afterwards
    sigma = np.zeros((1, S), np.float32)
    sigma[0, 0] = 1e6
    out = oracle.render(sigma, rgb, False, delta=delta)
    np.testing.assert_allclose(out[0], rgb[0, 0], rtol=0, atol=1e-8)


def torch_reference(t, sigma, rgb, white_bg, vsize=0.004):
    """The reference's clamp_t and render (epcq/epcq.py:508-530) in torch."""
    t_dist = t[:, 1:] - t[:, :-1]
    t_dist = torch.cat([t_dist, torch.full_like(t_dist[:, :1], vsize)], dim=-1)
    mask = torch.logical_or(t_dist < 1e-8, t_dist > 2 * vsize).to(t.dtype)
    deltas = t_dist * (1.0 - mask) + mask * vsize
    opacity = 1 - torch.exp(-sigma * deltas)
    acc = torch.cumprod(1. - opacity + 1e-10, dim=-1)
    acc = torch.cat([torch.ones((opacity.shape[0], 1), dtype=t.dtype, device=t.device), acc[..., :-1]], dim=-1)
    w = opacity * acc
    color = torch.sum(w.unsqueeze(-1) * rgb, dim=-2)
    bg = torch.ones_like(color) if white_bg else torch.zeros_like(color)
    return color + acc[:, [-1]] * bg


def random_rays(R, S, seed):
    """Sample depths whose gaps fall inside (1e-8, 2 vsize], above it, and at
    zero (a quarter of the rows repeat their first depth), with densities
    and colours."""
    rng = np.random.default_rng(seed)
    gaps = rng.uniform(0.0, 0.012, size=(R, S))
    t = (rng.uniform(2.0, 6.0, size=(R, 1)) + np.cumsum(gaps, axis=1)).astype(np.float32)
    if S > 1:
        t[: R // 4, 1] = t[: R // 4, 0]
    sigma = rng.exponential(20.0, size=(R, S)).astype(np.float32)
    rgb = rng.uniform(0, 1, size=(R, S, 3)).astype(np.float32)
    return t, sigma, rgb


def test_render_matches_reference_formula():
    t, sigma, rgb = random_rays(500, 16, 1)
    for white in (False, True):
        out = oracle.render(sigma, rgb, white, t=t)
        ref = torch_reference(torch.from_numpy(t).double(), torch.from_numpy(sigma).double(),
                              torch.from_numpy(rgb).double(), white).numpy()
        np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6)
        # t and its clamped intervals give the same bits
        assert np.array_equal(out, oracle.render(sigma, rgb, white, delta=oracle.clamp_t(t)))


def test_render_backward_against_autograd():
    t, sigma, rgb = random_rays(300, 12, 2)
    g = np.random.default_rng(3).normal(size=(300, 3)).astype(np.float32)
    for white in (False, True):
        gs, gc = oracle.render_backward(sigma, rgb, g, white, t=t)
        st = torch.from_numpy(sigma).double().requires_grad_(True)
        ct = torch.from_numpy(rgb).double().requires_grad_(True)
        out = torch_reference(torch.from_numpy(t).double(), st, ct, white)
        out.backward(torch.from_numpy(g).double())
        np.testing.assert_allclose(gc, ct.grad.numpy(), rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(gs, st.grad.numpy(), rtol=2e-4, atol=2e-6)
