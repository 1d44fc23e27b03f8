"""The volume-rendering epilogue on the GPU (clamp_t, render, and render's
backward through torch autograd) against the oracle bit for bit (same
operation order, same exponential) and against the reference's formulas in
torch fp32 within 1e-5 relative (torch's exp, cumprod and sum round
differently)."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from epcq.epcq import clamp_t, render
from oracle import oracle
from scenes import bits
from test_oracle_render import random_rays, torch_reference

pytestmark = pytest.mark.gpu


# S up to 51 takes the LDS-staged kernels (64-ray blocks), beyond that the
# direct ones; R = 4097 and 700 leave a partial last block.
SHAPES = [(1, 1), (1000, 16), (4097, 7), (300, 64), (700, 51), (700, 52), (128, 3)]


@pytest.mark.parametrize("R,S", SHAPES)
@pytest.mark.parametrize("white", [False, True])
def test_render_forward(dev, R, S, white):
    t, sigma, rgb = random_rays(R, S, 10 + S)
    td, sd, cd = (torch.from_numpy(a).to(dev) for a in (t, sigma, rgb))
    print(f"render R={R} S={S} white={white}", flush=True)
    delta = clamp_t(td)
    assert np.array_equal(bits(delta.cpu().numpy()), bits(oracle.clamp_t(t)))
    out = render(delta, sd, cd, white_bg=white)
    fused = C.render_rays(sd, cd, white_bg=white, t_val=td)
    torch.cuda.synchronize()
    ref = oracle.render(sigma, rgb, white, t=t)
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))
    assert np.array_equal(bits(fused.cpu().numpy()), bits(ref))
    tref = torch_reference(td, sd, cd, white)
    np.testing.assert_allclose(out.cpu().numpy(), tref.cpu().numpy(), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("white", [False, True])
def test_render_backward(dev, white):
    R, S = 2000, 16
    t, sigma, rgb = random_rays(R, S, 5)
    g = np.random.default_rng(6).normal(size=(R, 3)).astype(np.float32)
    td = torch.from_numpy(t).to(dev)
    sd = torch.from_numpy(sigma).to(dev).requires_grad_(True)
    cd = torch.from_numpy(rgb).to(dev).requires_grad_(True)
    print(f"render backward white={white}", flush=True)
    out = render(clamp_t(td), sd, cd, white_bg=white)
    out.backward(torch.from_numpy(g).to(dev))
    torch.cuda.synchronize()
    gs, gc = oracle.render_backward(sigma, rgb, g, white, t=t)
    assert np.array_equal(bits(sd.grad.cpu().numpy()), bits(gs))
    assert np.array_equal(bits(cd.grad.cpu().numpy()), bits(gc))
    # against autograd of the reference formulas (float64 on the CPU)
    st = torch.from_numpy(sigma).double().requires_grad_(True)
    ct = torch.from_numpy(rgb).double().requires_grad_(True)
    torch_reference(torch.from_numpy(t).double(), st, ct, white).backward(torch.from_numpy(g).double())
    np.testing.assert_allclose(cd.grad.cpu().numpy(), ct.grad.numpy(), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(sd.grad.cpu().numpy(), st.grad.numpy(), rtol=2e-4, atol=2e-6)


@pytest.mark.parametrize("R,S", SHAPES)
@pytest.mark.parametrize("fused", [False, True])
def test_render_backward_shapes(dev, R, S, fused):
    """render's backward with t (clamp_t fused) or with the intervals, on both
    kernel forms, bit for bit against the oracle."""
    if S == 1 and not fused:
        pytest.skip("clamp_t of one sample gives no interval column")
    t, sigma, rgb = random_rays(R, S, 30 + S)
    g = np.random.default_rng(31 + S).normal(size=(R, 3)).astype(np.float32)
    td, sd, cd, gd = (torch.from_numpy(a).to(dev) for a in (t, sigma, rgb, g))
    print(f"render backward R={R} S={S} fused={fused}", flush=True)
    if fused:
        gs, gc = C.render_rays_backward(sd, cd, gd, white_bg=True, t_val=td)
        ws, wc = oracle.render_backward(sigma, rgb, g, True, t=t)
    else:
        delta = oracle.clamp_t(t)
        gs, gc = C.render_rays_backward(sd, cd, gd, white_bg=True, deltas=torch.from_numpy(delta).to(dev))
        ws, wc = oracle.render_backward(sigma, rgb, g, True, delta=delta)
    torch.cuda.synchronize()
    assert np.array_equal(bits(gs.cpu().numpy()), bits(ws))
    assert np.array_equal(bits(gc.cpu().numpy()), bits(wc))


def test_render_argument_errors(dev):
    z = torch.zeros((4, 3), device=dev)
    with pytest.raises(RuntimeError, match="exactly one"):
        C.render_rays(z, torch.zeros((4, 3, 3), device=dev))
    with pytest.raises(RuntimeError, match="rgbs must have shape"):
        C.render_rays(z, torch.zeros((4, 3), device=dev), deltas=z)


@pytest.mark.parametrize("fused", [False, True])
def test_render_unaligned_rows(dev, fused):
    """Rows whose arrays start 4 bytes past a 16-byte boundary (views into a
    larger buffer) take the kernels that load one float at a time; the
    results are the same bit for bit."""
    R, S = 3000, 16
    t, sigma, rgb = random_rays(R, S, 41)
    g = np.random.default_rng(42).normal(size=(R, 3)).astype(np.float32)

    def shifted(a):
        buf = torch.empty(a.size + 1, device=dev)
        buf[1:] = torch.from_numpy(a.reshape(-1)).to(dev)
        return buf[1:].view(a.shape)

    td, sd, cd = shifted(t), shifted(sigma), shifted(rgb)
    assert sd.data_ptr() % 16 == 4
    gd = torch.from_numpy(g).to(dev)
    print(f"render unaligned fused={fused}", flush=True)
    if fused:
        out = C.render_rays(sd, cd, white_bg=True, t_val=td)
        gs, gc = C.render_rays_backward(sd, cd, gd, white_bg=True, t_val=td)
        ref = oracle.render(sigma, rgb, True, t=t)
        ws, wc = oracle.render_backward(sigma, rgb, g, True, t=t)
    else:
        delta = oracle.clamp_t(t)
        dd = shifted(delta)
        out = C.render_rays(sd, cd, white_bg=True, deltas=dd)
        gs, gc = C.render_rays_backward(sd, cd, gd, white_bg=True, deltas=dd)
        ref = oracle.render(sigma, rgb, True, delta=delta)
        ws, wc = oracle.render_backward(sigma, rgb, g, True, delta=delta)
    torch.cuda.synchronize()
    assert np.array_equal(bits(out.cpu().numpy()), bits(ref))
    assert np.array_equal(bits(gs.cpu().numpy()), bits(ws))
    assert np.array_equal(bits(gc.cpu().numpy()), bits(wc))
