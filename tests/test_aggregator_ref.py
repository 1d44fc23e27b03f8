"""The fp32 torch restatement of the Aggregator (the checker of the gfx950
aggregator kernels), pinned by closed forms: the positional encoding's order,
the parameter names and shapes a reference state dict carries, and outputs
that follow by hand when every weight is zero."""
import math

import pytest
import torch

from aggregator_ref import aggregator_reference, random_inputs
from epcq.epcq import Aggregator, positional_encoding


def test_positional_encoding_closed_form():
    x = torch.tensor([[0.5, -1.0]])
    got = positional_encoding(x, 2)
    want = [math.sin(0.5), math.cos(0.5), math.sin(1.0), math.cos(1.0),
            math.sin(-1.0), math.cos(-1.0), math.sin(-2.0), math.cos(-2.0)]
    assert got.shape == (1, 8)
    assert torch.allclose(got[0], torch.tensor(want), rtol=0, atol=1e-7)
    got = positional_encoding(x, 2, ori=True)
    want = [0.5, -1.0, math.sin(0.5), math.sin(1.0), math.sin(-1.0), math.sin(-2.0),
            math.cos(0.5), math.cos(1.0), math.cos(-1.0), math.cos(-2.0)]
    assert torch.allclose(got[0], torch.tensor(want), rtol=0, atol=1e-7)


def test_parameters_match_the_reference_layout():
    want = {"block1.0": (256, 284), "block1.2": (256, 256), "block3.0": (256, 263), "block3.2": (256, 256),
            "alpha_branch.0": (1, 256), "color_branch.0": (128, 280), "color_branch.2": (128, 128),
            "color_branch.4": (128, 128), "color_branch.6": (3, 128)}
    sd = Aggregator().state_dict()
    assert set(sd) == {f"{k}.{p}" for k in want for p in ("weight", "bias")}
    for k, shape in want.items():
        assert tuple(sd[f"{k}.weight"].shape) == shape and tuple(sd[f"{k}.bias"].shape) == shape[:1]
        assert (sd[f"{k}.bias"] == 0).all()  # xavier weights, zero biases
    # 284 = 7 F + 60 inputs: features, PE(F, 3), PE(6 distance terms, 5)
    assert 32 + 32 * 3 * 2 + 6 * 5 * 2 == 284 and 256 + 7 == 263 and 256 + 3 * 4 * 2 == 280


def test_zero_weights_closed_form():
    """All weights zero: every pair's alpha is softplus(ba - 1), the colour
    sigmoid(cb4) * 1.002 - 0.001, and a sample's alpha is softplus(ba - 1)
    times the sum of its weights, sum_k inv_k / (sum_k inv_k + 1e-6)."""
    mod = Aggregator()
    with torch.no_grad():
        for m in mod.modules():
            if isinstance(m, torch.nn.Linear):
                m.weight.zero_()
                m.bias.fill_(0.3)
        mod.alpha_branch[0].bias.fill_(2.0)
        mod.color_branch[6].bias.copy_(torch.tensor([-1.0, 0.0, 2.0]))
        pc, rgb, d, feat, sp, valid, rays = random_inputs(50, 6, seed=1)
        alpha, col = aggregator_reference(mod, pc, rgb, d, feat, sp, valid, rays)
    inv = (1.0 / ((pc.double() - sp.double()[:, None]).norm(dim=-1) + 1e-6)) * valid.double()
    wsum = inv.sum(-1) / (inv.sum(-1) + 1e-6)
    want_a = math.log1p(math.exp(1.0)) * wsum
    assert torch.allclose(alpha.double(), want_a, rtol=1e-6, atol=1e-7)
    assert (alpha[~valid.any(-1)] == 0).all()
    want_c = torch.sigmoid(torch.tensor([-1.0, 0.0, 2.0], dtype=torch.float64)) * 1.002 - 0.001
    assert torch.allclose(col.double(), want_c.expand_as(col), rtol=0, atol=1e-6)


def test_forward_guards():
    mod = Aggregator()
    with pytest.raises(RuntimeError, match="inference forward"):
        mod(*([None] * 5), None, torch.zeros((2, 2), dtype=torch.bool), None)
    with torch.no_grad(), pytest.raises(ValueError, match="mask_out shape error"):
        mod(*([None] * 5), None, torch.zeros(4, dtype=torch.bool), None)
