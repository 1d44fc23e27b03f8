"""The gfx950 aggregator kernels (bf16 matrix cores, fp32 accumulation)
against the fp32 torch restatement of Aggregator.forward.  Tolerance: the
point and colour MLPs take bf16 inputs and weights, so alphas agree within
5e-3 + 1e-2 |alpha| and colours within 5e-3 (a bf16 emulation of the same
roundings stays within 1.7e-3 on these inputs)."""
import numpy as np
import pytest
import torch

import epcq.csrc as C
from aggregator_ref import aggregator_reference, random_inputs
from epcq.epcq import Aggregator

pytestmark = pytest.mark.gpu


def module(dev, seed=0):
    torch.manual_seed(seed)
    mod = Aggregator()
    with torch.no_grad():
        for m in mod.modules():
            if isinstance(m, torch.nn.Linear):
                m.bias.normal_(0.0, 0.1)
    return mod.to(dev)


def check(mod, inputs, dev):
    dinp = [t.to(dev) for t in inputs]
    with torch.no_grad():
        ga, gc = C.aggregator_forward(mod.packed_params(), *dinp)
        torch.cuda.synchronize()
        ra, rc = aggregator_reference(mod.cpu(), *inputs)
        mod.to(dev)
    ga, gc = ga.cpu(), gc.cpu()
    assert torch.isfinite(ga).all() and torch.isfinite(gc).all()
    np.testing.assert_allclose(ga.numpy(), ra.numpy(), rtol=1e-2, atol=5e-3)
    np.testing.assert_allclose(gc.numpy(), rc.numpy(), rtol=0, atol=5e-3)
    empty = ~inputs[5].any(-1)
    assert (ga[empty] == 0).all()
    return ga, gc


@pytest.mark.parametrize("K", [8, 5, 1, 16, 32, 3])
def test_aggregator_vs_fp32_reference(dev, K):
    print(f"aggregator K={K}", flush=True)
    check(module(dev, K), random_inputs(2000 + 7 * K, K, seed=K), dev)


@pytest.mark.parametrize("M", [1, 7, 33, 65])
def test_aggregator_ragged_sample_counts(dev, M):
    print(f"aggregator M={M}", flush=True)
    check(module(dev), random_inputs(M, 8, seed=M, empty_every=3), dev)


def test_aggregator_no_samples_and_errors(dev):
    mod = module(dev)
    inp = [t.to(dev) for t in random_inputs(0, 8)]
    with torch.no_grad():
        a, c = C.aggregator_forward(mod.packed_params(), *inp)
    assert a.shape == (0,) and c.shape == (0, 3)
    bad = [t.to(dev) for t in random_inputs(10, 4, F=16)]
    with torch.no_grad(), pytest.raises(RuntimeError, match="pc_feat must have 32 features"):
        C.aggregator_forward(mod.packed_params(), *bad)
    inp = [t.to(dev) for t in random_inputs(10, 33)]
    with torch.no_grad(), pytest.raises(RuntimeError, match=r"num_point_cloud_per_sp must be in \[1, 32\]"):
        C.aggregator_forward(mod.packed_params(), *inp)


@pytest.mark.parametrize("layout", ["image", "rays"])
def test_module_forward_scatters_to_mask(dev, layout):
    """Aggregator.forward returns the reference's [rows, S] / [rows, S, 3]
    with zeros where mask_out is false, live samples in mask order."""
    mod = module(dev, 3)
    g = torch.Generator().manual_seed(4)
    shape = (6, 5, 4) if layout == "image" else (30, 4)
    mask = torch.rand(shape, generator=g) < 0.5
    M = int(mask.sum())
    inputs = random_inputs(M, 8, seed=5)
    with torch.no_grad():
        fa, fc = mod(*[t.to(dev) for t in inputs[:6]], mask.to(dev), inputs[6].to(dev))
        ga, gc = C.aggregator_forward(mod.packed_params(), *[t.to(dev) for t in inputs])
    rows = 30
    assert fa.shape == (rows, 4) and fc.shape == (rows, 4, 3)
    live = mask.reshape(-1).to(dev)
    assert torch.equal(fa.reshape(-1)[live], ga) and torch.equal(fc.reshape(-1, 3)[live], gc)
    assert (fa.reshape(-1)[~live] == 0).all() and (fc.reshape(-1, 3)[~live] == 0).all()


def test_repacks_after_a_parameter_update(dev):
    mod = module(dev, 6)
    inp = [t.to(dev) for t in random_inputs(300, 8, seed=6)]
    with torch.no_grad():
        a0, _ = C.aggregator_forward(mod.packed_params(), *inp)
        mod.alpha_branch[0].bias.add_(1.0)
        a1, _ = C.aggregator_forward(mod.packed_params(), *inp)
    live = inp[5].any(-1)
    assert (a1[live] > a0[live]).all()


def test_encodings_of_large_features(dev):
    """Features of magnitude 1024 to 4096 take the library's sincosf for the
    base frequency (and double-angle steps from it); the raw-feature columns
    of the first layer are zeroed so that the encodings carry the inputs."""
    mod = module(dev, 11)
    with torch.no_grad():
        mod.block1[0].weight[:, :32] = 0.0
    pc, rgb, d, feat, sp, valid, rays = random_inputs(500, 8, seed=12)
    g = torch.Generator().manual_seed(13)
    big = torch.rand(feat.shape, generator=g) * 3072 + 1024
    big = big * torch.where(torch.rand(feat.shape, generator=g) < 0.5, -1.0, 1.0)
    feat = torch.where(torch.rand((500, 1, 1), generator=g) < 0.5, big, feat).masked_fill(~valid[..., None], 0.0)
    assert (feat.abs() >= 1024).any()
    print("aggregator large features", flush=True)
    check(mod, (pc, rgb, d, feat, sp, valid, rays), dev)
