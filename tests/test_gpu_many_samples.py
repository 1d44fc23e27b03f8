"""More live samples per ray than the kernels keep in registers: a thick
volumetric cloud gives rays dozens of occupied depth slots, and a small
gamma lets each of them emit a sample (the reference emits up to S samples
for any S, epcq/csrc/epcq_render_topk.cu:184-203).  Dense drop-in and compact
layouts, both neighbour modes, one-lane and wave-wide forms, against the
oracle bit for bit."""
import numpy as np
import pytest

from epcq import synth
from test_gpu_sampling import View, check_dense, compact_vs_dense
from test_gpu_wave import heavy_threshold

pytestmark = pytest.mark.gpu


def volume_view():
    st = synth.Stream(77, 1)
    pts = st.uniform(3 * 150_000, -0.5, 0.5).astype(np.float32).reshape(-1, 3)
    cam = synth.look_at_camera([0.4, -3.0, 0.5], 48, 40, 60.0)
    return pts, cam


OPTS = dict(kernel_size=3, radius_threshold=0.03, t_intvl=0.01, num_sample_points_per_ray=48,
            num_point_cloud_per_sp=8, sdf2weight_gaussian_gamma=0.02, sdf2weight_gaussian_alpha=0.2)


@pytest.mark.parametrize("heavy", [0, 1])
@pytest.mark.parametrize("mode", ["compat", "knn"])
def test_many_live_samples(dev, mode, heavy):
    import epcq.csrc as C
    from epcq.epcq import RenderOptions
    pts, cam = volume_view()
    assert C.max_live_samples(RenderOptions(**OPTS), OPTS["radius_threshold"]) > 16
    print(f"many samples {mode} heavy={heavy}", flush=True)
    v = View(pts, cam, dev, **OPTS)
    with heavy_threshold(heavy):
        mask = check_dense(v, mode)
        out = compact_vs_dense(v, mode)
    n = mask.sum(1)
    assert n.max() > 16, n.max()  # rays really emit beyond the register-kept samples
    assert out["sample_t"].shape[1] > 16
