"""The boundary's error contract without a GPU: the reference's error texts
(epcq/csrc/include/util.hpp:4-9, epcq/csrc/epcq_render_topk.cu:317-323,
351-357) raised by the op module's host checks, and the C ABI's argument
checks, which return before any device work (so null device pointers are
safe here)."""
import ctypes

import pytest
import torch


def C():
    import epcq.csrc as C
    return C


def test_check_input_texts():
    mod = C()
    with pytest.raises(RuntimeError, match="^xyz_data must be a CUDA tensor$"):
        mod._check_input(torch.zeros(4, 3), "xyz_data")
    with pytest.raises(RuntimeError, match="^hashtable must be a CUDA tensor$"):
        mod._check_input(None, "hashtable")
    # the ops run these checks first: a CPU tensor never reaches the library
    spec = mod.NeuralPointsSpec()
    spec.xyz_data = torch.zeros(4, 3)
    with pytest.raises(RuntimeError, match="xyz_data must be a CUDA tensor"):
        spec.check()
    with pytest.raises(RuntimeError, match="point_idx_list must be a CUDA tensor"):
        mod.scatter_hashtable(torch.zeros(3, dtype=torch.int32), torch.zeros(3, dtype=torch.int32),
                              torch.zeros(3), (2, 2), False)
    with pytest.raises(RuntimeError, match="xyz_data must be a CUDA tensor"):
        mod.build_hashtable(torch.zeros(3, 3), torch.eye(4)[:3], torch.eye(3), (2, 2))


def test_hashtable_size_error(capsys):
    mod = C()
    good = torch.zeros(6, 2, dtype=torch.int32)
    assert mod._hashtable_view(good, 2, 3).shape == (2, 3, 2)
    assert mod._hashtable_view(torch.zeros(2, 3, 2, dtype=torch.int32), 2, 3).shape == (2, 3, 2)
    with pytest.raises(RuntimeError, match="^hashtable size error$"):
        mod._hashtable_view(torch.zeros(3, 3, 2, dtype=torch.int32), 2, 3)
    # the sizes are printed before the error, as the reference does
    assert "hashtable size: 3, 3, 2" in capsys.readouterr().out
    with pytest.raises(RuntimeError):
        mod._hashtable_view(torch.zeros(7, 2, dtype=torch.int32), 2, 3)  # cannot be viewed as [2, 3, 2]


def _cam(mod, w=4, h=4):
    return mod._Camera(ctypes.c_void_p(0), ctypes.c_void_p(0), 1.0, 1.0, 2.0, 2.0, w, h)


def _opts(mod, **kw):
    from epcq.epcq import RenderOptions
    return mod._options_struct(RenderOptions(**kw))


def _err(mod):
    return mod._lib.epcq_last_error().decode()


def test_kernel_size_too_large():
    mod = C()
    cam = _cam(mod)
    cam.c2w = ctypes.c_void_p(8)  # never dereferenced: the check comes first
    for k in (32, 40):
        rc = mod._lib.epcq_quick_query(None, 0, ctypes.byref(cam), ctypes.byref(_opts(mod, kernel_size=k)), None, None,
                                       None, 1, None, None)
        assert rc == 1
        assert _err(mod) == "kernel size is too large, can not be handled by current CUDA kernel"


def test_workspace_and_size_checks():
    mod = C()
    lib = mod._lib
    need = lib.epcq_build_hashtable_workspace_size(1000, 8, 8)
    assert need > 0
    rc = lib.epcq_build_hashtable(None, 1000, None, None, 8, 8, 0, None, None, None, None, need - 1, None)
    assert rc == 3 and _err(mod) == "build_hashtable: workspace too small"
    rc = lib.epcq_build_hashtable(None, -1, None, None, 8, 8, 0, None, None, None, None, 0, None)
    assert rc == 1 and "point count out of range" in _err(mod)
    rc = lib.epcq_build_hashtable(None, 10, None, None, 0, 8, 0, None, None, None, None, 0, None)
    assert rc == 1 and "image size must be positive" in _err(mod)
    assert lib.epcq_build_hashtable_workspace_size(-1, 8, 8) == 0
    rc = lib.epcq_scatter(None, None, None, 10, 4, 4, 0, None, None, None)
    assert rc == 1 and "bins must be positive" in _err(mod)
    rc = lib.epcq_scatter_hashtable(None, None, None, -5, 4, 4, 0, None, None, None, None, 0, None)
    assert rc == 1 and "point count out of range" in _err(mod)
    need = lib.epcq_scatter_hashtable_workspace_size(1000, 1)
    assert need == 8000 and lib.epcq_scatter_hashtable_workspace_size(1000, 0) == 0
    rc = lib.epcq_scatter_hashtable(None, None, None, 1000, 4, 4, 1, None, None, None, None, need - 1, None)
    assert rc == 3 and _err(mod) == "scatter_hashtable: workspace too small"
    cam = _cam(mod)
    cam.c2w = ctypes.c_void_p(8)
    rc = lib.epcq_quick_query(None, 0, ctypes.byref(cam), ctypes.byref(_opts(mod)), None, None, None, -1, None, None)
    assert rc == 1 and "negative ray count" in _err(mod)
    rc = lib.epcq_knn_sampling(None, 0, ctypes.byref(cam), ctypes.byref(_opts(mod, num_point_cloud_per_sp=33)),
                               None, None, None, None, 0, 1, 2, None, None, None, None, None, None, 0, 0, None, None,
                               None, 0, None)
    assert rc == 1 and "num_point_cloud_per_sp must be in [1, 32]" in _err(mod)
    rc = lib.epcq_knn_sampling(None, 0, ctypes.byref(cam), ctypes.byref(_opts(mod, num_point_cloud_per_sp=8)),
                               None, None, None, None, 0, 1, 0, None, None, None, None, None, None, 0, 0, None, None,
                               None, 0, None)
    assert rc == 1 and "s_max must be >= 1" in _err(mod)
    rc = lib.epcq_quick_sampling(None, 0, ctypes.byref(cam), ctypes.byref(_opts(mod)), None, None, 5, None, None,
                                 None, None, None, 0, None)
    assert rc == 1 and "unknown mode" in _err(mod)
    bad = _cam(mod, w=0)
    bad.c2w = ctypes.c_void_p(8)
    rc = lib.epcq_quick_sampling(None, 0, ctypes.byref(bad), ctypes.byref(_opts(mod)), None, None, 0, None, None,
                                 None, None, None, 0, None)
    assert rc == 1 and "image size must be positive" in _err(mod)
    rc = lib.epcq_idw_features(None, None, 0, 4, None, None, -1, 4, None, None, None)
    assert rc == 1 and "negative size" in _err(mod)


def test_batch_argument_checks():
    mod = C()
    lib = mod._lib
    h = ctypes.c_void_p(0)
    opts = _opts(mod, num_point_cloud_per_sp=8)
    rc = lib.epcq_batch_create(ctypes.byref(opts), 0, 2, 4, None, 0.0, None, None, 0, ctypes.byref(h))
    assert rc == 3 and "arena too small" in _err(mod)
    tu = mod._Tuning(16, 16, 1024, 2, 128)
    rc = lib.epcq_batch_create(ctypes.byref(opts), 0, 2, 4, None, 0.0, ctypes.byref(tu), ctypes.c_void_p(256),
                               1 << 20, ctypes.byref(h))
    assert rc == 1 and "unknown flags" in _err(mod)
    rc = lib.epcq_batch_create(ctypes.byref(opts), 7, 2, 4, None, 0.0, None, ctypes.c_void_p(256), 1 << 20,
                               ctypes.byref(h))
    assert rc == 1 and "unknown mode" in _err(mod)
    assert lib.epcq_batch_set_views(None, None, 0, None) == 1
