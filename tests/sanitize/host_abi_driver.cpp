// Exercises the host side of the HIP library's C ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer, without a GPU: workspace and arena sizing, the
// emission bound, squared thresholds, tuning validation, the batch's layout
// of view sets, and the argument checks of every entry point (which return
// before any device work, so null device pointers are never dereferenced).
// Exit status 0 means no sanitizer report and every check held.
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/epcq_abi.h"
#include "epcq_batch_kernels.h"

// Stand-ins for the sampler launch templates (instantiated in their own units
// in the library); the checks below return before any launch.
namespace epcq {
template <int MODE, int KMAX, int SMAX, int NW>
void launch_lay(const SampleOpts&, const SampleOut&, const ViewArgs&, const int*, int64_t, const TilePlan&,
                hipStream_t) {
    std::abort();
}
#define EPCQ_STAND_IN_NW(MODE, KMAX, NW)                                                                          \
    template void launch_lay<MODE, KMAX, 2, NW>(const SampleOpts&, const SampleOut&, const ViewArgs&,            \
                                                const int*, int64_t, const TilePlan&, hipStream_t);              \
    template void launch_lay<MODE, KMAX, kRegSamples, NW>(const SampleOpts&, const SampleOut&, const ViewArgs&,  \
                                                          const int*, int64_t, const TilePlan&, hipStream_t);
#define EPCQ_STAND_IN(MODE, KMAX)                                                                                 \
    EPCQ_STAND_IN_NW(MODE, KMAX, 1)                                                                              \
    EPCQ_STAND_IN_NW(MODE, KMAX, 4)                                                                              \
    template <>                                                                                                  \
    void launch_batch_sample<MODE, KMAX, 2>(int, const TilePlan&, BatchHeader*, const BatchView*, const int*,     \
                                            const SampleOpts&, hipStream_t) {                                    \
        std::abort();                                                                                            \
    }                                                                                                            \
    template <>                                                                                                  \
    void launch_batch_sample<MODE, KMAX, kRegSamples>(int, const TilePlan&, BatchHeader*, const BatchView*,       \
                                                      const int*, const SampleOpts&, hipStream_t) {              \
        std::abort();                                                                                            \
    }
EPCQ_STAND_IN(EPCQ_MODE_COMPAT, kCompatBuffer)
EPCQ_STAND_IN(EPCQ_MODE_KNN, 8)
EPCQ_STAND_IN(EPCQ_MODE_KNN, 16)
EPCQ_STAND_IN(EPCQ_MODE_KNN, 32)
}  // namespace epcq

namespace {

int g_fail = 0;
void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s (last error: %s)\n", what, epcq_last_error());
        g_fail = 1;
    }
}

epcq_options options(int k, float r, int S, int Kc) {
    epcq_options o{};
    o.kernel_size = k;
    o.t_intvl = 0.01f;
    o.radius_threshold = r;
    o.sigma_threshold = 1e-4f;
    o.background_brightness = 1.0f;
    o.stop_threshold = 1e-3f;
    o.num_sample_points_per_ray = S;
    o.num_point_cloud_per_sp = Kc;
    o.num_point_per_ray = 256;
    o.sdf2weight_gaussian_alpha = 0.2f;
    o.sdf2weight_gaussian_gamma = 0.9f;
    o.topK = 8;
    return o;
}

epcq_view_desc view(int64_t n, int W, int H) {
    epcq_view_desc d{};
    d.xyz = reinterpret_cast<const float*>(uintptr_t(4096));  // never dereferenced on the host
    d.n = n;
    d.width = W;
    d.height = H;
    d.fx = d.fy = 100.0f;
    d.cx = W / 2.0f;
    d.cy = H / 2.0f;
    return d;
}

}  // namespace

int main() {
    check(epcq_abi_version() == EPCQ_ABI_VERSION, "abi version");
    // sizes
    for (int64_t n : {int64_t(0), int64_t(1), int64_t(1000), int64_t(1) << 22})
        for (int W : {1, 7, 800, 1920})
            for (int H : {1, 5, 600, 1080}) {
                const size_t b = epcq_build_hashtable_workspace_size(n, H, W);
                check(b > 0 || n == 0, "build workspace");
                check(epcq_sampling_workspace_size(n, H, W) >= (size_t)W * H * 24, "sampling workspace");
            }
    check(epcq_build_hashtable_workspace_size(10, 0, 5) == 0, "workspace of an empty image");
    // emission bound and thresholds
    epcq_options o = options(5, 0.005f, 16, 8);
    check(epcq_max_live_samples(&o, 0.005f) == 2, "emission bound");
    o.sdf2weight_gaussian_gamma = 0.02f;
    o.num_sample_points_per_ray = 100;
    check(epcq_max_live_samples(&o, 0.005f) > 16, "large emission bound");
    for (float r : {0.0f, -1.0f, 1e-40f, 0.005f, 1.0f, 1e30f, INFINITY, NAN}) {
        const float s = epcq_squared_threshold(r);
        if (r > 0.0f && std::isfinite(r) && r < 1e19f) check(std::sqrt(s) >= r, "threshold reaches r");
    }
    // tuning
    epcq_tuning t{16, 16, 1024, EPCQ_TUNE_XCD_ORDER, 256}, back{};
    check(epcq_set_tuning(&t) == EPCQ_OK && epcq_get_tuning(&back) == EPCQ_OK && back.tile_w == 16, "tuning");
    epcq_tuning bad = t;
    bad.flags = 2;
    check(epcq_set_tuning(&bad) == EPCQ_ERR_INVALID, "unknown flags rejected");
    bad = t;
    bad.tile_w = 12;
    check(epcq_set_tuning(&bad) == EPCQ_ERR_INVALID, "tile width rejected");
    check(epcq_set_tuning(nullptr) == EPCQ_ERR_INVALID, "null tuning");
    // argument checks of the entry points
    epcq_camera cam{reinterpret_cast<const float*>(uintptr_t(4096)), nullptr, 1, 1, 2, 2, 4, 4};
    epcq_options big = options(40, 0.01f, 16, 8);
    check(epcq_quick_query(nullptr, 0, &cam, &big, nullptr, nullptr, nullptr, 1, nullptr, nullptr) ==
              EPCQ_ERR_INVALID && std::strstr(epcq_last_error(), "kernel size is too large"),
          "kernel size");
    const size_t need = epcq_build_hashtable_workspace_size(1000, 8, 8);
    check(epcq_build_hashtable(nullptr, 1000, nullptr, nullptr, 8, 8, 0, nullptr, nullptr, nullptr, nullptr, need - 1,
                               nullptr) == EPCQ_ERR_WORKSPACE,
          "workspace too small");
    check(epcq_build_hashtable_noalloc(nullptr, int64_t(40) << 20, nullptr, nullptr, 1080, 1920, 0, nullptr, nullptr,
                                       nullptr, nullptr) == EPCQ_ERR_WORKSPACE,
          "noalloc limit");
    check(epcq_scatter(nullptr, nullptr, nullptr, 10, 4, 4, 0, nullptr, nullptr, nullptr) == EPCQ_ERR_INVALID, "bins");
    check(epcq_knn_sampling(nullptr, 0, &cam, &o, nullptr, nullptr, nullptr, nullptr, 0, 7, 2, nullptr, nullptr,
                            nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, 0,
                            nullptr) == EPCQ_ERR_INVALID,
          "unknown mode");
    check(epcq_idw_features(nullptr, nullptr, 0, 4, nullptr, nullptr, -1, 4, nullptr, nullptr, nullptr) ==
              EPCQ_ERR_INVALID,
          "negative idw size");
    // batch layout: arena sizes of ragged view sets, creation on a stand-in arena
    std::vector<epcq_view_desc> views;
    const int res[5][2] = {{200, 150}, {400, 300}, {640, 512}, {800, 800}, {1600, 1200}};
    for (int i = 0; i < 12; ++i) views.push_back(view(int64_t(1) << (10 + i), res[i % 5][0], res[i % 5][1]));
    const epcq_options ob = options(5, 0.005f, 16, 8);
    for (int mode = 0; mode < 2; ++mode) {
        const size_t arena = epcq_batch_arena_bytes(views.data(), (int)views.size(), 16, &ob, mode, 2);
        size_t sum = epcq_batch_header_bytes(16);
        for (const auto& v : views) sum += epcq_batch_view_bytes(&v, &ob, mode, 2);
        check(arena > 0 && arena <= sum, "arena bound");
        epcq_batch* b = nullptr;
        check(epcq_batch_create(&ob, mode, 2, 16, nullptr, 0.0f, &t, reinterpret_cast<void*>(uintptr_t(1) << 20),
                                arena, &b) == EPCQ_OK && b,
              "batch create");
        epcq_batch_view_out vo{};
        check(epcq_batch_view_outputs(b, 0, &vo) == EPCQ_ERR_INVALID, "no views yet");
        check(epcq_batch_rays(b) == 0, "no rays yet");
        check(epcq_batch_run(b, nullptr) == EPCQ_ERR_INVALID, "run without views");
        epcq_batch_destroy(b);
    }
    check(epcq_batch_create(&ob, 0, 2, 4, nullptr, 0.0f, nullptr, nullptr, 0, nullptr) == EPCQ_ERR_INVALID,
          "null out");
    std::printf(g_fail ? "sanitized host ABI driver: FAILED\n" : "sanitized host ABI driver: ok\n");
    return g_fail;
}
