// Exercises every oracle function on small seeded scenes under AddressSanitizer
// and UndefinedBehaviorSanitizer (host code only), checking the invariants the
// tests rely on.  Exit status 0 means no sanitizer report and every check held.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../oracle/epcq_oracle.h"

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
double uniform() {
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

int g_fail = 0;
void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        g_fail = 1;
    }
}

orc_camera camera(int W, int H, float f) {
    // looking down +y from (0, -4, 0): OpenCV axes x right, y down, z forward
    orc_camera c{};
    const float R[9] = {1, 0, 0, 0, 0, -1, 0, 1, 0};  // world-to-camera rotation rows
    const float eye[3] = {0.0f, -4.0f, 0.0f};
    for (int r = 0; r < 3; ++r) {
        float t = 0.0f;
        for (int k = 0; k < 3; ++k) {
            c.w2c[4 * r + k] = R[3 * r + k];
            c.c2w[4 * r + k] = R[3 * k + r];
            t -= R[3 * r + k] * eye[k];
        }
        c.w2c[4 * r + 3] = t;
        c.c2w[4 * r + 3] = eye[r];
    }
    c.fx = c.fy = f;
    c.cx = W / 2.0f;
    c.cy = H / 2.0f;
    const float K[9] = {f, 0, c.cx, 0, f, c.cy, 0, 0, 1};
    for (int i = 0; i < 9; ++i) c.K[i] = K[i];
    c.width = W;
    c.height = H;
    return c;
}

orc_options options(int k, float r, int S, int Kc) {
    orc_options o{};
    o.kernel_size = k;
    o.t_intvl = 0.01f;
    o.radius_threshold = r;
    o.sigma_threshold = 1e-4f;
    o.background_brightness = 1.0f;
    o.stop_threshold = 1e-3f;
    o.num_sample_points_per_ray = S;
    o.num_point_cloud_per_sp = Kc;
    o.num_point_per_ray = 64;
    o.sdf2weight_gaussian_alpha = 0.2f;
    o.sdf2weight_gaussian_gamma = 0.9f;
    o.topK = 8;
    return o;
}

void scene(int64_t n, int W, int H, float f, int k, float r, int S, int Kc) {
    std::vector<float> xyz(3 * n), radius(n);
    for (int64_t i = 0; i < n; ++i) {
        // points near the unit sphere, some behind the camera and off-image
        double x = uniform() * 2 - 1, y = uniform() * 2 - 1, z = uniform() * 2 - 1;
        const double s = std::sqrt(x * x + y * y + z * z) + 1e-9;
        const double rad = (i % 17 == 0) ? 6.0 : 1.0 + 0.01 * (uniform() - 0.5);
        xyz[3 * i] = (float)(x / s * rad);
        xyz[3 * i + 1] = (float)(y / s * rad);
        xyz[3 * i + 2] = (float)(z / s * rad);
        radius[i] = (float)(r * (0.5 + uniform()));
    }
    const orc_camera cam = camera(W, H, f);
    const int64_t P = (int64_t)W * H;
    for (int sort = 0; sort < 2; ++sort) {
        std::vector<int32_t> idx(n), hash(2 * P);
        std::vector<float> depth(n);
        check(orc_pts2plane(xyz.data(), n, cam.w2c, cam.K, H, W, sort, idx.data(), depth.data(), hash.data()) == 0,
              "pts2plane");
        int64_t total = 0, run = 0;
        for (int64_t p = 0; p < P; ++p) {
            const int s = hash[2 * p], c = hash[2 * p + 1];
            if (c < 0) {
                check(s == -1, "empty pixel (-1, -1)");
                continue;
            }
            check(s == run, "row-major starts");
            run += c;
            total += c;
        }
        for (int64_t i = total; i < n; ++i) check(idx[i] == -1, "tail -1");
        std::vector<int32_t> keys(n);
        std::vector<float> dk(n);
        check(orc_project(xyz.data(), n, cam.w2c, cam.K, H, W, keys.data(), dk.data()) == 0, "project");
        std::vector<int32_t> ks(n), is(n);
        std::vector<float> ds(n);
        check(orc_sort_by_key(keys.data(), dk.data(), n, ks.data(), is.data(), ds.data()) == 0, "sort_by_key");
        for (int64_t i = 1; i < n; ++i) check(ks[i - 1] <= ks[i], "sorted keys");
        std::vector<int32_t> si(n), sh(2 * P);
        std::vector<float> sd(n);
        check(orc_scatter_hashtable(is.data(), ks.data(), ds.data(), n, H, W, sort, si.data(), sd.data(), sh.data()) ==
                  0,
              "scatter_hashtable");
        check(si == idx && sh == hash, "pts2plane = project + sort + scatter_hashtable");
        const int bins = 3;
        std::vector<float> bd(P * bins, -1.0f);
        std::vector<int32_t> bi(P * bins, -1);
        check(orc_scatter(is.data(), ks.data(), ds.data(), n, H, W, bins, bd.data(), bi.data()) == 0, "scatter");
        for (int mode = 0; mode < 2; ++mode) {
            for (int pr = 0; pr < 2; ++pr) {
                const orc_options o = options(k, r, S, Kc);
                std::vector<int32_t> rays;
                for (int64_t q = -2; q < P + 2; q += 3) rays.push_back((int32_t)q);  // includes invalid ids
                for (int listed = 0; listed < 2; ++listed) {
                    const int64_t R = listed ? (int64_t)rays.size() : P;
                    std::vector<float> sp(R * S * 3), st(R * S), nbt(R * S * Kc), nbd(R * S * Kc);
                    std::vector<int32_t> pc(R * S * Kc);
                    std::vector<uint8_t> mask(R * S);
                    check(orc_quick_sampling(xyz.data(), n, &cam, &o, idx.data(), hash.data(),
                                             pr ? radius.data() : nullptr, mode, listed ? rays.data() : nullptr, R,
                                             sp.data(), pc.data(), st.data(), mask.data(), nbt.data(), nbd.data()) == 0,
                          "quick_sampling");
                    for (int64_t i = 0; i < R * S * Kc; ++i) check(pc[i] >= -1 && pc[i] < n, "neighbour ids");
                    if (mode == 1 && pr == 0 && listed == 1) {
                        std::vector<float> w(R * S * Kc), agg(R * S * 4);
                        std::vector<float> feat(n * 4);
                        for (auto& v : feat) v = (float)uniform();
                        check(orc_idw(xyz.data(), feat.data(), 0, 4, sp.data(), pc.data(), R * S, Kc, w.data(),
                                      agg.data()) == 0,
                              "idw");
                    }
                }
            }
        }
        const orc_options oq = options(k, r, S, Kc);
        std::vector<int32_t> rays;
        for (int64_t q = -3; q < P + 3; q += 2) rays.push_back((int32_t)q);
        std::vector<int32_t> qo(rays.size() * oq.num_point_per_ray);
        check(orc_quick_query(xyz.data(), n, &cam, &oq, idx.data(), hash.data(), rays.data(), (int64_t)rays.size(),
                              qo.data()) == 0,
              "quick_query");
    }
}

}  // namespace

int main() {
    orc_set_threads(2);
    for (int i = -5; i <= 5; ++i) check(std::isfinite(orc_expf((float)i * 17.0f)) || i * 17 > 88, "expf");
    for (int k = 1; k <= 7; ++k)
        for (int l = 0; l < k * k; ++l) {
            int32_t du, dv;
            orc_neighbor_offset(k, l, &du, &dv);
            check(du > -k && du < k && dv > -k && dv < k, "neighbour offsets");
        }
    scene(3000, 40, 30, 50.0f, 5, 0.05f, 16, 8);
    scene(2000, 17, 9, 20.0f, 3, 0.1f, 4, 3);
    scene(1500, 23, 19, 30.0f, 7, 0.08f, 8, 12);
    scene(1, 5, 5, 5.0f, 5, 0.5f, 16, 8);
    scene(0, 4, 3, 5.0f, 5, 0.5f, 16, 8);
    std::printf(g_fail ? "sanitized oracle driver: FAILED\n" : "sanitized oracle driver: ok\n");
    return g_fail;
}
