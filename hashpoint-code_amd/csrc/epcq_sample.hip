// Primary-surface search, K-nearest sampling and radius query for gfx950.
//
//   epcq_quick_sampling  dense drop-in layout of quick_sampling
//                        (epcq/csrc/epcq_render_topk.cu:10-220,341-417)
//   epcq_knn_sampling    compact layout, optional per-point radius, KNN
//                        neighbours with ray depth / distance, fused IDW
//   epcq_quick_query     radius query for a ray list (:222-339)
//   epcq_idw_features    IDW weights + weighted features (epcq/epcq.py:434-452)
//
// Full-image sampling runs one 256-lane block per tile of rays.  The block
// first stages its tile's halo region -- the hash entries and bucket cones of
// the (tw + k - 1) x (th + k - 1) neighbour pixels and, row segment by row
// segment (contiguous in the row-major point lists), the coordinates and
// indices of their points as float4 -- into LDS with coalesced loads; every
// candidate walk of every ray then reads LDS.  A tile whose region holds more
// points than the LDS capacity reads global memory instead, in the same
// block.  Ray lists (no tile structure) read global memory.  The per-ray
// logic is in epcq_sampler.h.
#include <cmath>
#include <mutex>

#include "epcq_common.h"
#include "epcq_sampler.h"

namespace epcq {
// Defined in epcq_sample_kernels.h, instantiated in epcq_sample_inst_*.hip
// (NW = 1) and epcq_sample_wide_*.hip (NW = 4).
template <int MODE, int KMAX, int SMAX, int NW>
void launch_lay(const SampleOpts& o, const SampleOut& out, const ViewArgs& view, const int* rays, int64_t R,
                const TilePlan& tp, hipStream_t st);
}  // namespace epcq

extern "C" int32_t epcq_max_live_samples(const epcq_options* opt, float r_max);

namespace epcq {

float squared_threshold(float r) {
    if (!(r > 0.0f)) return 0.0f;  // no distance is below r (NaN r included)
    if (std::isinf(r)) return INFINITY;
    float s = r * r;
    if (std::sqrt(s) >= r) {
        while (s > 0.0f && std::sqrt(std::nextafter(s, 0.0f)) >= r) s = std::nextafter(s, 0.0f);
    } else {
        while (!(std::sqrt(s) >= r)) s = std::nextafter(s, INFINITY);
    }
    return s;
}

TilePlan plan_tiles(int k, int tile_w, int tile_h, int cap, int flags, bool point_radius) {
    TilePlan p{};
    p.tw = tile_w;
    p.th = tile_h;
    p.xcd = (flags & EPCQ_TUNE_XCD_ORDER) != 0;
    p.rw_max = p.tw + k - 1;
    p.rh_max = p.th + k - 1;
    // per point: float4 + radius; per region pixel: cone_a, hash, cone_b; per row: lo, base
    const size_t fixed =
        align_up((size_t)p.rw_max * p.rh_max * (16 + 8 + 8) + (size_t)(2 * p.rh_max + 1) * 4, 16);
    // 64 KB: a larger stage leaves one block per CU, slower on every workload measured
    const size_t limit = 64 * 1024;
    size_t c = (size_t)std::max(cap, 0);
    const size_t per = point_radius ? 20 : 16;
    if (fixed + c * per > limit) c = fixed < limit ? (limit - fixed) / per : 0;
    p.cap = (int)c;
    p.lds_bytes = c * per + fixed;
    return p;
}

int check_tuning(const epcq_tuning& t, const char* who) {
    const std::string w(who);
    const int rays = t.tile_w * t.tile_h;
    if (t.tile_w < 1 || t.tile_w > 64 || (t.tile_w & (t.tile_w - 1)) || t.tile_h < 1 ||
        (rays != 64 && rays != 128 && rays != 256))
        return fail(EPCQ_ERR_INVALID, w + ": tile_w must be a power of two <= 64 and tile_w * tile_h 64, 128 or 256");
    if (t.tile_cap < 0) return fail(EPCQ_ERR_INVALID, w + ": negative tile_cap");
    if (t.heavy < 0) return fail(EPCQ_ERR_INVALID, w + ": negative heavy");
    if (t.list_wave < 0) return fail(EPCQ_ERR_INVALID, w + ": negative list_wave");
    if (t.flags & ~EPCQ_TUNE_XCD_ORDER) return fail(EPCQ_ERR_INVALID, w + ": unknown flags");
    return EPCQ_OK;
}

namespace {

constexpr int kBlock = 256;

std::mutex g_tuning_mu;
epcq_tuning g_tuning{32, 8, 1024, 0, kHeavyCandidates, kListWaveRays};

epcq_tuning current_tuning() {
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    return g_tuning;
}

// Pixel-ordered copy of the points, one thread per list position: (x, y,
// z, index bits) and, with per-point radii, the radius's squared threshold.
// Positions holding no valid index (the -1 tail) are left alone; the hash
// entries never point there.
__global__ __launch_bounds__(kBlock) void point_copy_kernel(const int* __restrict__ sorted_idx,
                                                            const float* __restrict__ xyz,
                                                            const float* __restrict__ radius, int64_t n,
                                                            float4* __restrict__ pts, float* __restrict__ prad) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int id = sorted_idx[i];
    if (id < 0 || id >= n) return;
    const float* q = xyz + 3 * (int64_t)id;
    pts[i] = make_float4(q[0], q[1], q[2], __int_as_float(id));
    if (prad) prad[i] = point_threshold(radius[id]);
}

// Bucket cones, one thread per pixel: from the pixel-ordered copy pts_in
// when given, else gathering through sorted_idx / xyz (and then writing the
// copy into pts, when not null).
__global__ __launch_bounds__(kBlock) void bucket_cone_kernel(const int2* __restrict__ hash,
                                                             const int* __restrict__ sorted_idx,
                                                             const float* __restrict__ xyz,
                                                             const float4* __restrict__ pts_in,
                                                             const float* __restrict__ radius, float r0,
                                                             const float* __restrict__ c2w, int64_t P,
                                                             float4* __restrict__ cone_a, float2* __restrict__ cone_b,
                                                             float4* __restrict__ pts, float* __restrict__ prad) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < P) bucket_cone_body(hash, sorted_idx, xyz, pts_in, radius, r0, c2w, p, cone_a, cone_b, pts, prad);
}

// hashtable_query_for_nearby_point (:222-283): one lane per ray, hits of each
// neighbour bucket (at most 32) appended in lane order.
__global__ __launch_bounds__(kBlock) void query_kernel(const float* __restrict__ xyz,
                                                       const int* __restrict__ sorted_idx,
                                                       const int2* __restrict__ hash,
                                                       const float* __restrict__ c2w, float fx, float fy,
                                                       float cx, float cy, int W, int H, int k, int ko,
                                                       float r0, const int* __restrict__ rays, int64_t R,
                                                       int cap, int* __restrict__ out) {
    const int64_t ri = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (ri >= R) return;
    int* row = out + ri * cap;
    int total = 0;
    const int64_t q = rays[ri];
    if (q >= 0 && q < (int64_t)W * H) {
        const int u = (int)(q % W), v = (int)(q / W);
        float m[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) m[i] = c2w[i];
        Ray ray;
        make_ray(m, fx, fy, cx, cy, u, v, ray);
        const int L = k * k;
        for (int l = 0; l < L && total < cap; ++l) {
            const int2 e = hash[neighbor_pixel(k, ko, l, u, v, W, H)];
            if (e.y < 0) continue;
            int kept = 0;
            for (int j = 0; j < e.y && kept < kQueryLaneHits; ++j) {
                const int id = sorted_idx[e.x + j];
                const float* p = xyz + 3 * (int64_t)id;
                float t, dist;
                point_ray(ray, p[0], p[1], p[2], t, dist);
                if (dist < r0) {
                    if (total < cap) row[total] = id;
                    ++total;
                    ++kept;
                }
            }
        }
    }
    for (int j = total; j < cap; ++j) row[j] = -1;
}

__global__ __launch_bounds__(kBlock) void idw_kernel(const float* __restrict__ xyz, const void* feat,
                                                     int feat_bf16, int F, const float* __restrict__ sp_xyz,
                                                     const int* __restrict__ nb_idx, int64_t M, int K,
                                                     float* __restrict__ w, float* __restrict__ agg) {
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= M) return;
    const float s0 = sp_xyz[3 * m], s1 = sp_xyz[3 * m + 1], s2 = sp_xyz[3 * m + 2];
    const int* ids = nb_idx + m * K;
    float* wo = w + m * K;
    float den = 0.0f;
    for (int j = 0; j < K; ++j) {
        const int id = ids[j];
        float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
        if (id >= 0) {
            p0 = xyz[3 * (int64_t)id];
            p1 = xyz[3 * (int64_t)id + 1];
            p2 = xyz[3 * (int64_t)id + 2];
        }
        const float d = norm3(p0 - s0, p1 - s1, p2 - s2);
        const float inv = (1.0f / (d + 1e-6f)) * (id >= 0 ? 1.0f : 0.0f);
        wo[j] = inv;
        den = den + inv;
    }
    den = den + 1e-6f;
    for (int j = 0; j < K; ++j) wo[j] = wo[j] / den;
    float* ao = agg + m * (int64_t)F;
    for (int f = 0; f < F; ++f) {
        float acc = 0.0f;
        for (int j = 0; j < K; ++j) {
            const int id = ids[j];
            if (id < 0) continue;
            float fv;
            if (feat_bf16) {
                fv = __uint_as_float((uint32_t)static_cast<const uint16_t*>(feat)[(int64_t)id * F + f] << 16);
            } else {
                fv = static_cast<const float*>(feat)[(int64_t)id * F + f];
            }
            acc = acc + wo[j] * fv;
        }
        ao[f] = acc;
    }
}

int check_common(const epcq_camera* cam, const epcq_options* opt, const char* who) {
    if (!cam || !opt) return fail(EPCQ_ERR_INVALID, std::string(who) + ": null camera or options");
    if (cam->width <= 0 || cam->height <= 0)
        return fail(EPCQ_ERR_INVALID, std::string(who) + ": image size must be positive");
    if ((int64_t)cam->width * cam->height >= (int64_t(1) << 30))
        return fail(EPCQ_ERR_INVALID, std::string(who) + ": image too large");
    if (opt->kernel_size < 1) return fail(EPCQ_ERR_INVALID, std::string(who) + ": kernel_size must be >= 1");
    if (opt->kernel_size > 64) return fail(EPCQ_ERR_INVALID, std::string(who) + ": kernel_size must be <= 64");
    if (!cam->c2w) return fail(EPCQ_ERR_INVALID, std::string(who) + ": null c2w");
    return EPCQ_OK;
}

SampleOpts make_opts(const float* xyz, const epcq_options* opt, const float* radius) {
    SampleOpts o{};
    o.k = opt->kernel_size;
    o.ko = (opt->kernel_size + 1) / 2;
    o.kdiv = (65536 + opt->kernel_size - 1) / opt->kernel_size;
    o.L = opt->kernel_size * opt->kernel_size;
    o.r0 = opt->radius_threshold;
    o.r2lim = squared_threshold(opt->radius_threshold);
    o.radius = radius;
    o.t_intvl = opt->t_intvl;
    o.alpha_p = opt->sdf2weight_gaussian_alpha;
    o.gamma = opt->sdf2weight_gaussian_gamma;
    o.S = opt->num_sample_points_per_ray;
    o.Kc = opt->num_point_cloud_per_sp;
    o.xyz = xyz;
    const epcq_tuning tu = current_tuning();
    o.heavy = tu.heavy;
    o.list_wave = tu.list_wave;
    return o;
}

ViewArgs make_view(const epcq_camera* cam, const int* sorted_idx, const int* hash) {
    return ViewArgs{cam->c2w, cam->fx, cam->fy, cam->cx, cam->cy, cam->width, cam->height,
                    reinterpret_cast<const int2*>(hash), sorted_idx, nullptr, nullptr, nullptr, nullptr, nullptr};
}

// Workspace layout of the samplers: cone_a [P] float4 | cone_b [P] float2 |
// pixel-ordered points [n] float4 | pixel-ordered radii [n] float.
size_t cone_bytes(int64_t P) { return align_up((size_t)P * 16, 256) + align_up((size_t)P * 8, 256); }

// Bucket cones and the pixel-ordered point copy in the workspace (when it is
// large enough: then o.cull is set and view reads them).  They cost a pass
// over every pixel and point of the view, so a ray list (rays >= 0 rays)
// whose windows visit fewer than P / kConeVisitRatio buckets in all skips
// them.  Measured: a 4096-ray batch of a 1M-point 800x800 view with k = 5
// (visits 0.16 P) samples in 0.069 ms without them against 0.080 ms with,
// and both ways meet near 0.24 P; on a 20M-point 1080p view with k = 15 they
// meet near 0.2 P.  The outputs are the same either way.
constexpr int64_t kConeVisitRatio = 4;

void prepare_bounds(SampleOpts& o, ViewArgs& view, int64_t n, const epcq_camera* cam, const int* sorted_idx,
                    const int* hash, void* workspace, size_t workspace_bytes, int64_t rays, hipStream_t st) {
    const int64_t P = (int64_t)cam->width * cam->height;
    o.cull = 0;
    if (!workspace || workspace_bytes < cone_bytes(P)) return;
    if (rays >= 0 && rays * o.L * kConeVisitRatio < P) return;
    char* w = static_cast<char*>(workspace);
    float4* ca = reinterpret_cast<float4*>(w);
    float2* cb = reinterpret_cast<float2*>(w + align_up((size_t)P * 16, 256));
    const size_t pts_off = cone_bytes(P), rad_off = pts_off + align_up((size_t)n * 16, 256);
    float4* pts = workspace_bytes >= rad_off ? reinterpret_cast<float4*>(w + pts_off) : nullptr;
    float* prad = o.radius && workspace_bytes >= rad_off + (size_t)n * 4 ? reinterpret_cast<float*>(w + rad_off)
                                                                        : nullptr;
    if (pts) {
        // the copy first, in parallel over list positions, then the cones from it
        if (n > 0)
            point_copy_kernel<<<dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0, st>>>(sorted_idx, o.xyz,
                                                                                          o.radius, n, pts, prad);
        bucket_cone_kernel<<<dim3((unsigned)ceil_div(P, kBlock)), dim3(kBlock), 0, st>>>(
            reinterpret_cast<const int2*>(hash), nullptr, nullptr, pts, o.radius, o.r0, cam->c2w, P, ca, cb, nullptr,
            nullptr);
    } else {
        bucket_cone_kernel<<<dim3((unsigned)ceil_div(P, kBlock)), dim3(kBlock), 0, st>>>(
            reinterpret_cast<const int2*>(hash), sorted_idx, o.xyz, nullptr, o.radius, o.r0, cam->c2w, P, ca, cb,
            nullptr, nullptr);
    }
    o.cull = 1;
    view.cone_a = ca;
    view.cone_b = cb;
    view.pts = pts;
    view.prad = prad;
}

// Windows of more than 64 lanes (k > 8) take the kernels with 4-word lane
// sets (exact culling up to k = 16; larger windows visit every lane).
template <int MODE, int KMAX>
void launch_smax(const SampleOpts& o, const SampleOut& out, const ViewArgs& view, const int* rays, int64_t R,
                 const TilePlan& tp, int bound, hipStream_t st) {
    const bool wide = o.L > 64;
    if (bound <= 2) {
        if (wide) launch_lay<MODE, KMAX, 2, 4>(o, out, view, rays, R, tp, st);
        else launch_lay<MODE, KMAX, 2, 1>(o, out, view, rays, R, tp, st);
    } else {
        if (wide) launch_lay<MODE, KMAX, kRegSamples, 4>(o, out, view, rays, R, tp, st);
        else launch_lay<MODE, KMAX, kRegSamples, 1>(o, out, view, rays, R, tp, st);
    }
}

// `bound`: samples of one ray that are stored (kernels with a bound of 2 keep
// two sample depths in registers; the others keep kRegSamples and re-read
// later ones from the outputs).
int launch_sample(const SampleOpts& o, const SampleOut& out, const ViewArgs& view, const int* rays, int64_t R,
                  int mode, int bound, hipStream_t st, const char* who) {
    if (R == 0) return EPCQ_OK;
    const epcq_tuning tu = current_tuning();
    const TilePlan tp = plan_tiles(o.k, tu.tile_w, tu.tile_h, tu.tile_cap, tu.flags, o.radius != nullptr);
    if (mode == EPCQ_MODE_COMPAT) {
        launch_smax<EPCQ_MODE_COMPAT, kCompatBuffer>(o, out, view, rays, R, tp, bound, st);
    } else if (o.Kc <= 8) {
        launch_smax<EPCQ_MODE_KNN, 8>(o, out, view, rays, R, tp, bound, st);
    } else if (o.Kc <= 16) {
        launch_smax<EPCQ_MODE_KNN, 16>(o, out, view, rays, R, tp, bound, st);
    } else {
        launch_smax<EPCQ_MODE_KNN, 32>(o, out, view, rays, R, tp, bound, st);
    }
    return check_launch(who);
}

}  // namespace
}  // namespace epcq

using namespace epcq;

extern "C" int epcq_set_tuning(const epcq_tuning* t) {
    if (!t) return fail(EPCQ_ERR_INVALID, "set_tuning: null");
    if (int e = check_tuning(*t, "set_tuning")) return e;
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    g_tuning = *t;
    return EPCQ_OK;
}

extern "C" int epcq_get_tuning(epcq_tuning* t) {
    if (!t) return fail(EPCQ_ERR_INVALID, "get_tuning: null");
    *t = current_tuning();
    return EPCQ_OK;
}

extern "C" float epcq_squared_threshold(float r) { return squared_threshold(r); }

extern "C" int32_t epcq_max_live_samples(const epcq_options* opt, float r_max) {
    // Every visited slot (count >= 2) lowers the transmittance by
    // alpha = gamma * exp(-mean / alpha_p) with mean < r_max, and a sample can
    // only be emitted while the transmittance is >= 1e-3.  Walk the worst case
    // (smallest alpha, shrunk by a margin that covers rounding) in double.
    if (!opt) return 0;
    const int S = opt->num_sample_points_per_ray;
    if (S <= 0) return 0;
    const double g = opt->sdf2weight_gaussian_gamma, ap = opt->sdf2weight_gaussian_alpha;
    if (!(g > 0.0) || !(ap > 0.0) || !(r_max >= 0.0f)) return S;
    const double amin = g * std::exp(-(double)r_max * (1.0 + 1e-3) / ap) * (1.0 - 1e-4);
    if (!(amin > 0.0)) return S;
    double T = 1.0;
    int visits = 0;
    while (T >= 1e-3 * (1.0 - 1e-4) && visits < S) {
        T -= amin;
        ++visits;
    }
    return visits;
}

extern "C" size_t epcq_sampling_workspace_size(int64_t n, int32_t height, int32_t width) {
    if (height <= 0 || width <= 0 || n < 0) return 0;
    return cone_bytes((int64_t)height * width) + align_up((size_t)n * 16, 256) + (size_t)n * 4;
}

extern "C" int epcq_quick_sampling(const float* xyz, int64_t n, const epcq_camera* cam, const epcq_options* opt,
                                   const int32_t* sorted_idx, const int32_t* hash, int32_t mode, float* sp_xyz,
                                   int32_t* pc_idx, float* sp_t, uint8_t* mask, void* workspace,
                                   size_t workspace_bytes, void* stream) {
    if (int e = check_common(cam, opt, "quick_sampling")) return e;
    if (mode != EPCQ_MODE_COMPAT && mode != EPCQ_MODE_KNN)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: unknown mode");
    if (opt->num_sample_points_per_ray < 0 || opt->num_point_cloud_per_sp < 0)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: negative output size");
    if (mode == EPCQ_MODE_KNN && opt->num_point_cloud_per_sp > 32)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: num_point_cloud_per_sp > 32 in KNN mode");
    if (opt->num_sample_points_per_ray == 0) return EPCQ_OK;
    const int bound = epcq_max_live_samples(opt, opt->radius_threshold);
    SampleOpts o = make_opts(xyz, opt, nullptr);
    SampleOut out{};
    out.sp_xyz = sp_xyz;
    out.pc_idx = pc_idx;
    out.sp_t = sp_t;
    out.mask = mask;
    hipStream_t st = as_stream(stream);
    ViewArgs view = make_view(cam, sorted_idx, hash);
    prepare_bounds(o, view, n, cam, sorted_idx, hash, workspace, workspace_bytes, -1, st);
    return launch_sample(o, out, view, nullptr, (int64_t)cam->width * cam->height, mode, bound, st,
                         "quick_sampling");
}

extern "C" int epcq_quick_sampling_rays(const float* xyz, int64_t n, const epcq_camera* cam,
                                        const epcq_options* opt, const int32_t* sorted_idx, const int32_t* hash,
                                        int32_t mode, const int32_t* rays, int64_t R, float* sp_xyz, int32_t* pc_idx,
                                        float* sp_t, uint8_t* mask, void* workspace, size_t workspace_bytes,
                                        void* stream) {
    if (int e = check_common(cam, opt, "quick_sampling")) return e;
    if (mode != EPCQ_MODE_COMPAT && mode != EPCQ_MODE_KNN)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: unknown mode");
    if (opt->num_sample_points_per_ray < 0 || opt->num_point_cloud_per_sp < 0)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: negative output size");
    if (mode == EPCQ_MODE_KNN && opt->num_point_cloud_per_sp > 32)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: num_point_cloud_per_sp > 32 in KNN mode");
    if (R < 0 || (R > 0 && !rays)) return fail(EPCQ_ERR_INVALID, "quick_sampling: bad ray list");
    if (opt->num_sample_points_per_ray == 0 || R == 0) return EPCQ_OK;
    const int bound = epcq_max_live_samples(opt, opt->radius_threshold);
    SampleOpts o = make_opts(xyz, opt, nullptr);
    SampleOut out{};
    out.sp_xyz = sp_xyz;
    out.pc_idx = pc_idx;
    out.sp_t = sp_t;
    out.mask = mask;
    hipStream_t st = as_stream(stream);
    ViewArgs view = make_view(cam, sorted_idx, hash);
    prepare_bounds(o, view, n, cam, sorted_idx, hash, workspace, workspace_bytes, R, st);
    return launch_sample(o, out, view, rays, R, mode, bound, st, "quick_sampling");
}

extern "C" int epcq_knn_sampling(const float* xyz, int64_t n, const epcq_camera* cam, const epcq_options* opt,
                                 const int32_t* sorted_idx, const int32_t* hash, const float* radius,
                                 const int32_t* rays, int64_t R, int32_t mode, int32_t s_max, int32_t* n_samples,
                                 float* sample_t, int32_t* nb_idx, float* nb_t, float* nb_dist, const void* feat,
                                 int32_t feat_bf16, int32_t F, float* w, float* agg, void* workspace,
                                 size_t workspace_bytes, void* stream) {
    if (int e = check_common(cam, opt, "knn_sampling")) return e;
    if (mode != EPCQ_MODE_COMPAT && mode != EPCQ_MODE_KNN)
        return fail(EPCQ_ERR_INVALID, "knn_sampling: unknown mode");
    if (opt->num_point_cloud_per_sp < 1 || opt->num_point_cloud_per_sp > 32)
        return fail(EPCQ_ERR_INVALID, "knn_sampling: num_point_cloud_per_sp must be in [1, 32]");
    if (s_max < 1) return fail(EPCQ_ERR_INVALID, "knn_sampling: s_max must be >= 1");
    if (!n_samples || !sample_t || !nb_idx) return fail(EPCQ_ERR_INVALID, "knn_sampling: null output");
    if (feat && (F < 1 || !w || !agg)) return fail(EPCQ_ERR_INVALID, "knn_sampling: bad feature arguments");
    if (rays && R < 0) return fail(EPCQ_ERR_INVALID, "knn_sampling: negative ray count");
    // with per-point radii the caller's s_max stands for the bound
    const int bound = radius ? s_max : std::min<int>(epcq_max_live_samples(opt, opt->radius_threshold), s_max);
    SampleOpts o = make_opts(xyz, opt, radius);
    o.feat = feat;
    o.feat_bf16 = feat_bf16;
    o.F = feat ? F : 0;
    o.feat_vec = feat && F % 4 == 0 && reinterpret_cast<uintptr_t>(feat) % (feat_bf16 ? 8 : 16) == 0;
    SampleOut out{};
    out.s_max = s_max;
    out.n_samples = n_samples;
    out.sample_t = sample_t;
    out.nb_idx = nb_idx;
    out.nb_t = nb_t;
    out.nb_dist = nb_dist;
    out.w = w;
    out.agg = agg;
    const int64_t nr = rays ? R : (int64_t)cam->width * cam->height;
    hipStream_t st = as_stream(stream);
    ViewArgs view = make_view(cam, sorted_idx, hash);
    prepare_bounds(o, view, n, cam, sorted_idx, hash, workspace, workspace_bytes, rays ? R : -1, st);
    return launch_sample(o, out, view, rays, nr, mode, bound, st, "knn_sampling");
}

extern "C" int epcq_quick_query(const float* xyz, int64_t n, const epcq_camera* cam, const epcq_options* opt,
                                const int32_t* sorted_idx, const int32_t* hash, const int32_t* rays, int64_t R,
                                int32_t* out, void* stream) {
    if (int e = check_common(cam, opt, "quick_query_for_nearby_point")) return e;
    // the reference's launch limit (epcq_render_topk.cu:317-323)
    const int k2 = opt->kernel_size * opt->kernel_size;
    const int need = (k2 + 31) / 32 * 32;
    int threads = 128;
    if (threads < need) threads = (need + 127) / 128 * 128;
    if (threads >= 1024)
        return fail(EPCQ_ERR_INVALID, "kernel size is too large, can not be handled by current CUDA kernel");
    if (opt->num_point_per_ray < 0) return fail(EPCQ_ERR_INVALID, "quick_query_for_nearby_point: negative num_point_per_ray");
    if (R < 0) return fail(EPCQ_ERR_INVALID, "quick_query_for_nearby_point: negative ray count");
    (void)n;
    if (R == 0 || opt->num_point_per_ray == 0) return EPCQ_OK;
    query_kernel<<<dim3((unsigned)ceil_div(R, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(
        xyz, sorted_idx, reinterpret_cast<const int2*>(hash), cam->c2w, cam->fx, cam->fy, cam->cx, cam->cy,
        cam->width, cam->height, opt->kernel_size, (opt->kernel_size + 1) / 2, opt->radius_threshold, rays, R,
        opt->num_point_per_ray, out);
    return check_launch("quick_query_for_nearby_point");
}

extern "C" int epcq_idw_features(const float* xyz, const void* feat, int32_t feat_bf16, int32_t F,
                                 const float* sp_xyz, const int32_t* nb_idx, int64_t M, int32_t K, float* w,
                                 float* agg, void* stream) {
    if (M < 0 || K < 0 || F < 0) return fail(EPCQ_ERR_INVALID, "idw_features: negative size");
    if (M == 0) return EPCQ_OK;
    idw_kernel<<<dim3((unsigned)ceil_div(M, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(
        xyz, feat, feat_bf16, F, sp_xyz, nb_idx, M, K, w, agg);
    return check_launch("idw_features");
}
