// Primary-surface search, K-nearest sampling and radius query for gfx950.
//
//   epcq_quick_sampling  dense drop-in layout of quick_sampling
//                        (epcq/csrc/epcq_render_topk.cu:10-220,341-417)
//   epcq_knn_sampling    compact layout, optional per-point radius, KNN
//                        neighbours with ray depth / distance, fused IDW
//   epcq_quick_query     radius query for a ray list (:222-339)
//   epcq_idw_features    IDW weights + weighted features (epcq/epcq.py:434-452)
//
// One lane owns one ray.  The 64 rays of a wave are 64 consecutive pixels of
// a row, whose k x k windows overlap almost entirely, so the hash entries,
// point indices and coordinates they gather are shared through L1/L2 and the
// wave stays converged.  A ray's hits are never stored: the candidate walk is
// repeated instead (tmin/tmax pass, slot pass, one pass per live sample), so
// per-lane state stays in registers and nothing spills.
//
// Slot selection keeps a table of the 8 lowest occupied depth slots with
// their distance sums, accumulated in the canonical (neighbour lane, bucket
// position) order.  When more than 8 slots are occupied below the slot where
// the front-to-back walk stops, the slot pass is repeated for the next
// window of slots.  Results equal a full 128-slot histogram.
#include <climits>

#include "epcq_common.h"
#include "epcq_device.h"

namespace epcq {
namespace {

constexpr int kBlock = 256;
constexpr int kSlotTable = 8;

struct SampleArgs {
    const float* xyz;
    const int* sorted_idx;
    const int2* hash;
    const float* radius;  // per-point radius or nullptr
    const float* c2w;
    float fx, fy, cx, cy;
    int W, H;
    int k, ko, L;  // window size, ceil(k/2), k*k
    float r0, t_intvl, alpha_p, gamma;
    int S, Kc;
    const int* rays;  // nullptr = all pixels in order
    int64_t R;
    // dense drop-in layout
    float* sp_xyz;
    int* pc_idx;
    float* sp_t;
    uint8_t* mask;
    // compact layout
    int s_max;
    int* n_samples;
    float* sample_t;
    int* nb_idx;
    float* nb_t;
    float* nb_dist;
    const void* feat;
    int feat_bf16, F;
    float* w;
    float* agg;
};

__device__ __forceinline__ float radius_of(const SampleArgs& a, int id) {
    return a.radius ? a.radius[id] : a.r0;
}

// Every point of the k*k (clamped) neighbour buckets whose point-to-ray
// distance is below its radius, at most 48 per bucket, in lane order then
// bucket order (quick_sampling_kernel :125-155).
template <class Fn>
__device__ __forceinline__ void for_each_ray_hit(const SampleArgs& a, const Ray& ray, int u, int v, Fn&& fn) {
    for (int l = 0; l < a.L; ++l) {
        const int h = neighbor_pixel(a.k, a.ko, l, u, v, a.W, a.H);
        const int2 e = a.hash[h];
        if (e.y <= 0) continue;
        int kept = 0;
        for (int j = 0; j < e.y && kept < kSampleLaneHits; ++j) {
            const int id = a.sorted_idx[e.x + j];
            const float* p = a.xyz + 3 * (int64_t)id;
            float t, dist;
            point_ray(ray, p[0], p[1], p[2], t, dist);
            if (dist < radius_of(a, id)) {
                ++kept;
                fn(t, dist);
            }
        }
    }
}

__device__ __forceinline__ bool knn_less(float d, int id, float bd, int bi) {
    return d < bd || (d == bd && id < bi);
}

// Neighbours of one sample point; writes Kc indices (and, in the compact
// layout, ray depths / distances / IDW outputs) for sample slot `so`.
template <int MODE, int KMAX>
__device__ void sample_neighbours(const SampleArgs& a, const Ray& ray, int u, int v, const float sp[3],
                                  int64_t ri, int so, bool dense) {
    float bd[KMAX];
    int bi[KMAX];
    int bsize = 0;
    if (MODE == EPCQ_MODE_COMPAT) {
        // quick_search_nearby_pc (:39-74): 8-entry buffer that tracks its
        // minimum while filling, then replaces its tracked entry.
        int maxi = 0;
        float maxd = 1e3f;
#pragma unroll
        for (int b = 0; b < KMAX; ++b) {
            bd[b] = 0.0f;
            bi[b] = -1;
        }
        for (int l = 0; l < a.L; ++l) {
            const int h = neighbor_pixel(a.k, a.ko, l, u, v, a.W, a.H);
            const int2 e = a.hash[h];
            if (e.y < 0) continue;
            for (int j = 0; j < e.y; ++j) {
                const int id = a.sorted_idx[e.x + j];
                const float* p = a.xyz + 3 * (int64_t)id;
                const float d = norm3(sp[0] - p[0], sp[1] - p[1], sp[2] - p[2]);
                if (!(d < radius_of(a, id))) continue;
                if (bsize < KMAX) {
#pragma unroll
                    for (int b = 0; b < KMAX; ++b)
                        if (b == bsize) {
                            bd[b] = d;
                            bi[b] = id;
                        }
                    if (maxd > d) {
                        maxd = d;
                        maxi = bsize;
                    }
                    ++bsize;
                } else if (d < maxd) {
#pragma unroll
                    for (int b = 0; b < KMAX; ++b)
                        if (b == maxi) {
                            bd[b] = d;
                            bi[b] = id;
                        }
                    maxd = d;
#pragma unroll
                    for (int b = 0; b < KMAX; ++b) {
                        if (bd[b] > maxd) {
                            maxd = bd[b];
                            maxi = b;
                        }
                    }
                }
            }
        }
    } else {
        // The KMAX nearest by (distance, index) over the distinct window
        // pixels, kept sorted; the first Kc <= KMAX of them are the Kc nearest.
#pragma unroll
        for (int b = 0; b < KMAX; ++b) {
            bd[b] = __builtin_inff();
            bi[b] = INT_MAX;
        }
        const int dv0 = 1 - a.ko, dv1 = a.k - a.ko;
        const int du0 = -(a.ko - 1), du1 = a.k - a.ko;
        const int va = max(0, v + dv0), vz = min(a.H - 1, v + dv1);
        const int ua = max(0, u + du0), uz = min(a.W - 1, u + du1);
        for (int vb = va; vb <= vz; ++vb) {
            for (int ub = ua; ub <= uz; ++ub) {
                const int2 e = a.hash[vb * a.W + ub];
                for (int j = 0; j < e.y; ++j) {
                    const int id = a.sorted_idx[e.x + j];
                    const float* p = a.xyz + 3 * (int64_t)id;
                    const float d = norm3(sp[0] - p[0], sp[1] - p[1], sp[2] - p[2]);
                    if (!(d < radius_of(a, id))) continue;
                    if (!knn_less(d, id, bd[KMAX - 1], bi[KMAX - 1])) continue;
                    bool placed = false;
#pragma unroll
                    for (int b = KMAX - 1; b > 0; --b) {
                        if (!placed) {
                            if (knn_less(d, id, bd[b - 1], bi[b - 1])) {
                                bd[b] = bd[b - 1];
                                bi[b] = bi[b - 1];
                            } else {
                                bd[b] = d;
                                bi[b] = id;
                                placed = true;
                            }
                        }
                    }
                    if (!placed) {
                        bd[0] = d;
                        bi[0] = id;
                    }
                    if (bsize < KMAX) ++bsize;
                }
            }
        }
    }

    const int nkeep = bsize < a.Kc ? bsize : a.Kc;
    if (dense) {
        int* out = a.pc_idx + (ri * a.S + so) * (int64_t)a.Kc;
#pragma unroll
        for (int b = 0; b < KMAX; ++b)
            if (b < a.Kc) out[b] = b < nkeep ? bi[b] : -1;
        for (int b = KMAX; b < a.Kc; ++b) out[b] = -1;
        return;
    }
    const int64_t row = (ri * a.s_max + so) * (int64_t)a.Kc;
    float den = 0.0f;
    float inv[KMAX];
#pragma unroll
    for (int b = 0; b < KMAX; ++b) {
        inv[b] = 0.0f;
        if (b < a.Kc) {
            const bool valid = b < nkeep;
            a.nb_idx[row + b] = valid ? bi[b] : -1;
            if (a.nb_t || a.nb_dist) {
                float t = 0.0f, dist = 0.0f;
                if (valid) {
                    const float* p = a.xyz + 3 * (int64_t)bi[b];
                    point_ray(ray, p[0], p[1], p[2], t, dist);
                }
                if (a.nb_t) a.nb_t[row + b] = t;
                if (a.nb_dist) a.nb_dist[row + b] = dist;
            }
            if (valid) {
                // (1 / (d + 1e-6)) * mask, epcq.py:436
                inv[b] = 1.0f / (bd[b] + 1e-6f);
                den = den + inv[b];
            }
        }
    }
    for (int b = KMAX; b < a.Kc; ++b) {
        a.nb_idx[row + b] = -1;
        if (a.nb_t) a.nb_t[row + b] = 0.0f;
        if (a.nb_dist) a.nb_dist[row + b] = 0.0f;
    }
    if (!a.feat) return;
    den = den + 1e-6f;
    float* wo = a.w + row;
#pragma unroll
    for (int b = 0; b < KMAX; ++b) {
        if (b < a.Kc) wo[b] = inv[b] / den;
    }
    for (int b = KMAX; b < a.Kc; ++b) wo[b] = 0.0f;
    float* ao = a.agg + (ri * a.s_max + so) * (int64_t)a.F;
    for (int f0 = 0; f0 < a.F; f0 += 4) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < KMAX; ++b) {
            if (b < nkeep) {
                const float wb = inv[b] / den;
                const int64_t base = (int64_t)bi[b] * a.F + f0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (f0 + c < a.F) {
                        float fv;
                        if (a.feat_bf16) {
                            const uint16_t hb = static_cast<const uint16_t*>(a.feat)[base + c];
                            fv = __uint_as_float((uint32_t)hb << 16);
                        } else {
                            fv = static_cast<const float*>(a.feat)[base + c];
                        }
                        acc[c] = acc[c] + wb * fv;
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (f0 + c < a.F) ao[f0 + c] = acc[c];
    }
}

template <int MODE, int KMAX>
__global__ __launch_bounds__(kBlock) void sample_kernel(SampleArgs a) {
    const int64_t ri = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (ri >= a.R) return;
    const bool dense = a.sp_xyz != nullptr;
    const int64_t q = a.rays ? (int64_t)a.rays[ri] : ri;
    const int64_t P = (int64_t)a.W * a.H;
    int ns = 0;
    if (q >= 0 && q < P) {
        const int u = (int)(q % a.W), v = (int)(q / a.W);
        float m[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) m[i] = a.c2w[i];
        Ray ray;
        make_ray(m, a.fx, a.fy, a.cx, a.cy, u, v, ray);

        // Step 1: tmin / tmax over the hits (WarpReduce min / max, :156-161).
        float tmin = 1e3f, tmax = -1e3f;
        for_each_ray_hit(a, ray, u, v, [&](float t, float) {
            if (tmin > t) tmin = t;
            if (tmax < t) tmax = t;
        });

        if (!(tmin > tmax)) {
            // Steps 2-3: slot sums and the front-to-back walk (:168-219).
            float T = 1.0f;
            int lo = 0;
            bool done = false;
            while (!done) {
                int tslot[kSlotTable];
                float tsum[kSlotTable], tcnt[kSlotTable];
#pragma unroll
                for (int e = 0; e < kSlotTable; ++e) {
                    tslot[e] = INT_MAX;
                    tsum[e] = 0.0f;
                    tcnt[e] = 0.0f;
                }
                int nt = 0;
                int limit = kMaxSlots;
                for_each_ray_hit(a, ray, u, v, [&](float t, float dist) {
                    const float qf = (t - tmin) / a.t_intvl;
                    if (!(qf > -1.0f && qf < (float)kMaxSlots)) return;
                    const int s = (int)qf;
                    if (s < lo || s >= limit) return;
                    bool found = false;
#pragma unroll
                    for (int e = 0; e < kSlotTable; ++e) {
                        if (tslot[e] == s) {
                            tsum[e] = tsum[e] + dist;
                            tcnt[e] = tcnt[e] + 1.0f;
                            found = true;
                        }
                    }
                    if (found) return;
                    if (nt < kSlotTable) {
#pragma unroll
                        for (int e = 0; e < kSlotTable; ++e)
                            if (e == nt) {
                                tslot[e] = s;
                                tsum[e] = dist;
                                tcnt[e] = 1.0f;
                            }
                        ++nt;
                        return;
                    }
                    int mslot = -1, mx = 0;
#pragma unroll
                    for (int e = 0; e < kSlotTable; ++e)
                        if (tslot[e] > mslot) {
                            mslot = tslot[e];
                            mx = e;
                        }
                    if (s > mslot) {
                        limit = min(limit, s);
                        return;
                    }
                    limit = mslot;  // the evicted slot and everything above wait for the next pass
#pragma unroll
                    for (int e = 0; e < kSlotTable; ++e)
                        if (e == mx) {
                            tslot[e] = s;
                            tsum[e] = dist;
                            tcnt[e] = 1.0f;
                        }
                });
                // walk this window's slots in ascending order
                int prev = lo - 1;
                for (int it = 0; it < nt && !done; ++it) {
                    int best = INT_MAX;
                    float bsum = 0.0f, bcnt = 0.0f;
#pragma unroll
                    for (int e = 0; e < kSlotTable; ++e)
                        if (tslot[e] > prev && tslot[e] < best) {
                            best = tslot[e];
                            bsum = tsum[e];
                            bcnt = tcnt[e];
                        }
                    prev = best;
                    if (bcnt < 2.0f) continue;
                    if (T < 1e-3f || ns >= a.S) {
                        done = true;
                        break;
                    }
                    const float sdf = bsum / (bcnt + 1e-6f);
                    const float alpha = a.gamma * expf_cw(-sdf / a.alpha_p);
                    const float wgt = alpha * T;
                    if (wgt > 1e-3f) {
                        const float ts = tmin + (float)best * a.t_intvl;
                        float sp[3];
                        sp[0] = ray.o[0] + ts * ray.d[0];
                        sp[1] = ray.o[1] + ts * ray.d[1];
                        sp[2] = ray.o[2] + ts * ray.d[2];
                        if (dense) {
                            const int64_t o = ri * a.S + ns;
                            a.sp_xyz[3 * o + 0] = sp[0];
                            a.sp_xyz[3 * o + 1] = sp[1];
                            a.sp_xyz[3 * o + 2] = sp[2];
                            a.sp_t[o] = ts;
                            a.mask[o] = 1;
                            sample_neighbours<MODE, KMAX>(a, ray, u, v, sp, ri, ns, true);
                        } else if (ns < a.s_max) {
                            a.sample_t[ri * a.s_max + ns] = ts;
                            sample_neighbours<MODE, KMAX>(a, ray, u, v, sp, ri, ns, false);
                        }
                        ++ns;
                    }
                    T -= alpha;
                }
                if (!done) {
                    if (limit >= kMaxSlots) break;
                    lo = limit;
                }
            }
        }
    }

    // Everything not written above: zeros / -1 / false.
    if (dense) {
        for (int s = ns; s < a.S; ++s) {
            const int64_t o = ri * a.S + s;
            a.sp_xyz[3 * o + 0] = 0.0f;
            a.sp_xyz[3 * o + 1] = 0.0f;
            a.sp_xyz[3 * o + 2] = 0.0f;
            a.sp_t[o] = 0.0f;
            a.mask[o] = 0;
            int* out = a.pc_idx + o * a.Kc;
            for (int b = 0; b < a.Kc; ++b) out[b] = -1;
        }
    } else {
        a.n_samples[ri] = ns;
        for (int s = ns; s < a.s_max; ++s) {
            const int64_t o = ri * a.s_max + s;
            a.sample_t[o] = 0.0f;
            for (int b = 0; b < a.Kc; ++b) {
                a.nb_idx[o * a.Kc + b] = -1;
                if (a.nb_t) a.nb_t[o * a.Kc + b] = 0.0f;
                if (a.nb_dist) a.nb_dist[o * a.Kc + b] = 0.0f;
                if (a.feat) a.w[o * a.Kc + b] = 0.0f;
            }
            if (a.feat)
                for (int f = 0; f < a.F; ++f) a.agg[o * a.F + f] = 0.0f;
        }
    }
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
    if (!cam->c2w) return fail(EPCQ_ERR_INVALID, std::string(who) + ": null c2w");
    return EPCQ_OK;
}

SampleArgs make_args(const float* xyz, const epcq_camera* cam, const epcq_options* opt, const int* sorted_idx,
                     const int* hash, const float* radius) {
    SampleArgs a{};
    a.xyz = xyz;
    a.sorted_idx = sorted_idx;
    a.hash = reinterpret_cast<const int2*>(hash);
    a.radius = radius;
    a.c2w = cam->c2w;
    a.fx = cam->fx;
    a.fy = cam->fy;
    a.cx = cam->cx;
    a.cy = cam->cy;
    a.W = cam->width;
    a.H = cam->height;
    a.k = opt->kernel_size;
    a.ko = (opt->kernel_size + 1) / 2;
    a.L = opt->kernel_size * opt->kernel_size;
    a.r0 = opt->radius_threshold;
    a.t_intvl = opt->t_intvl;
    a.alpha_p = opt->sdf2weight_gaussian_alpha;
    a.gamma = opt->sdf2weight_gaussian_gamma;
    a.S = opt->num_sample_points_per_ray;
    a.Kc = opt->num_point_cloud_per_sp;
    return a;
}

int launch_sample(const SampleArgs& a, int mode, hipStream_t st, const char* who) {
    if (a.R == 0) return EPCQ_OK;
    const dim3 grid((unsigned)ceil_div(a.R, kBlock)), block(kBlock);
    if (mode == EPCQ_MODE_COMPAT) {
        sample_kernel<EPCQ_MODE_COMPAT, kCompatBuffer><<<grid, block, 0, st>>>(a);
    } else if (a.Kc <= 8) {
        sample_kernel<EPCQ_MODE_KNN, 8><<<grid, block, 0, st>>>(a);
    } else if (a.Kc <= 16) {
        sample_kernel<EPCQ_MODE_KNN, 16><<<grid, block, 0, st>>>(a);
    } else {
        sample_kernel<EPCQ_MODE_KNN, 32><<<grid, block, 0, st>>>(a);
    }
    return check_launch(who);
}

}  // namespace
}  // namespace epcq

using namespace epcq;

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

extern "C" int epcq_quick_sampling(const float* xyz, int64_t n, const epcq_camera* cam, const epcq_options* opt,
                                   const int32_t* sorted_idx, const int32_t* hash, int32_t mode, float* sp_xyz,
                                   int32_t* pc_idx, float* sp_t, uint8_t* mask, void* stream) {
    if (int e = check_common(cam, opt, "quick_sampling")) return e;
    if (mode != EPCQ_MODE_COMPAT && mode != EPCQ_MODE_KNN)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: unknown mode");
    if (opt->num_sample_points_per_ray < 0 || opt->num_point_cloud_per_sp < 0)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: negative output size");
    if (mode == EPCQ_MODE_KNN && opt->num_point_cloud_per_sp > 32)
        return fail(EPCQ_ERR_INVALID, "quick_sampling: num_point_cloud_per_sp > 32 in KNN mode");
    (void)n;
    SampleArgs a = make_args(xyz, cam, opt, sorted_idx, hash, nullptr);
    a.R = (int64_t)cam->width * cam->height;
    a.sp_xyz = sp_xyz;
    a.pc_idx = pc_idx;
    a.sp_t = sp_t;
    a.mask = mask;
    if (a.S == 0) return EPCQ_OK;
    return launch_sample(a, mode, as_stream(stream), "quick_sampling");
}

extern "C" int epcq_knn_sampling(const float* xyz, int64_t n, const epcq_camera* cam, const epcq_options* opt,
                                 const int32_t* sorted_idx, const int32_t* hash, const float* radius,
                                 const int32_t* rays, int64_t R, int32_t mode, int32_t s_max, int32_t* n_samples,
                                 float* sample_t, int32_t* nb_idx, float* nb_t, float* nb_dist, const void* feat,
                                 int32_t feat_bf16, int32_t F, float* w, float* agg, void* stream) {
    if (int e = check_common(cam, opt, "knn_sampling")) return e;
    if (mode != EPCQ_MODE_COMPAT && mode != EPCQ_MODE_KNN)
        return fail(EPCQ_ERR_INVALID, "knn_sampling: unknown mode");
    if (opt->num_point_cloud_per_sp < 1 || opt->num_point_cloud_per_sp > 32)
        return fail(EPCQ_ERR_INVALID, "knn_sampling: num_point_cloud_per_sp must be in [1, 32]");
    if (s_max < 1) return fail(EPCQ_ERR_INVALID, "knn_sampling: s_max must be >= 1");
    if (!n_samples || !sample_t || !nb_idx) return fail(EPCQ_ERR_INVALID, "knn_sampling: null output");
    if (feat && (F < 1 || !w || !agg)) return fail(EPCQ_ERR_INVALID, "knn_sampling: bad feature arguments");
    (void)n;
    SampleArgs a = make_args(xyz, cam, opt, sorted_idx, hash, radius);
    a.rays = rays;
    a.R = rays ? R : (int64_t)cam->width * cam->height;
    if (a.R < 0) return fail(EPCQ_ERR_INVALID, "knn_sampling: negative ray count");
    a.s_max = s_max;
    a.n_samples = n_samples;
    a.sample_t = sample_t;
    a.nb_idx = nb_idx;
    a.nb_t = nb_t;
    a.nb_dist = nb_dist;
    a.feat = feat;
    a.feat_bf16 = feat_bf16;
    a.F = feat ? F : 0;
    a.w = w;
    a.agg = agg;
    return launch_sample(a, mode, as_stream(stream), "knn_sampling");
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
