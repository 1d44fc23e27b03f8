// CPU oracle for the epcq searching + sampling path.  TEST INFRASTRUCTURE
// ONLY: the product never links or calls this file.  Semantics and the
// canonical choices it fixes are described in epcq_oracle.h.
//
// Build: g++ -O2 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
#include "epcq_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

// Constants of the reference (epcq/csrc/include/cuda_util.cuh:12-19).
constexpr int kQueryLaneHits = 32;   // POINT_BUFFER_SIZE
constexpr int kCompatBuffer = 8;     // POINT_BUFFER_SIZE8
constexpr int kSampleLaneHits = 48;  // POINT_BUFFER_SIZE48
constexpr int kMaxSlots = 128;       // MAX_SAMPLE_POINTS_NUM

inline float pow2i(int m) {  // exact 2^m for m in [-126, 127]
    uint32_t b = (uint32_t)(m + 127) << 23;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

inline float dot3(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

inline float norm3(float x, float y, float z) {
    float s = x * x;
    s = s + y * y;
    s = s + z * z;
    return sqrtf(s);
}

// calculate_neighbor_uv (cuda_util.cuh:382-388) in integer form:
// ceil((l+1)/k) == (l+k)/k and ceil(k/2.0f) == (k+1)/2 for the int ranges used.
inline void neighbor(int k, int lane, int u, int v, int W, int H, int* ub, int* vb) {
    const int ko = (k + 1) / 2;
    const int dv = (lane + k) / k - ko;
    const int du = (lane % k) - (ko - 1);
    int uu = u + du, vv = v + dv;
    uu = uu < 0 ? 0 : (uu > W - 1 ? W - 1 : uu);
    vv = vv < 0 ? 0 : (vv > H - 1 ? H - 1 : vv);
    *ub = uu;
    *vb = vv;
}

struct Ray {
    float o[3];
    float d[3];
    float dd;
};

inline void make_ray(const orc_camera* cam, int u, int v, Ray* r) {
    orc_ray(cam, u, v, r->o, r->d);
    r->dd = dot3(r->d, r->d);
}

// SingleRaySpec::PointRayInterset (cuda_util.cuh:215-236).
inline void point_ray(const Ray& r, const float* p, float* t_out, float* dist_out) {
    float w[3] = {p[0] - r.o[0], p[1] - r.o[1], p[2] - r.o[2]};
    float t = dot3(w, r.d) / r.dd;
    float q0 = r.o[0] + t * r.d[0];
    float q1 = r.o[1] + t * r.d[1];
    float q2 = r.o[2] + t * r.d[2];
    *t_out = t;
    *dist_out = norm3(p[0] - q0, p[1] - q1, p[2] - q2);
}

// binary_search (cuda_util.cuh:558-579): leftmost index of target, or -1.
inline int64_t leftmost(const int32_t* data, int64_t n, int32_t target) {
    int64_t left = 0, right = n - 1, result = -1;
    while (left <= right) {
        int64_t mid = (left + right) / 2;
        if (data[mid] == target) {
            result = mid;
            right = mid - 1;
        } else if (data[mid] < target) {
            left = mid + 1;
        } else {
            right = mid - 1;
        }
    }
    return result;
}

struct Hit {
    float t, dist;
};

}  // namespace

extern "C" {

void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

// exp(x): Cody-Waite reduction x = n*ln2 + r with ln2 split in two parts,
// degree-7 Taylor polynomial of e^r on |r| <= ln2/2 in Horner form with
// fused multiply-adds, then exact scaling by 2^n in two steps.
float orc_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541015625f) return 0.0f;
    const float fn = rintf(x * 1.44269502162933349609375f);
    const int n = (int)fn;
    float r = fmaf(-fn, 0.693145751953125f, x);
    r = fmaf(-fn, 1.428606765330187045e-06f, r);
    float p = 1.98412698412698412698e-4f;
    p = fmaf(p, r, 1.38888888888888888889e-3f);
    p = fmaf(p, r, 8.33333333333333333333e-3f);
    p = fmaf(p, r, 4.16666666666666666667e-2f);
    p = fmaf(p, r, 1.66666666666666666667e-1f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const int n1 = n / 2;
    const int n2 = n - n1;
    return (p * pow2i(n1)) * pow2i(n2);
}

void orc_neighbor_offset(int32_t k, int32_t lane, int32_t* du, int32_t* dv) {
    const int ko = (k + 1) / 2;
    *dv = (lane + k) / k - ko;
    *du = (lane % k) - (ko - 1);
}

// cam2world_ray (cuda_util.cuh:338-354): pixel centre in double, rounded to
// float; direction = c2w[:3,:3] (x, y, 1); normalised.
void orc_ray(const orc_camera* cam, int32_t u, int32_t v, float* origin, float* dir) {
    const float x = (float)(((double)u + 0.5 - (double)cam->cx) / (double)cam->fx);
    const float y = (float)(((double)v + 0.5 - (double)cam->cy) / (double)cam->fy);
    const float z = 1.0f;
    const float* m = cam->c2w;
    float d0 = m[0] * x;
    d0 = d0 + m[1] * y;
    d0 = d0 + m[2] * z;
    float d1 = m[4] * x;
    d1 = d1 + m[5] * y;
    d1 = d1 + m[6] * z;
    float d2 = m[8] * x;
    d2 = d2 + m[9] * y;
    d2 = d2 + m[10] * z;
    float s = d0 * d0;
    s = s + d1 * d1;
    s = s + d2 * d2;
    const float rn = 1.0f / sqrtf(s);
    dir[0] = d0 * rn;
    dir[1] = d1 * rn;
    dir[2] = d2 * rn;
    origin[0] = m[3];
    origin[1] = m[7];
    origin[2] = m[11];
}

void orc_point_ray(const float* origin, const float* dir, const float* p, float* t, float* dist) {
    Ray r;
    std::memcpy(r.o, origin, 12);
    std::memcpy(r.d, dir, 12);
    r.dd = dot3(r.d, r.d);
    point_ray(r, p, t, dist);
}

// NeuralPoints._world2cam / _cam2image / _convetpixel (epcq.py:251-274).
int orc_project(const float* xyz, int64_t n, const float* w, const float* K, int32_t H, int32_t W,
                int32_t* key, float* depth) {
    const int32_t sentinel = H * W;
    const float fW = (float)W, fH = (float)H;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        float X = w[0] * x;
        X = X + w[1] * y;
        X = X + w[2] * z;
        X = X + w[3];
        float Y = w[4] * x;
        Y = Y + w[5] * y;
        Y = Y + w[6] * z;
        Y = Y + w[7];
        float Z = w[8] * x;
        Z = Z + w[9] * y;
        Z = Z + w[10] * z;
        Z = Z + w[11];
        float sx = K[0] * X;
        sx = sx + K[1] * Y;
        sx = sx + K[2] * Z;
        float sy = K[3] * X;
        sy = sy + K[4] * Y;
        sy = sy + K[5] * Z;
        float sz = K[6] * X;
        sz = sz + K[7] * Y;
        sz = sz + K[8] * Z;
        sx = sx / sz;
        sy = sy / sz;
        // int32 truncation toward zero, then 0 <= u < W and 0 <= v < H; the
        // comparisons below are that test without converting out-of-range
        // floats (NaN fails them all and lands off-image).
        const bool in = (sx > -1.0f) && (sx < fW) && (sy > -1.0f) && (sy < fH);
        key[i] = in ? ((int32_t)sx + (int32_t)sy * W) : sentinel;
        depth[i] = Z;
    }
    return 0;
}

int orc_sort_by_key(const int32_t* key, const float* depth, int64_t n, int32_t* key_sorted,
                    int32_t* idx_sorted, float* depth_sorted) {
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    for (int64_t i = 0; i < n; ++i) {
        key_sorted[i] = key[perm[i]];
        idx_sorted[i] = perm[i];
        depth_sorted[i] = depth[perm[i]];
    }
    return 0;
}

// scatter_hashtable (rasterization.cu:7-34, 130-171).  The run walk is
// bounded by n (the reference reads past the end when the last pixel's run
// reaches it).
int orc_scatter_hashtable(const int32_t* point_idx, const int32_t* keys, const float* depth, int64_t n,
                          int32_t H, int32_t W, int32_t sort, int32_t* out_idx, float* out_depth,
                          int32_t* out_hash) {
    const int64_t P = (int64_t)H * W;
    for (int64_t i = 0; i < n; ++i) {
        out_idx[i] = -1;
        out_depth[i] = -1.0f;
    }
    for (int64_t p = 0; p < 2 * P; ++p) out_hash[p] = -1;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t p = 0; p < P; ++p) {
        const int64_t first = leftmost(keys, n, (int32_t)p);
        if (first < 0) continue;
        int64_t end = first;
        while (end < n && keys[end] == (int32_t)p) {
            out_idx[end] = point_idx[end];
            out_depth[end] = depth[end];
            ++end;
        }
        const int cnt = (int)(end - first);
        if (sort) {  // bubble_sort<float>(data, index, size, ascending), cuda_util.cuh:291-313
            float* d = out_depth + first;
            int32_t* ix = out_idx + first;
            for (int i = 0; i < cnt; ++i) {
                bool swapped = false;
                for (int j = 0; j < cnt - i - 1; ++j) {
                    if (d[j] > d[j + 1]) {
                        std::swap(d[j], d[j + 1]);
                        std::swap(ix[j], ix[j + 1]);
                        swapped = true;
                    }
                }
                if (!swapped) break;
            }
        }
        out_hash[2 * p] = (int32_t)first;
        out_hash[2 * p + 1] = cnt;
    }
    return 0;
}

// NeuralPoints.pts2plane (epcq.py:279-295) with the canonical stable sort.
int orc_pts2plane(const float* xyz, int64_t n, const float* w2c, const float* K, int32_t H, int32_t W,
                  int32_t sort, int32_t* out_idx, float* out_depth, int32_t* out_hash) {
    const int64_t P = (int64_t)H * W;
    std::vector<int32_t> key(n), ks(n), is(n);
    std::vector<float> dep(n), ds(n);
    orc_project(xyz, n, w2c, K, H, W, key.data(), dep.data());
    // stable counting sort over keys in [0, P]
    std::vector<int64_t> off(P + 2, 0);
    for (int64_t i = 0; i < n; ++i) off[key[i] + 1]++;
    for (int64_t p = 0; p <= P; ++p) off[p + 1] += off[p];
    for (int64_t i = 0; i < n; ++i) {
        const int64_t dst = off[key[i]]++;
        ks[dst] = key[i];
        is[dst] = (int32_t)i;
        ds[dst] = dep[i];
    }
    return orc_scatter_hashtable(is.data(), ks.data(), ds.data(), n, H, W, sort, out_idx, out_depth, out_hash);
}

// scatter (rasterization.cu:36-128): B nearest by depth per pixel, ascending.
// Negative depths are skipped (the reference loops forever on them).
int orc_scatter(const int32_t* point_idx, const int32_t* keys, const float* depth, int64_t n, int32_t H,
                int32_t W, int32_t B, float* out_depth, int32_t* out_idx) {
    const int64_t P = (int64_t)H * W;
    for (int64_t i = 0; i < P * B; ++i) {
        out_depth[i] = -1.0f;
        out_idx[i] = -1;
    }
#pragma omp parallel
    {
        std::vector<float> qd(B > 0 ? B : 1);
        std::vector<int32_t> qi(B > 0 ? B : 1);
#pragma omp for schedule(dynamic, 1024)
        for (int64_t p = 0; p < P; ++p) {
            int64_t i = leftmost(keys, n, (int32_t)p);
            if (i < 0 || B <= 0) continue;
            int qs = 0, qmax = -1;
            float qmaxz = -1000.0f;
            for (; i < n && keys[i] == (int32_t)p; ++i) {
                const float d = depth[i];
                if (d < 0) continue;
                const int32_t id = point_idx[i];
                if (qs < B) {
                    qd[qs] = d;
                    qi[qs] = id;
                    if (d > qmaxz) {
                        qmaxz = d;
                        qmax = qs;
                    }
                    ++qs;
                } else if (d < qmaxz) {
                    qd[qmax] = d;
                    qi[qmax] = id;
                    qmaxz = d;
                    for (int j = 0; j < B; ++j) {
                        if (qd[j] > qmaxz) {
                            qmaxz = qd[j];
                            qmax = j;
                        }
                    }
                }
            }
            for (int a = 0; a < qs; ++a) {  // bubble_sort<PixSpec> ascending
                bool swapped = false;
                for (int j = 0; j < qs - a - 1; ++j) {
                    if (qd[j] > qd[j + 1]) {
                        std::swap(qd[j], qd[j + 1]);
                        std::swap(qi[j], qi[j + 1]);
                        swapped = true;
                    }
                }
                if (!swapped) break;
            }
            for (int j = 0; j < qs; ++j) {
                out_depth[p * B + j] = qd[j];
                out_idx[p * B + j] = qi[j];
            }
        }
    }
    return 0;
}

// quick_sampling_kernel (epcq_render_topk.cu:82-220) followed by
// quick_search_nearby_pc (epcq_render_topk.cu:10-81) or the canonical KNN.
int orc_quick_sampling(const float* xyz, int64_t n, const orc_camera* cam, const orc_options* opt,
                       const int32_t* sorted_idx, const int32_t* hash, const float* radius, int32_t mode,
                       const int32_t* rays, int64_t R, float* sp_xyz, int32_t* pc_idx, float* sp_t,
                       uint8_t* mask, float* nb_t, float* nb_dist) {
    (void)n;
    const int H = cam->height, W = cam->width;
    const int64_t P = (int64_t)H * W;
    const int64_t nrays = rays ? R : P;
    const int S = opt->num_sample_points_per_ray;
    const int Kc = opt->num_point_cloud_per_sp;
    const int k = opt->kernel_size;
    const int L = k * k;
    const float r0 = opt->radius_threshold;

    for (int64_t i = 0; i < nrays * S; ++i) {
        sp_t[i] = 0.0f;
        mask[i] = 0;
        sp_xyz[3 * i] = sp_xyz[3 * i + 1] = sp_xyz[3 * i + 2] = 0.0f;
    }
    for (int64_t i = 0; i < nrays * S * Kc; ++i) {
        pc_idx[i] = -1;
        if (nb_t) nb_t[i] = 0.0f;
        if (nb_dist) nb_dist[i] = 0.0f;
    }

#pragma omp parallel
    {
        std::vector<Hit> hits;
        std::vector<std::pair<float, int32_t>> cand;
        float sum[kMaxSlots], cnt[kMaxSlots];
#pragma omp for schedule(dynamic, 256)
        for (int64_t ri = 0; ri < nrays; ++ri) {
            const int64_t q = rays ? rays[ri] : ri;
            if (q < 0 || q >= P) continue;
            const int u = (int)(q % W), v = (int)(q / W);
            Ray ray;
            make_ray(cam, u, v, &ray);

            // Step 1: per-lane hits (first <= 48 per bucket) and tmin / tmax.
            hits.clear();
            float tmin = 1e3f, tmax = -1e3f;
            for (int l = 0; l < L; ++l) {
                int ub, vb;
                neighbor(k, l, u, v, W, H, &ub, &vb);
                const int64_t h = (int64_t)vb * W + ub;
                const int start = hash[2 * h], c = hash[2 * h + 1];
                if (c <= 0) continue;
                int kept = 0;
                for (int j = 0; j < c && kept < kSampleLaneHits; ++j) {
                    const int32_t id = sorted_idx[start + j];
                    float t, dist;
                    point_ray(ray, xyz + 3 * (int64_t)id, &t, &dist);
                    const float rr = radius ? radius[id] : r0;
                    if (dist < rr) {
                        if (tmin > t) tmin = t;
                        if (tmax < t) tmax = t;
                        hits.push_back({t, dist});
                        ++kept;
                    }
                }
            }
            if (tmin > tmax) continue;  // empty ray

            // Step 2: slot histogram, canonical (lane, bucket) summation order.
            for (int s = 0; s < kMaxSlots; ++s) sum[s] = cnt[s] = 0.0f;
            for (const Hit& hh : hits) {
                const float qf = (hh.t - tmin) / opt->t_intvl;
                if (!(qf > -1.0f && qf < (float)kMaxSlots)) continue;
                const int s = (int)qf;
                sum[s] = sum[s] + hh.dist;
                cnt[s] = cnt[s] + 1.0f;
            }
            // Step 3: front-to-back slot selection (lane 0 loop, :181-219).
            float T = 1.0f;
            int ns = 0;
            float* o_xyz = sp_xyz + ri * S * 3;
            float* o_t = sp_t + ri * S;
            uint8_t* o_m = mask + ri * S;
            for (int s = 0; s < kMaxSlots; ++s) {
                if (cnt[s] < 2.0f) continue;
                if (T < 1e-3f) break;
                if (ns >= S) break;
                const float sdf = sum[s] / (cnt[s] + 1e-6f);
                const float alpha = opt->sdf2weight_gaussian_gamma * orc_expf(-sdf / opt->sdf2weight_gaussian_alpha);
                const float wgt = alpha * T;
                if (wgt > 1e-3f) {
                    const float t = tmin + (float)s * opt->t_intvl;
                    o_xyz[3 * ns + 0] = ray.o[0] + t * ray.d[0];
                    o_xyz[3 * ns + 1] = ray.o[1] + t * ray.d[1];
                    o_xyz[3 * ns + 2] = ray.o[2] + t * ray.d[2];
                    o_t[ns] = t;
                    o_m[ns] = 1;
                    ++ns;
                }
                T -= alpha;
            }

            // Step 4: neighbours of every emitted sample.
            for (int j = 0; j < ns; ++j) {
                const float* sp = o_xyz + 3 * j;
                int32_t* out = pc_idx + (ri * S + j) * Kc;
                if (mode == ORC_MODE_COMPAT) {
                    float bd[kCompatBuffer];
                    int32_t bi[kCompatBuffer];
                    int bsize = 0, maxi = 0;
                    float maxd = 1e3f;
                    for (int l = 0; l < L; ++l) {
                        int ub, vb;
                        neighbor(k, l, u, v, W, H, &ub, &vb);
                        const int64_t h = (int64_t)vb * W + ub;
                        const int start = hash[2 * h], c = hash[2 * h + 1];
                        if (c < 0) continue;
                        for (int e = start; e < start + c; ++e) {
                            const int32_t id = sorted_idx[e];
                            const float* p = xyz + 3 * (int64_t)id;
                            const float d = norm3(sp[0] - p[0], sp[1] - p[1], sp[2] - p[2]);
                            const float rr = radius ? radius[id] : r0;
                            if (!(d < rr)) continue;
                            if (bsize < kCompatBuffer) {
                                bd[bsize] = d;
                                bi[bsize] = id;
                                if (maxd > d) {  // tracks the minimum while filling (:54-57)
                                    maxd = d;
                                    maxi = bsize;
                                }
                                ++bsize;
                            } else if (d < maxd) {
                                bd[maxi] = d;
                                bi[maxi] = id;
                                maxd = d;
                                for (int b = 0; b < kCompatBuffer; ++b) {
                                    if (bd[b] > maxd) {
                                        maxd = bd[b];
                                        maxi = b;
                                    }
                                }
                            }
                        }
                    }
                    const int nw = bsize < Kc ? bsize : Kc;
                    for (int b = 0; b < nw; ++b) out[b] = bi[b];
                } else {
                    // canonical KNN over the distinct pixels of the window
                    const int ko = (k + 1) / 2;
                    const int dv0 = 1 - ko, dv1 = (L - 1 + k) / k - ko;
                    const int du0 = -(ko - 1), du1 = (k - 1) - (ko - 1);
                    const int va = std::max(0, v + dv0), vz = std::min(H - 1, v + dv1);
                    const int ua = std::max(0, u + du0), uz = std::min(W - 1, u + du1);
                    cand.clear();
                    for (int vb = va; vb <= vz; ++vb) {
                        for (int ub = ua; ub <= uz; ++ub) {
                            const int64_t h = (int64_t)vb * W + ub;
                            const int start = hash[2 * h], c = hash[2 * h + 1];
                            for (int e = start; e < start + c; ++e) {
                                const int32_t id = sorted_idx[e];
                                const float* p = xyz + 3 * (int64_t)id;
                                const float d = norm3(sp[0] - p[0], sp[1] - p[1], sp[2] - p[2]);
                                const float rr = radius ? radius[id] : r0;
                                if (d < rr) cand.push_back({d, id});
                            }
                        }
                    }
                    std::sort(cand.begin(), cand.end());
                    const int nw = (int)std::min<size_t>(cand.size(), (size_t)Kc);
                    for (int b = 0; b < nw; ++b) {
                        out[b] = cand[b].second;
                        float t, dist;
                        point_ray(ray, xyz + 3 * (int64_t)cand[b].second, &t, &dist);
                        if (nb_t) nb_t[(ri * S + j) * Kc + b] = t;
                        if (nb_dist) nb_dist[(ri * S + j) * Kc + b] = dist;
                    }
                }
            }
        }
    }
    return 0;
}

// hashtable_query_for_nearby_point (epcq_render_topk.cu:222-339); output in
// lane order instead of atomic-arrival order.
int orc_quick_query(const float* xyz, int64_t n, const orc_camera* cam, const orc_options* opt,
                    const int32_t* sorted_idx, const int32_t* hash, const int32_t* rays, int64_t R,
                    int32_t* out) {
    (void)n;
    const int H = cam->height, W = cam->width;
    const int64_t P = (int64_t)H * W;
    const int cap = opt->num_point_per_ray;
    const int k = opt->kernel_size;
    const int L = k * k;
    const float r0 = opt->radius_threshold;
    for (int64_t i = 0; i < R * cap; ++i) out[i] = -1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ri = 0; ri < R; ++ri) {
        const int64_t q = rays[ri];
        if (q < 0 || q >= P) continue;
        const int u = (int)(q % W), v = (int)(q / W);
        Ray ray;
        make_ray(cam, u, v, &ray);
        int total = 0;
        for (int l = 0; l < L && total < cap; ++l) {
            int ub, vb;
            neighbor(k, l, u, v, W, H, &ub, &vb);
            const int64_t h = (int64_t)vb * W + ub;
            const int start = hash[2 * h], c = hash[2 * h + 1];
            if (c < 0) continue;
            int kept = 0;
            for (int j = 0; j < c && kept < kQueryLaneHits; ++j) {
                const int32_t id = sorted_idx[start + j];
                float t, dist;
                point_ray(ray, xyz + 3 * (int64_t)id, &t, &dist);
                if (dist < r0) {
                    if (total < cap) out[ri * cap + total] = id;
                    ++total;
                    ++kept;
                }
            }
        }
    }
    return 0;
}

// FeatureQuery gathers + Aggregator.calculate_weight (epcq.py:345-387,
// 434-437, distance at 449-452), sums in index order.
int orc_idw(const float* xyz, const void* feat, int32_t feat_bf16, int32_t F, const float* sp_xyz,
            const int32_t* nb_idx, int64_t M, int32_t K, float* w, float* agg) {
#pragma omp parallel
    {
        std::vector<float> inv(K > 0 ? K : 1);
#pragma omp for schedule(static)
        for (int64_t m = 0; m < M; ++m) {
            const float* sp = sp_xyz + 3 * m;
            const int32_t* ids = nb_idx + m * K;
            float den = 0.0f;
            for (int j = 0; j < K; ++j) {
                const int32_t id = ids[j];
                float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
                if (id >= 0) {
                    p0 = xyz[3 * (int64_t)id];
                    p1 = xyz[3 * (int64_t)id + 1];
                    p2 = xyz[3 * (int64_t)id + 2];
                }
                const float d = norm3(p0 - sp[0], p1 - sp[1], p2 - sp[2]);
                inv[j] = (1.0f / (d + 1e-6f)) * (id >= 0 ? 1.0f : 0.0f);
                den = den + inv[j];
            }
            den = den + 1e-6f;
            float* wo = w + m * K;
            for (int j = 0; j < K; ++j) wo[j] = inv[j] / den;
            float* ao = agg + m * (int64_t)F;
            for (int f = 0; f < F; ++f) ao[f] = 0.0f;
            for (int j = 0; j < K; ++j) {
                const int32_t id = ids[j];
                if (id < 0) continue;
                for (int f = 0; f < F; ++f) {
                    float fv;
                    if (feat_bf16) {
                        const uint16_t b = ((const uint16_t*)feat)[(int64_t)id * F + f];
                        const uint32_t bits = (uint32_t)b << 16;
                        std::memcpy(&fv, &bits, 4);
                    } else {
                        fv = ((const float*)feat)[(int64_t)id * F + f];
                    }
                    ao[f] = ao[f] + wo[j] * fv;
                }
            }
        }
    }
    return 0;
}


/* ---- volume-rendering epilogue: clamp_t and render (epcq/epcq.py:508-530) ----
 * One ray at a time, samples front to back, every operation of the
 * reference's formulas in its order, one rounding each:
 *   clamp_t  t_dist = t[k+1] - t[k] (vsize appended for the last), mask =
 *            (t_dist < 1e-8) | (t_dist > 2 vsize), t_dist (1 - mask) + mask vsize
 *   render   opacity = 1 - exp(-sigma delta); acc = cumprod(1 - opacity +
 *            1e-10) shifted by one sample (acc_0 = 1); colour = sum of
 *            opacity acc rgb, plus acc_{S-1} bg.
 * exp is orc_expf.  The backward is the derivative of these formulas with
 * respect to sigma and rgb, in the order the device kernel evaluates it:
 *   d rgb_k   = opacity_k acc_k g
 *   d sigma_k = delta_k e_k (acc_k (g . rgb_k) - (R_k + [k < S-1] acc_{S-1} (g . bg)) / a_k)
 * with e_k = exp(-sigma_k delta_k), a_k = 1 - opacity_k + 1e-10, and R_k =
 * total - (running prefix up to k) of opacity_i acc_i (g . rgb_i). */
static float orc_delta(const float* t, const float* delta, int32_t S, int32_t k, float vsize) {
    if (delta) return delta[k];
    const float d = k + 1 < S ? t[k + 1] - t[k] : vsize;
    const float m = (d < 1e-8f || d > 2.0f * vsize) ? 1.0f : 0.0f;
    return d * (1.0f - m) + m * vsize;
}

int orc_clamp_t(const float* t, int64_t R, int32_t S, float vsize, float* delta) {
    if (S == 1) return 0;  // the reference's t_dist[:, :1] of an empty slice: no interval
    for (int64_t r = 0; r < R; ++r)
        for (int32_t k = 0; k < S; ++k) delta[r * S + k] = orc_delta(t + r * S, nullptr, S, k, vsize);
    return 0;
}

int orc_render(const float* t, const float* delta, float vsize, const float* sigma, const float* rgb, int64_t R,
               int32_t S, int32_t white_bg, float* out) {
    const float bg = white_bg ? 1.0f : 0.0f;
    for (int64_t r = 0; r < R; ++r) {
        const float* tr = t ? t + r * S : nullptr;
        const float* dr = delta ? delta + r * S : nullptr;
        float acc = 1.0f, acc_last = 1.0f, c[3] = {0.0f, 0.0f, 0.0f};
        const int32_t live = (t && S == 1) ? 0 : S;  // S = 1 from t: no interval, no contribution
        for (int32_t k = 0; k < live; ++k) {
            const float dk = orc_delta(tr, dr, S, k, vsize);
            const float op = 1.0f - orc_expf(-sigma[r * S + k] * dk);
            const float w = op * acc;
            for (int ch = 0; ch < 3; ++ch) c[ch] = c[ch] + w * rgb[(r * S + k) * 3 + ch];
            acc_last = acc;
            acc = acc * ((1.0f - op) + 1e-10f);
        }
        for (int ch = 0; ch < 3; ++ch) out[3 * r + ch] = c[ch] + acc_last * bg;
    }
    return 0;
}

int orc_render_backward(const float* t, const float* delta, float vsize, const float* sigma, const float* rgb,
                        const float* grad_out, int64_t R, int32_t S, int32_t white_bg, float* grad_sigma,
                        float* grad_rgb) {
    const float bg = white_bg ? 1.0f : 0.0f;
    std::vector<float> dk(S > 0 ? S : 1), op(S > 0 ? S : 1), acc(S > 0 ? S : 1), gc(S > 0 ? S : 1);
    for (int64_t r = 0; r < R; ++r) {
        const float* tr = t ? t + r * S : nullptr;
        const float* dr = delta ? delta + r * S : nullptr;
        const float* g = grad_out + 3 * r;
        if (t && S == 1) {
            grad_sigma[r] = 0.0f;
            for (int ch = 0; ch < 3; ++ch) grad_rgb[3 * r + ch] = 0.0f;
            continue;
        }
        float a = 1.0f, total = 0.0f;
        for (int32_t k = 0; k < S; ++k) {
            dk[k] = orc_delta(tr, dr, S, k, vsize);
            op[k] = 1.0f - orc_expf(-sigma[r * S + k] * dk[k]);
            const float* ck = rgb + (r * S + k) * 3;
            float dot = g[0] * ck[0];
            dot = dot + g[1] * ck[1];
            dot = dot + g[2] * ck[2];
            gc[k] = dot;
            acc[k] = a;
            total = total + op[k] * a * gc[k];
            a = a * ((1.0f - op[k]) + 1e-10f);
        }
        const float acc_last = S > 0 ? acc[S - 1] : 1.0f;
        const float gbg = ((g[0] + g[1]) + g[2]) * bg;
        float prefix = 0.0f;
        for (int32_t k = 0; k < S; ++k) {
            const float e = orc_expf(-sigma[r * S + k] * dk[k]);
            const float opk = 1.0f - e;
            const float ak = (1.0f - opk) + 1e-10f;
            const float wk = opk * acc[k];
            prefix = prefix + wk * gc[k];
            const float after = (total - prefix) + (k + 1 < S ? acc_last * gbg : 0.0f);
            grad_sigma[r * S + k] = dk[k] * e * (acc[k] * gc[k] - after / ak);
            for (int ch = 0; ch < 3; ++ch) grad_rgb[(r * S + k) * 3 + ch] = wk * g[ch];
        }
    }
    return 0;
}

}  // extern "C"
