// Per-ray sampling core shared by the per-view and batched sampling kernels.
//
// One lane owns one ray.  run_ray() performs, for the ray through pixel
// (u, v):
//   1. the hit walk of quick_sampling_kernel (epcq/csrc/epcq_render_topk.cu
//      :125-161): every point of the k*k clamped neighbour buckets whose
//      point-to-ray distance is below its radius, at most 48 per bucket, in
//      lane order then bucket order, giving tmin / tmax;
//   2. the 0.01-wide depth-slot histogram and the front-to-back walk that
//      emits up to S samples (:168-219);
//   3. for each emitted sample, its neighbours: the reference's 8-entry
//      buffer (quick_search_nearby_pc, :39-79) or the K nearest by
//      (distance, index), with optional inverse-distance weighted features
//      (epcq/epcq.py:434-437,449-452).
// Hits are never stored: the candidate walk is repeated for each step, so all
// per-lane state stays in registers.  The walks are: tmin / tmax; slot
// occupancy (once / at least twice) with the distance sums of the lowest
// occupied slots; a refill walk only when the front-to-back walk reaches a
// slot the table does not hold; one neighbour walk per two sample points.
// Sums are accumulated in the canonical (lane, bucket position) order, so the
// result equals a full 128-slot histogram.  A conservative cross-product test
// rejects candidates that are certainly outside the radius before the exact
// projection.
//
// Candidates come from a source `Src` with
//   int2   bucket(ub, vb)  (start, count) of pixel (ub, vb), count -1 if empty
//   float4 point(j)        (x, y, z, index bits) of list position j
// either global memory (SrcGlobal) or an LDS copy of the tile's halo region
// (SrcLds); each kernel instantiates exactly one, so LDS reads stay ds_read
// instructions.
//
// The hit test sqrtf(e2) < r is evaluated exactly as e2 < lim, where lim is
// the smallest float whose correctly rounded square root is >= r (sqrt is
// monotone): r2lim for a uniform radius, and with per-point radii each
// point's own threshold, stored with the pixel-ordered points.  Square roots
// are only taken for hits.
#pragma once

#include <climits>
#include <type_traits>

#include "epcq_device.h"

namespace epcq {

constexpr int kSlotTable = 8;
// Sample depths a ray keeps in registers for its neighbour walks (kernels with
// a larger sample bound re-read the later ones from the outputs they wrote).
constexpr int kRegSamples = 16;
// Rays whose hit lanes hold more candidates than this are run by a whole wave
// when at most kHeavyMinority rays of their wave are (or when they hold 4x as
// many).
constexpr int kHeavyCandidates = 128;
constexpr int kHeavyMinority = 16;
// Ray lists of at most this many rays are sampled one wave per ray: a list of
// 4096 rays of a 1M-point 800x800 view takes 0.08 ms against 0.29 ms one
// lane per ray, the two forms meet near 65536 rays there, and views with
// larger windows favour the wave form further.
constexpr int kListWaveRays = 65536;

struct SampleOpts {
    int k, ko, L;         // window size, ceil(k/2), k*k
    int kdiv;             // ceil(65536 / k): l / k == (l * kdiv) >> 16 for lanes l < 64
    float r0, r2lim;      // uniform radius and its exact squared threshold
    const float* radius;  // per-point radius or nullptr
    float t_intvl, alpha_p, gamma;
    int S, Kc;
    const float* xyz;
    const void* feat;
    int feat_bf16, F;
    int feat_vec;         // F % 4 == 0 and feat aligned for 4-feature loads
    int cull;             // skip neighbour buckets by their bounding spheres
    int heavy;            // a ray whose hit lanes hold more candidates is run by its whole wave (0: never)
    int list_wave;        // ray lists of at most this many rays run one wave per ray (0: never)
};

struct SampleOut {
    // dense drop-in layout [R,S,...] (sp_xyz != nullptr)
    float* sp_xyz;
    int* pc_idx;
    float* sp_t;
    uint8_t* mask;
    // compact layout [R,s_max,...]
    int s_max;
    int* n_samples;
    float* sample_t;
    int* nb_idx;
    float* nb_t;
    float* nb_dist;
    float* w;
    float* agg;
};

// Output layouts, fixed per kernel instantiation: the dense drop-in arrays,
// the compact arrays, or the compact arrays with inverse-distance weighted
// features.
constexpr int kLayDense = 0;
constexpr int kLayCompact = 1;
constexpr int kLayFeat = 2;

// Smallest float s with sqrtf(s) >= r, so that sqrtf(e2) < r <=> e2 < s
// (sqrt is monotone); the device twin of squared_threshold.  0 for r <= 0 or
// NaN (no distance is below r), +inf for r = +inf.
__device__ __forceinline__ float point_threshold(float r) {
    if (!(r > 0.0f)) return 0.0f;
    if (r == __builtin_inff()) return r;
    float s = r * r;
    if (sqrtf(s) >= r) {
        while (s > 0.0f && sqrtf(nextafterf(s, 0.0f)) >= r) s = nextafterf(s, 0.0f);
    } else {
        while (!(sqrtf(s) >= r)) s = nextafterf(s, __builtin_inff());
    }
    return s;
}

// Candidate sources.  Per-pixel bucket cone (see bucket_cone_body): ca =
// (axis, P), cb = (Q, B); point(j) = (x, y, z, index bits) of list position
// j; rad(j, p) = the squared threshold of its radius (point_threshold).  PR:
// radii are per point (SampleOpts::radius);
// a kernel without them carries neither the radius pointers nor the test.
template <bool PR>
struct SrcGlobal {
    static constexpr bool kPR = PR;
    static constexpr bool kReadPast = false;  // see SrcLds
    const int2* hash;
    const int* sorted_idx;
    const float* xyz;
    const float4* pts;   // pixel-ordered (x, y, z, index bits) or nullptr (gather)
    const float* prad;   // pixel-ordered radius thresholds or nullptr (gather)
    const float* radius;  // per-point radii by index (when prad is null)
    const float4* cone_a;
    const float2* cone_b;
    int W;
    __device__ __forceinline__ int2 bucket(int ub, int vb) const { return hash[vb * W + ub]; }
    __device__ __forceinline__ float4 cone4(int ub, int vb) const { return cone_a[vb * W + ub]; }
    __device__ __forceinline__ float2 cone2(int ub, int vb) const { return cone_b[vb * W + ub]; }
    __device__ __forceinline__ float4 point(int j) const {
        if (pts) return pts[j];
        const int id = sorted_idx[j];
        const float* p = xyz + 3 * (int64_t)id;
        return make_float4(p[0], p[1], p[2], __int_as_float(id));
    }
    __device__ __forceinline__ float rad(int j, const float4& p) const {
        return prad ? prad[j] : point_threshold(radius[__float_as_int(p.w)]);
    }
};

template <bool PR>
struct SrcLds {
    static constexpr bool kPR = PR;
    // Point j = e.x + e.y, one past a bucket, may be read (and not used): a
    // tile is staged only when its halo holds at most `cap` points, the
    // buckets' rebased ranges lie within [0, total), and the cap-entry point
    // array is followed by the cone array in the same allocation, so index
    // total <= cap stays inside it.  The walks then read ahead without a test.
    static constexpr bool kReadPast = true;
    const int2* hash;
    const float4* pts;
    const float* prad;  // staged radius thresholds (per-point radius only)
    const float4* cone_a;
    const float2* cone_b;
    int ra, va, RW;
    __device__ __forceinline__ float rad(int j, const float4&) const { return prad[j]; }
    __device__ __forceinline__ int2 bucket(int ub, int vb) const { return hash[(vb - va) * RW + (ub - ra)]; }
    __device__ __forceinline__ float4 cone4(int ub, int vb) const { return cone_a[(vb - va) * RW + (ub - ra)]; }
    __device__ __forceinline__ float2 cone2(int ub, int vb) const { return cone_b[(vb - va) * RW + (ub - ra)]; }
    __device__ __forceinline__ float4 point(int j) const { return pts[j]; }
};

// A tile whose halo points do not fit the LDS: the halo's hash entries (with
// their list positions in the view, not rebased) and cones are in LDS, the
// points come from global memory (the pixel-ordered copy, or a gather by
// index when there is none).
template <bool PR>
struct SrcHybrid {
    static constexpr bool kPR = PR;
    static constexpr bool kReadPast = false;  // see SrcLds
    const int2* hash;  // LDS, halo region
    const float4* cone_a;
    const float2* cone_b;
    int ra, va, RW;
    const float4* pts;  // global
    const float* prad;
    const int* sorted_idx;
    const float* xyz;
    const float* radius;
    __device__ __forceinline__ int2 bucket(int ub, int vb) const { return hash[(vb - va) * RW + (ub - ra)]; }
    __device__ __forceinline__ float4 cone4(int ub, int vb) const { return cone_a[(vb - va) * RW + (ub - ra)]; }
    __device__ __forceinline__ float2 cone2(int ub, int vb) const { return cone_b[(vb - va) * RW + (ub - ra)]; }
    __device__ __forceinline__ float4 point(int j) const {
        if (pts) return pts[j];
        const int id = sorted_idx[j];
        const float* p = xyz + 3 * (int64_t)id;
        return make_float4(p[0], p[1], p[2], __int_as_float(id));
    }
    __device__ __forceinline__ float rad(int j, const float4& p) const {
        return prad ? prad[j] : point_threshold(radius[__float_as_int(p.w)]);
    }
};

// PointRayInterset up to the squared distance: t and |p - (o + t d)|^2.
__device__ __forceinline__ void ray_t_e2(const Ray& r, const float4& p, float& t, float& e2) {
    const float w0 = p.x - r.o[0], w1 = p.y - r.o[1], w2 = p.z - r.o[2];
    t = dot3(w0, w1, w2, r.d[0], r.d[1], r.d[2]) / r.dd;
    const float q0 = r.o[0] + t * r.d[0];
    const float q1 = r.o[1] + t * r.d[1];
    const float q2 = r.o[2] + t * r.d[2];
    e2 = sqnorm3(p.x - q0, p.y - q1, p.z - q2);
}

// Bucket culling by view cones.  Every ray of a view starts at the same
// origin o, so a bucket is summarised by the cone of its points seen from o:
// unit axis a, rho <= |p - o| and phi >= the angle between the line through p
// and the axis, for every point p of the bucket.  For a ray with direction d
// at (line) angle alpha to the axis, each point is at distance
// |p - o| sin(alpha - phi) >= rho (sin alpha cos phi - cos alpha sin phi)
// from the ray.  The cone stores P = rho cos phi and Q = rho sin phi rounded
// towards that bound, and B = the bucket's largest radius plus 4e-6 rho; the
// ray side's cos / sin errors (a few ulp of rho) are inside the 4e-6 rho, and
// the rounded distances the exact tests compute differ from the exact
// distance by at most ~eps (18 |p - o| + 4 |o| + 1) (the rounding of t, of the
// foot point o + t d and of the norm, with |d| = 1 up to an ulp), which the
// slack of 4e-6 rho over 18 eps rho and the origin margin
// 3e-6 (1 + |o|_1) cover for
// every |p - o| >= rho.  So a bucket clears the ray by `margin` -- cone gap
// (s P - c Q) / |d| - B > margin, with c = |d . a| and s = |d x a| -- only if
// no point of the bucket passes the exact test.  cone_clears evaluates the
// equivalent s^2 P^2 > (c Q + (B + margin) |d|)^2 (for P > 0) without square
// roots or divisions; its few extra roundings stay far inside the slack.
__device__ __forceinline__ bool cone_clears(const Ray& ray, const float4& ca, const float2& cb, float margin) {
    if (!(ca.w > 0.0f)) return false;
    const float c = fabsf(dot3(ray.d[0], ray.d[1], ray.d[2], ca.x, ca.y, ca.z));
    const float x0 = ray.d[1] * ca.z - ray.d[2] * ca.y;
    const float x1 = ray.d[2] * ca.x - ray.d[0] * ca.z;
    const float x2 = ray.d[0] * ca.y - ray.d[1] * ca.x;
    const float sq = sqnorm3(x0, x1, x2);
    const float x = c * cb.x + (cb.y + margin) * ray.nd;
    return sq * (ca.w * ca.w) > x * x;
}

// dist < radius for a squared distance, exact (see the file comment): lim is
// the squared threshold of the radius (o.r2lim, or the point's own with
// per-point radii), so no square root is taken.
__device__ __forceinline__ bool within(float e2, float lim) { return e2 < lim; }

// Squared threshold of the radius of the candidate at list position j (point
// p): with per-point radii the sources hold each point's threshold.
template <class Src>
__device__ __forceinline__ float point_limit(const SampleOpts& o, const Src& src, int j, const float4& p) {
    if constexpr (Src::kPR) return src.rad(j, p);
    else return o.r2lim;
}

// Lane sets of one ray: bit l for window lane l, in NW 64-bit words (NW = 1
// for windows of at most 64 lanes, k <= 8; NW = 4 for up to 256, k <= 16).
// Windows larger than 64 * NW lanes carry no sets: every lane is visited.
// The lane -> (row, column) split l / k = (l * kdiv) >> 16 is exact for every
// l < 256 and k <= 16 (the rounding of kdiv = ceil(65536 / k) adds less than
// 256 / 65536 < 1 / k to the quotient).
template <int NW>
struct LaneSet {
    uint64_t w[NW];

    __device__ __forceinline__ static LaneSet none() {
        LaneSet s;
#pragma unroll
        for (int c = 0; c < NW; ++c) s.w[c] = 0ull;
        return s;
    }
    __device__ __forceinline__ static LaneSet full() {
        LaneSet s;
#pragma unroll
        for (int c = 0; c < NW; ++c) s.w[c] = ~0ull;
        return s;
    }
    __device__ __forceinline__ bool any() const {
        uint64_t x = 0ull;
#pragma unroll
        for (int c = 0; c < NW; ++c) x |= w[c];
        return x != 0ull;
    }
    __device__ __forceinline__ LaneSet operator|(const LaneSet& b) const {
        LaneSet s;
#pragma unroll
        for (int c = 0; c < NW; ++c) s.w[c] = w[c] | b.w[c];
        return s;
    }
    __device__ __forceinline__ LaneSet operator&(const LaneSet& b) const {
        LaneSet s;
#pragma unroll
        for (int c = 0; c < NW; ++c) s.w[c] = w[c] & b.w[c];
        return s;
    }
    __device__ __forceinline__ LaneSet minus(const LaneSet& b) const {  // this & ~b
        LaneSet s;
#pragma unroll
        for (int c = 0; c < NW; ++c) s.w[c] = w[c] & ~b.w[c];
        return s;
    }
    // bits [lo, lo + n) |= bits of x (n <= 64), lo < 64 * NW
    __device__ __forceinline__ void or_bits(uint64_t x, int lo) {
        if constexpr (NW == 1) {
            w[0] |= x << lo;
        } else {
            const int word = lo >> 6, sh = lo & 63;
#pragma unroll
            for (int c = 0; c < NW; ++c) {
                if (c == word) w[c] |= x << sh;
                if (c == word + 1 && sh > 0) w[c] |= x >> (64 - sh);
            }
        }
    }
    __device__ __forceinline__ void set(int l) { or_bits(1ull, l); }
};

// Calls fn(l) for the set lanes in order.
template <int NW, class Fn>
__device__ __forceinline__ void for_each_set_lane(const LaneSet<NW>& set, Fn&& fn) {
#pragma unroll
    for (int c = 0; c < NW; ++c) {
        uint64_t m = set.w[c];
        while (m) {
            const int l = 64 * c + __ffsll((unsigned long long)m) - 1;
            m &= m - 1ull;
            fn(l);
        }
    }
}

// Calls fn(ub, vb, l) for the k*k neighbour lanes l in order: lane l is
// (du, dv) = (l % k - (ko - 1), (l + k) / k - ko), i.e. rows dv = 1 - ko ..
// k - ko, columns du = 1 - ko .. k - ko, clamped to the image
// (calculate_neighbor_uv, cuda_util.cuh:382-388).
template <class Fn>
__device__ __forceinline__ void for_each_lane(const SampleOpts& o, int u, int v, int W, int H, Fn&& fn) {
    int l = 0;
    for (int dv = 1 - o.ko; dv <= o.k - o.ko; ++dv) {
        const int vb = min(max(v + dv, 0), H - 1);
        for (int du = 1 - o.ko; du <= o.k - o.ko; ++du) {
            const int ub = min(max(u + du, 0), W - 1);
            fn(ub, vb, l);
            ++l;
        }
    }
}

// Calls fn(ub, vb, l) for the lanes set in `mask`, in lane order (every lane
// when the window has no lane sets).
template <int NW, class Fn>
__device__ __forceinline__ void for_each_masked_lane(const SampleOpts& o, int u, int v, int W, int H,
                                                     const LaneSet<NW>& mask, Fn&& fn) {
    if (o.L > 64 * NW) {
        for_each_lane(o, u, v, W, H, fn);
        return;
    }
    for_each_set_lane(mask, [&](int l) {
        const int r = (l * o.kdiv) >> 16;
        const int ub = min(max(u + (l - r * o.k) - (o.ko - 1), 0), W - 1);
        const int vb = min(max(v + r - (o.ko - 1), 0), H - 1);
        fn(ub, vb, l);
    });
}

// Lanes whose bucket is non-empty (emask) and, of those, the ones whose cone
// does not clear the ray by `margin` (returned; every lane when culling is
// off).  Only the returned lanes can hold hits; `cand` counts their points
// (0 for windows without lane sets).
template <int NW, class Src>
__device__ __forceinline__ LaneSet<NW> lane_mask(const SampleOpts& o, const Src& src, const Ray& ray, int u, int v,
                                                 int W, int H, float margin, LaneSet<NW>& emask, int& cand) {
    cand = 0;
    if (o.L > 64 * NW) {
        emask = LaneSet<NW>::full();
        return LaneSet<NW>::full();
    }
    LaneSet<NW> m = LaneSet<NW>::none(), e = LaneSet<NW>::none();
    int c = 0;
    for_each_lane(o, u, v, W, H, [&](int ub, int vb, int l) {
        const int n = src.bucket(ub, vb).y;
        if (n <= 0) return;
        e.set(l);
        if (o.cull && cone_clears(ray, src.cone4(ub, vb), src.cone2(ub, vb), margin)) return;
        m.set(l);
        c += n;
    });
    emask = e;
    cand = c;
    return m;
}

// The lanes of `cand` whose cone does not clear the ray by `margin`.
template <int NW, class Src>
__device__ __forceinline__ LaneSet<NW> recheck_lanes(const SampleOpts& o, const Src& src, const Ray& ray, int u,
                                                     int v, int W, int H, float margin, const LaneSet<NW>& cand) {
    LaneSet<NW> m = LaneSet<NW>::none();
    for_each_masked_lane(o, u, v, W, H, cand, [&](int ub, int vb, int l) {
        if (!cone_clears(ray, src.cone4(ub, vb), src.cone2(ub, vb), margin)) m.set(l);
    });
    return m;
}

// Lanes whose unclamped window position lies in the image: one lane per
// distinct window pixel.
template <int NW>
__device__ __forceinline__ LaneSet<NW> inside_lanes(const SampleOpts& o, int u, int v, int W, int H) {
    uint64_t cols = 0;
    for (int c = 0; c < o.k; ++c) {
        const int uu = u + c - (o.ko - 1);
        if (uu >= 0 && uu < W) cols |= 1ull << c;
    }
    LaneSet<NW> m = LaneSet<NW>::none();
    for (int r = 0; r < o.k; ++r) {
        const int vv = v + r - (o.ko - 1);
        if (vv >= 0 && vv < H) m.or_bits(cols, r * o.k);
    }
    return m;
}

// Calls fn(t, e2) for each hit of the ray: the points of its neighbour
// buckets within their radius, at most 48 per bucket, in lane then bucket
// order (quick_sampling_kernel :125-155), skipping the lanes off in `mask`
// (empty or culled buckets, which hold no hit).  The next point is read
// before the current one is tested.
template <int NW, class Src, class Fn>
__device__ __forceinline__ void for_each_ray_hit(const SampleOpts& o, const Src& src, const Ray& ray, int u, int v,
                                                 int W, int H, const LaneSet<NW>& mask, Fn&& fn) {
    for_each_masked_lane(o, u, v, W, H, mask, [&](int ub, int vb, int) {
        const int2 e = src.bucket(ub, vb);
        if (e.y <= 0) return;
        int kept = 0;
        float4 nxt = src.point(e.x);
#pragma unroll 2
        for (int j = 0; j < e.y && kept < kSampleLaneHits; ++j) {
            const float4 p = nxt;
            if (Src::kReadPast || j + 1 < e.y) nxt = src.point(e.x + j + 1);
            const float lim = point_limit(o, src, e.x + j, p);
            float t, e2;
            ray_t_e2(ray, p, t, e2);
            if (within(e2, lim)) {
                ++kept;
                fn(t, e2);
            }
        }
    });
}

// 128-bit set of depth slots, one named word per 32 slots (kept in registers).
struct SlotMask {
    uint32_t w0, w1, w2, w3;
};

// Bit `s` as the four words of a SlotMask (0 <= s < 128), without branches.
__device__ __forceinline__ SlotMask slot_bit(int s) {
    const uint32_t bit = 1u << (s & 31);
    const int w = s >> 5;
    return SlotMask{w == 0 ? bit : 0u, w == 1 ? bit : 0u, w == 2 ? bit : 0u, w == 3 ? bit : 0u};
}

__device__ __forceinline__ uint32_t mask_from(uint32_t word, int lo) {
    return lo >= 32 ? 0u : (lo > 0 ? word & ~((1u << lo) - 1u) : word);
}

// First slot >= from in m, or kMaxSlots.
__device__ __forceinline__ int next_slot(const SlotMask& m, int from) {
    const uint32_t a = mask_from(m.w0, from), b = mask_from(m.w1, from - 32);
    const uint32_t c = mask_from(m.w2, from - 64), d = mask_from(m.w3, from - 96);
    if (a) return __ffs(a) - 1;
    if (b) return 32 + __ffs(b) - 1;
    if (c) return 64 + __ffs(c) - 1;
    if (d) return 96 + __ffs(d) - 1;
    return kMaxSlots;
}

__device__ __forceinline__ bool knn_less(float d, int id, float bd, int bi) {
    return d < bd || (d == bd && id < bi);
}

// Neighbour buffer of one sample point.
template <int MODE, int KMAX>
struct NbBuf {
    float bd[KMAX];
    int bi[KMAX];
    int bsize;
    int maxi;
    float maxd;

    __device__ __forceinline__ void init() {
        bsize = 0;
        maxi = 0;
        maxd = 1e3f;
#pragma unroll
        for (int b = 0; b < KMAX; ++b) {
            bd[b] = MODE == EPCQ_MODE_COMPAT ? 0.0f : __builtin_inff();
            bi[b] = MODE == EPCQ_MODE_COMPAT ? -1 : INT_MAX;
        }
    }

    __device__ __forceinline__ void add(float d, int id) {
        if (MODE == EPCQ_MODE_COMPAT) {
            // quick_search_nearby_pc (:50-73): while filling it tracks the
            // minimum; once full a closer point replaces the tracked entry
            // and the maximum is tracked from then on.
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
        } else {
            // sorted insertion: the KMAX nearest by (distance, index)
            if (!knn_less(d, id, bd[KMAX - 1], bi[KMAX - 1])) return;
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
};

// Stores value into the n 4-byte words at p: single words up to a 16-byte
// boundary, then 16 bytes at a time, then single words.
__device__ __forceinline__ void fill_words(uint32_t* p, int n, uint32_t value) {
    while (n > 0 && (reinterpret_cast<uintptr_t>(p) & 15)) *p++ = value, --n;
    for (; n >= 4; n -= 4, p += 4) *reinterpret_cast<uint4*>(p) = make_uint4(value, value, value, value);
    while (n > 0) *p++ = value, --n;
}

// The reference's 8-entry buffer (quick_search_nearby_pc, :50-73) for the
// one-lane sampler when only the neighbour indices are output: the indices
// go straight to the sample's output row as the buffer takes them (entry q
// of the buffer is column q of the row), and only the distances stay in
// registers, as a shift register while the buffer fills (r[n - 1 - q] holds
// entry q, newest first), so no update depends on a per-lane index and
// nothing branches on one.  Same semantics as NbBuf in compat mode: while
// filling, the minimum is tracked; once full, a closer point replaces the
// tracked entry, after which the maximum is tracked.
template <int KMAX>
struct CompatStream {
    float r[KMAX];
    int n;      // entries taken (<= KMAX)
    int maxi;   // buffer position of the tracked entry
    float maxd;
    int* row;   // the sample's output row, kc columns
    int kc;

    __device__ __forceinline__ void init(int* out_row, int out_kc) {
        n = 0;
        maxi = 0;
        maxd = 1e3f;
        row = out_row;
        kc = out_kc;
#pragma unroll
        for (int b = 0; b < KMAX; ++b) r[b] = 0.0f;
    }

    __device__ __forceinline__ void add(float d, int id) {
        if (n < KMAX) {
#pragma unroll
            for (int b = KMAX - 1; b > 0; --b) r[b] = r[b - 1];
            r[0] = d;
            if (n < kc) row[n] = id;
            if (maxd > d) {
                maxd = d;
                maxi = n;
            }
            ++n;
        } else if (d < maxd) {
            const int reg = KMAX - 1 - maxi;
#pragma unroll
            for (int b = 0; b < KMAX; ++b) r[b] = b == reg ? d : r[b];
            if (maxi < kc) row[maxi] = id;
            maxd = d;
            // the maximum, first position on ties (registers in reverse)
#pragma unroll
            for (int q = 0; q < KMAX; ++q) {
                const float x = r[KMAX - 1 - q];
                const bool gt = x > maxd;
                maxd = gt ? x : maxd;
                maxi = gt ? q : maxi;
            }
        }
    }

    // -1 in the row's columns the buffer left empty.
    __device__ __forceinline__ void finish() const {
        if (n < kc) fill_words(reinterpret_cast<uint32_t*>(row + n), kc - n, 0xffffffffu);
    }
};

// One walk over the candidates of the window feeding the buffers of the
// first `ns` (<= NS) sample points sp[], over the lanes set in `mask`
// (buckets that may hold a point within reach of a sample point).  Compat
// mode walks the k*k lanes with their border duplicates, like
// quick_search_nearby_pc (:39-47); KNN mode walks the distinct window pixels
// (the caller restricts `mask` to inside_lanes).
template <int MODE, int NS, int NW, class Src, class Buf>
__device__ __forceinline__ void neighbour_pass(const SampleOpts& o, const Src& src, int u, int v, int W, int H,
                                               const float (&sp)[NS][3], int ns, const LaneSet<NW>& mask,
                                               Buf (&buf)[NS]) {
    auto visit = [&](const float4& p, int jp) {
        const int id = __float_as_int(p.w);
        const float lim = point_limit(o, src, jp, p);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s < ns) {
                const float d2 = sqnorm3(sp[s][0] - p.x, sp[s][1] - p.y, sp[s][2] - p.z);
                if (within(d2, lim)) buf[s].add(sqrtf(d2), id);
            }
        }
    };
    auto bucket = [&](int ub, int vb, int) {
        const int2 e = src.bucket(ub, vb);
        if (e.y <= 0) return;
        float4 nxt = src.point(e.x);
        for (int j = 0; j < e.y; ++j) {
            const float4 p = nxt;
            if (Src::kReadPast || j + 1 < e.y) nxt = src.point(e.x + j + 1);
            visit(p, e.x + j);
        }
    };
    if (MODE == EPCQ_MODE_COMPAT || o.L <= 64 * NW) {
        for_each_masked_lane(o, u, v, W, H, mask, bucket);
    } else {
        const int ua = max(0, u - (o.ko - 1)), uz = min(W - 1, u + o.k - o.ko);
        const int va = max(0, v + 1 - o.ko), vz = min(H - 1, v + o.k - o.ko);
        for (int vb = va; vb <= vz; ++vb)
            for (int ub = ua; ub <= uz; ++ub) bucket(ub, vb, 0);
    }
}

// Writes the neighbour outputs of sample slot `so` of ray `ri`.
template <int MODE, int KMAX, int LAY>
__device__ __forceinline__ void write_neighbours(const SampleOpts& o, const SampleOut& out, const Ray& ray,
                                                 const NbBuf<MODE, KMAX>& nb, int64_t ri, int so) {
    const int nkeep = nb.bsize < o.Kc ? nb.bsize : o.Kc;
    if constexpr (LAY == kLayDense) {
        int* dst = out.pc_idx + (ri * o.S + so) * (int64_t)o.Kc;
#pragma unroll
        for (int b = 0; b < KMAX; ++b)
            if (b < o.Kc) dst[b] = b < nkeep ? nb.bi[b] : -1;
        for (int b = KMAX; b < o.Kc; ++b) dst[b] = -1;
        return;
    }
    // compact rows; neighbour depth and distance in KNN mode only
    const int64_t row = (ri * out.s_max + so) * (int64_t)o.Kc;
    float den = 0.0f;
    float inv[KMAX];
#pragma unroll
    for (int b = 0; b < KMAX; ++b) {
        inv[b] = 0.0f;
        if (b < o.Kc) {
            const bool valid = b < nkeep;
            out.nb_idx[row + b] = valid ? nb.bi[b] : -1;
            if constexpr (MODE == EPCQ_MODE_KNN) {
                if (out.nb_t || out.nb_dist) {
                    float t = 0.0f, dist = 0.0f;
                    if (valid) {
                        const float* p = o.xyz + 3 * (int64_t)nb.bi[b];
                        float e2;
                        ray_t_e2(ray, make_float4(p[0], p[1], p[2], 0.0f), t, e2);
                        dist = sqrtf(e2);
                    }
                    if (out.nb_t) out.nb_t[row + b] = t;
                    if (out.nb_dist) out.nb_dist[row + b] = dist;
                }
            }
            if constexpr (LAY == kLayFeat) {
                if (valid) {
                    inv[b] = 1.0f / (nb.bd[b] + 1e-6f);  // (1 / (d + 1e-6)) * mask, epcq.py:436
                    den = den + inv[b];
                }
            }
        }
    }
    for (int b = KMAX; b < o.Kc; ++b) {
        out.nb_idx[row + b] = -1;
        if constexpr (MODE == EPCQ_MODE_KNN) {
            if (out.nb_t) out.nb_t[row + b] = 0.0f;
            if (out.nb_dist) out.nb_dist[row + b] = 0.0f;
        }
    }
    if constexpr (LAY == kLayFeat) {
        den = den + 1e-6f;
        float* wo = out.w + row;
#pragma unroll
        for (int b = 0; b < KMAX; ++b)
            if (b < o.Kc) wo[b] = inv[b] / den;
        for (int b = KMAX; b < o.Kc; ++b) wo[b] = 0.0f;
        float* ao = out.agg + (ri * out.s_max + so) * (int64_t)o.F;
        for (int f0 = 0; f0 < o.F; f0 += 4) {
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int b = 0; b < KMAX; ++b) {
                if (b < nkeep) {
                    const float wb = inv[b] / den;
                    const int64_t base = (int64_t)nb.bi[b] * o.F + f0;
                    float fv[4];
                    if (o.feat_vec) {  // four features in one load
                        if (o.feat_bf16) {
                            const uint2 h = *reinterpret_cast<const uint2*>(static_cast<const uint16_t*>(o.feat) + base);
                            fv[0] = __uint_as_float(h.x << 16);
                            fv[1] = __uint_as_float(h.x & 0xffff0000u);
                            fv[2] = __uint_as_float(h.y << 16);
                            fv[3] = __uint_as_float(h.y & 0xffff0000u);
                        } else {
                            const float4 q = *reinterpret_cast<const float4*>(static_cast<const float*>(o.feat) + base);
                            fv[0] = q.x;
                            fv[1] = q.y;
                            fv[2] = q.z;
                            fv[3] = q.w;
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            fv[c] = 0.0f;
                            if (f0 + c < o.F) {
                                if (o.feat_bf16) {
                                    const uint16_t hb = static_cast<const uint16_t*>(o.feat)[base + c];
                                    fv[c] = __uint_as_float((uint32_t)hb << 16);
                                } else {
                                    fv[c] = static_cast<const float*>(o.feat)[base + c];
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (f0 + c < o.F) acc[c] = acc[c] + wb * fv[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (f0 + c < o.F) ao[f0 + c] = acc[c];
        }
    }
}

// Writes emitted sample number ns of ray ri (depth ts along the ray).
template <int LAY>
__device__ __forceinline__ void write_sample(const SampleOpts& o, const SampleOut& out, const Ray& ray, int64_t ri,
                                             int ns, float ts) {
    if constexpr (LAY == kLayDense) {
        const int64_t at = ri * o.S + ns;
        out.sp_xyz[3 * at + 0] = ray.o[0] + ts * ray.d[0];
        out.sp_xyz[3 * at + 1] = ray.o[1] + ts * ray.d[1];
        out.sp_xyz[3 * at + 2] = ray.o[2] + ts * ray.d[2];
        out.sp_t[at] = ts;
        out.mask[at] = 1;
    } else if (ns < out.s_max) {
        out.sample_t[ri * out.s_max + ns] = ts;
    }
}

// Every output element of ray ri not written for its ns emitted samples:
// zeros / -1 / false, and the sample count.
template <int MODE, int LAY>
__device__ __forceinline__ void write_tail(const SampleOpts& o, const SampleOut& out, int64_t ri, int ns) {
    if constexpr (LAY == kLayDense) {
        // slots [ns, S) of the ray are contiguous in each output
        const int64_t a0 = ri * o.S + ns;
        const int n = o.S - ns;
        fill_words(reinterpret_cast<uint32_t*>(out.sp_xyz + 3 * a0), 3 * n, 0u);
        fill_words(reinterpret_cast<uint32_t*>(out.sp_t + a0), n, 0u);
        fill_words(reinterpret_cast<uint32_t*>(out.pc_idx + a0 * o.Kc), n * o.Kc, 0xffffffffu);
        // the mask bytes: single bytes up to a 4-byte boundary, then words
        uint8_t* m = out.mask + a0;
        int nb = n;
        while (nb > 0 && (reinterpret_cast<uintptr_t>(m) & 3)) *m++ = 0, --nb;
        fill_words(reinterpret_cast<uint32_t*>(m), nb >> 2, 0u);
        m += nb & ~3;
        for (int s = 0; s < (nb & 3); ++s) m[s] = 0;
    } else {
        out.n_samples[ri] = ns;
        if (ns >= out.s_max) return;
        // slots [ns, s_max) of the ray are contiguous in each output
        const int64_t a0 = ri * out.s_max + ns;
        const int n = out.s_max - ns;
        fill_words(reinterpret_cast<uint32_t*>(out.sample_t + a0), n, 0u);
        fill_words(reinterpret_cast<uint32_t*>(out.nb_idx + a0 * o.Kc), n * o.Kc, 0xffffffffu);
        if constexpr (MODE == EPCQ_MODE_KNN) {
            if (out.nb_t) fill_words(reinterpret_cast<uint32_t*>(out.nb_t + a0 * o.Kc), n * o.Kc, 0u);
            if (out.nb_dist) fill_words(reinterpret_cast<uint32_t*>(out.nb_dist + a0 * o.Kc), n * o.Kc, 0u);
        }
        if constexpr (LAY == kLayFeat) {
            fill_words(reinterpret_cast<uint32_t*>(out.w + a0 * o.Kc), n * o.Kc, 0u);
            fill_words(reinterpret_cast<uint32_t*>(out.agg + a0 * o.F), n * o.F, 0u);
        }
    }
}

// Depth of emitted sample q of ray ri as this thread wrote it to the outputs
// (for samples beyond the ones kept in registers).
template <int LAY>
__device__ __forceinline__ float stored_sample_t(const SampleOpts& o, const SampleOut& out, int64_t ri, int q) {
    if constexpr (LAY == kLayDense) return out.sp_t[ri * o.S + q];
    else return out.sample_t[ri * out.s_max + q];
}

// Samples and neighbours of one ray; writes every output element of row ri.
// SMAX bounds the samples one ray can emit (epcq_max_live_samples).
// With may_defer, a ray whose hit lanes hold more than o.heavy candidates is
// left untouched and false returned: the caller then runs it with
// run_ray_wave, which spreads its candidates over a whole wave.
template <int MODE, int KMAX, int SMAX, int LAY, int NW, class Src>
__device__ __forceinline__ bool run_ray(const SampleOpts& o, const SampleOut& out, const Src& src, const Ray& ray,
                                        int u, int v, int W, int H, int64_t ri, bool valid_ray, bool may_defer) {
    constexpr int NS = KMAX <= 8 ? 2 : 1;  // sample points served by one neighbour walk
    constexpr bool dense = LAY == kLayDense;
    int ns = 0;
    float ts_arr[SMAX];
#pragma unroll
    for (int q = 0; q < SMAX; ++q) ts_arr[q] = 0.0f;
    // lanes whose bucket may hold a hit (a ray with none is empty)
    LaneSet<NW> emask = LaneSet<NW>::none();
    int cand = 0;
    const LaneSet<NW> hmask =
        valid_ray ? lane_mask<NW>(o, src, ray, u, v, W, H, ray.omargin, emask, cand) : LaneSet<NW>::none();
    if (may_defer && o.heavy > 0) {
        // only the wave's outliers: a ray above `heavy` while at most
        // kHeavyMinority of the wave's rays are, or one above 4x `heavy` (a
        // wave whose rays are all dense is faster one ray per lane)
        const int above = __popcll(__ballot(cand > o.heavy));
        if (cand > o.heavy && (above <= kHeavyMinority || cand > 4 * o.heavy)) return false;
    }
    if (valid_ray && hmask.any()) {
        // Step 1: tmin / tmax over the hits (WarpReduce min / max, :156-161).
        float tmin = 1e3f, tmax = -1e3f;
        for_each_ray_hit(o, src, ray, u, v, W, H, hmask, [&](float t, float) {
            if (tmin > t) tmin = t;
            if (tmax < t) tmax = t;
        });
        if (!(tmin > tmax)) {
            // Steps 2-3: slot sums and the front-to-back walk (:168-219).
            // One pass marks the slots hit once and at least twice (only the
            // latter take part in the walk) and sums the lowest occupied
            // slots in a table; when the walk reaches a slot the table does
            // not hold, a refill pass sums exactly the next table-full of
            // slots with at least two hits.
            SlotMask seen1{0u, 0u, 0u, 0u}, seen2{0u, 0u, 0u, 0u};
            int tslot[kSlotTable];
            float tsum[kSlotTable], tcnt[kSlotTable];
#pragma unroll
            for (int e = 0; e < kSlotTable; ++e) {
                tslot[e] = INT_MAX;
                tsum[e] = 0.0f;
                tcnt[e] = 0.0f;
            }
            int nt = 0;
            int limit = kMaxSlots;  // the table holds complete sums of every occupied slot below this
            for_each_ray_hit(o, src, ray, u, v, W, H, hmask, [&](float t, float e2) {
                const float qf = (t - tmin) / o.t_intvl;
                if (!(qf > -1.0f && qf < (float)kMaxSlots)) return;
                const int s = (int)qf;
                const SlotMask b = slot_bit(s);
                seen2.w0 |= seen1.w0 & b.w0;
                seen2.w1 |= seen1.w1 & b.w1;
                seen2.w2 |= seen1.w2 & b.w2;
                seen2.w3 |= seen1.w3 & b.w3;
                seen1.w0 |= b.w0;
                seen1.w1 |= b.w1;
                seen1.w2 |= b.w2;
                seen1.w3 |= b.w3;
                if (s >= limit) return;
                const float dist = sqrtf(e2);
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
                limit = mslot;  // the evicted slot and everything above leave the table
#pragma unroll
                for (int e = 0; e < kSlotTable; ++e)
                    if (e == mx) {
                        tslot[e] = s;
                        tsum[e] = dist;
                        tcnt[e] = 1.0f;
                    }
            });
            float T = 1.0f;
            int prev = -1;
            while (true) {
                // the walk's end does not depend on the next slot's sums, so
                // it is tested before a refill could be needed for them
                if (T < 1e-3f || ns >= o.S) break;
                const int s = next_slot(seen2, prev + 1);
                if (s >= kMaxSlots) break;
                prev = s;
                if (s >= limit) {
                    // refill: the next kSlotTable slots with >= 2 hits, from s
                    int from = s;
#pragma unroll
                    for (int e = 0; e < kSlotTable; ++e) {
                        tslot[e] = from < kMaxSlots ? from : INT_MAX;
                        tsum[e] = 0.0f;
                        tcnt[e] = 0.0f;
                        from = from < kMaxSlots ? next_slot(seen2, from + 1) : kMaxSlots;
                    }
                    limit = from;
                    const int first = s;
                    for_each_ray_hit(o, src, ray, u, v, W, H, hmask, [&](float t, float e2) {
                        const float qf = (t - tmin) / o.t_intvl;
                        if (!(qf > -1.0f && qf < (float)kMaxSlots)) return;
                        const int q = (int)qf;
                        if (q < first || q >= limit) return;
                        const float dist = sqrtf(e2);
#pragma unroll
                        for (int e = 0; e < kSlotTable; ++e) {
                            if (tslot[e] == q) {
                                tsum[e] = tsum[e] + dist;
                                tcnt[e] = tcnt[e] + 1.0f;
                            }
                        }
                    });
                }
                float bsum = 0.0f, bcnt = 0.0f;
#pragma unroll
                for (int e = 0; e < kSlotTable; ++e)
                    if (tslot[e] == s) {
                        bsum = tsum[e];
                        bcnt = tcnt[e];
                    }
                const float sdf = bsum / (bcnt + 1e-6f);
                const float alpha = o.gamma * expf_cw(-sdf / o.alpha_p);
                const float wgt = alpha * T;
                if (wgt > 1e-3f) {
                    const float ts = tmin + (float)s * o.t_intvl;
                    write_sample<LAY>(o, out, ray, ri, ns, ts);
#pragma unroll
                    for (int q = 0; q < SMAX; ++q)
                        if (q == ns) ts_arr[q] = ts;
                    ++ns;
                }
                T -= alpha;
            }
        }
        // Step 4: neighbours of the emitted samples, NS per walk.  Kernels
        // with a sample bound above 2 keep kRegSamples depths in registers
        // and re-read later ones from the outputs this thread wrote.
        const int nstore = min(SMAX < kRegSamples ? min(ns, SMAX) : ns, dense ? o.S : out.s_max);
        // compat mode without features streams the indices to the outputs
        constexpr bool stream = MODE == EPCQ_MODE_COMPAT && LAY != kLayFeat;
        using Buf = typename std::conditional<stream, CompatStream<KMAX>, NbBuf<MODE, KMAX>>::type;
        for (int g = 0; g < nstore; g += NS) {
            float pend[NS][3];
            Buf buf[NS];
            const int cnt = min(NS, nstore - g);
            float tabs = 0.0f;
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                float tq = 0.0f;
#pragma unroll
                for (int e = 0; e < SMAX; ++e)
                    if (e == g + q) tq = ts_arr[e];
                if (SMAX >= kRegSamples && g + q >= SMAX && q < cnt) tq = stored_sample_t<LAY>(o, out, ri, g + q);
                pend[q][0] = ray.o[0] + tq * ray.d[0];
                pend[q][1] = ray.o[1] + tq * ray.d[1];
                pend[q][2] = ray.o[2] + tq * ray.d[2];
                if (q < cnt) tabs = fmaxf(tabs, fabsf(tq));
                if constexpr (stream) {
                    // a buffer beyond the stored samples (q >= cnt) never takes a point
                    const int so = min(g + q, nstore - 1);
                    int* row = dense ? out.pc_idx + (ri * o.S + so) * (int64_t)o.Kc
                                     : out.nb_idx + (ri * out.s_max + so) * (int64_t)o.Kc;
                    buf[q].init(row, o.Kc);
                } else {
                    buf[q].init();
                }
            }
            // A sample point o + t d rounds to within ~eps (2 |t| |d|_1 + |o|_1)
            // of the ray and |sp - p| to a few ulp, so a bucket whose cone
            // clears the ray by that much more holds no neighbour.
            const float dl1 = fabsf(ray.d[0]) + fabsf(ray.d[1]) + fabsf(ray.d[2]);
            LaneSet<NW> nmask = hmask;
            if (o.cull && o.L <= 64 * NW)
                nmask = nmask | recheck_lanes<NW>(o, src, ray, u, v, W, H, ray.omargin + 6e-7f * tabs * dl1,
                                                  emask.minus(hmask));
            if (MODE != EPCQ_MODE_COMPAT && o.L <= 64 * NW) nmask = nmask & inside_lanes<NW>(o, u, v, W, H);
            neighbour_pass<MODE, NS>(o, src, u, v, W, H, pend, cnt, nmask, buf);
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                if (q < cnt) {
                    if constexpr (stream) buf[q].finish();
                    else write_neighbours<MODE, KMAX, LAY>(o, out, ray, buf[q], ri, g + q);
                }
            }
        }
    }

    // Everything not written above: zeros / -1 / false.
    write_tail<MODE, LAY>(o, out, ri, ns);
    return true;
}

// ---- wave-cooperative form ---------------------------------------------------
// run_ray_wave computes exactly what run_ray computes for one ray, with all 64
// lanes of the wave working on it: lane l < L tests window lane l, and each
// walk reads the candidates of a bucket 64 at a time, one per lane.  The
// order-dependent steps take the candidates in walk order (set bits of a
// ballot, lowest lane first):
//   - the 48-hit cap of a bucket keeps the hits of rank < 48 among its hits;
//   - slot sums: lane q accumulates slots q and q + 64, each in walk order;
//   - neighbour buffers: every lane holds the same buffer state and adds the
//     accepted candidates one at a time, in walk order.
// tmin / tmax are plain min / max reductions (a +-0 tie cannot change any
// output: t - tmin and tmin + s * t_intvl round the same either way).  Every
// lane must call it with the same ray, pixel and row (wave-uniform control
// flow); lane 0 writes the outputs.

__device__ __forceinline__ float lane_bcast(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ int lane_bcast(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

__device__ __forceinline__ float wave_min(float x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float y = __shfl_xor(x, o);
        x = y < x ? y : x;
    }
    return x;
}
__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float y = __shfl_xor(x, o);
        x = y > x ? y : x;
    }
    return x;
}

// write_neighbours of the feature layout with the wave's lanes sharing the
// work (the buffer is the same in every lane): lane b writes neighbour b,
// lane f the weighted feature f, each value computed exactly as
// write_neighbours computes it.
template <int MODE, int KMAX>
__device__ __forceinline__ void write_neighbours_wave(const SampleOpts& o, const SampleOut& out, const Ray& ray,
                                                      const NbBuf<MODE, KMAX>& nb, int64_t ri, int so) {
    constexpr int LAY = kLayFeat;
    const int lane = __lane_id();
    const int nkeep = nb.bsize < o.Kc ? nb.bsize : o.Kc;
    const int64_t row = (ri * out.s_max + so) * (int64_t)o.Kc;
    // the weights 1 / (d + 1e-6) (epcq.py:436) are recomputed where used (the
    // same value each time) rather than held in an array
    float den = 0.0f;
    if constexpr (LAY == kLayFeat) {
#pragma unroll
        for (int b = 0; b < KMAX; ++b)
            if (b < nkeep) den = den + 1.0f / (nb.bd[b] + 1e-6f);
        den = den + 1e-6f;
    }
    for (int b = lane; b < o.Kc; b += kWave) {
        int id = -1;
        float myd = 0.0f;
#pragma unroll
        for (int c = 0; c < KMAX; ++c)
            if (c == b) {
                id = nb.bi[c];
                myd = nb.bd[c];
            }
        const bool valid = b < KMAX && b < nkeep;
        out.nb_idx[row + b] = valid ? id : -1;
        if constexpr (MODE == EPCQ_MODE_KNN) {
            if (out.nb_t || out.nb_dist) {
                float t = 0.0f, dist = 0.0f;
                if (valid) {
                    const float* p = o.xyz + 3 * (int64_t)id;
                    float e2;
                    ray_t_e2(ray, make_float4(p[0], p[1], p[2], 0.0f), t, e2);
                    dist = sqrtf(e2);
                }
                if (out.nb_t) out.nb_t[row + b] = t;
                if (out.nb_dist) out.nb_dist[row + b] = dist;
            }
        }
        if constexpr (LAY == kLayFeat) out.w[row + b] = valid ? (1.0f / (myd + 1e-6f)) / den : 0.0f;
    }
    if constexpr (LAY == kLayFeat) {
        float* ao = out.agg + (ri * out.s_max + so) * (int64_t)o.F;
        for (int f = lane; f < o.F; f += kWave) {
            float acc = 0.0f;
#pragma unroll
            for (int b = 0; b < KMAX; ++b) {
                if (b < nkeep) {
                    const float wb = (1.0f / (nb.bd[b] + 1e-6f)) / den;
                    const int64_t at = (int64_t)nb.bi[b] * o.F + f;
                    const float fv = o.feat_bf16
                                         ? __uint_as_float((uint32_t) static_cast<const uint16_t*>(o.feat)[at] << 16)
                                         : static_cast<const float*>(o.feat)[at];
                    acc = acc + wb * fv;
                }
            }
            ao[f] = acc;
        }
    }
}

// write_tail with the wave's lanes sharing the slots: lane l writes elements
// l, l + 64, ... of each output range the tail covers.
template <int MODE, int LAY>
__device__ __forceinline__ void write_tail_wave(const SampleOpts& o, const SampleOut& out, int64_t ri, int ns) {
    const int lane = __lane_id();
    if constexpr (LAY == kLayDense) {
        const int64_t a0 = ri * o.S + ns;
        const int n = o.S - ns;  // empty slots
        for (int i = lane; i < 3 * n; i += kWave) out.sp_xyz[3 * a0 + i] = 0.0f;
        for (int i = lane; i < n; i += kWave) {
            out.sp_t[a0 + i] = 0.0f;
            out.mask[a0 + i] = 0;
        }
        for (int i = lane; i < n * o.Kc; i += kWave) out.pc_idx[a0 * o.Kc + i] = -1;
    } else {
        if (lane == 0) out.n_samples[ri] = ns;
        const int per = out.s_max - ns;  // empty slots, o.Kc entries each
        for (int i = lane; i < per * o.Kc; i += kWave) {
            const int64_t at = (ri * out.s_max + ns + i / o.Kc) * o.Kc + i % o.Kc;
            out.nb_idx[at] = -1;
            if constexpr (MODE == EPCQ_MODE_KNN) {
                if (out.nb_t) out.nb_t[at] = 0.0f;
                if (out.nb_dist) out.nb_dist[at] = 0.0f;
            }
            if constexpr (LAY == kLayFeat) out.w[at] = 0.0f;
        }
        for (int s = ns + lane; s < out.s_max; s += kWave) out.sample_t[ri * out.s_max + s] = 0.0f;
        if constexpr (LAY == kLayFeat)
            for (int i = lane; i < per * o.F; i += kWave) out.agg[(ri * out.s_max + ns) * o.F + i] = 0.0f;
    }
}

// SHARED_TAIL: every layout's tail is written by the whole wave (kernels that
// also hold the one-lane form keep lane 0 writing it outside the feature
// layout: sharing it there costs the one-lane form registers).
template <int MODE, int KMAX, int SMAX, int LAY, int NW, bool SHARED_TAIL = false, class Src>
__device__ __forceinline__ void run_ray_wave(const SampleOpts& o, const SampleOut& out, const Src& src, const Ray& ray,
                                             int u, int v, int W, int H, int64_t ri) {
    constexpr int NS = KMAX <= 8 ? 2 : 1;
    const int lane = __lane_id();
    // window lanes 64 c + lane (c < NW): their buckets and cone tests (same
    // margins as run_ray)
    int2 e[NW];
    int ubl[NW], vbl[NW];
    LaneSet<NW> emask, hmask;
#pragma unroll
    for (int c = 0; c < NW; ++c) {
        const int l = 64 * c + lane;
        e[c] = make_int2(0, -1);
        ubl[c] = 0;
        vbl[c] = 0;
        bool nonempty = false, maybe = false;
        if (l < o.L) {
            const int r = (l * o.kdiv) >> 16;
            ubl[c] = min(max(u + (l - r * o.k) - (o.ko - 1), 0), W - 1);
            vbl[c] = min(max(v + r - (o.ko - 1), 0), H - 1);
            e[c] = src.bucket(ubl[c], vbl[c]);
            nonempty = e[c].y > 0;
            maybe = nonempty &&
                    !(o.cull && cone_clears(ray, src.cone4(ubl[c], vbl[c]), src.cone2(ubl[c], vbl[c]), ray.omargin));
        }
        emask.w[c] = __ballot(nonempty);
        hmask.w[c] = __ballot(maybe);
    }

    // Calls bucket_fn(ex, ey) on every lane for each window lane set in
    // `set`, in lane order (the bucket's start and count, wave-uniform).
    auto for_each_bucket = [&](const LaneSet<NW>& set, auto&& bucket_fn) {
#pragma unroll
        for (int c = 0; c < NW; ++c) {
            uint64_t m = set.w[c];
            while (m) {
                const int x = __ffsll((unsigned long long)m) - 1;
                m &= m - 1ull;
                bucket_fn(lane_bcast(e[c].x, x), lane_bcast(e[c].y, x));
            }
        }
    };

    // fn(keep, t, e2) on every lane for each 64-candidate chunk of the hit
    // walk, `keep` set on the lanes holding a hit that the 48 cap keeps
    auto hit_walk = [&](auto&& fn) {
        for_each_bucket(hmask, [&](int ex, int ey) {
            int kept = 0;
            for (int c0 = 0; c0 < ey && kept < kSampleLaneHits; c0 += kWave) {
                const int j = c0 + lane;
                bool hit = false;
                float t = 0.0f, e2 = 0.0f;
                if (j < ey) {
                    const float4 p = src.point(ex + j);
                    const float lim = point_limit(o, src, ex + j, p);
                    ray_t_e2(ray, p, t, e2);
                    hit = within(e2, lim);
                }
                const uint64_t hb = __ballot(hit);
                const bool keep = hit && kept + __popcll(hb & lanemask_lt()) < kSampleLaneHits;
                fn(keep, t, e2);
                kept += __popcll(hb);
            }
        });
    };

    int ns = 0;
    float ts_arr[SMAX];
#pragma unroll
    for (int q = 0; q < SMAX; ++q) ts_arr[q] = 0.0f;
    if (hmask.any()) {
        // Step 1: tmin / tmax
        float tmin = 1e3f, tmax = -1e3f;
        hit_walk([&](bool keep, float t, float) {
            if (keep) {
                if (tmin > t) tmin = t;
                if (tmax < t) tmax = t;
            }
        });
        tmin = wave_min(tmin);
        tmax = wave_max(tmax);
        if (!(tmin > tmax)) {
            // Step 2: slot sums in walk order; lane q owns slots q and q + 64
            float sum0 = 0.0f, sum1 = 0.0f, cnt0 = 0.0f, cnt1 = 0.0f;
            hit_walk([&](bool keep, float t, float e2) {
                int sl = 0;
                float dist = 0.0f;
                bool ok = false;
                if (keep) {
                    const float qf = (t - tmin) / o.t_intvl;
                    ok = qf > -1.0f && qf < (float)kMaxSlots;
                    if (ok) {
                        sl = (int)qf;
                        dist = sqrtf(e2);
                    }
                }
                uint64_t b = __ballot(ok);
                while (b) {
                    const int x = __ffsll((unsigned long long)b) - 1;
                    b &= b - 1ull;
                    const int sx = lane_bcast(sl, x);
                    const float dx = lane_bcast(dist, x);
                    if ((sx & (kWave - 1)) == lane) {
                        if (sx < kWave) {
                            sum0 = sum0 + dx;
                            cnt0 = cnt0 + 1.0f;
                        } else {
                            sum1 = sum1 + dx;
                            cnt1 = cnt1 + 1.0f;
                        }
                    }
                }
            });
            // Step 3: front-to-back walk over the slots with >= 2 hits
            const uint64_t two0 = __ballot(cnt0 >= 2.0f), two1 = __ballot(cnt1 >= 2.0f);
            float T = 1.0f;
            for (int half = 0; half < 2; ++half) {
                uint64_t b = half == 0 ? two0 : two1;
                bool stop = false;
                while (b) {
                    const int x = __ffsll((unsigned long long)b) - 1;
                    b &= b - 1ull;
                    const int sl = half * kWave + x;
                    const float bsum = lane_bcast(half == 0 ? sum0 : sum1, x);
                    const float bcnt = lane_bcast(half == 0 ? cnt0 : cnt1, x);
                    if (T < 1e-3f || ns >= o.S) {
                        stop = true;
                        break;
                    }
                    const float sdf = bsum / (bcnt + 1e-6f);
                    const float alpha = o.gamma * expf_cw(-sdf / o.alpha_p);
                    const float wgt = alpha * T;
                    if (wgt > 1e-3f) {
                        const float ts = tmin + (float)sl * o.t_intvl;
                        if (lane == 0) write_sample<LAY>(o, out, ray, ri, ns, ts);
#pragma unroll
                        for (int q = 0; q < SMAX; ++q)
                            if (q == ns) ts_arr[q] = ts;
                        ++ns;
                    }
                    T -= alpha;
                }
                if (stop) break;
            }
        }
        // Step 4: neighbours of the emitted samples, NS per walk
        constexpr bool dense = LAY == kLayDense;
        const int nstore = min(SMAX < kRegSamples ? min(ns, SMAX) : ns, dense ? o.S : out.s_max);
        for (int g = 0; g < nstore; g += NS) {
            float pend[NS][3];
            NbBuf<MODE, KMAX> buf[NS];
            const int cnt = min(NS, nstore - g);
            float tabs = 0.0f;
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                float tq = 0.0f;
#pragma unroll
                for (int x = 0; x < SMAX; ++x)
                    if (x == g + q) tq = ts_arr[x];
                if (SMAX >= kRegSamples && g + q >= SMAX && q < cnt) {
                    // lane 0 wrote the sample; it reads it back for the wave
                    float back = 0.0f;
                    if (lane == 0) back = stored_sample_t<LAY>(o, out, ri, g + q);
                    tq = lane_bcast(back, 0);
                }
                pend[q][0] = ray.o[0] + tq * ray.d[0];
                pend[q][1] = ray.o[1] + tq * ray.d[1];
                pend[q][2] = ray.o[2] + tq * ray.d[2];
                if (q < cnt) tabs = fmaxf(tabs, fabsf(tq));
                buf[q].init();
            }
            const float dl1 = fabsf(ray.d[0]) + fabsf(ray.d[1]) + fabsf(ray.d[2]);
            LaneSet<NW> nmask = hmask;
            if (o.cull) {
                const LaneSet<NW> culled = emask.minus(hmask);
#pragma unroll
                for (int c = 0; c < NW; ++c) {
                    bool keep_lane = false;
                    if ((culled.w[c] >> lane) & 1ull)
                        keep_lane = !cone_clears(ray, src.cone4(ubl[c], vbl[c]), src.cone2(ubl[c], vbl[c]),
                                                 ray.omargin + 6e-7f * tabs * dl1);
                    nmask.w[c] |= __ballot(keep_lane);
                }
            }
            if (MODE != EPCQ_MODE_COMPAT) nmask = nmask & inside_lanes<NW>(o, u, v, W, H);
            for_each_bucket(nmask, [&](int ex, int ey) {
                for (int c0 = 0; c0 < ey; c0 += kWave) {
                    const int j = c0 + lane;
                    float dq[NS];
                    bool aq[NS];
                    int id = 0;
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        dq[q] = 0.0f;
                        aq[q] = false;
                    }
                    if (j < ey) {
                        const float4 p = src.point(ex + j);
                        id = __float_as_int(p.w);
                        const float lim = point_limit(o, src, ex + j, p);
#pragma unroll
                        for (int q = 0; q < NS; ++q) {
                            if (q < cnt) {
                                const float d2 = sqnorm3(pend[q][0] - p.x, pend[q][1] - p.y, pend[q][2] - p.z);
                                aq[q] = within(d2, lim);
                                if (aq[q]) dq[q] = sqrtf(d2);
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        uint64_t b = __ballot(aq[q]);
                        while (b) {
                            const int x = __ffsll((unsigned long long)b) - 1;
                            b &= b - 1ull;
                            buf[q].add(lane_bcast(dq[q], x), lane_bcast(id, x));
                        }
                    }
                }
            });
            // feature rows are wide enough to share out; the other layouts
            // write from lane 0 (sharing them slows the one-lane form's code)
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                if (q < cnt) {
                    if constexpr (LAY == kLayFeat) write_neighbours_wave<MODE, KMAX>(o, out, ray, buf[q], ri, g + q);
                    else if (lane == 0) write_neighbours<MODE, KMAX, LAY>(o, out, ray, buf[q], ri, g + q);
                }
            }
        }
    }
    if constexpr (LAY == kLayFeat || SHARED_TAIL) write_tail_wave<MODE, LAY>(o, out, ri, ns);
    else if (lane == 0) write_tail<MODE, LAY>(o, out, ri, ns);
}

// Smallest float s with sqrtf(s) >= r, so that sqrtf(e2) < r <=> e2 < s.
float squared_threshold(float r);

// Maps block b of nb to a tile index so that the blocks the hardware places
// on one XCD (b, b + 8, b + 16, ...) take consecutive tiles and share halos
// in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb / 8, rem = nb % 8;
    const int x = b % 8, j = b / 8;
    return x * per + min(x, rem) + j;
}

// Tile geometry of the LDS-staged sampler: a tw x th block of rays stages the
// hash entries and points of its halo region (tw + k - 1) x (th + k - 1).
struct TilePlan {
    int tw, th;      // tile size in rays (tw * th == block size)
    int xcd;         // XCD-contiguous tile order
    int cap;         // staged points per tile
    int rw_max, rh_max;
    size_t lds_bytes;
};

// Geometry and inputs of one view for the samplers.
struct ViewArgs {
    const float* c2w;  // device, >= 3 rows x 4
    float fx, fy, cx, cy;
    int W, H;
    const int2* hash;
    const int* sorted_idx;
    const float4* cone_a;  // bucket cones (bucket_cone_body) or nullptr
    const float2* cone_b;
    const float4* pts;     // pixel-ordered points (bucket_cone_body) or nullptr
    const float* prad;     // pixel-ordered radius thresholds (per-point radius) or nullptr
    const float* xyz;      // the view's cloud (nullptr: SampleOpts::xyz)
};

__device__ __forceinline__ void view_ray(const ViewArgs& v, int u, int vv, Ray& ray) {
    float m[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = v.c2w[i];
    make_ray(m, v.fx, v.fy, v.cx, v.cy, u, vv, ray);
}

// LDS of one tile: points [cap] | cone_a [rw*rh] | hash [rw*rh] | cone_b
// [rw*rh] | rowlo | rowbase.
struct TileSmem {
    float4* pts;
    float4* ca;
    int2* hash;
    float2* cb;
    int* rowlo;
    int* rowbase;
    float* rad;  // [cap] radius thresholds of the staged points (per-point radius)
    int* total;
    int* heavy;  // [kMaxTileRays] deferred rays (thread index in the tile), then their count
};

constexpr int kMaxTileRays = 256;  // threads (rays) of a tile block

__device__ __forceinline__ TileSmem carve_tile_smem(float4* dyn, const TilePlan& tp, int* total, int* heavy) {
    TileSmem m;
    const int rr = tp.rw_max * tp.rh_max;
    m.pts = dyn;
    m.ca = dyn + tp.cap;
    m.hash = reinterpret_cast<int2*>(m.ca + rr);
    m.cb = reinterpret_cast<float2*>(m.hash + rr);
    m.rowlo = reinterpret_cast<int*>(m.cb + rr);
    m.rowbase = m.rowlo + tp.rh_max;
    m.rad = reinterpret_cast<float*>(m.rowbase + tp.rh_max + 1);
    m.total = total;
    m.heavy = heavy;
    return m;
}

// LDS plan of the tiled samplers; radii are staged only with per-point radius.
TilePlan plan_tiles(int k, int tile_w, int tile_h, int cap, int flags, bool point_radius);

// Halo region of the tile with top-left ray (u0, v0): pixels [ra, ra+RW) x
// [va, va+RH) of the image.
__device__ __forceinline__ void tile_region(int W, int H, int k, int ko, int u0, int v0, int tw, int th, int& ra,
                                            int& va, int& RW, int& RH) {
    ra = max(0, u0 - (ko - 1));
    const int rz = min(W - 1, u0 + tw - 1 + (k - ko));
    va = max(0, v0 + 1 - ko);
    const int vz = min(H - 1, v0 + th - 1 + (k - ko));
    RW = rz - ra + 1;
    RH = vz - va + 1;
}

// Stages the halo region of the tile with top-left ray (u0, v0) into LDS:
// its hash entries (rebased to the staged list), bucket cones (when
// culling) and points.  Returns the number of staged points, or -1 if the
// region holds more than `cap` points (the tile then reads global memory).
// Block-wide; every thread of the block must call it.
__device__ __forceinline__ int stage_tile(const ViewArgs& view, const float* __restrict__ xyz,
                                          const float* __restrict__ radius, bool cull, int k, int ko, int u0, int v0,
                                          int tw, int th, int cap, const TileSmem& sm, int& ra, int& va, int& RW) {
    int RH;
    const int W = view.W;
    tile_region(W, view.H, k, ko, u0, v0, tw, th, ra, va, RW, RH);
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    for (int i = tid; i < RW * RH; i += nthr) {
        const int64_t g = (int64_t)(va + i / RW) * W + ra + i % RW;
        sm.hash[i] = view.hash[g];
        if (cull) {
            sm.ca[i] = view.cone_a[g];
            sm.cb[i] = view.cone_b[g];
        }
    }
    __syncthreads();
    for (int r = tid; r < RH; r += nthr) {
        int lo = 0, hi = 0;
        bool found = false;
        for (int c = 0; c < RW; ++c) {
            const int2 e = sm.hash[r * RW + c];
            if (e.y > 0) {
                if (!found) lo = e.x;
                found = true;
                hi = e.x + e.y;
            }
        }
        sm.rowlo[r] = lo;
        sm.rowbase[r] = hi - lo;
    }
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int r = 0; r < RH; ++r) {
            const int len = sm.rowbase[r];
            sm.rowbase[r] = run;
            run += len;
        }
        sm.rowbase[RH] = run;
        *sm.total = run;
    }
    __syncthreads();
    const int total = *sm.total;
    if (total > cap) return -1;
    for (int i = tid; i < total; i += nthr) {
        int lo = 0, hi = RH;  // last row r with rowbase[r] <= i
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sm.rowbase[mid] <= i) lo = mid;
            else hi = mid;
        }
        const int gi = sm.rowlo[lo] + (i - sm.rowbase[lo]);
        int id;
        if (view.pts) {
            const float4 q = view.pts[gi];
            sm.pts[i] = q;
            id = __float_as_int(q.w);
        } else {
            id = view.sorted_idx[gi];
            const float* p = xyz + 3 * (int64_t)id;
            sm.pts[i] = make_float4(p[0], p[1], p[2], __int_as_float(id));
        }
        if (radius) sm.rad[i] = view.prad ? view.prad[gi] : point_threshold(radius[id]);
    }
    for (int i = tid; i < RW * RH; i += nthr) {
        const int2 e = sm.hash[i];
        if (e.y > 0) {
            const int r = i / RW;
            sm.hash[i].x = sm.rowbase[r] + (e.x - sm.rowlo[r]);
        }
    }
    __syncthreads();
    return total;
}

// Appends the calling thread to the block's queue of deferred (heavy) rays:
// heavy[0..n) entries, n in heavy[kMaxTileRays].
__device__ __forceinline__ void defer_ray(int* heavy, int who) {
    const int at = atomicAdd(&heavy[kMaxTileRays], 1);
    heavy[at] = who;
}

// The block's deferred rays of a tile with top-left ray (u0, v0), each run by
// one whole wave (run_ray_wave); waves take every nw-th entry.  The queue is
// read after a barrier; the caller clears it before the next tile.
template <int MODE, int KMAX, int SMAX, int LAY, int NW, class Src>
__device__ __forceinline__ void run_deferred(const SampleOpts& o, const SampleOut& out, const ViewArgs& view,
                                             const Src& src, int tw, int u0, int v0, int64_t ray_off,
                                             const int* heavy) {
    const int n = heavy[kMaxTileRays];
    const int nw = (int)blockDim.x / kWave;
    for (int h = (int)threadIdx.x / kWave; h < n; h += nw) {
        const int t = heavy[h];
        const int u = u0 + t % tw, v = v0 + t / tw;
        Ray ray;
        view_ray(view, u, v, ray);
        run_ray_wave<MODE, KMAX, SMAX, LAY, NW>(o, out, src, ray, u, v, view.W, view.H,
                                            ray_off + (int64_t)v * view.W + u);
    }
}

// One tile of one view: stage its halo region and sample its rays.  Output
// row of ray (u, v) is ray_off + v * W + u.  Block-wide; no early returns, so
// persistent kernels can call it in a loop.  Rays with many candidates are
// deferred and then run one per wave.
// Returns false (block-uniform) when the tile's region did not fit the LDS;
// the caller then samples the tile from global memory (sample_tile_global).
template <int MODE, int KMAX, int SMAX, int LAY, bool PR, int NW = 1>
__device__ __forceinline__ bool sample_tile_body(const SampleOpts& o0, const SampleOut& out, const ViewArgs& view,
                                                 const TilePlan& tp, int tile, int tiles_x, int64_t ray_off,
                                                 const TileSmem& sm) {
    SampleOpts o = o0;
    if (view.xyz) o.xyz = view.xyz;
    const int u0 = (tile % tiles_x) * tp.tw;
    const int v0 = (tile / tiles_x) * tp.th;
    int ra, va, RW;
    if (threadIdx.x == 0) sm.heavy[kMaxTileRays] = 0;  // stage_tile's barriers publish it
    const int staged =
        stage_tile(view, o.xyz, PR ? o.radius : nullptr, o.cull, o.k, o.ko, u0, v0, tp.tw, tp.th, tp.cap, sm, ra, va,
                   RW);
    const int u = u0 + (int)threadIdx.x % tp.tw;
    const int v = v0 + (int)threadIdx.x / tp.tw;
    const SrcLds<PR> src{sm.hash, sm.pts, sm.rad, sm.ca, sm.cb, ra, va, RW};
    if (staged >= 0 && u < view.W && v < view.H) {
        Ray ray;
        view_ray(view, u, v, ray);
        if (!run_ray<MODE, KMAX, SMAX, LAY, NW>(o, out, src, ray, u, v, view.W, view.H,
                                            ray_off + (int64_t)v * view.W + u, true, true))
            defer_ray(sm.heavy, (int)threadIdx.x);
    }
    __syncthreads();
    if (staged >= 0) run_deferred<MODE, KMAX, SMAX, LAY, NW>(o, out, view, src, tp.tw, u0, v0, ray_off, sm.heavy);
    __syncthreads();  // the LDS is reused by the block's next tile
    return staged >= 0;
}

// Every ray of a tile whose halo points did not fit the LDS; called right
// after sample_tile_body returned false, so the LDS still holds the halo's
// hash entries (not rebased) and cones, and its deferred-ray queue is empty.
// Points are read from global memory.  Block-wide.
template <int MODE, int KMAX, int SMAX, int LAY, bool PR, int NW = 1>
__device__ __forceinline__ void sample_tile_global(const SampleOpts& o0, const SampleOut& out, const ViewArgs& view,
                                                   const TilePlan& tp, int tile, int tiles_x, int64_t ray_off,
                                                   const TileSmem& sm) {
    SampleOpts o = o0;
    if (view.xyz) o.xyz = view.xyz;
    const int u0 = (tile % tiles_x) * tp.tw, v0 = (tile / tiles_x) * tp.th;
    const int u = u0 + (int)threadIdx.x % tp.tw;
    const int v = v0 + (int)threadIdx.x / tp.tw;
    int ra, va, RW, RH;
    tile_region(view.W, view.H, o.k, o.ko, u0, v0, tp.tw, tp.th, ra, va, RW, RH);
    const SrcHybrid<PR> src{sm.hash, sm.ca, sm.cb, ra, va, RW, view.pts, view.prad, view.sorted_idx, o.xyz,
                            o.radius};
    if (u < view.W && v < view.H) {
        Ray ray;
        view_ray(view, u, v, ray);
        if (!run_ray<MODE, KMAX, SMAX, LAY, NW>(o, out, src, ray, u, v, view.W, view.H,
                                            ray_off + (int64_t)v * view.W + u, true, true))
            defer_ray(sm.heavy, (int)threadIdx.x);
    }
    __syncthreads();
    run_deferred<MODE, KMAX, SMAX, LAY, NW>(o, out, view, src, tp.tw, u0, v0, ray_off, sm.heavy);
}

// Cone of pixel p's points seen from the view origin o (see cone_gap).  The
// points come from pts_in (pixel-ordered (x, y, z, index bits), list position
// i at pts_in[i - pts_base]) when given, else through sorted_idx / xyz, in
// which case they are also copied in list order into pts (when not null);
// with per-point radii, the radii are copied in list order into prad (when
// not null).  One thread per pixel; two passes over the bucket: the axis
// (the normalised sum of the unit directions, each oriented to the first
// one's side, since cones bound lines), then the largest angle to it.  A
// bucket holding a point at o or a non-finite coordinate gets a cone that
// never culls.
__device__ __forceinline__ void bucket_cone_body(const int2* __restrict__ hash, const int* __restrict__ sorted_idx,
                                                 const float* __restrict__ xyz, const float4* __restrict__ pts_in,
                                                 const float* __restrict__ radius, float r0,
                                                 const float* __restrict__ c2w, int64_t p,
                                                 float4* __restrict__ cone_a, float2* __restrict__ cone_b,
                                                 float4* __restrict__ pts, float* __restrict__ prad,
                                                 int pts_base = 0) {
    const int2 e = hash[p];
    const float never = __builtin_inff();
    // an empty pixel's cone is never read (every reader tests the bucket's
    // count first), so it is not written
    if (e.y <= 0) return;
    auto point = [&](int j) {
        if (pts_in) return pts_in[e.x + j - pts_base];
        const int id = sorted_idx[e.x + j];
        const float* q = xyz + 3 * (int64_t)id;
        return make_float4(q[0], q[1], q[2], __int_as_float(id));
    };
    const float o0 = c2w[3], o1 = c2w[7], o2 = c2w[11];
    float rho = never, rb = r0;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    bool ok = true;
    for (int j = 0; j < e.y; ++j) {
        const float4 q = point(j);
        const int id = __float_as_int(q.w);
        if (!pts_in && pts) pts[e.x + j] = q;
        if (radius) {
            const float rj = radius[id];
            if (prad) prad[e.x + j] = point_threshold(rj);
            rb = j == 0 ? rj : fmaxf(rb, rj);
        }
        const float w0 = q.x - o0, w1 = q.y - o1, w2 = q.z - o2;
        const float x = sqnorm3(w0, w1, w2);
        if (!(x > 1e-30f && x < never)) {  // at the origin, tiny or not finite: a cone that never culls
            ok = false;
            continue;
        }
        // unit direction from the hardware reciprocal square root (1 ulp):
        // the cone is a bound, not a result, and its slack covers the few
        // ulp this leaves in n and u
        const float inv = __builtin_amdgcn_rsqf(x);
        const float n = x * inv;
        const float u0 = w0 * inv, u1 = w1 * inv, u2 = w2 * inv;
        if (j == 0) {
            f0 = u0;
            f1 = u1;
            f2 = u2;
        }
        const float sg = dot3(u0, u1, u2, f0, f1, f2) < 0.0f ? -1.0f : 1.0f;
        s0 = s0 + sg * u0;
        s1 = s1 + sg * u1;
        s2 = s2 + sg * u2;
        rho = fminf(rho, n);
    }
    const float sn = sqrtf(sqnorm3(s0, s1, s2));
    if (!ok || !(sn > 0.0f) || !(rb <= never)) {
        cone_a[p] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
        cone_b[p] = make_float2(0.0f, never);
        return;
    }
    const float a0 = s0 / sn, a1 = s1 / sn, a2 = s2 / sn;
    // the largest sine is the square root of the largest squared one (the
    // rounded square root is monotonic), so one root per bucket
    float cmin = 1.0f, s2max = 0.0f;
    for (int j = 0; j < e.y; ++j) {
        const float4 q = pts_in ? pts_in[e.x + j - pts_base] : (pts ? pts[e.x + j] : point(j));
        const float w0 = q.x - o0, w1 = q.y - o1, w2 = q.z - o2;
        const float inv = __builtin_amdgcn_rsqf(sqnorm3(w0, w1, w2));
        const float u0 = w0 * inv, u1 = w1 * inv, u2 = w2 * inv;
        cmin = fminf(cmin, fabsf(dot3(u0, u1, u2, a0, a1, a2)));
        const float x0 = u1 * a2 - u2 * a1, x1 = u2 * a0 - u0 * a2, x2 = u0 * a1 - u1 * a0;
        s2max = fmaxf(s2max, sqnorm3(x0, x1, x2));
    }
    const float smax = sqrtf(s2max);
    // rho, cos phi rounded down, sin phi up: each carries a few ulp of
    // error (reciprocal square root, products, sums), well inside 2e-6
    const float rl = rho * (1.0f - 2e-6f);
    cone_a[p] = make_float4(a0, a1, a2, rl * (cmin - 2e-6f));
    cone_b[p] = make_float2(rl * (smax + 2e-6f), rb * (1.0f + 1e-6f) + 4e-6f * rho);
}

}  // namespace epcq
