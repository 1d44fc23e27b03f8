// Device-side geometry and math for the epcq gfx950 kernels.
//
// Every floating-point expression here is written in the order the canonical
// definition fixes (left-to-right sums, no contraction: the library is built
// with -ffp-contract=off), so the results are bit-identical to a sequential
// IEEE binary32 evaluation on any host.  Division and sqrtf are the correctly
// rounded HIP defaults.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace epcq {

// Limits of the reference kernels (epcq/csrc/include/cuda_util.cuh:12-19).
constexpr int kQueryLaneHits = 32;   // POINT_BUFFER_SIZE: hits kept per bucket, radius query
constexpr int kCompatBuffer = 8;     // POINT_BUFFER_SIZE8: neighbour buffer per sample
constexpr int kSampleLaneHits = 48;  // POINT_BUFFER_SIZE48: hits kept per bucket, sampling
constexpr int kMaxSlots = 128;       // MAX_SAMPLE_POINTS_NUM: depth slots per ray

constexpr int kWave = 64;

struct Ray {
    float o[3];
    float d[3];
    float dd;       // dot(d, d), the divisor of PointRayInterset
    float omargin;  // 3e-6 (1 + |o|_1): origin part of the rejection margin
    float nd;       // sqrt(dd)
};

__device__ __forceinline__ float dot3(float a0, float a1, float a2, float b0, float b1, float b2) {
    float s = a0 * b0;
    s = s + a1 * b1;
    s = s + a2 * b2;
    return s;
}

__device__ __forceinline__ float norm3(float x, float y, float z) {
    float s = x * x;
    s = s + y * y;
    s = s + z * z;
    return sqrtf(s);
}

__device__ __forceinline__ float sqnorm3(float x, float y, float z) {
    float s = x * x;
    s = s + y * y;
    s = s + z * z;
    return s;
}

// cam2world_ray (cuda_util.cuh:338-354): the pixel centre is formed in double
// precision and rounded to float, direction = c2w[:3,:3] (x, y, 1), normalised
// by one reciprocal square root.
__device__ __forceinline__ void make_ray(const float* __restrict__ m, float fx, float fy, float cx,
                                         float cy, int u, int v, Ray& r) {
    const float x = (float)(((double)u + 0.5 - (double)cx) / (double)fx);
    const float y = (float)(((double)v + 0.5 - (double)cy) / (double)fy);
    const float z = 1.0f;
    float d0 = m[0] * x;
    d0 = d0 + m[1] * y;
    d0 = d0 + m[2] * z;
    float d1 = m[4] * x;
    d1 = d1 + m[5] * y;
    d1 = d1 + m[6] * z;
    float d2 = m[8] * x;
    d2 = d2 + m[9] * y;
    d2 = d2 + m[10] * z;
    const float rn = 1.0f / sqrtf(sqnorm3(d0, d1, d2));
    r.d[0] = d0 * rn;
    r.d[1] = d1 * rn;
    r.d[2] = d2 * rn;
    r.o[0] = m[3];
    r.o[1] = m[7];
    r.o[2] = m[11];
    r.dd = dot3(r.d[0], r.d[1], r.d[2], r.d[0], r.d[1], r.d[2]);
    r.omargin = 3e-6f * (1.0f + fabsf(r.o[0]) + fabsf(r.o[1]) + fabsf(r.o[2]));
    r.nd = sqrtf(r.dd);
}

// SingleRaySpec::PointRayInterset (cuda_util.cuh:215-236): t along the ray and
// the distance from p to its foot point o + t d.
__device__ __forceinline__ void point_ray(const Ray& r, float px, float py, float pz, float& t,
                                          float& dist) {
    const float w0 = px - r.o[0], w1 = py - r.o[1], w2 = pz - r.o[2];
    t = dot3(w0, w1, w2, r.d[0], r.d[1], r.d[2]) / r.dd;
    const float q0 = r.o[0] + t * r.d[0];
    const float q1 = r.o[1] + t * r.d[1];
    const float q2 = r.o[2] + t * r.d[2];
    dist = norm3(px - q0, py - q1, pz - q2);
}

// calculate_neighbor_uv (cuda_util.cuh:382-388) in integer form:
// ceil((l+1)/k) = (l+k)/k and ceil(k/2) = (k+1)/2.  Clamped to the image, so
// border rays visit the edge buckets more than once, like the reference.
__device__ __forceinline__ int neighbor_pixel(int k, int ko, int lane, int u, int v, int W, int H) {
    const int dv = (lane + k) / k - ko;
    const int du = (lane % k) - (ko - 1);
    const int uu = min(max(u + du, 0), W - 1);
    const int vv = min(max(v + dv, 0), H - 1);
    return vv * W + uu;
}

__device__ __forceinline__ float pow2i(int m) { return __int_as_float((m + 127) << 23); }

// exp(x) from its definition, not a vendor intrinsic: Cody-Waite reduction
// x = n ln2 + r with ln2 split in a short high part and a low part, e^r by its
// degree-7 Taylor polynomial (Horner, fused multiply-adds; truncation error
// below 1 ulp on |r| <= ln2/2), then exact scaling by 2^n in two steps so
// results in the subnormal range round once.
__device__ __forceinline__ float expf_cw(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    const float fn = __builtin_rintf(x * 1.44269502162933349609375f);
    const int n = (int)fn;
    float r = __builtin_fmaf(-fn, 0.693145751953125f, x);
    r = __builtin_fmaf(-fn, 1.428606765330187045e-06f, r);
    float p = 1.98412698412698412698e-4f;
    p = __builtin_fmaf(p, r, 1.38888888888888888889e-3f);
    p = __builtin_fmaf(p, r, 8.33333333333333333333e-3f);
    p = __builtin_fmaf(p, r, 4.16666666666666666667e-2f);
    p = __builtin_fmaf(p, r, 1.66666666666666666667e-1f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const int n1 = n / 2;
    const int n2 = n - n1;
    return (p * pow2i(n1)) * pow2i(n2);
}

// Lanes of this wave whose `key` equals mine (among lanes with `active`),
// from one ballot per key bit.
__device__ __forceinline__ uint64_t match_lanes(uint32_t key, int bits, bool active) {
    uint64_t m = __ballot(active);
    for (int b = 0; b < bits; ++b) {
        const bool bit = (key >> b) & 1u;
        const uint64_t bal = __ballot(active && bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

}  // namespace epcq
