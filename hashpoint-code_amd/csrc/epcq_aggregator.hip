// The reference renderer's Aggregator (epcq/epcq.py:406-507) for inference on
// gfx950: the point MLP over every (sample, neighbour) pair, the inverse-
// distance aggregation of its alpha and features per sample, and the colour
// MLP, on the bf16 matrix cores with fp32 accumulation.
//
// Per pair (m, k) with a valid neighbour (the reference's, per chunk of rows):
//   x1 = [f, PE(f, 3), PE([p - s, p*p - s*s], 5)]             284 = 7F + 60, F = 32
//   h2 = lrelu(W2 lrelu(W1 x1 + b1) + b2)                     block1
//   x3 = [h2, rgb, dir - d, dir . d]                          263
//   h4 = lrelu(W4 lrelu(W3 x3 + b3) + b4)                     block3
//   a  = softplus(wa . h4 + ba - 1)                           alpha_branch, alpha_act(alphas - 1)
// with f the point's feature, p its position, s the sample, d the ray
// direction, PE(x, n) = [sin(x_i 2^j), cos(x_i 2^j)] in (i, j) order
// (positional_encoding, epcq.py:389-404), lrelu slope 0.01, softplus with
// threshold 20.  Per sample m, with w_k = (1/(|p_k - s| + 1e-6)) v_k /
// (sum_k (1/(|p_k - s| + 1e-6)) v_k + 1e-6) (calculate_weight, :434-437):
//   alpha = sum_k w_k a_k                                     [M]
//   xc    = [sum_k w_k h4_k, PE(d, 4)]                        280
//   rgb   = sigmoid(C4 lrelu(C3 lrelu(C2 lrelu(C1 xc + c1) + c2) + c3) + c4) * 1.002 - 0.001
//
// Layout.  Pairs are rows r = m * Kp + k of a virtual [M, Kp] grid, Kp the
// power of two >= K (rows with k >= K are invalid), so that 32 consecutive
// rows -- one wave's strip -- hold whole samples.  A wave runs its strip
// through all four point layers with the activations in its own LDS region
// (32 rows of bf16), each layer a 32 x N tile sequence of
// v_mfma_f32_32x32x16_bf16 whose B operand (the weight rows, [N, K] bf16
// row-major, zero-padded to a multiple of 16 inputs) streams from L2; the
// epilogue adds the bias, applies the activation and writes bf16 back over
// the layer's input rows.  The strip then reduces its samples in LDS and
// writes each sample's colour-branch input (bf16, 288 wide) to the
// workspace; a second kernel runs the colour MLP over strips of 32 samples.
// Nothing is accumulated across waves, so results do not depend on
// scheduling.
#include <math.h>

#include <type_traits>

#include "epcq_common.h"

namespace epcq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWave = 64;
constexpr int kStrip = 32;              // rows per wave
constexpr int kWavesPerBlock = 4;          // colour kernel
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kPairWaves = 8;              // pair kernel: 256 rows share each weight chunk
constexpr int kPairBlock = kWave * kPairWaves;
constexpr int kLdsStride = 296;         // bf16 per LDS row (288 + 8: rows start on different banks)
constexpr int kFeat = 32;               // F: block1 takes 7F + 60 = 284 inputs
constexpr int kX1 = 288, kH = 256, kX3 = 272, kXc = 288, kC = 128;

// Round to nearest even (the hardware conversion, v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// Two floats rounded to bf16 by one packed conversion, lo in the low half.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

// sin and cos of x for the positional encodings, whose results are rounded
// to bf16: one range reduction to [-1/2, 1/2] revolution and the hardware
// sine and cosine.  The rounding of x / 2 pi bounds the phase error by
// |x| 2^-24 (6e-5 rad at |x| = 1024), far below the bf16 step of 2^-8 near
// 1; the library's sincosf beyond.
// CHECKED false: the caller has established |x| < 1024 (no branch here).
template <bool CHECKED = true>
__device__ __forceinline__ void pe_sincos(float x, float& s, float& c) {
    if (!CHECKED || fabsf(x) < 1024.0f) {
        float r = x * 0.15915494309189535f;
        r = r - rintf(r);
        s = __builtin_amdgcn_sinf(r);
        c = __builtin_amdgcn_cosf(r);
    } else {
        sincosf(x, &s, &c);
    }
}

// sin and cos of 2x from those of x (double-angle formulas).  Each step at
// most quadruples the absolute error of its inputs and adds a rounding; the
// encodings take at most two steps from a pe_sincos value, which keeps them
// within 16 times its error, far below the bf16 step they are rounded to.
__device__ __forceinline__ void pe_double(float s, float c, float& s2, float& c2) {
    s2 = 2.0f * s * c;
    c2 = c * c - s * s;
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// v for v > 0, else 0.01 v, as the larger of the two: one multiply and one
// v_max_f32, written out because fmaxf of a matrix-core result adds a
// NaN-quieting max of its own.  Equal to the compare-and-select form for every
// finite v, +-0 and +-inf (a NaN stays a NaN).
__device__ __forceinline__ float lrelu(float v) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(v), "v"(0.01f * v));
    return r;
}

// torch.nn.Softplus(beta=1, threshold=20)
__device__ __forceinline__ float softplus(float v) { return v > 20.0f ? v : log1pf(expf(v)); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    bf16x8 v;
    __builtin_memcpy(&v, p, sizeof(v));
    return v;
}

// One layer over the wave's strip: out[32, 32 NT] = act(in[32, 16 KS] W^T + b)
// with W [32 NT, 16 KS] bf16 row-major.  `io` is the strip's LDS rows; the
// output overwrites columns [0, 32 NT) of the same rows once every input has
// been read.  The 32x32x16 operands: lane l holds A[l % 32][8 (l / 32) + v]
// and B[8 (l / 32) + v][l % 32], v < 8; accumulator register r of lane l is
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
template <int KS, int NT, bool LEAKY>
__device__ __forceinline__ void strip_layer(uint16_t* io, const uint16_t* __restrict__ W,
                                            const float* __restrict__ bias) {
    const int lane = threadIdx.x % kWave;
    const int row = lane & 31;
    const int kh = (lane >> 5) * 8;
    constexpr int K = 16 * KS;
    // every accumulator of tile t in this lane is output column 32 t + row:
    // it starts at that column's bias
    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f32x16)(bias[t * 32 + row]);
    const uint16_t* a_row = io + row * kLdsStride + kh;
    const uint16_t* w_row = W + (size_t)row * K + kh;
    // the weight operands of steps ks + 1 and ks + 2 load while step ks
    // multiplies (they come from L2: two steps cover its latency)
    bf16x8 b0[NT], b1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        b0[t] = load8(w_row + (size_t)t * 32 * K);
        b1[t] = load8(w_row + (size_t)t * 32 * K + (KS > 1 ? 16 : 0));
    }
#pragma unroll 1
    for (int ks = 0; ks < KS; ++ks) {
        const bf16x8 a = load8(a_row + ks * 16);
        const int kn = (ks + 2 < KS ? ks + 2 : KS - 1) * 16;
        bf16x8 bn[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) bn[t] = load8(w_row + (size_t)t * 32 * K + kn);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0[t], acc[t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            b0[t] = b1[t];
            b1[t] = bn[t];
        }
    }
    const int r0 = (lane >> 5) * 4;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int c = t * 32 + row;
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            // rows rr and rr + 1 of column c: one conversion for both
            const int rr = (r >> 2) * 8 + r0 + (r & 3);
            float v0 = acc[t][r], v1 = acc[t][r + 1];
            if (LEAKY) v0 = lrelu(v0), v1 = lrelu(v1);
            const uint32_t p = pack_bf16(v0, v1);
            io[rr * kLdsStride + c] = (uint16_t)p;
            io[(rr + 1) * kLdsStride + c] = (uint16_t)(p >> 16);
        }
    }
}

// strip_layer for the pair kernel's 256-wide layers, with the weight rows
// staged through LDS for the whole block: step ks of the K loop multiplies
// the chunk W[:, 16 ks, 16 ks + 16) (8 KB), which the block's 512 threads
// copy from L2 into s_w, 16 bytes each, two steps ahead through registers.
// Every wave of the block calls it (it holds block barriers); `active`
// waves multiply, the others only help with the copies.
template <int KS>
__device__ __forceinline__ void block_layer(uint16_t* io, uint16_t* s_w, const uint16_t* __restrict__ W,
                                            const float* __restrict__ bias, bool active) {
    constexpr int K = 16 * KS;
    constexpr int NT = kH / 32;
    static_assert(KS >= 2, "two chunks in flight");
    static_assert(kPairBlock == 2 * kH, "one 16-byte piece of the chunk per thread");
    const int tid = threadIdx.x;
    const int lane = tid % kWave;
    const int row = lane & 31;
    const int kh = (lane >> 5) * 8;
    const uint16_t* wsrc = W + (size_t)(tid >> 1) * K + (tid & 1) * 8;
    uint4 c0, c1;
    __builtin_memcpy(&c0, wsrc, 16);
    __builtin_memcpy(&c1, wsrc + 16, 16);
    f32x16 acc[NT];  // seeded with the bias of the lane's output column, as in strip_layer
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f32x16)(bias[t * 32 + row]);
    const uint16_t* a_row = io + row * kLdsStride + kh;
    // unrolled so that the two chunks in flight swap registers instead of
    // being copied each step (four steps measured faster than two)
#pragma unroll 4
    for (int ks = 0; ks < KS; ++ks) {
        __syncthreads();  // the previous chunk has been read
        __builtin_memcpy(s_w + tid * 8, &c0, 16);
        c0 = c1;
        if (ks + 2 < KS) __builtin_memcpy(&c1, wsrc + (ks + 2) * 16, 16);
        __syncthreads();
        if (active) {
            const bf16x8 a = load8(a_row + ks * 16);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const bf16x8 b = load8(s_w + (t * 32 + row) * 16 + kh);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
    if (!active) return;
    // Rows rr and rr + 1 of column c are converted together; each even lane
    // then swaps halves with the odd lane beside it (column c + 1), so that
    // the even lane stores row rr and the odd lane row rr + 1 of the column
    // pair (c & ~1, c | 1) as one 32-bit word.
    const int r0 = (lane >> 5) * 4;
    const bool odd = lane & 1;
    const uint32_t sel = odd ? 0x03020706u : 0x05040100u;  // bytes of (neighbour, own)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int c = t * 32 + (row & ~1);
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int rr = (r >> 2) * 8 + r0 + (r & 3) + (odd ? 1 : 0);
            const uint32_t p = pack_bf16(lrelu(acc[t][r]), lrelu(acc[t][r + 1]));
            const uint32_t q = (uint32_t)__builtin_amdgcn_mov_dpp((int)p, 0xB1, 0xF, 0xF, false);
            *reinterpret_cast<uint32_t*>(io + rr * kLdsStride + c) = __builtin_amdgcn_perm(q, p, sel);
        }
    }
}

struct PairArgs {
    const float* pc_xyz;   // [M, K, 3]
    const float* pc_rgb;   // [M, K, 3]
    const float* pc_dir;   // [M, K, 3]
    const float* pc_feat;  // [M, K, 32]
    const float* sp_xyz;   // [M, 3]
    const uint8_t* valid;  // [M, K]
    const float* rays_d;   // [M, 3]
    int64_t M;
    int K, kp_log2;
    epcq_aggregator_params p;
    float* out_alpha;      // [M]
    uint16_t* xc;          // [M, 288] bf16 colour-branch inputs
};

__global__ __launch_bounds__(kPairBlock) void agg_pair_kernel(PairArgs a) {
    __shared__ __align__(16) uint16_t s_rows[kPairWaves][kStrip * kLdsStride];
    __shared__ __align__(16) uint16_t s_w[kH * 16];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    uint16_t* io = s_rows[wave];
    const int Kp = 1 << a.kp_log2;
    const int64_t rows_total = (int64_t)a.M << a.kp_log2;
    if ((int64_t)blockIdx.x * kPairWaves * kStrip >= rows_total) return;  // whole block past the end
    const int64_t row0 = ((int64_t)blockIdx.x * kPairWaves + wave) * kStrip;
    const bool live_strip = row0 < rows_total;  // waves past the end still take part in the copies
    const int row = lane & 31;
    const int half = lane >> 5;
    const int64_t r = row0 + row;
    const int64_t m = r >> a.kp_log2;
    const int k = (int)(r & (Kp - 1));
    const bool in_range = m < a.M;
    const bool valid = in_range && k < a.K && a.valid[m * a.K + k] != 0;
    const int64_t pair = m * a.K + k;

    // ---- x1 rows: lane half 0 takes features 0..15, half 1 features 16..31
    // and the distance terms; invalid rows are zero
    float px = 0.0f, py = 0.0f, pz = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (in_range) {
        sx = a.sp_xyz[3 * m];
        sy = a.sp_xyz[3 * m + 1];
        sz = a.sp_xyz[3 * m + 2];
    }
    if (valid) {
        px = a.pc_xyz[3 * pair];
        py = a.pc_xyz[3 * pair + 1];
        pz = a.pc_xyz[3 * pair + 2];
    }
    uint16_t* xr = io + row * kLdsStride;
    float4 fv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        fv[q] = valid ? reinterpret_cast<const float4*>(a.pc_feat + pair * kFeat + half * 16)[q]
                      : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float* fs = reinterpret_cast<const float*>(fv);
    {
        uint4 w0 = make_uint4(pack_bf16(fs[0], fs[1]), pack_bf16(fs[2], fs[3]), pack_bf16(fs[4], fs[5]),
                              pack_bf16(fs[6], fs[7]));
        uint4 w1 = make_uint4(pack_bf16(fs[8], fs[9]), pack_bf16(fs[10], fs[11]), pack_bf16(fs[12], fs[13]),
                              pack_bf16(fs[14], fs[15]));
        __builtin_memcpy(xr + half * 16, &w0, 16);
        __builtin_memcpy(xr + half * 16 + 8, &w1, 16);
    }
    // dist = p - s and distpow = p^2 - s^2; invalid pairs have p = 0 as the
    // reference's zero-filled gathers
    const float df[6] = {px - sx, py - sy, pz - sz, px * px - sx * sx, py * py - sy * sy, pz * pz - sz * sz};
    auto encode = [&](auto checked) {
        constexpr bool C = decltype(checked)::value;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            // PE(f, 3) of feature d = half * 16 + i: 6 values at 32 + 6 d
            float sn[3], cs[3];
            pe_sincos<C>(fs[i], sn[0], cs[0]);
            pe_double(sn[0], cs[0], sn[1], cs[1]);
            pe_double(sn[1], cs[1], sn[2], cs[2]);
            uint32_t* o = reinterpret_cast<uint32_t*>(xr + kFeat + 6 * (half * 16 + i));
            o[0] = pack_bf16(sn[0], cs[0]);
            o[1] = pack_bf16(sn[1], cs[1]);
            o[2] = pack_bf16(sn[2], cs[2]);
        }
        if (half) {
            uint32_t* o = reinterpret_cast<uint32_t*>(xr + 224);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                // frequencies 1 and 4 directly; 2, 8 and 16 by doubling
                float sn[5], cs[5];
                pe_sincos<C>(df[i], sn[0], cs[0]);
                pe_sincos<C>(df[i] * 4.0f, sn[2], cs[2]);
                pe_double(sn[0], cs[0], sn[1], cs[1]);
                pe_double(sn[2], cs[2], sn[3], cs[3]);
                pe_double(sn[3], cs[3], sn[4], cs[4]);
#pragma unroll
                for (int j = 0; j < 5; ++j) o[i * 5 + j] = pack_bf16(sn[j], cs[j]);
            }
            o[30] = 0;  // columns 284..287
            o[31] = 0;
        }
    };
    // one test for the wave: when every argument of its encodings is below
    // 1024 in magnitude (as it is for any features of ordinary scale) they
    // all take the hardware path without a branch each
    bool large = false;
#pragma unroll
    for (int i = 0; i < 16; ++i) large = large || !(fabsf(fs[i]) < 1024.0f);
    if (half) {
#pragma unroll
        for (int i = 0; i < 6; ++i) large = large || !(fabsf(df[i] * 4.0f) < 1024.0f);
    }
    if (__ballot(large) == 0) encode(std::false_type{});
    else encode(std::true_type{});
    const bool any_valid = __ballot(valid) != 0;
    block_layer<kX1 / 16>(io, s_w, static_cast<const uint16_t*>(a.p.w1), a.p.b1, any_valid);
    block_layer<kH / 16>(io, s_w, static_cast<const uint16_t*>(a.p.w2), a.p.b2, any_valid);
    if (any_valid) {
        // x3 = [h2, rgb, dir - d, dir . d, 0...]
        if (half == 0) {
            float rgb[3] = {0.0f, 0.0f, 0.0f}, dir[3] = {0.0f, 0.0f, 0.0f}, rd[3] = {0.0f, 0.0f, 0.0f};
            if (valid) {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    rgb[i] = a.pc_rgb[3 * pair + i];
                    dir[i] = a.pc_dir[3 * pair + i];
                }
            }
            if (in_range) {
#pragma unroll
                for (int i = 0; i < 3; ++i) rd[i] = a.rays_d[3 * m + i];
            }
            float dot = dir[0] * rd[0];
            dot = dot + dir[1] * rd[1];
            dot = dot + dir[2] * rd[2];
            xr[kH + 0] = to_bf16(rgb[0]);
            xr[kH + 1] = to_bf16(rgb[1]);
            xr[kH + 2] = to_bf16(rgb[2]);
            xr[kH + 3] = to_bf16(dir[0] - rd[0]);
            xr[kH + 4] = to_bf16(dir[1] - rd[1]);
            xr[kH + 5] = to_bf16(dir[2] - rd[2]);
            xr[kH + 6] = to_bf16(dot);
        } else {
#pragma unroll
            for (int c = kH + 7; c < kX3; ++c) xr[c] = 0;
        }
    }
    block_layer<kX3 / 16>(io, s_w, static_cast<const uint16_t*>(a.p.w3), a.p.b3, any_valid);
    block_layer<kH / 16>(io, s_w, static_cast<const uint16_t*>(a.p.w4), a.p.b4, any_valid);
    if (!live_strip) return;
    if (!any_valid)
        for (int c = half * 128; c < half * 128 + 128; ++c) xr[c] = 0;

    // ---- per-pair alpha: lane halves take 128 columns each
    float part = 0.0f;
    {
        const float4* wa4 = reinterpret_cast<const float4*>(a.p.wa + half * 128);
        const uint16_t* hr = xr + half * 128;
#pragma unroll 8
        for (int q = 0; q < 32; ++q) {
            const float4 wq = wa4[q];
            uint2 h;
            __builtin_memcpy(&h, hr + 4 * q, sizeof(h));
            // fused multiply-adds (one rounding each; the module's fp32 sums are
            // matched within the bf16 tolerance either way)
            part = __builtin_fmaf(wq.x, __uint_as_float(h.x << 16), part);
            part = __builtin_fmaf(wq.y, __uint_as_float(h.x & 0xffff0000u), part);
            part = __builtin_fmaf(wq.z, __uint_as_float(h.y << 16), part);
            part = __builtin_fmaf(wq.w, __uint_as_float(h.y & 0xffff0000u), part);
        }
    }
    const float other = __shfl_xor(part, 32);
    const float dotv = (half ? other + part : part + other) + a.p.ba[0];
    const float alpha_k = softplus(dotv - 1.0f);

    // ---- inverse-distance weights of the strip's samples (rows of a sample
    // are lanes [s Kp, s Kp + Kp) of each half)
    float dx = px - sx, dy = py - sy, dz = pz - sz;
    float nrm = dx * dx;
    nrm = nrm + dy * dy;
    nrm = nrm + dz * dz;
    const float inv = valid ? 1.0f / (sqrtf(nrm) + 1e-6f) : 0.0f;
    const int base = (row & ~(Kp - 1)) + half * 32;
    float sum = 0.0f;
    for (int j = 0; j < Kp; ++j) sum += __shfl(inv, base + j);
    const float w = inv / (sum + 1e-6f);

    // ---- per sample: alpha, weighted feature sum and PE(d, 4) -> xc row
    const int ns = kStrip >> a.kp_log2;
    for (int s = 0; s < ns; ++s) {
        const int64_t ms = (row0 >> a.kp_log2) + s;
        float acc_a = 0.0f;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int j = 0; j < Kp; ++j) {
            // every lane reads the same source row (a wave-uniform lane), so
            // the values come from registers by readlane, with no LDS round
            // trip, and the branch is uniform
            const int src = s * Kp + j;
            const float wj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), src));
            const float aj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(alpha_k), src));
            const int vj = __builtin_amdgcn_readlane((int)valid, src);
            if (!vj) continue;
            acc_a = __builtin_fmaf(wj, aj, acc_a);
            uint2 h;
            __builtin_memcpy(&h, io + src * kLdsStride + lane * 4, sizeof(h));
            acc[0] = __builtin_fmaf(wj, __uint_as_float(h.x << 16), acc[0]);
            acc[1] = __builtin_fmaf(wj, __uint_as_float(h.x & 0xffff0000u), acc[1]);
            acc[2] = __builtin_fmaf(wj, __uint_as_float(h.y << 16), acc[2]);
            acc[3] = __builtin_fmaf(wj, __uint_as_float(h.y & 0xffff0000u), acc[3]);
        }
        if (ms >= a.M) continue;
        uint16_t* xo = a.xc + ms * kXc;
        const uint2 packed = make_uint2((uint32_t)to_bf16(acc[0]) | ((uint32_t)to_bf16(acc[1]) << 16),
                                        (uint32_t)to_bf16(acc[2]) | ((uint32_t)to_bf16(acc[3]) << 16));
        __builtin_memcpy(xo + lane * 4, &packed, sizeof(packed));
        if (lane < 12) {
            const int i = lane / 4, j = lane % 4;
            float sn, cs;
            pe_sincos(a.rays_d[3 * ms + i] * (float)(1 << j), sn, cs);
            reinterpret_cast<uint32_t*>(xo + kH)[lane] = pack_bf16(sn, cs);
        } else if (lane < 20) {
            xo[kH + 24 + (lane - 12)] = 0;
        }
        if (lane == 0) a.out_alpha[ms] = acc_a;
    }
}

struct ColorArgs {
    const uint16_t* xc;  // [M, 288]
    int64_t M;
    epcq_aggregator_params p;
    float* out_rgb;      // [M, 3]
};

__global__ __launch_bounds__(kBlock) void agg_color_kernel(ColorArgs a) {
    __shared__ __align__(16) uint16_t s_rows[kWavesPerBlock][kStrip * kLdsStride];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    uint16_t* io = s_rows[wave];
    const int64_t m0 = ((int64_t)blockIdx.x * kWavesPerBlock + wave) * kStrip;
    if (m0 >= a.M) return;
    const int row = lane & 31;
    const int half = lane >> 5;
    const int64_t m = m0 + row;
    // stage the strip's inputs (rows past M are zero)
    uint16_t* xr = io + row * kLdsStride;
    for (int c = half * 8; c < kXc; c += 16) {
        bf16x8 v = bf16x8{};
        if (m < a.M) v = load8(a.xc + m * kXc + c);
        __builtin_memcpy(xr + c, &v, sizeof(v));
    }
    strip_layer<kXc / 16, kC / 32, true>(io, static_cast<const uint16_t*>(a.p.c1), a.p.cb1);
    strip_layer<kC / 16, kC / 32, true>(io, static_cast<const uint16_t*>(a.p.c2), a.p.cb2);
    strip_layer<kC / 16, kC / 32, true>(io, static_cast<const uint16_t*>(a.p.c3), a.p.cb3);
    // last layer (128 -> 3) in fp32, halves take 64 columns each
    float part[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
        const int c = half * 64 + 4 * q;
        uint2 h;
        __builtin_memcpy(&h, xr + c, sizeof(h));
        const float hv[4] = {__uint_as_float(h.x << 16), __uint_as_float(h.x & 0xffff0000u),
                             __uint_as_float(h.y << 16), __uint_as_float(h.y & 0xffff0000u)};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float4 wq = *reinterpret_cast<const float4*>(a.p.c4 + j * kC + c);
            part[j] += wq.x * hv[0];
            part[j] += wq.y * hv[1];
            part[j] += wq.z * hv[2];
            part[j] += wq.w * hv[3];
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float other = __shfl_xor(part[j], 32);
        const float v = (half ? other + part[j] : part[j] + other) + a.p.cb4[j];
        const float sig = 1.0f / (1.0f + expf(-v));
        if (half == 0 && m < a.M) a.out_rgb[3 * m + j] = sig * (1.0f + 2.0f * 0.001f) - 0.001f;
    }
}

}  // namespace
}  // namespace epcq

using namespace epcq;

extern "C" size_t epcq_aggregator_workspace_size(int64_t M) {
    if (M <= 0) return 0;
    return (size_t)M * kXc * sizeof(uint16_t);
}

extern "C" int epcq_aggregator_forward(const float* pc_xyz, const float* pc_rgb, const float* pc_dir,
                                       const float* pc_feat, const float* sp_xyz, const uint8_t* valid,
                                       const float* rays_d, int64_t M, int32_t K, int32_t F,
                                       const epcq_aggregator_params* params, float* out_alpha, float* out_rgb,
                                       void* workspace, size_t workspace_bytes, void* stream) {
    if (M < 0) return fail(EPCQ_ERR_INVALID, "aggregator: negative sample count");
    if (K < 1 || K > 32) return fail(EPCQ_ERR_INVALID, "aggregator: num_point_cloud_per_sp must be in [1, 32]");
    if (F != kFeat)
        return fail(EPCQ_ERR_INVALID, "aggregator: pc_feat must have 32 features (block1 takes 7F + 60 = 284 inputs)");
    if (!params) return fail(EPCQ_ERR_INVALID, "aggregator: missing parameters");
    if (M == 0) return EPCQ_OK;
    const size_t need = epcq_aggregator_workspace_size(M);
    if (workspace_bytes < need || !workspace) return fail(EPCQ_ERR_WORKSPACE, "aggregator: workspace too small");
    if ((reinterpret_cast<uintptr_t>(workspace) & 15u) != 0)
        return fail(EPCQ_ERR_INVALID, "aggregator: workspace must be 16-byte aligned");
    hipStream_t st = as_stream(stream);
    PairArgs pa;
    pa.pc_xyz = pc_xyz;
    pa.pc_rgb = pc_rgb;
    pa.pc_dir = pc_dir;
    pa.pc_feat = pc_feat;
    pa.sp_xyz = sp_xyz;
    pa.valid = valid;
    pa.rays_d = rays_d;
    pa.M = M;
    pa.K = K;
    pa.kp_log2 = ceil_log2(K);
    pa.p = *params;
    pa.out_alpha = out_alpha;
    pa.xc = static_cast<uint16_t*>(workspace);
    const int64_t strips = ceil_div(M << pa.kp_log2, kStrip);
    const int64_t blocks = ceil_div(strips, kPairWaves);
    if (blocks > INT32_MAX) return fail(EPCQ_ERR_INVALID, "aggregator: too many samples");
    agg_pair_kernel<<<dim3((unsigned)blocks), dim3(kPairBlock), 0, st>>>(pa);
    if (int e = check_launch("aggregator/pairs")) return e;
    ColorArgs ca;
    ca.xc = pa.xc;
    ca.M = M;
    ca.p = *params;
    ca.out_rgb = out_rgb;
    const int64_t cblocks = ceil_div(ceil_div(M, kStrip), kWavesPerBlock);
    agg_color_kernel<<<dim3((unsigned)cblocks), dim3(kBlock), 0, st>>>(ca);
    return check_launch("aggregator/colour");
}
