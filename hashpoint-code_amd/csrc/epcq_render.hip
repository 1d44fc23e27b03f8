// Volume-rendering epilogue of the reference renderer for gfx950: clamp_t and
// render (epcq/epcq.py:508-530, also utils/utils.py:422-467), as used after
// the Aggregator (opt/render_image.py:101-109), and render's backward.
//
// One lane per ray walks its S samples front to back in registers:
//   delta_i = t_{i+1} - t_i (vsize for the last), replaced by vsize where
//             below 1e-8 or above 2 vsize                   (clamp_t)
//   op_i    = 1 - exp(-sigma_i delta_i)
//   T_0 = 1, T_{i+1} = T_i (1 - op_i + 1e-10)               (cumprod, shifted)
//   rgb     = sum_i op_i T_i c_i + T_{S-1} bg               (bg = 1 or 0)
// Note the background weight is T_{S-1}, the transmittance before the last
// sample, exactly as the reference's acc_transmission[:, -1] after its
// one-sample shift.  With S = 1 the reference's clamp_t returns no interval
// at all (its appended vsize column is cut from an empty slice), so no
// sample contributes and the colour is the background: the fused form (t
// given) does the same, and clamp_t writes nothing.  Every operation is the reference's, one rounding each,
// in this order; the exponential is the library's own (expf_cw).  The
// backward pass recomputes the walk and forms, per sample k (g = dL/drgb):
//   dL/dc_k     = op_k T_k g
//   dL/dsigma_k = delta_k e_k (T_k (g . c_k) - (R_k + [k < S-1] T_{S-1} (g . bg)) / a_k)
// with e_k = exp(-sigma_k delta_k), a_k = 1 - op_k + 1e-10 and R_k the sum of
// op_i T_i (g . c_i) over i > k (taken as the total minus the running
// prefix).  Rows are [R, S] (t, delta, sigma), [R, S, 3] (rgb), [R, 3] (out).
#include "epcq_common.h"
#include "epcq_device.h"

namespace epcq {
namespace {

constexpr int kBlock = 256;

// clamp_t of sample k of a row of t.
__device__ __forceinline__ float clamp_delta(const float* __restrict__ t, int S, int k, float vsize) {
    const float d = k + 1 < S ? t[k + 1] - t[k] : vsize;
    const float m = (d < 1e-8f || d > 2.0f * vsize) ? 1.0f : 0.0f;
    return d * (1.0f - m) + m * vsize;
}


// clamp_t of a sample whose depth tk and next depth tk1 are in registers
// (last: the row's last sample, whose interval is vsize); the same
// operations as clamp_delta.
__device__ __forceinline__ float clamp_delta_v(float tk, float tk1, bool last, float vsize) {
    const float d = last ? vsize : tk1 - tk;
    const float m = (d < 1e-8f || d > 2.0f * vsize) ? 1.0f : 0.0f;
    return d * (1.0f - m) + m * vsize;
}

// Samples [k0, k0 + n) of a row (n <= kChunk) in registers: their
// intervals, densities and colours.  All loads of the chunk are issued before
// any is used, so a lane waits for memory once per chunk, not once per sample.
constexpr int kChunk = 16;
struct RowChunk {
    float d[kChunk], s[kChunk], c[3 * kChunk];
};

__device__ __forceinline__ void load_chunk(const float* __restrict__ tr, const float* __restrict__ dr,
                                           const float* __restrict__ sr, const float* __restrict__ cr, int S,
                                           int k0, int n, float vsize, RowChunk& ch) {
    float tv[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
        if (j < n) {
            ch.s[j] = sr[k0 + j];
            ch.c[3 * j] = cr[3 * (k0 + j)];
            ch.c[3 * j + 1] = cr[3 * (k0 + j) + 1];
            ch.c[3 * j + 2] = cr[3 * (k0 + j) + 2];
            if (dr) ch.d[j] = dr[k0 + j];
            else tv[j] = tr[k0 + j];
        }
    }
    if (!dr) {
        const float tnext = k0 + n < S ? tr[k0 + n] : 0.0f;  // the depth after the chunk
#pragma unroll
        for (int j = 0; j < kChunk; ++j)
            if (j < n) ch.d[j] = clamp_delta_v(tv[j], j + 1 < n ? tv[min(j + 1, kChunk - 1)] : tnext, k0 + j + 1 >= S, vsize);
    }
}

__global__ __launch_bounds__(kBlock) void clamp_t_kernel(const float* __restrict__ t, int64_t R, int S, float vsize,
                                                         float* __restrict__ delta) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= R * S) return;
    const int64_t r = i / S;
    const int k = (int)(i - r * S);
    delta[i] = clamp_delta(t + r * S, S, k, vsize);
}

__global__ __launch_bounds__(kBlock) void render_kernel(const float* __restrict__ t, const float* __restrict__ delta,
                                                        float vsize, const float* __restrict__ sigma,
                                                        const float* __restrict__ rgb, int64_t R, int S, float bg,
                                                        float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= R) return;
    const float* tr = t ? t + r * S : nullptr;
    const float* dr = delta ? delta + r * S : nullptr;
    const float* sr = sigma + r * S;
    const float* cr = rgb + r * S * 3;
    float T = 1.0f, Tlast = 1.0f;
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    const int live = (t && S == 1) ? 0 : S;
    for (int k0 = 0; k0 < live; k0 += kChunk) {
        const int n = min(kChunk, live - k0);
        RowChunk ch;
        load_chunk(tr, dr, sr, cr, S, k0, n, vsize, ch);
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            if (j < n) {
                const float op = 1.0f - expf_cw(-ch.s[j] * ch.d[j]);
                const float w = op * T;
                c0 = c0 + w * ch.c[3 * j];
                c1 = c1 + w * ch.c[3 * j + 1];
                c2 = c2 + w * ch.c[3 * j + 2];
                Tlast = T;
                T = T * ((1.0f - op) + 1e-10f);
            }
        }
    }
    out[3 * r] = c0 + Tlast * bg;
    out[3 * r + 1] = c1 + Tlast * bg;
    out[3 * r + 2] = c2 + Tlast * bg;
}

__global__ __launch_bounds__(kBlock) void render_backward_kernel(
    const float* __restrict__ t, const float* __restrict__ delta, float vsize, const float* __restrict__ sigma,
    const float* __restrict__ rgb, const float* __restrict__ grad_out, int64_t R, int S, float bg,
    float* __restrict__ grad_sigma, float* __restrict__ grad_rgb) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= R) return;
    const float* tr = t ? t + r * S : nullptr;
    const float* dr = delta ? delta + r * S : nullptr;
    const float* sr = sigma + r * S;
    const float* cr = rgb + r * S * 3;
    const float g0 = grad_out[3 * r], g1 = grad_out[3 * r + 1], g2 = grad_out[3 * r + 2];
    if (t && S == 1) {  // no sample contributes (see above)
        grad_sigma[r] = 0.0f;
        grad_rgb[3 * r] = grad_rgb[3 * r + 1] = grad_rgb[3 * r + 2] = 0.0f;
        return;
    }
    // pass 1: the total of op_i T_i (g . c_i) and T_{S-1}
    float T = 1.0f, Tlast = 1.0f, total = 0.0f;
    for (int k0 = 0; k0 < S; k0 += kChunk) {
        const int n = min(kChunk, S - k0);
        RowChunk ch;
        load_chunk(tr, dr, sr, cr, S, k0, n, vsize, ch);
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            if (j < n) {
                const float op = 1.0f - expf_cw(-ch.s[j] * ch.d[j]);
                const float gc = dot3(g0, g1, g2, ch.c[3 * j], ch.c[3 * j + 1], ch.c[3 * j + 2]);
                total = total + op * T * gc;
                Tlast = T;
                T = T * ((1.0f - op) + 1e-10f);
            }
        }
    }
    const float gbg = (g0 + g1 + g2) * bg;
    // pass 2: per-sample gradients with the suffix sums
    T = 1.0f;
    float prefix = 0.0f;
    for (int k0 = 0; k0 < S; k0 += kChunk) {
        const int n = min(kChunk, S - k0);
        RowChunk ch;
        load_chunk(tr, dr, sr, cr, S, k0, n, vsize, ch);
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            if (j < n) {
                const int k = k0 + j;
                const float dk = ch.d[j];
                const float e = expf_cw(-ch.s[j] * dk);
                const float op = 1.0f - e;
                const float a = (1.0f - op) + 1e-10f;
                const float gc = dot3(g0, g1, g2, ch.c[3 * j], ch.c[3 * j + 1], ch.c[3 * j + 2]);
                const float wk = op * T;
                prefix = prefix + wk * gc;
                const float after = (total - prefix) + (k + 1 < S ? Tlast * gbg : 0.0f);
                grad_sigma[r * S + k] = dk * e * (T * gc - after / a);
                grad_rgb[(r * S + k) * 3] = wk * g0;
                grad_rgb[(r * S + k) * 3 + 1] = wk * g1;
                grad_rgb[(r * S + k) * 3 + 2] = wk * g2;
                T = T * a;
            }
        }
    }
}

int check_render(const float* t, const float* delta, int64_t R, int32_t S, const char* who) {
    if (R < 0 || S < 0) return fail(EPCQ_ERR_INVALID, std::string(who) + ": negative size");
    if (R > 0 && S > 0 && !t == !delta)
        return fail(EPCQ_ERR_INVALID, std::string(who) + ": give exactly one of t and delta");
    return EPCQ_OK;
}

}  // namespace
}  // namespace epcq

using namespace epcq;

extern "C" int epcq_clamp_t(const float* t, int64_t R, int32_t S, float vsize, float* delta, void* stream) {
    if (R < 0 || S < 0) return fail(EPCQ_ERR_INVALID, "clamp_t: negative size");
    if (R * S == 0 || S == 1) return EPCQ_OK;  // S = 1: no interval (see above)
    clamp_t_kernel<<<dim3((unsigned)ceil_div(R * S, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(t, R, S, vsize,
                                                                                                    delta);
    return check_launch("clamp_t");
}

extern "C" int epcq_render_rays(const float* t, const float* delta, float vsize, const float* sigma, const float* rgb,
                                int64_t R, int32_t S, int32_t white_bg, float* out_rgb, void* stream) {
    if (int e = check_render(t, delta, R, S, "render")) return e;
    if (R == 0) return EPCQ_OK;
    render_kernel<<<dim3((unsigned)ceil_div(R, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(
        t, delta, vsize, sigma, rgb, R, S, white_bg ? 1.0f : 0.0f, out_rgb);
    return check_launch("render");
}

extern "C" int epcq_render_rays_backward(const float* t, const float* delta, float vsize, const float* sigma,
                                         const float* rgb, const float* grad_rgb_out, int64_t R, int32_t S,
                                         int32_t white_bg, float* grad_sigma, float* grad_rgb, void* stream) {
    if (int e = check_render(t, delta, R, S, "render_backward")) return e;
    if (R == 0 || S == 0) return EPCQ_OK;
    render_backward_kernel<<<dim3((unsigned)ceil_div(R, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(
        t, delta, vsize, sigma, rgb, grad_rgb_out, R, S, white_bg ? 1.0f : 0.0f, grad_sigma, grad_rgb);
    return check_launch("render_backward");
}
