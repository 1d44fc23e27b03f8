// Volume-rendering epilogue of the reference renderer for gfx950: clamp_t and
// render (epcq/epcq.py:508-530, also utils/utils.py:422-467), as used after
// the Aggregator (opt/render_image.py:101-109), and render's backward.
//
// One lane per ray walks its S samples front to back:
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
// given) does the same, and clamp_t writes nothing.  Every operation is the
// reference's, one rounding each, in this order; the exponential is the
// library's own (expf_cw).  The
// backward pass recomputes the walk and forms, per sample k (g = dL/drgb):
//   dL/dc_k     = op_k T_k g
//   dL/dsigma_k = delta_k e_k (T_k (g . c_k) - (R_k + [k < S-1] T_{S-1} (g . bg)) / a_k)
// with e_k = exp(-sigma_k delta_k), a_k = 1 - op_k + 1e-10 and R_k the sum of
// op_i T_i (g . c_i) over i > k (taken as the total minus the running
// prefix).  Rows are [R, S] (t, delta, sigma), [R, S, 3] (rgb), [R, 3] (out).
#include <initializer_list>

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

// Samples [k0, k0 + n) of a row (n <= CH) in registers: their
// intervals, densities and colours.  All loads of the chunk are issued before
// any is used, so a lane waits for memory once per chunk, not once per sample.
// CH = kChunk for rows in global memory; the LDS-staged forms use the
// smaller kLdsChunk and kLdsChunkBwd (LDS answers quickly; the sizes are the
// measured best of 4 and 8 for each walk).
constexpr int kChunk = 16;
constexpr int kLdsChunk = 8;      // forward walk
constexpr int kLdsChunkBwd = 4;   // backward walks (more live values per sample)
template <int CH>
struct RowChunk {
    float d[CH], s[CH], c[3 * CH];
};

template <int CH>
__device__ __forceinline__ void load_chunk(const float* tr, const float* dr, const float* sr, const float* cr, int S,
                                           int k0, int n, float vsize, RowChunk<CH>& ch) {
    float tv[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
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
        for (int j = 0; j < CH; ++j)
            if (j < n)
                ch.d[j] = clamp_delta_v(tv[j], j + 1 < n ? tv[min(j + 1, CH - 1)] : tnext, k0 + j + 1 >= S, vsize);
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

// The forward walk of one row whose samples sit at tr (or dr), sr and cr:
// global memory or LDS rows.  Returns the ray's colour in o0..o2.
template <int CH>
__device__ __forceinline__ void render_row(const float* tr, const float* dr, const float* sr, const float* cr, int S,
                                           float vsize, float bg, float& o0, float& o1, float& o2) {
    float T = 1.0f, Tlast = 1.0f;
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    const int live = (tr && S == 1) ? 0 : S;
    for (int k0 = 0; k0 < live; k0 += CH) {
        const int n = min(CH, live - k0);
        RowChunk<CH> ch;
        load_chunk(tr, dr, sr, cr, S, k0, n, vsize, ch);
#pragma unroll
        for (int j = 0; j < CH; ++j) {
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
    o0 = c0 + Tlast * bg;
    o1 = c1 + Tlast * bg;
    o2 = c2 + Tlast * bg;
}

// The backward walk of one row for the colour gradient g0..g2: writes
// dL/dsigma to gs[k] and dL/dc to gc[3k..3k+2].  gs and gc may be the rows sr
// and cr themselves: sample k's gradients are written after its inputs are
// read, and no later sample reads them.
// With CACHE (rows in LDS, writable): pass 1 leaves each sample's interval in
// the t row (t given) and exp(-sigma delta) in the sigma row, and pass 2
// reads them back instead of recomputing them (the same values: a stored
// float reloads with its bits).  gs and gc may be the rows sr and cr
// themselves: sample k's gradients are written after its inputs are read,
// and no later sample reads them.
template <int CH, bool CACHE = false>
__device__ __forceinline__ void render_row_backward(const float* tr, const float* dr, const float* sr, const float* cr,
                                                    int S, float vsize, float bg, float g0, float g1, float g2,
                                                    float* gs, float* gc) {
    if (tr && S == 1) {  // no sample contributes (see above)
        gs[0] = 0.0f;
        gc[0] = gc[1] = gc[2] = 0.0f;
        return;
    }
    // pass 1: the total of op_i T_i (g . c_i) and T_{S-1}
    float T = 1.0f, Tlast = 1.0f, total = 0.0f;
    for (int k0 = 0; k0 < S; k0 += CH) {
        const int n = min(CH, S - k0);
        RowChunk<CH> ch;
        load_chunk(tr, dr, sr, cr, S, k0, n, vsize, ch);
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (j < n) {
                const float e = expf_cw(-ch.s[j] * ch.d[j]);
                const float op = 1.0f - e;
                const float gc = dot3(g0, g1, g2, ch.c[3 * j], ch.c[3 * j + 1], ch.c[3 * j + 2]);
                total = total + op * T * gc;
                Tlast = T;
                T = T * ((1.0f - op) + 1e-10f);
                if constexpr (CACHE) {
                    // the chunk's t were read before any is replaced, and the
                    // next chunk's first t (its own and this chunk's tnext)
                    // is not replaced until that chunk is processed
                    if (tr) const_cast<float*>(tr)[k0 + j] = ch.d[j];
                    const_cast<float*>(sr)[k0 + j] = e;
                }
            }
        }
    }
    const float gbg = (g0 + g1 + g2) * bg;
    // pass 2: per-sample gradients with the suffix sums
    T = 1.0f;
    float prefix = 0.0f;
    const float* drow = CACHE && tr ? tr : dr;  // the intervals, when cached or given
    for (int k0 = 0; k0 < S; k0 += CH) {
        const int n = min(CH, S - k0);
        RowChunk<CH> ch;
        if constexpr (CACHE) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                if (j < n) {
                    ch.d[j] = drow[k0 + j];
                    ch.s[j] = sr[k0 + j];  // exp(-sigma delta)
                    ch.c[3 * j] = cr[3 * (k0 + j)];
                    ch.c[3 * j + 1] = cr[3 * (k0 + j) + 1];
                    ch.c[3 * j + 2] = cr[3 * (k0 + j) + 2];
                }
            }
        } else {
            load_chunk(tr, dr, sr, cr, S, k0, n, vsize, ch);
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (j < n) {
                const int k = k0 + j;
                const float dk = ch.d[j];
                const float e = CACHE ? ch.s[j] : expf_cw(-ch.s[j] * dk);
                const float op = 1.0f - e;
                const float a = (1.0f - op) + 1e-10f;
                const float gck = dot3(g0, g1, g2, ch.c[3 * j], ch.c[3 * j + 1], ch.c[3 * j + 2]);
                const float wk = op * T;
                prefix = prefix + wk * gck;
                const float after = (total - prefix) + (k + 1 < S ? Tlast * gbg : 0.0f);
                gs[k] = dk * e * (T * gck - after / a);
                gc[3 * k] = wk * g0;
                gc[3 * k + 1] = wk * g1;
                gc[3 * k + 2] = wk * g2;
                T = T * a;
            }
        }
    }
}

// Direct forms: each lane reads its own row from global memory (any S).
__global__ __launch_bounds__(kBlock) void render_kernel(const float* __restrict__ t, const float* __restrict__ delta,
                                                        float vsize, const float* __restrict__ sigma,
                                                        const float* __restrict__ rgb, int64_t R, int S, float bg,
                                                        float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= R) return;
    float o0, o1, o2;
    render_row<kChunk>(t ? t + r * S : nullptr, delta ? delta + r * S : nullptr, sigma + r * S, rgb + r * S * 3, S,
                       vsize, bg, o0, o1, o2);
    out[3 * r] = o0;
    out[3 * r + 1] = o1;
    out[3 * r + 2] = o2;
}

__global__ __launch_bounds__(kBlock) void render_backward_kernel(
    const float* __restrict__ t, const float* __restrict__ delta, float vsize, const float* __restrict__ sigma,
    const float* __restrict__ rgb, const float* __restrict__ grad_out, int64_t R, int S, float bg,
    float* __restrict__ grad_sigma, float* __restrict__ grad_rgb) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= R) return;
    render_row_backward<kChunk>(t ? t + r * S : nullptr, delta ? delta + r * S : nullptr, sigma + r * S,
                                rgb + r * S * 3, S, vsize, bg, grad_out[3 * r], grad_out[3 * r + 1],
                                grad_out[3 * r + 2], grad_sigma + r * S, grad_rgb + r * S * 3);
}

// Staged forms, for rows that fit: a one-wave block takes 64 consecutive rays,
// copies their rows of t (or delta), sigma and rgb into LDS with coalesced
// loads (the wave's lanes read consecutive floats), walks each ray from LDS,
// and (backward) writes the gradients back the same way.  The LDS rows have an
// odd stride, so the 64 lanes reading sample k of their rows hit 64 different
// banks.  The arithmetic is the direct form's, operation for operation.
constexpr int kStageRays = 64;
constexpr int kStageUnroll = 16;                  // loads of t and of sigma per lane and round
constexpr size_t kStageLdsMax = 64 * 1024;        // beyond this, the direct form

__host__ __device__ constexpr int odd_stride(int w) { return w | 1; }

__host__ __device__ constexpr size_t stage_lds_bytes(int S) {
    return (size_t)kStageRays * sizeof(float) * (2 * odd_stride(S) + odd_stride(3 * S));
}

// Advances a lane's (row, column) in rows of width w by a fixed number of flat
// elements, q rows and m columns of it (m < w), with a carry.
__device__ __forceinline__ void step_rc(int& row, int& col, int q, int m, int w) {
    col += m;
    row += q;
    if (col >= w) col -= w, ++row;
}

// Copies the block's nr rows of t (or delta) and sigma (width S) and of rgb
// (width 3S), contiguous in global memory from a, b and c, into LDS rows of
// stride ws and wc.  Lane l loads the flat elements l, l + 64, ...; a round
// issues kStageUnroll loads of a and b and 3 kStageUnroll of c before it
// stores any, so a wave waits for memory once per round (once in all for
// S <= 16).
__device__ __forceinline__ void stage_rows(const float* __restrict__ a, const float* __restrict__ b,
                                           const float* __restrict__ c, float* sa, float* sb, float* sc, int nr,
                                           int S, int ws, int wc) {
    constexpr int U = kStageUnroll;
    const int lane = threadIdx.x, n1 = nr * S, n3 = 3 * n1, w3 = 3 * S;
    const int q1 = kStageRays / S, m1 = kStageRays - q1 * S;
    const int q3 = kStageRays / w3, m3 = kStageRays - q3 * w3;
    int r1 = lane / S, c1 = lane - r1 * S;
    int r3 = lane / w3, c3 = lane - r3 * w3;
    // round k covers elements [k 64 U, (k + 1) 64 U) of a and b and three
    // times that span of c: the rounds of a and b cover c as well
    for (int base = 0; base < n1; base += kStageRays * U) {
        const int i0 = base + lane, j0 = 3 * base + lane;
        float va[U], vb[U], vc[3 * U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i0 + kStageRays * u < n1) {
                va[u] = a[i0 + kStageRays * u];
                vb[u] = b[i0 + kStageRays * u];
            }
        }
#pragma unroll
        for (int u = 0; u < 3 * U; ++u)
            if (j0 + kStageRays * u < n3) vc[u] = c[j0 + kStageRays * u];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i0 + kStageRays * u < n1) {
                sa[r1 * ws + c1] = va[u];
                sb[r1 * ws + c1] = vb[u];
                step_rc(r1, c1, q1, m1, S);
            }
        }
#pragma unroll
        for (int u = 0; u < 3 * U; ++u) {
            if (j0 + kStageRays * u < n3) {
                sc[r3 * wc + c3] = vc[u];
                step_rc(r3, c3, q3, m3, w3);
            }
        }
    }
}

// Writes the nr LDS rows of stride ws back to the contiguous rows of width w
// at g, the wave's lanes storing consecutive floats.
__device__ __forceinline__ void unstage_rows(float* __restrict__ g, const float* s, int nr, int w, int ws) {
    const int lane = threadIdx.x, total = nr * w;
    const int q = kStageRays / w, m = kStageRays - q * w;
    int row = lane / w, col = lane - row * w;
    for (int i = lane; i < total; i += kStageRays) {
        g[i] = s[row * ws + col];
        step_rc(row, col, q, m, w);
    }
}

__global__ __launch_bounds__(kStageRays) void render_staged_kernel(const float* __restrict__ t,
                                                                   const float* __restrict__ delta, float vsize,
                                                                   const float* __restrict__ sigma,
                                                                   const float* __restrict__ rgb, int64_t R, int S,
                                                                   float bg, float* __restrict__ out) {
    extern __shared__ float s_rows[];
    const int ws = odd_stride(S), wc = odd_stride(3 * S);
    float* st = s_rows;
    float* ss = st + kStageRays * ws;
    float* sc = ss + kStageRays * ws;
    const int64_t r0 = (int64_t)blockIdx.x * kStageRays;
    const int nr = (int)min<int64_t>(kStageRays, R - r0);
    stage_rows((t ? t : delta) + r0 * S, sigma + r0 * S, rgb + r0 * 3 * S, st, ss, sc, nr, S, ws, wc);
    __syncthreads();
    const int l = threadIdx.x;
    if (l >= nr) return;
    float o0, o1, o2;
    render_row<kLdsChunk>(t ? st + l * ws : nullptr, t ? nullptr : st + l * ws, ss + l * ws, sc + l * wc, S, vsize,
                          bg, o0, o1, o2);
    const int64_t r = r0 + l;
    out[3 * r] = o0;
    out[3 * r + 1] = o1;
    out[3 * r + 2] = o2;
}

__global__ __launch_bounds__(kStageRays) void render_backward_staged_kernel(
    const float* __restrict__ t, const float* __restrict__ delta, float vsize, const float* __restrict__ sigma,
    const float* __restrict__ rgb, const float* __restrict__ grad_out, int64_t R, int S, float bg,
    float* __restrict__ grad_sigma, float* __restrict__ grad_rgb) {
    extern __shared__ float s_rows[];
    const int ws = odd_stride(S), wc = odd_stride(3 * S);
    float* st = s_rows;
    float* ss = st + kStageRays * ws;
    float* sc = ss + kStageRays * ws;
    const int64_t r0 = (int64_t)blockIdx.x * kStageRays;
    const int nr = (int)min<int64_t>(kStageRays, R - r0);
    stage_rows((t ? t : delta) + r0 * S, sigma + r0 * S, rgb + r0 * 3 * S, st, ss, sc, nr, S, ws, wc);
    __syncthreads();
    const int l = threadIdx.x;
    if (l < nr) {
        const int64_t r = r0 + l;
        // the gradients replace the row's sigma and rgb in LDS
        render_row_backward<kLdsChunkBwd, true>(t ? st + l * ws : nullptr, t ? nullptr : st + l * ws, ss + l * ws,
                                       sc + l * wc, S, vsize, bg, grad_out[3 * r], grad_out[3 * r + 1],
                                       grad_out[3 * r + 2], ss + l * ws, sc + l * wc);
    }
    __syncthreads();
    unstage_rows(grad_sigma + r0 * S, ss, nr, S, ws);
    unstage_rows(grad_rgb + r0 * 3 * S, sc, nr, 3 * S, wc);
}

// Staged forms with 16-byte loads, for rows of at most kVecS samples whose
// arrays start 16-byte aligned: all of a 64-ray group's loads fit one round
// of registers (80 per lane), a quarter as many load instructions as the
// one-float form issues.
constexpr int kVecS = 16;

struct GroupRegs {
    float4 a[kVecS / 4], b[kVecS / 4], c[3 * kVecS / 4];
};

// Loads of one group's rows (nr rays of S <= kVecS samples), 16 bytes per
// lane and instruction: lane l takes the flat elements 4 (l + 64 u) .. +3 of
// each span (the spans start 16-byte aligned; the host checks the base
// pointers).  When the span's length is a multiple of 4 (always for a full
// group) a vector lies wholly inside the span or wholly past it; otherwise
// (WHOLE false, a short last group) the vector at the end is read element by
// element.
template <bool WHOLE>
__device__ __forceinline__ void load_span4(const float* __restrict__ g, int n, float4& v, int i) {
    if (WHOLE || i + 3 < n) {
        if (i < n) v = *reinterpret_cast<const float4*>(g + i);
    } else {
        if (i < n) v.x = g[i];
        if (i + 1 < n) v.y = g[i + 1];
        if (i + 2 < n) v.z = g[i + 2];
    }
}

template <bool WHOLE>
__device__ __forceinline__ void group_load_t(const float* __restrict__ a, const float* __restrict__ b,
                                             const float* __restrict__ c, int n1, GroupRegs& v) {
    const int lane = threadIdx.x, n3 = 3 * n1;
#pragma unroll
    for (int u = 0; u < kVecS / 4; ++u) {
        load_span4<WHOLE>(a, n1, v.a[u], 4 * (lane + kStageRays * u));
        load_span4<WHOLE>(b, n1, v.b[u], 4 * (lane + kStageRays * u));
    }
#pragma unroll
    for (int u = 0; u < 3 * kVecS / 4; ++u) load_span4<WHOLE>(c, n3, v.c[u], 4 * (lane + kStageRays * u));
}

__device__ __forceinline__ void group_load(const float* __restrict__ a, const float* __restrict__ b,
                                           const float* __restrict__ c, int nr, int S, GroupRegs& v) {
    const int n1 = nr * S;
    if ((n1 & 3) == 0) group_load_t<true>(a, b, c, n1, v);
    else group_load_t<false>(a, b, c, n1, v);
}

// One loaded vector into LDS rows of width w and stride ws; (row, col) is its
// first element's place and advances by one per element.
template <bool WHOLE>
__device__ __forceinline__ void store4(const float4& v, float* s, int i, int n, int row, int col, int w, int ws) {
    const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (WHOLE || i + k < n) s[row * ws + col] = e[k];
        if (++col == w) col = 0, ++row;
    }
}

// The loaded group into LDS rows of stride ws and wc.  Consecutive vectors of
// a lane are 256 elements apart: (row, col) advances by (256 / w, 256 % w).
template <bool WHOLE>
__device__ __forceinline__ void group_store_t(const GroupRegs& v, float* sa, float* sb, float* sc, int n1, int S,
                                              int ws, int wc) {
    constexpr int kSpan = 4 * kStageRays;
    const int lane = threadIdx.x, n3 = 3 * n1, w3 = 3 * S;
    const int q1 = kSpan / S, m1 = kSpan - q1 * S;
    const int q3 = kSpan / w3, m3 = kSpan - q3 * w3;
    int r1 = (4 * lane) / S, c1 = 4 * lane - r1 * S;
    int r3 = (4 * lane) / w3, c3 = 4 * lane - r3 * w3;
#pragma unroll
    for (int u = 0; u < kVecS / 4; ++u) {
        const int i = 4 * (lane + kStageRays * u);
        if (i < n1) {
            store4<WHOLE>(v.a[u], sa, i, n1, r1, c1, S, ws);
            store4<WHOLE>(v.b[u], sb, i, n1, r1, c1, S, ws);
        }
        step_rc(r1, c1, q1, m1, S);
    }
#pragma unroll
    for (int u = 0; u < 3 * kVecS / 4; ++u) {
        const int i = 4 * (lane + kStageRays * u);
        if (i < n3) store4<WHOLE>(v.c[u], sc, i, n3, r3, c3, w3, wc);
        step_rc(r3, c3, q3, m3, w3);
    }
}

__device__ __forceinline__ void group_store(const GroupRegs& v, float* sa, float* sb, float* sc, int nr, int S, int ws,
                                            int wc) {
    const int n1 = nr * S;
    if ((n1 & 3) == 0) group_store_t<true>(v, sa, sb, sc, n1, S, ws, wc);
    else group_store_t<false>(v, sa, sb, sc, n1, S, ws, wc);
}

// The mirror of group_store for the backward pass's gradients: lane l
// gathers the flat elements 4 (l + 64 u) .. +3 of the LDS rows (stride ws,
// width S; stride wc, width 3S) and writes them with 16-byte stores to the
// group's spans ga and gc (element by element at a span end that is not a
// multiple of 4).
template <bool WHOLE>
__device__ __forceinline__ void unstore4(float* __restrict__ g, const float* s, int i, int n, int row, int col, int w,
                                         int ws) {
    float e[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        e[k] = (WHOLE || i + k < n) ? s[row * ws + col] : 0.0f;
        if (++col == w) col = 0, ++row;
    }
    if (WHOLE || i + 3 < n) {
        *reinterpret_cast<float4*>(g + i) = make_float4(e[0], e[1], e[2], e[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (i + k < n) g[i + k] = e[k];
    }
}

template <bool WHOLE>
__device__ __forceinline__ void group_unstore_t(float* __restrict__ ga, float* __restrict__ gc, const float* sa,
                                                const float* sc, int n1, int S, int ws, int wc) {
    constexpr int kSpan = 4 * kStageRays;
    const int lane = threadIdx.x, n3 = 3 * n1, w3 = 3 * S;
    const int q1 = kSpan / S, m1 = kSpan - q1 * S;
    const int q3 = kSpan / w3, m3 = kSpan - q3 * w3;
    int r1 = (4 * lane) / S, c1 = 4 * lane - r1 * S;
    int r3 = (4 * lane) / w3, c3 = 4 * lane - r3 * w3;
#pragma unroll
    for (int u = 0; u < kVecS / 4; ++u) {
        const int i = 4 * (lane + kStageRays * u);
        if (i < n1) unstore4<WHOLE>(ga, sa, i, n1, r1, c1, S, ws);
        step_rc(r1, c1, q1, m1, S);
    }
#pragma unroll
    for (int u = 0; u < 3 * kVecS / 4; ++u) {
        const int i = 4 * (lane + kStageRays * u);
        if (i < n3) unstore4<WHOLE>(gc, sc, i, n3, r3, c3, w3, wc);
        step_rc(r3, c3, q3, m3, w3);
    }
}

__device__ __forceinline__ void group_unstore(float* ga, float* gc, const float* sa, const float* sc, int nr, int S,
                                              int ws, int wc) {
    const int n1 = nr * S;
    if ((n1 & 3) == 0) group_unstore_t<true>(ga, gc, sa, sc, n1, S, ws, wc);
    else group_unstore_t<false>(ga, gc, sa, sc, n1, S, ws, wc);
}

__global__ __launch_bounds__(kStageRays) void render_vec_kernel(const float* __restrict__ t,
                                                                 const float* __restrict__ delta, float vsize,
                                                                 const float* __restrict__ sigma,
                                                                 const float* __restrict__ rgb, int64_t R, int S,
                                                                 float bg, float* __restrict__ out) {
    extern __shared__ float s_rows[];
    const int ws = odd_stride(S), wc = odd_stride(3 * S);
    float* st = s_rows;
    float* ss = st + kStageRays * ws;
    float* sc = ss + kStageRays * ws;
    const float* tt = t ? t : delta;
    const int64_t g = blockIdx.x;
    const int nr = (int)min<int64_t>(kStageRays, R - g * kStageRays);
    GroupRegs v;
    group_load(tt + g * kStageRays * S, sigma + g * kStageRays * S, rgb + g * kStageRays * 3 * S, nr, S, v);
    group_store(v, st, ss, sc, nr, S, ws, wc);
    __syncthreads();
    const int l = threadIdx.x;
    if (l >= nr) return;
    float o0, o1, o2;
    render_row<kLdsChunk>(t ? st + l * ws : nullptr, t ? nullptr : st + l * ws, ss + l * ws, sc + l * wc, S, vsize,
                          bg, o0, o1, o2);
    const int64_t r = g * kStageRays + l;
    out[3 * r] = o0;
    out[3 * r + 1] = o1;
    out[3 * r + 2] = o2;
}

__global__ __launch_bounds__(kStageRays) void render_backward_vec_kernel(
    const float* __restrict__ t, const float* __restrict__ delta, float vsize, const float* __restrict__ sigma,
    const float* __restrict__ rgb, const float* __restrict__ grad_out, int64_t R, int S, float bg,
    float* __restrict__ grad_sigma, float* __restrict__ grad_rgb) {
    extern __shared__ float s_rows[];
    const int ws = odd_stride(S), wc = odd_stride(3 * S);
    float* st = s_rows;
    float* ss = st + kStageRays * ws;
    float* sc = ss + kStageRays * ws;
    const float* tt = t ? t : delta;
    const int64_t g = blockIdx.x, r0 = g * kStageRays;
    const int nr = (int)min<int64_t>(kStageRays, R - r0);
    GroupRegs v;
    group_load(tt + r0 * S, sigma + r0 * S, rgb + r0 * 3 * S, nr, S, v);
    group_store(v, st, ss, sc, nr, S, ws, wc);
    __syncthreads();
    const int l = threadIdx.x;
    if (l < nr) {
        const int64_t r = r0 + l;
        render_row_backward<kLdsChunkBwd, true>(t ? st + l * ws : nullptr, t ? nullptr : st + l * ws, ss + l * ws,
                                       sc + l * wc, S, vsize, bg, grad_out[3 * r], grad_out[3 * r + 1],
                                       grad_out[3 * r + 2], ss + l * ws, sc + l * wc);
    }
    __syncthreads();
    group_unstore(grad_sigma + r0 * S, grad_rgb + r0 * 3 * S, ss, sc, nr, S, ws, wc);
}

bool aligned16(std::initializer_list<const float*> ps) {
    for (const float* p : ps)
        if (reinterpret_cast<uintptr_t>(p) & 15) return false;
    return true;
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
    if (S > 0 && S <= kVecS && aligned16({t ? t : delta, sigma, rgb}))
        render_vec_kernel<<<dim3((unsigned)ceil_div(R, kStageRays)), dim3(kStageRays), stage_lds_bytes(S),
                             as_stream(stream)>>>(
            t, delta, vsize, sigma, rgb, R, S, white_bg ? 1.0f : 0.0f, out_rgb);
    else if (S > 0 && stage_lds_bytes(S) <= kStageLdsMax)
        render_staged_kernel<<<dim3((unsigned)ceil_div(R, kStageRays)), dim3(kStageRays), stage_lds_bytes(S),
                               as_stream(stream)>>>(t, delta, vsize, sigma, rgb, R, S, white_bg ? 1.0f : 0.0f, out_rgb);
    else
        render_kernel<<<dim3((unsigned)ceil_div(R, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(
            t, delta, vsize, sigma, rgb, R, S, white_bg ? 1.0f : 0.0f, out_rgb);
    return check_launch("render");
}

extern "C" int epcq_render_rays_backward(const float* t, const float* delta, float vsize, const float* sigma,
                                         const float* rgb, const float* grad_rgb_out, int64_t R, int32_t S,
                                         int32_t white_bg, float* grad_sigma, float* grad_rgb, void* stream) {
    if (int e = check_render(t, delta, R, S, "render_backward")) return e;
    if (R == 0 || S == 0) return EPCQ_OK;
    if (S <= kVecS && aligned16({t ? t : delta, sigma, rgb, grad_sigma, grad_rgb}))
        render_backward_vec_kernel<<<dim3((unsigned)ceil_div(R, kStageRays)), dim3(kStageRays), stage_lds_bytes(S),
                                      as_stream(stream)>>>(
            t, delta, vsize, sigma, rgb, grad_rgb_out, R, S, white_bg ? 1.0f : 0.0f, grad_sigma, grad_rgb);
    else if (stage_lds_bytes(S) <= kStageLdsMax)
        render_backward_staged_kernel<<<dim3((unsigned)ceil_div(R, kStageRays)), dim3(kStageRays),
                                        stage_lds_bytes(S), as_stream(stream)>>>(
            t, delta, vsize, sigma, rgb, grad_rgb_out, R, S, white_bg ? 1.0f : 0.0f, grad_sigma, grad_rgb);
    else
        render_backward_kernel<<<dim3((unsigned)ceil_div(R, kBlock)), dim3(kBlock), 0, as_stream(stream)>>>(
            t, delta, vsize, sigma, rgb, grad_rgb_out, R, S, white_bg ? 1.0f : 0.0f, grad_sigma, grad_rgb);
    return check_launch("render_backward");
}
