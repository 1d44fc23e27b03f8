// Device bodies and launch plan of the fused hash-table build, shared by the
// per-view entry point (epcq_build.hip) and the batched multi-view kernels
// (epcq_batch.hip).  See epcq_build.hip for the algorithm.
#pragma once

#include <algorithm>

#include "epcq_common.h"
#include "epcq_device.h"

namespace epcq {

constexpr int kBuildBlock = 256;
constexpr int kBuildWaves = kBuildBlock / kWave;
constexpr int kMaxBinBits = 13;    // <= 8192 coarse bins (32 KB of LDS cursors)
constexpr int kMaxTileShift = 10;  // <= 1024 pixels per bin unless the image needs more
constexpr int kBinPoints = 2048;   // target points per bin (dense views get smaller bins)
constexpr int kSegBlocks = 64;     // blocks per column-scan segment

struct ProjParams {
    float w[12];
    float K[9];
    int H, W;
    int P;
};

__device__ __forceinline__ void load_proj(const float* __restrict__ w2c, const float* __restrict__ K,
                                          int H, int W, ProjParams& pp) {
#pragma unroll
    for (int i = 0; i < 12; ++i) pp.w[i] = w2c[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) pp.K[i] = K[i];
    pp.H = H;
    pp.W = W;
    pp.P = H * W;
}

// _world2cam / _cam2image / _convetpixel (epcq/epcq.py:251-274): camera depth
// and row-major pixel key, P for points whose truncated screen coordinate
// falls outside the image.
__device__ __forceinline__ int project_key(const ProjParams& pp, float x, float y, float z, float& depth) {
    const float* w = pp.w;
    const float* K = pp.K;
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
    depth = Z;
    // trunc-toward-zero then 0 <= u < W, 0 <= v < H, without converting
    // out-of-range values (NaN fails every comparison)
    const bool in = (sx > -1.0f) && (sx < (float)pp.W) && (sy > -1.0f) && (sy < (float)pp.H);
    return in ? ((int)sx + (int)sy * pp.W) : pp.P;
}

// Exclusive scan of one int per thread over the 256-thread block.
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave, int& total) {
    const int lane = __lane_id();
    const int wave = threadIdx.x / kWave;
    int x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) s_wave[wave] = x;
    __syncthreads();
    int wbase = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kBuildWaves; ++w) {
        const int t = s_wave[w];
        if (w < wave) wbase += t;
        total += t;
    }
    __syncthreads();
    return wbase + x - v;
}

// Exclusive scan of a[0..len) in LDS (len a power of two), in place.
__device__ void lds_exclusive_scan(int* a, int len, int* s_wave) {
    if (len < kBuildBlock) {
        const int v = (int)threadIdx.x < len ? a[threadIdx.x] : 0;
        int total;
        const int ex = block_exclusive_scan(v, s_wave, total);
        if ((int)threadIdx.x < len) a[threadIdx.x] = ex;
        __syncthreads();
        return;
    }
    const int per = len / kBuildBlock;
    int* seg = a + threadIdx.x * per;
    int sum = 0;
    for (int j = 0; j < per; ++j) sum += seg[j];
    int total;
    int run = block_exclusive_scan(sum, s_wave, total);
    for (int j = 0; j < per; ++j) {
        const int v = seg[j];
        seg[j] = run;
        run += v;
    }
    __syncthreads();
}

// Appends this round's keyed elements to per-key cursors in LDS in lane order
// within a wave and wave order within the block; returns each active lane's
// position (undefined for inactive lanes).
__device__ __forceinline__ int ordered_append(int* cursor, uint32_t key, int bits, bool active) {
    const uint64_t m = match_lanes(key, bits, active);
    const int lane = __lane_id();
    const int wave = threadIdx.x / kWave;
    const int leader = m ? (__ffsll((unsigned long long)m) - 1) : 0;
    const int rank = __popcll(m & lanemask_lt());
    int base = 0;
    for (int w = 0; w < kBuildWaves; ++w) {
        if (wave == w && active && lane == leader) {
            base = cursor[key];
            cursor[key] = base + __popcll(m);
        }
        __syncthreads();
    }
    base = __shfl(base, leader);
    return base + rank;
}


// ---- depth order of a pixel run (the reference's stable bubble sort,
// cuda_util.cuh:291-313, with its strict '>' comparison) ------------------------

// Unsigned key with the order of the float comparison: negative values
// reversed, and -0 equal to +0.  Depths of on-image points are finite (a
// non-finite depth makes the screen coordinate NaN or infinite).
__device__ __forceinline__ uint32_t depth_key(float d) {
    const uint32_t u = d == 0.0f ? 0u : __float_as_uint(d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

constexpr int kSmallRun = 32;  // runs up to this length are sorted by one thread

// Stable insertion sort of a short run by one thread.
__device__ __forceinline__ void thread_sort_run(float* d, int* ix, int c) {
    for (int a = 1; a < c; ++a) {
        const float kd = d[a];
        const int ki = ix[a];
        int j = a - 1;
        while (j >= 0 && d[j] > kd) {
            d[j + 1] = d[j];
            ix[j + 1] = ix[j];
            --j;
        }
        d[j + 1] = kd;
        ix[j + 1] = ki;
    }
}

// Stable sort of a long run (d, ix)[0, c) by the whole block: a least-
// significant-digit radix sort of depth_key, 8 bits a pass, through the
// scratch (sd, sx)[0, c); a pass whose digit is the same for the whole run
// moves nothing.  O(c) per pass where the insertion sort is O(c^2).
// s_dig: 256 LDS cursors.  Every thread of the block calls it with the same c.
__device__ void block_sort_run(float* d, int* ix, float* sd, int* sx, int c, int* s_dig, int* s_wave) {
    float* src_d = d;
    int* src_i = ix;
    float* dst_d = sd;
    int* dst_i = sx;
    for (int shift = 0; shift < 32; shift += 8) {
        for (int k = threadIdx.x; k < 256; k += kBuildBlock) s_dig[k] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < c; i += kBuildBlock) atomicAdd(&s_dig[(depth_key(src_d[i]) >> shift) & 255u], 1);
        __syncthreads();
        const bool uniform = s_dig[(depth_key(src_d[0]) >> shift) & 255u] == c;
        __syncthreads();
        if (uniform) continue;
        lds_exclusive_scan(s_dig, 256, s_wave);
        for (int r0 = 0; r0 < c; r0 += kBuildBlock) {
            const int i = r0 + threadIdx.x;
            const bool active = i < c;
            const float v = active ? src_d[i] : 0.0f;
            const int id = active ? src_i[i] : 0;
            const uint32_t dig = active ? (depth_key(v) >> shift) & 255u : 0u;
            const int pos = ordered_append(s_dig, dig, 8, active);
            if (active) {
                dst_d[pos] = v;
                dst_i[pos] = id;
            }
        }
        __syncthreads();
        float* td = src_d;
        src_d = dst_d;
        dst_d = td;
        int* ti = src_i;
        src_i = dst_i;
        dst_i = ti;
    }
    if (src_d != d) {
        for (int i = threadIdx.x; i < c; i += kBuildBlock) {
            d[i] = src_d[i];
            ix[i] = src_i[i];
        }
        __syncthreads();
    }
}

// Sorts the pixel runs of `tile` pixels by depth: run f has count cnt[f] and
// ends at base + end[f] of (d, ix); (sd, sx) are scratch at the same
// positions.  Short runs one per thread, long runs by the whole block, one
// after another in pixel order.
__device__ void sort_pixel_runs(int tile, const int* cnt, const int* end, int64_t base, float* d, int* ix,
                                float* sd, int* sx, int* s_wave) {
    __shared__ int s_dig[256];
    __shared__ uint64_t s_hot[kBuildWaves];
    for (int f = threadIdx.x; f < tile; f += kBuildBlock) {
        const int c = cnt[f];
        if (c < 2 || c > kSmallRun) continue;
        const int64_t s = base + end[f] - c;
        thread_sort_run(d + s, ix + s, c);
    }
    for (int f0 = 0; f0 < tile; f0 += kBuildBlock) {
        const int f = f0 + threadIdx.x;
        const uint64_t m = __ballot(f < tile && cnt[f] > kSmallRun);
        if (__lane_id() == 0) s_hot[threadIdx.x / kWave] = m;
        __syncthreads();
        for (int w = 0; w < kBuildWaves; ++w) {
            uint64_t mm = s_hot[w];
            while (mm) {
                const int g = f0 + w * kWave + (__ffsll((unsigned long long)mm) - 1);
                mm &= mm - 1;
                const int c = cnt[g];
                const int64_t s = base + end[g] - c;
                block_sort_run(d + s, ix + s, sd + s, sx + s, c, s_dig, s_wave);
            }
        }
        __syncthreads();
    }
}

// Everything one view's build reads and writes.
struct BuildView {
    const float* xyz;
    int64_t n;
    const float* w2c;  // >= 3 rows x 4, row-major
    const float* K;    // 3 x 3, row-major
    int H, W, P;
    int ts, nbins, bin_bits;
    int64_t chunk;
    int nblk, nseg;
    int sort;
    int rel_cursors;  // hist holds cursors relative to the bin's start (bstart is added when binning)
    int* hist;    // [nblk][nbins] counts, then write cursors
    int* part;    // [nseg][nbins]
    int* bstart;  // [nbins + 1]
    int* btotal;  // [nbins]
    int* tidx;    // [n] binned point index (null when tpt is used)
    float4* tpt;  // [n] binned (x, y, z, index bits) records, or null
    float* tdep;  // [n] binned depth
    uint16_t* tfine;  // [n] pixel within bin (null with tpt: recomputed from the record)
    int* out_idx;     // [n]
    float* out_depth; // [n] (may be null: not written)
    int* out_hash;    // [P, 2]
    float4* out_pts;  // [n] pixel-ordered (x, y, z, index bits) copy (with tpt), or null
};

// Launch geometry of the build for (n, H, W) and its workspace layout.
struct BuildPlan {
    int64_t P;
    int ts;         // log2 pixels per coarse bin
    int nbins;
    int bin_bits;   // bits of a coarse bin id
    int64_t chunk;  // points per block in count / bin
    int nblk;
    int nseg;
    size_t off_hist, off_part, off_start, off_total, off_tidx, off_tdep, off_tfine, bytes;
    bool points;
};

// points: the bin pass carries each point's coordinates with its index (a
// 16-byte record) so that the fine pass can write the pixel-ordered point
// copy without gathering from xyz; no depth is binned then.  bin_points: the
// points a bin of a dense view is sized to hold (the fine pass gives each bin
// one block).
// Largest block x bin count matrix the per-view build scans with one block
// per bin (larger ones take the segmented scan).
constexpr int64_t kColScanEntries = int64_t(1) << 20;

inline BuildPlan plan_build(int64_t n, int H, int W, bool points = false, int bin_points = kBinPoints) {
    BuildPlan bp{};
    bp.P = (int64_t)H * W;
    // bins of at most 2^kMaxTileShift pixels, fewer for dense views so that a
    // bin holds about kBinPoints points or less (the fine pass gives each bin
    // one block), but never more than 2^kMaxBinBits bins
    const int64_t q = (int64_t)bin_points * bp.P / std::max<int64_t>(n, 1);
    int ts_dense = 0;
    while (ts_dense < kMaxTileShift && (int64_t(2) << ts_dense) <= q) ++ts_dense;
    bp.ts = std::max(ceil_log2(bp.P) - kMaxBinBits, ts_dense);
    bp.nbins = (int)ceil_div(bp.P, int64_t(1) << bp.ts);
    bp.bin_bits = std::max(1, ceil_log2(bp.nbins));
    // about 2048 blocks at most, each a multiple of 4096 points
    const int64_t unit = 4096;
    const int64_t units = std::max<int64_t>(1, ceil_div(ceil_div(n, unit), 2048));
    bp.chunk = unit * units;
    bp.nblk = (int)std::max<int64_t>(1, ceil_div(n, bp.chunk));
    bp.nseg = (int)ceil_div(bp.nblk, kSegBlocks);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes, 256);
        return at;
    };
    bp.off_hist = take((size_t)bp.nblk * bp.nbins * 4);
    bp.off_part = take((size_t)bp.nseg * bp.nbins * 4);
    bp.off_start = take((size_t)(bp.nbins + 1) * 4);
    bp.off_total = take((size_t)bp.nbins * 4);
    bp.off_tidx = take((size_t)n * (points ? 16 : 4));
    bp.off_tdep = take(points ? 0 : (size_t)n * 4);
    bp.points = points;
    bp.off_tfine = take(points ? 0 : (size_t)n * 2);  // records recompute their pixel
    bp.bytes = off;
    return bp;
}

// BuildView of plan bp over workspace ws.
inline BuildView make_build_view(const BuildPlan& bp, const float* xyz, int64_t n, const float* w2c,
                                 const float* K, int H, int W, int sort, void* ws, int* out_idx,
                                 float* out_depth, int* out_hash, float4* out_pts = nullptr) {
    char* w = static_cast<char*>(ws);
    BuildView v{};
    v.xyz = xyz;
    v.n = n;
    v.w2c = w2c;
    v.K = K;
    v.H = H;
    v.W = W;
    v.P = (int)bp.P;
    v.ts = bp.ts;
    v.nbins = bp.nbins;
    v.bin_bits = bp.bin_bits;
    v.chunk = bp.chunk;
    v.nblk = bp.nblk;
    v.nseg = bp.nseg;
    v.sort = sort;
    v.hist = reinterpret_cast<int*>(w + bp.off_hist);
    v.part = reinterpret_cast<int*>(w + bp.off_part);
    v.bstart = reinterpret_cast<int*>(w + bp.off_start);
    v.btotal = reinterpret_cast<int*>(w + bp.off_total);
    v.tidx = bp.points ? nullptr : reinterpret_cast<int*>(w + bp.off_tidx);
    v.tpt = bp.points ? reinterpret_cast<float4*>(w + bp.off_tidx) : nullptr;
    v.tdep = bp.points ? nullptr : reinterpret_cast<float*>(w + bp.off_tdep);
    v.tfine = bp.points ? nullptr : reinterpret_cast<uint16_t*>(w + bp.off_tfine);
    v.out_idx = out_idx;
    v.out_depth = out_depth;
    v.out_hash = out_hash;
    v.out_pts = bp.points ? out_pts : nullptr;
    return v;
}

// ---- 1. count: block `blk` histograms its chunk over the coarse bins --------
__device__ __forceinline__ void count_body(const BuildView& bv, int blk, int* s_hist) {
    ProjParams pp;
    load_proj(bv.w2c, bv.K, bv.H, bv.W, pp);
    for (int b = threadIdx.x; b < bv.nbins; b += kBuildBlock) s_hist[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blk * bv.chunk;
    const int64_t end = min(base + bv.chunk, bv.n);
    const float* xyz = bv.xyz;
#pragma unroll 4
    for (int64_t i = base + threadIdx.x; i < end; i += kBuildBlock) {
        float depth;
        const int key = project_key(pp, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], depth);
        if (key < pp.P) atomicAdd(&s_hist[key >> bv.ts], 1);
    }
    __syncthreads();
    int* row = bv.hist + (int64_t)blk * bv.nbins;
    for (int b = threadIdx.x; b < bv.nbins; b += kBuildBlock) row[b] = s_hist[b];
    __syncthreads();
}

// ---- 2. scan -------------------------------------------------------------------
// part[seg][b] = sum of hist[blk][b] over the blocks of segment seg.
__device__ __forceinline__ void col_partial_body(const BuildView& bv, int b, int seg) {
    if (b >= bv.nbins) return;
    const int b0 = seg * kSegBlocks, b1 = min(b0 + kSegBlocks, bv.nblk);
    int s = 0;
    for (int blk = b0; blk < b1; ++blk) s += bv.hist[(int64_t)blk * bv.nbins + b];
    bv.part[(int64_t)seg * bv.nbins + b] = s;
}

// Bin b's total over the segments; part[seg][b] becomes the exclusive
// prefix over the segments (one thread per bin).
__device__ __forceinline__ void bin_total_body(const BuildView& bv, int b) {
    if (b >= bv.nbins) return;
    int run = 0;
    for (int seg = 0; seg < bv.nseg; ++seg) {
        const int v = bv.part[(int64_t)seg * bv.nbins + b];
        bv.part[(int64_t)seg * bv.nbins + b] = run;
        run += v;
    }
    bv.btotal[b] = run;
}

// One block: bin starts, the exclusive scan of the bin totals (N_in at the
// end).
__device__ __forceinline__ void bin_scan_body(const BuildView& bv, int* s_wave, int* s_carry) {
    const int nbins = bv.nbins;
    if (threadIdx.x == 0) *s_carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nbins; b0 += kBuildBlock) {
        const int b = b0 + threadIdx.x;
        const int tot = b < nbins ? bv.btotal[b] : 0;
        const int carry = *s_carry;
        int total;
        const int ex = block_exclusive_scan(tot, s_wave, total);
        if (b < nbins) bv.bstart[b] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) *s_carry = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) bv.bstart[nbins] = *s_carry;
    __syncthreads();
}

// One block per bin b: hist[blk][b] <- the exclusive prefix of bin b's
// column over the blocks (the cursor relative to the bin's start), btotal[b]
// <- the column's total.  Thread t scans a run of ceil(nblk / 256) blocks.
__device__ __forceinline__ void col_scan_body(const BuildView& bv, int b, int* s_wave) {
    const int per = (bv.nblk + kBuildBlock - 1) / kBuildBlock;
    const int j0 = (int)threadIdx.x * per, j1 = min(j0 + per, bv.nblk);
    int sum = 0;
    for (int j = j0; j < j1; ++j) sum += bv.hist[(int64_t)j * bv.nbins + b];
    int total;
    int run = block_exclusive_scan(sum, s_wave, total);
    for (int j = j0; j < j1; ++j) {
        const int64_t at = (int64_t)j * bv.nbins + b;
        const int v = bv.hist[at];
        bv.hist[at] = run;
        run += v;
    }
    if (threadIdx.x == 0) bv.btotal[b] = total;
}

// hist[blk][b] <- absolute write cursor of block blk in bin b.
__device__ __forceinline__ void col_apply_body(const BuildView& bv, int b, int seg) {
    if (b >= bv.nbins) return;
    const int b0 = seg * kSegBlocks, b1 = min(b0 + kSegBlocks, bv.nblk);
    int run = bv.bstart[b] + bv.part[(int64_t)seg * bv.nbins + b];
    for (int blk = b0; blk < b1; ++blk) {
        const int64_t at = (int64_t)blk * bv.nbins + b;
        const int v = bv.hist[at];
        bv.hist[at] = run;
        run += v;
    }
}

// ---- 3. bin: block `blk` appends its chunk's points to their bins -------------
__device__ __forceinline__ void bin_body(const BuildView& bv, int blk, int* s_cur) {
    ProjParams pp;
    load_proj(bv.w2c, bv.K, bv.H, bv.W, pp);
    const int* row = bv.hist + (int64_t)blk * bv.nbins;
    for (int b = threadIdx.x; b < bv.nbins; b += kBuildBlock)
        s_cur[b] = row[b] + (bv.rel_cursors ? bv.bstart[b] : 0);
    __syncthreads();
    const int64_t base = (int64_t)blk * bv.chunk;
    const int64_t end = min(base + bv.chunk, bv.n);
    const uint32_t fmask = (1u << bv.ts) - 1u;
    const float* xyz = bv.xyz;
    for (int64_t r0 = base; r0 < end; r0 += kBuildBlock) {
        const int64_t i = r0 + threadIdx.x;
        float depth = 0.0f, x = 0.0f, y = 0.0f, z = 0.0f;
        int key = pp.P;
        if (i < end) {
            x = xyz[3 * i];
            y = xyz[3 * i + 1];
            z = xyz[3 * i + 2];
            key = project_key(pp, x, y, z, depth);
        }
        const bool active = key < pp.P;
        const uint32_t bin = active ? ((uint32_t)key >> bv.ts) : 0u;
        const int pos = ordered_append(s_cur, bin, bv.bin_bits, active);
        if (active) {
            if (bv.tpt) bv.tpt[pos] = make_float4(x, y, z, __int_as_float((int)i));
            else bv.tidx[pos] = (int)i;
            if (bv.out_depth) bv.tdep[pos] = depth;  // depths only when they are an output
            if (!bv.tpt) bv.tfine[pos] = (uint16_t)((uint32_t)key & fmask);
        }
    }
    __syncthreads();
}

// ---- 4. fine: bin `b` of the view: pixel counts, hash entries, stable
// placement, optional depth sort, and its share of the off-image tail ----------
__device__ __forceinline__ void fine_body(const BuildView& bv, int b, int* s_mem, int* s_wave) {
    const int ts = bv.ts;
    const int tile = 1 << ts;
    int* s_cnt = s_mem;
    int* s_cur = s_mem + tile;
    const int bs = bv.bstart[b];
    const int tot = bv.btotal[b];
    // pixel within the bin of binned point i: stored by the bin pass, or, for
    // binned (x, y, z, index) records, recomputed from the coordinates by the
    // same projection (bit-identical to the bin pass's: the same operations
    // in the same order), which saves the bin pass a second scattered write
    // per point
    ProjParams pp;
    if (bv.tpt) load_proj(bv.w2c, bv.K, bv.H, bv.W, pp);
    const uint32_t fmask = (uint32_t)tile - 1u;
    auto fine_of = [&](int i, float4& q) -> uint32_t {
        if (!bv.tpt) return bv.tfine[i];
        q = bv.tpt[i];
        float depth;
        return (uint32_t)project_key(pp, q.x, q.y, q.z, depth) & fmask;
    };
    for (int f = threadIdx.x; f < tile; f += kBuildBlock) s_cnt[f] = 0;
    __syncthreads();
    for (int i = bs + threadIdx.x; i < bs + tot; i += kBuildBlock) {
        float4 q;
        atomicAdd(&s_cnt[fine_of(i, q)], 1);
    }
    __syncthreads();
    for (int f = threadIdx.x; f < tile; f += kBuildBlock) s_cur[f] = s_cnt[f];
    __syncthreads();
    lds_exclusive_scan(s_cur, tile, s_wave);
    const int64_t p0 = (int64_t)b << ts;
    for (int f = threadIdx.x; f < tile; f += kBuildBlock) {
        const int64_t p = p0 + f;
        if (p < bv.P) {
            const int c = s_cnt[f];
            bv.out_hash[2 * p] = c ? bs + s_cur[f] : -1;
            bv.out_hash[2 * p + 1] = c ? c : -1;
        }
    }
    __syncthreads();
    float* out_depth = bv.out_depth;
    for (int r0 = bs; r0 < bs + tot; r0 += kBuildBlock) {
        const int i = r0 + threadIdx.x;
        const bool active = i < bs + tot;
        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const uint32_t f = active ? fine_of(i, q) : 0u;
        const int pos = ordered_append(s_cur, f, ts, active);
        if (active) {
            if (bv.tpt) {
                bv.out_idx[bs + pos] = __float_as_int(q.w);
                bv.out_pts[bs + pos] = q;
            } else {
                bv.out_idx[bs + pos] = bv.tidx[i];
                if (out_depth) out_depth[bs + pos] = bv.tdep[i];
            }
        }
    }
    if (bv.sort && out_depth) {
        __syncthreads();
        // each pixel run by depth; the binned index and depth arrays of this
        // bin are free once placed and serve as the long runs' scratch
        sort_pixel_runs(tile, s_cnt, s_cur, bs, out_depth, bv.out_idx, bv.tdep, bv.tidx, s_wave);
    }
    // off-image tail: positions [N_in, n) hold -1 like the reference's
    // torch::full(-1) outputs; bin b fills every nbins-th run of 256
    const int64_t n_in = bv.bstart[bv.nbins];
    for (int64_t i = n_in + (int64_t)b * kBuildBlock + threadIdx.x; i < bv.n;
         i += (int64_t)bv.nbins * kBuildBlock) {
        bv.out_idx[i] = -1;
        if (out_depth) out_depth[i] = -1.0f;
    }
    __syncthreads();
}

}  // namespace epcq
