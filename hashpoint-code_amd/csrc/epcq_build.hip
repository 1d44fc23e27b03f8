// Per-pixel hash-table build for gfx950.
//
//   epcq_build_hashtable   fused pts2plane (epcq/epcq.py:279-295)
//   epcq_scatter_hashtable scatter_hashtable on pixel-sorted input
//                          (epcq/csrc/rasterization.cu:7-34,130-171)
//   epcq_scatter           per-pixel B nearest by depth
//                          (epcq/csrc/rasterization.cu:36-128)
//
// The fused build is a two-level stable counting sort on the pixel key that
// never uses global atomics:
//   1. count    each 256-thread block projects a chunk of points and builds an
//               LDS histogram over coarse bins (runs of 2^ts consecutive
//               row-major pixels);
//   2. scan     per-bin column scan of the block histograms (three small
//               kernels) gives every (block, bin) a write cursor and every
//               bin its start in the output;
//   3. bin      blocks re-project their chunk and append each point to its
//               bin; lanes with equal bins are matched with ballots and waves
//               take their turn in order, so each bin receives its points in
//               increasing point index;
//   4. fine     one block per bin counts its pixels in LDS, scans them into
//               the hash table, and places the points with the same ordered
//               ballot scheme; optional per-pixel stable sort by depth; the
//               off-image tail of the output is filled with -1.
// The output order is therefore (pixel, point index), or (pixel, depth, point
// index) with sort, independent of scheduling.
#include "epcq_common.h"
#include "epcq_device.h"

namespace epcq {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxBinBits = 13;   // <= 8192 coarse bins (32 KB of LDS cursors)
constexpr int kMinTileShift = 10; // >= 1024 pixels per bin
constexpr int kSegBlocks = 64;    // blocks per column-scan segment

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
    for (int w = 0; w < kWavesPerBlock; ++w) {
        const int t = s_wave[w];
        if (w < wave) wbase += t;
        total += t;
    }
    __syncthreads();
    return wbase + x - v;
}

// Exclusive scan of a[0..len) in LDS (len a multiple of 256), in place.
__device__ void lds_exclusive_scan(int* a, int len, int* s_wave) {
    const int per = len / kBlock;
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
    for (int w = 0; w < kWavesPerBlock; ++w) {
        if (wave == w && active && lane == leader) {
            base = cursor[key];
            cursor[key] = base + __popcll(m);
        }
        __syncthreads();
    }
    base = __shfl(base, leader);
    return base + rank;
}

// ---- 1. count --------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void build_count_kernel(const float* __restrict__ xyz, int64_t n,
                                                             const float* __restrict__ w2c,
                                                             const float* __restrict__ K, int H, int W,
                                                             int ts, int nbins, int64_t chunk,
                                                             int* __restrict__ hist) {
    extern __shared__ int s_hist[];
    ProjParams pp;
    load_proj(w2c, K, H, W, pp);
    for (int b = threadIdx.x; b < nbins; b += kBlock) s_hist[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(base + chunk, n);
#pragma unroll 4
    for (int64_t i = base + threadIdx.x; i < end; i += kBlock) {
        float depth;
        const int key = project_key(pp, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], depth);
        if (key < pp.P) atomicAdd(&s_hist[key >> ts], 1);
    }
    __syncthreads();
    int* row = hist + (int64_t)blockIdx.x * nbins;
    for (int b = threadIdx.x; b < nbins; b += kBlock) row[b] = s_hist[b];
}

// ---- 2. scan ---------------------------------------------------------------
// part[seg][b] = sum of hist[blk][b] over the blocks of segment seg.
__global__ __launch_bounds__(kBlock) void col_partial_kernel(const int* __restrict__ hist, int nblk,
                                                             int nbins, int* __restrict__ part) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const int seg = blockIdx.y;
    if (b >= nbins) return;
    const int b0 = seg * kSegBlocks, b1 = min(b0 + kSegBlocks, nblk);
    int s = 0;
    for (int blk = b0; blk < b1; ++blk) s += hist[(int64_t)blk * nbins + b];
    part[(int64_t)seg * nbins + b] = s;
}

// One block: bin totals, bin starts (exclusive scan over bins, N_in at the
// end) and absolute segment offsets part[seg][b].
__global__ __launch_bounds__(kBlock) void bin_scan_kernel(int* __restrict__ part, int nseg, int nbins,
                                                          int* __restrict__ bin_start,
                                                          int* __restrict__ bin_total) {
    __shared__ int s_wave[kWavesPerBlock];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    for (int b0 = 0; b0 < nbins; b0 += kBlock) {
        const int b = b0 + threadIdx.x;
        int tot = 0;
        if (b < nbins) {
            for (int seg = 0; seg < nseg; ++seg) {
                const int v = part[(int64_t)seg * nbins + b];
                part[(int64_t)seg * nbins + b] = tot;
                tot += v;
            }
        }
        __syncthreads();
        const int carry = s_carry;
        int total;
        const int ex = block_exclusive_scan(tot, s_wave, total);
        if (b < nbins) {
            const int start = carry + ex;
            bin_start[b] = start;
            bin_total[b] = tot;
            for (int seg = 0; seg < nseg; ++seg) part[(int64_t)seg * nbins + b] += start;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) bin_start[nbins] = s_carry;
}

// hist[blk][b] <- absolute write cursor of block blk in bin b.
__global__ __launch_bounds__(kBlock) void col_apply_kernel(int* __restrict__ hist, int nblk, int nbins,
                                                           const int* __restrict__ part) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const int seg = blockIdx.y;
    if (b >= nbins) return;
    const int b0 = seg * kSegBlocks, b1 = min(b0 + kSegBlocks, nblk);
    int run = part[(int64_t)seg * nbins + b];
    for (int blk = b0; blk < b1; ++blk) {
        const int64_t at = (int64_t)blk * nbins + b;
        const int v = hist[at];
        hist[at] = run;
        run += v;
    }
}

// ---- 3. bin ----------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void build_bin_kernel(const float* __restrict__ xyz, int64_t n,
                                                           const float* __restrict__ w2c,
                                                           const float* __restrict__ K, int H, int W,
                                                           int ts, int nbins, int bin_bits, int64_t chunk,
                                                           const int* __restrict__ cursors,
                                                           int* __restrict__ tmp_idx,
                                                           float* __restrict__ tmp_depth,
                                                           uint16_t* __restrict__ tmp_fine) {
    extern __shared__ int s_cur[];
    ProjParams pp;
    load_proj(w2c, K, H, W, pp);
    const int* row = cursors + (int64_t)blockIdx.x * nbins;
    for (int b = threadIdx.x; b < nbins; b += kBlock) s_cur[b] = row[b];
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(base + chunk, n);
    const uint32_t fmask = (1u << ts) - 1u;
    for (int64_t r0 = base; r0 < end; r0 += kBlock) {
        const int64_t i = r0 + threadIdx.x;
        float depth = 0.0f;
        int key = pp.P;
        if (i < end) key = project_key(pp, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], depth);
        const bool active = key < pp.P;
        const uint32_t bin = active ? ((uint32_t)key >> ts) : 0u;
        const int pos = ordered_append(s_cur, bin, bin_bits, active);
        if (active) {
            tmp_idx[pos] = (int)i;
            tmp_depth[pos] = depth;
            tmp_fine[pos] = (uint16_t)((uint32_t)key & fmask);
        }
    }
}

// ---- 4. fine ---------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void build_fine_kernel(const int* __restrict__ bin_start,
                                                            const int* __restrict__ bin_total, int nbins,
                                                            int ts, int P, int64_t n, int sort,
                                                            const int* __restrict__ tmp_idx,
                                                            const float* __restrict__ tmp_depth,
                                                            const uint16_t* __restrict__ tmp_fine,
                                                            int* __restrict__ out_idx,
                                                            float* __restrict__ out_depth,
                                                            int* __restrict__ out_hash) {
    extern __shared__ int s_mem[];
    __shared__ int s_wave[kWavesPerBlock];
    const int tile = 1 << ts;
    int* s_cnt = s_mem;
    int* s_cur = s_mem + tile;
    const int b = blockIdx.x;
    const int bs = bin_start[b];
    const int tot = bin_total[b];
    for (int f = threadIdx.x; f < tile; f += kBlock) s_cnt[f] = 0;
    __syncthreads();
    for (int i = bs + threadIdx.x; i < bs + tot; i += kBlock) atomicAdd(&s_cnt[tmp_fine[i]], 1);
    __syncthreads();
    for (int f = threadIdx.x; f < tile; f += kBlock) s_cur[f] = s_cnt[f];
    __syncthreads();
    lds_exclusive_scan(s_cur, tile, s_wave);
    const int64_t p0 = (int64_t)b << ts;
    for (int f = threadIdx.x; f < tile; f += kBlock) {
        const int64_t p = p0 + f;
        if (p < P) {
            const int c = s_cnt[f];
            out_hash[2 * p] = c ? bs + s_cur[f] : -1;
            out_hash[2 * p + 1] = c ? c : -1;
        }
    }
    __syncthreads();
    for (int r0 = bs; r0 < bs + tot; r0 += kBlock) {
        const int i = r0 + threadIdx.x;
        const bool active = i < bs + tot;
        const uint32_t f = active ? tmp_fine[i] : 0u;
        const int pos = ordered_append(s_cur, f, ts, active);
        if (active) {
            out_idx[bs + pos] = tmp_idx[i];
            out_depth[bs + pos] = tmp_depth[i];
        }
    }
    if (sort) {
        __syncthreads();
        // stable insertion sort of each pixel run by depth (same result as the
        // reference's stable bubble sort, cuda_util.cuh:291-313)
        for (int f = threadIdx.x; f < tile; f += kBlock) {
            const int c = s_cnt[f];
            if (c < 2) continue;
            const int s = bs + s_cur[f] - c;
            float* d = out_depth + s;
            int* ix = out_idx + s;
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
    }
    // off-image tail: positions [N_in, n) hold -1 like the reference's
    // torch::full(-1) outputs
    const int64_t n_in = bin_start[nbins];
    for (int64_t i = n_in + (int64_t)b * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        out_idx[i] = -1;
        out_depth[i] = -1.0f;
    }
}

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t count, int value) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---- scatter_hashtable on pre-sorted keys -------------------------------------
// Run starts and ends of the sorted key array; hash[2k+1] temporarily holds
// the run end.
__global__ __launch_bounds__(kBlock) void sh_mark_kernel(const int* __restrict__ point_idx,
                                                         const int* __restrict__ keys,
                                                         const float* __restrict__ depth, int64_t n,
                                                         int P, int* __restrict__ out_idx,
                                                         float* __restrict__ out_depth,
                                                         int* __restrict__ hash) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int key = keys[i];
    const bool valid = key >= 0 && key < P;
    out_idx[i] = valid ? point_idx[i] : -1;
    out_depth[i] = valid ? depth[i] : -1.0f;
    if (!valid) return;
    if (i == 0 || keys[i - 1] != key) hash[2 * (int64_t)key] = (int)i;
    if (i == n - 1 || keys[i + 1] != key) hash[2 * (int64_t)key + 1] = (int)(i + 1);
}

__global__ __launch_bounds__(kBlock) void sh_fix_kernel(int* __restrict__ hash, int P, int sort,
                                                        int* __restrict__ out_idx,
                                                        float* __restrict__ out_depth) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const int s = hash[2 * p];
    if (s < 0) return;
    const int e = hash[2 * p + 1];
    const int c = e - s;
    hash[2 * p + 1] = c;
    if (!sort || c < 2) return;
    float* d = out_depth + s;
    int* ix = out_idx + s;
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

// ---- scatter (per-pixel B nearest) -------------------------------------------
// One thread per pixel keeps the reference's max-replacement queue in its own
// output row (rasterization.cu:44-83).
__global__ __launch_bounds__(kBlock) void scatter_kernel(const int* __restrict__ point_idx,
                                                         const int* __restrict__ keys,
                                                         const float* __restrict__ depth, int64_t n,
                                                         int P, int B, float* __restrict__ out_depth,
                                                         int* __restrict__ out_idx) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    float* qd = out_depth + p * B;
    int* qi = out_idx + p * B;
    for (int j = 0; j < B; ++j) {
        qd[j] = -1.0f;
        qi[j] = -1;
    }
    // leftmost occurrence of key p (binary_search, cuda_util.cuh:558-579)
    int64_t lo = 0, hi = n - 1, first = -1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) / 2;
        const int km = keys[mid];
        if (km == (int)p) {
            first = mid;
            hi = mid - 1;
        } else if (km < (int)p) {
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    if (first < 0) return;
    int qs = 0, qmax = -1;
    float qmaxz = -1000.0f;
    for (int64_t i = first; i < n && keys[i] == (int)p; ++i) {
        const float d = depth[i];
        if (d < 0) continue;
        const int id = point_idx[i];
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
    for (int a = 0; a < qs; ++a) {
        bool swapped = false;
        for (int j = 0; j < qs - a - 1; ++j) {
            if (qd[j] > qd[j + 1]) {
                const float td = qd[j];
                qd[j] = qd[j + 1];
                qd[j + 1] = td;
                const int ti = qi[j];
                qi[j] = qi[j + 1];
                qi[j + 1] = ti;
                swapped = true;
            }
        }
        if (!swapped) break;
    }
}

// Launch geometry of the fused build for (n, H, W).
struct BuildPlan {
    int64_t P;
    int ts;         // log2 pixels per coarse bin
    int nbins;
    int bin_bits;   // bits of a coarse bin id
    int64_t chunk;  // points per block in count / bin
    int nblk;
    int nseg;
    // workspace layout (byte offsets)
    size_t off_hist, off_part, off_start, off_total, off_tidx, off_tdep, off_tfine, bytes;
};

BuildPlan plan_build(int64_t n, int H, int W) {
    BuildPlan bp{};
    bp.P = (int64_t)H * W;
    bp.ts = std::max(kMinTileShift, ceil_log2(bp.P) - kMaxBinBits);
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
    bp.off_tidx = take((size_t)n * 4);
    bp.off_tdep = take((size_t)n * 4);
    bp.off_tfine = take((size_t)n * 2);
    bp.bytes = off;
    return bp;
}

int fill_launch(int* p, int64_t count, int value, hipStream_t st) {
    if (count <= 0) return EPCQ_OK;
    const int64_t blocks = std::min<int64_t>(ceil_div(count, kBlock), 65536);
    fill_i32_kernel<<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(p, count, value);
    return check_launch("fill");
}

}  // namespace
}  // namespace epcq

using namespace epcq;

extern "C" size_t epcq_build_hashtable_workspace_size(int64_t n, int32_t height, int32_t width) {
    if (n < 0 || height <= 0 || width <= 0) return 0;
    return plan_build(n, height, width).bytes;
}

extern "C" int epcq_build_hashtable(const float* xyz, int64_t n, const float* w2c, const float* K,
                                    int32_t height, int32_t width, int32_t sort, int32_t* out_idx,
                                    float* out_depth, int32_t* out_hash, void* workspace,
                                    size_t workspace_bytes, void* stream) {
    if (height <= 0 || width <= 0) return fail(EPCQ_ERR_INVALID, "build_hashtable: image size must be positive");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(EPCQ_ERR_INVALID, "build_hashtable: point count out of range");
    if ((int64_t)height * width >= (int64_t(1) << 31) / 2)
        return fail(EPCQ_ERR_INVALID, "build_hashtable: image too large");
    const BuildPlan bp = plan_build(n, height, width);
    if (workspace_bytes < bp.bytes || (bp.bytes && !workspace))
        return fail(EPCQ_ERR_WORKSPACE, "build_hashtable: workspace too small");
    hipStream_t st = as_stream(stream);
    if (n == 0) return fill_launch(out_hash, 2 * bp.P, -1, st);
    char* ws = static_cast<char*>(workspace);
    int* hist = reinterpret_cast<int*>(ws + bp.off_hist);
    int* part = reinterpret_cast<int*>(ws + bp.off_part);
    int* bstart = reinterpret_cast<int*>(ws + bp.off_start);
    int* btotal = reinterpret_cast<int*>(ws + bp.off_total);
    int* tidx = reinterpret_cast<int*>(ws + bp.off_tidx);
    float* tdep = reinterpret_cast<float*>(ws + bp.off_tdep);
    uint16_t* tfine = reinterpret_cast<uint16_t*>(ws + bp.off_tfine);
    const size_t lds_bins = (size_t)bp.nbins * 4;

    build_count_kernel<<<dim3(bp.nblk), dim3(kBlock), lds_bins, st>>>(xyz, n, w2c, K, height, width, bp.ts,
                                                                      bp.nbins, bp.chunk, hist);
    if (int e = check_launch("build_hashtable/count")) return e;
    const dim3 cgrid((unsigned)ceil_div(bp.nbins, kBlock), (unsigned)bp.nseg);
    col_partial_kernel<<<cgrid, dim3(kBlock), 0, st>>>(hist, bp.nblk, bp.nbins, part);
    if (int e = check_launch("build_hashtable/col_partial")) return e;
    bin_scan_kernel<<<dim3(1), dim3(kBlock), 0, st>>>(part, bp.nseg, bp.nbins, bstart, btotal);
    if (int e = check_launch("build_hashtable/bin_scan")) return e;
    col_apply_kernel<<<cgrid, dim3(kBlock), 0, st>>>(hist, bp.nblk, bp.nbins, part);
    if (int e = check_launch("build_hashtable/col_apply")) return e;
    build_bin_kernel<<<dim3(bp.nblk), dim3(kBlock), lds_bins, st>>>(xyz, n, w2c, K, height, width, bp.ts,
                                                                    bp.nbins, bp.bin_bits, bp.chunk, hist,
                                                                    tidx, tdep, tfine);
    if (int e = check_launch("build_hashtable/bin")) return e;
    const size_t lds_fine = (size_t)2 * (size_t(1) << bp.ts) * 4;
    build_fine_kernel<<<dim3(bp.nbins), dim3(kBlock), lds_fine, st>>>(bstart, btotal, bp.nbins, bp.ts,
                                                                      (int)bp.P, n, sort, tidx, tdep, tfine,
                                                                      out_idx, out_depth, out_hash);
    return check_launch("build_hashtable/fine");
}

extern "C" int epcq_scatter_hashtable(const int32_t* point_idx, const int32_t* keys, const float* depth,
                                      int64_t n, int32_t height, int32_t width, int32_t sort,
                                      int32_t* out_idx, float* out_depth, int32_t* out_hash, void* stream) {
    if (height <= 0 || width <= 0) return fail(EPCQ_ERR_INVALID, "scatter_hashtable: image size must be positive");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(EPCQ_ERR_INVALID, "scatter_hashtable: point count out of range");
    const int64_t P = (int64_t)height * width;
    if (P >= (int64_t(1) << 31) / 2) return fail(EPCQ_ERR_INVALID, "scatter_hashtable: image too large");
    hipStream_t st = as_stream(stream);
    if (int e = fill_launch(out_hash, 2 * P, -1, st)) return e;
    if (n == 0) return EPCQ_OK;
    sh_mark_kernel<<<dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0, st>>>(point_idx, keys, depth, n, (int)P,
                                                                                out_idx, out_depth, out_hash);
    if (int e = check_launch("scatter_hashtable/mark")) return e;
    sh_fix_kernel<<<dim3((unsigned)ceil_div(P, kBlock)), dim3(kBlock), 0, st>>>(out_hash, (int)P, sort, out_idx,
                                                                               out_depth);
    return check_launch("scatter_hashtable/fix");
}

extern "C" int epcq_scatter(const int32_t* point_idx, const int32_t* keys, const float* depth, int64_t n,
                            int32_t height, int32_t width, int32_t bins, float* out_depth, int32_t* out_idx,
                            void* stream) {
    if (height <= 0 || width <= 0) return fail(EPCQ_ERR_INVALID, "scatter: image size must be positive");
    if (bins <= 0) return fail(EPCQ_ERR_INVALID, "scatter: bins must be positive");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(EPCQ_ERR_INVALID, "scatter: point count out of range");
    const int64_t P = (int64_t)height * width;
    if (P * bins >= (int64_t(1) << 40)) return fail(EPCQ_ERR_INVALID, "scatter: output too large");
    hipStream_t st = as_stream(stream);
    scatter_kernel<<<dim3((unsigned)ceil_div(P, kBlock)), dim3(kBlock), 0, st>>>(point_idx, keys, depth, n, (int)P,
                                                                                bins, out_depth, out_idx);
    return check_launch("scatter");
}
