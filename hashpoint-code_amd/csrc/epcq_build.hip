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
//   2. scan     small count matrices: one block per bin scans its column of
//               the block histograms (each (block, bin) gets its cursor
//               within the bin, each bin its total), then one block scans the
//               totals into the bins' starts in the output, which the bin
//               pass adds.  Large ones: segment sums, bin totals, bin starts,
//               then absolute cursors (four small kernels over rows of bins);
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
#include <mutex>
#include <string>

#include "epcq_hashbuild.h"

namespace epcq {
namespace {

constexpr int kBlock = kBuildBlock;

__global__ __launch_bounds__(kBlock) void build_count_kernel(BuildView bv) {
    extern __shared__ int s_hist[];
    count_body(bv, blockIdx.x, s_hist);
}

__global__ __launch_bounds__(kBlock) void col_scan_kernel(BuildView bv) {
    __shared__ int s_wave[kBuildWaves];
    col_scan_body(bv, blockIdx.x, s_wave);
}

__global__ __launch_bounds__(kBlock) void col_partial_kernel(BuildView bv) {
    col_partial_body(bv, blockIdx.x * kBlock + threadIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(kBlock) void bin_total_kernel(BuildView bv) {
    bin_total_body(bv, blockIdx.x * kBlock + threadIdx.x);
}

__global__ __launch_bounds__(kBlock) void col_apply_kernel(BuildView bv) {
    col_apply_body(bv, blockIdx.x * kBlock + threadIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(kBlock) void bin_scan_kernel(BuildView bv) {
    __shared__ int s_wave[kBuildWaves];
    __shared__ int s_carry;
    bin_scan_body(bv, s_wave, &s_carry);
}

__global__ __launch_bounds__(kBlock) void build_bin_kernel(BuildView bv) {
    extern __shared__ int s_cur[];
    bin_body(bv, blockIdx.x, s_cur);
}

__global__ __launch_bounds__(kBlock) void build_fine_kernel(BuildView bv) {
    extern __shared__ int s_mem[];
    __shared__ int s_wave[kBuildWaves];
    fine_body(bv, blockIdx.x, s_mem, s_wave);
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

// Run ends become counts; with `sort` each run is ordered by depth, short
// runs by one thread, long ones by the block (scratch (sd, sx)[n]).
__global__ __launch_bounds__(kBlock) void sh_fix_kernel(int* __restrict__ hash, int P, int sort,
                                                        int* __restrict__ out_idx, float* __restrict__ out_depth,
                                                        float* __restrict__ sd, int* __restrict__ sx) {
    __shared__ int s_cnt[kBlock];
    __shared__ int s_end[kBlock];
    __shared__ int s_wave[kBuildWaves];
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int c = 0, e = 0;
    if (p < P) {
        const int s = hash[2 * p];
        if (s >= 0) {
            e = hash[2 * p + 1];
            c = e - s;
            hash[2 * p + 1] = c;
        }
    }
    if (!sort) return;
    s_cnt[threadIdx.x] = c;
    s_end[threadIdx.x] = e;
    __syncthreads();
    sort_pixel_runs(kBlock, s_cnt, s_end, 0, out_depth, out_idx, sd, sx, s_wave);
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

namespace epcq {
namespace {

int build_hashtable_impl(const float* xyz, int64_t n, const float* w2c, const float* K, int32_t height, int32_t width,
                         int32_t sort, int32_t* out_idx, float* out_depth, int32_t* out_hash, void* workspace,
                         size_t workspace_bytes, void* stream, const char* who) {
    const std::string w(who);
    if (height <= 0 || width <= 0) return fail(EPCQ_ERR_INVALID, w + ": image size must be positive");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(EPCQ_ERR_INVALID, w + ": point count out of range");
    if ((int64_t)height * width >= (int64_t(1) << 31) / 2) return fail(EPCQ_ERR_INVALID, w + ": image too large");
    const BuildPlan bp = plan_build(n, height, width);
    if (workspace_bytes < bp.bytes || (bp.bytes && !workspace))
        return fail(EPCQ_ERR_WORKSPACE, w + ": workspace too small");
    hipStream_t st = as_stream(stream);
    if (n == 0) return fill_launch(out_hash, 2 * bp.P, -1, st);
    BuildView bv = make_build_view(bp, xyz, n, w2c, K, height, width, sort, workspace, out_idx, out_depth,
                                   out_hash);
    // One block per bin scanning its column is one launch in place of three and
    // wins on small block x bin count matrices (1M points at 800x800, 245 x
    // 1250: 6 against 31 us); its strided column reads lose on large ones (20M
    // points at 1920x1080, 1628 x 8100: 424 against 47 us), which take the
    // segmented scan.
    const bool one_pass = (int64_t)bp.nblk * bp.nbins <= kColScanEntries;
    bv.rel_cursors = one_pass ? 1 : 0;  // the one-pass scan leaves cursors relative to the bin's start
    const size_t lds_bins = (size_t)bp.nbins * 4;
    build_count_kernel<<<dim3(bp.nblk), dim3(kBlock), lds_bins, st>>>(bv);
    if (int e = check_launch("build_hashtable/count")) return e;
    if (one_pass) {
        col_scan_kernel<<<dim3(bp.nbins), dim3(kBlock), 0, st>>>(bv);
        if (int e = check_launch("build_hashtable/col_scan")) return e;
        bin_scan_kernel<<<dim3(1), dim3(kBlock), 0, st>>>(bv);
        if (int e = check_launch("build_hashtable/bin_scan")) return e;
    } else {
        const dim3 cgrid((unsigned)ceil_div(bp.nbins, kBlock), (unsigned)bp.nseg);
        col_partial_kernel<<<cgrid, dim3(kBlock), 0, st>>>(bv);
        if (int e = check_launch("build_hashtable/col_partial")) return e;
        bin_total_kernel<<<dim3(cgrid.x), dim3(kBlock), 0, st>>>(bv);
        if (int e = check_launch("build_hashtable/bin_total")) return e;
        bin_scan_kernel<<<dim3(1), dim3(kBlock), 0, st>>>(bv);
        if (int e = check_launch("build_hashtable/bin_scan")) return e;
        col_apply_kernel<<<cgrid, dim3(kBlock), 0, st>>>(bv);
        if (int e = check_launch("build_hashtable/col_apply")) return e;
    }
    build_bin_kernel<<<dim3(bp.nblk), dim3(kBlock), lds_bins, st>>>(bv);
    if (int e = check_launch("build_hashtable/bin")) return e;
    const size_t lds_fine = (size_t)2 * (size_t(1) << bp.ts) * 4;
    build_fine_kernel<<<dim3(bp.nbins), dim3(kBlock), lds_fine, st>>>(bv);
    return check_launch("build_hashtable/fine");
}

}  // namespace
}  // namespace epcq

extern "C" int epcq_build_hashtable(const float* xyz, int64_t n, const float* w2c, const float* K,
                                    int32_t height, int32_t width, int32_t sort, int32_t* out_idx,
                                    float* out_depth, int32_t* out_hash, void* workspace,
                                    size_t workspace_bytes, void* stream) {
    return build_hashtable_impl(xyz, n, w2c, K, height, width, sort, out_idx, out_depth, out_hash, workspace,
                                workspace_bytes, stream, "build_hashtable");
}

// ---- build_hashtable without allocation --------------------------------------
// The workspace of epcq_build_hashtable_noalloc is this file-scope device
// array, one copy per GPU for the life of the process.  Host code never
// passes `scratch` itself to a kernel (on the host that name is a stand-in,
// not a device address): it passes the address hipGetSymbolAddress returns
// for the current device, looked up once per device and kept in
// g_scratch_ptr.  Every build kernel writes each workspace word before it
// reads it, so nothing needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[EPCQ_NOALLOC_SCRATCH_BYTES];

namespace epcq {
namespace {

constexpr size_t kNoallocScratchBytes = EPCQ_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up on its first use).
int scratch_address(void** out, const char* who) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return check_launch(who);
    if (dev < 0 || dev >= kMaxDevices) return fail(EPCQ_ERR_INVALID, std::string(who) + ": device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(EPCQ_ERR_HIP, std::string(who) + ": hipGetSymbolAddress: " + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return EPCQ_OK;
}

}  // namespace
}  // namespace epcq

extern "C" int epcq_build_hashtable_noalloc(const float* xyz, int64_t n, const float* w2c, const float* K,
                                            int32_t height, int32_t width, int32_t sort, int32_t* out_idx,
                                            float* out_depth, int32_t* out_hash, void* stream) {
    const char* who = "build_hashtable_noalloc";
    if (height > 0 && width > 0 && n >= 0 && n < (int64_t(1) << 31) &&
        (int64_t)height * width < (int64_t(1) << 31) / 2) {
        const size_t need = plan_build(n, height, width).bytes;
        if (need > kNoallocScratchBytes)
            return fail(EPCQ_ERR_WORKSPACE, std::string(who) + ": needs " + std::to_string(need) +
                                                " bytes of scratch, more than its " +
                                                std::to_string(kNoallocScratchBytes) + " (too many points for this image)");
    }
    void* ws = nullptr;
    if (int e = scratch_address(&ws, who)) return e;
    return build_hashtable_impl(xyz, n, w2c, K, height, width, sort, out_idx, out_depth, out_hash, ws,
                                kNoallocScratchBytes, stream, who);
}

extern "C" size_t epcq_scatter_hashtable_workspace_size(int64_t n, int32_t sort) {
    if (n <= 0 || !sort) return 0;
    return (size_t)n * (sizeof(float) + sizeof(int32_t));
}

extern "C" int epcq_scatter_hashtable(const int32_t* point_idx, const int32_t* keys, const float* depth,
                                      int64_t n, int32_t height, int32_t width, int32_t sort,
                                      int32_t* out_idx, float* out_depth, int32_t* out_hash, void* workspace,
                                      size_t workspace_bytes, void* stream) {
    if (height <= 0 || width <= 0) return fail(EPCQ_ERR_INVALID, "scatter_hashtable: image size must be positive");
    if (n < 0 || n >= (int64_t(1) << 31)) return fail(EPCQ_ERR_INVALID, "scatter_hashtable: point count out of range");
    const int64_t P = (int64_t)height * width;
    if (P >= (int64_t(1) << 31) / 2) return fail(EPCQ_ERR_INVALID, "scatter_hashtable: image too large");
    const size_t need = epcq_scatter_hashtable_workspace_size(n, sort);
    if (workspace_bytes < need || (need && !workspace))
        return fail(EPCQ_ERR_WORKSPACE, "scatter_hashtable: workspace too small");
    hipStream_t st = as_stream(stream);
    if (int e = fill_launch(out_hash, 2 * P, -1, st)) return e;
    if (n == 0) return EPCQ_OK;
    sh_mark_kernel<<<dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0, st>>>(point_idx, keys, depth, n, (int)P,
                                                                                out_idx, out_depth, out_hash);
    if (int e = check_launch("scatter_hashtable/mark")) return e;
    float* sd = need ? static_cast<float*>(workspace) : nullptr;
    int* sx = need ? reinterpret_cast<int*>(static_cast<float*>(workspace) + n) : nullptr;
    sh_fix_kernel<<<dim3((unsigned)ceil_div(P, kBlock)), dim3(kBlock), 0, st>>>(out_hash, (int)P, sort, out_idx,
                                                                               out_depth, sd, sx);
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
