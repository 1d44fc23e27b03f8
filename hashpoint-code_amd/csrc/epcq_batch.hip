// Batched multi-view search + sampling: many views of uneven resolution over
// one or more resident point clouds in one persistent pass, replayable as a
// HIP graph.
//
// Everything lives in one caller-supplied device arena:
//   header | view descriptors | item prefix tables | per-view build
//   workspace, hash table, point lists, bucket bounds and compact outputs.
// epcq_batch_set_views() lays the views out in the arena and uploads the
// descriptors (a few KB, stream-ordered); every kernel is persistent -- a
// fixed grid strides over work items whose count it reads from the header --
// so a graph captured once replays any view set that fits the arena.
//
// Pipeline (all views at once, one launch per stage):
//   reset -> count -> col_partial -> bin_total -> bin_scan -> col_apply -> bin -> fine
//   (the fused hash-table build of epcq_build.hip, per (view, chunk) / (view,
//   bin) item; bin carries 16-byte point records, and fine also writes the
//   bin's pixel-ordered point copy and its pixels' bucket cones) ->
//   sample (one (view, tile) item at a time from an atomic work counter,
//   heaviest views first; a tile whose halo exceeds the LDS reads global
//   memory; a wave's outlier rays are run wave-wide after the tile).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "epcq_batch_kernels.h"

EPCQ_EXTERN_BATCH(EPCQ_MODE_COMPAT, epcq::kCompatBuffer)
EPCQ_EXTERN_BATCH(EPCQ_MODE_KNN, 8)
EPCQ_EXTERN_BATCH(EPCQ_MODE_KNN, 16)
EPCQ_EXTERN_BATCH(EPCQ_MODE_KNN, 32)

extern "C" int32_t epcq_max_live_samples(const epcq_options* opt, float r_max);

namespace epcq {

float squared_threshold(float r);
int check_tuning(const epcq_tuning& t, const char* who);

namespace {

__global__ void batch_reset_kernel(BatchHeader* h) { h->counter = 0; }

// Calls fn(view, chunk) for this block's share of the (view, chunk) items of
// its XCD: the views the XCD was given (xlist) and their chunk prefix
// (xprefix); the launch's grid is a multiple of kXcds.
template <class Fn>
__device__ __forceinline__ void for_each_xcd_chunk(const int* __restrict__ tables, int M, Fn&& fn) {
    const int x = (int)blockIdx.x % kXcds, loc = (int)blockIdx.x / kXcds, nloc = (int)gridDim.x / kXcds;
    const int* xprefix = tables + (kStages + x) * M;
    const int* xlist = tables + (kStages + kXcds + x) * M;
    const int nv = xlist[M - 1];
    const int n_items = xprefix[nv];
    for (int item = loc; item < n_items; item += nloc) {
        const int j = find_view(xprefix, nv, item);
        fn(xlist[j], item - xprefix[j]);
    }
}

__global__ __launch_bounds__(kBuildBlock) void batch_count_kernel(const BatchView* __restrict__ views,
                                                                  const int* __restrict__ tables, int M) {
    __shared__ int s_hist[kMaxBins];
    for_each_xcd_chunk(tables, M, [&](int v, int chunk) { count_body(views[v].b, chunk, s_hist); });
}

__global__ __launch_bounds__(kBuildBlock) void batch_colpart_kernel(const BatchHeader* __restrict__ h,
                                                                    const BatchView* __restrict__ views,
                                                                    const int* __restrict__ prefix, int apply) {
    const int n_items = h->items[kCol], V = h->V;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int v = find_view(prefix, V, item);
        const BatchView& bv = views[v];
        const int local = item - prefix[v];
        const int g = local % bv.col_groups, seg = local / bv.col_groups;
        if (apply) col_apply_body(bv.b, g * kBuildBlock + threadIdx.x, seg);
        else col_partial_body(bv.b, g * kBuildBlock + threadIdx.x, seg);
    }
}

__global__ __launch_bounds__(kBuildBlock) void batch_bintotal_kernel(const BatchHeader* __restrict__ h,
                                                                     const BatchView* __restrict__ views,
                                                                     const int* __restrict__ prefix) {
    const int n_items = h->items[kTot], V = h->V;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int v = find_view(prefix, V, item);
        bin_total_body(views[v].b, (item - prefix[v]) * kBuildBlock + threadIdx.x);
    }
}

__global__ __launch_bounds__(kBuildBlock) void batch_binscan_kernel(const BatchHeader* __restrict__ h,
                                                                    const BatchView* __restrict__ views) {
    __shared__ int s_wave[kBuildWaves];
    __shared__ int s_carry;
    const int V = h->V;
    for (int v = blockIdx.x; v < V; v += gridDim.x) bin_scan_body(views[v].b, s_wave, &s_carry);
}

__global__ __launch_bounds__(kBuildBlock) void batch_bin_kernel(const BatchView* __restrict__ views,
                                                                const int* __restrict__ tables, int M) {
    __shared__ int s_cur[kMaxBins];
    for_each_xcd_chunk(tables, M, [&](int v, int chunk) { bin_body(views[v].b, chunk, s_cur); });
}

// Cone of pixel p (as bucket_cone_body) computed by a group of G = 2^gs
// adjacent lanes, each taking every G-th point of the bucket from s_pts (list
// position i at s_pts[i - base]); for bins of fewer pixels than threads.  The
// orientation reference is the bucket's first point, as in the one-thread
// form; sums meet in a butterfly (every lane ends with the same bits), minima
// and maxima are exact in any order.  Every lane of the wave calls this
// (shuffles); `live` is false for lanes without a pixel.
__device__ __forceinline__ void bucket_cone_group(const int2* __restrict__ hash, const float4* s_pts, int base,
                                                  const float* __restrict__ radius, float r0,
                                                  const float* __restrict__ c2w, int64_t p, bool live, int gs,
                                                  float4* __restrict__ cone_a, float2* __restrict__ cone_b) {
    const int G = 1 << gs, sub = __lane_id() & (G - 1);
    const int2 e = live ? hash[p] : make_int2(0, 0);
    const float never = __builtin_inff();
    const float o0 = c2w[3], o1 = c2w[7], o2 = c2w[11];
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    if (e.y > 0) {
        const float4 q = s_pts[e.x - base];
        const float w0 = q.x - o0, w1 = q.y - o1, w2 = q.z - o2;
        const float x = sqnorm3(w0, w1, w2);
        if (x > 1e-30f && x < never) {
            const float inv = __builtin_amdgcn_rsqf(x);
            f0 = w0 * inv;
            f1 = w1 * inv;
            f2 = w2 * inv;
        }
    }
    float rho = never, rb = radius ? __builtin_nanf("") : r0;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    int bad = 0;
    for (int j = sub; j < e.y; j += G) {
        const float4 q = s_pts[e.x + j - base];
        if (radius) rb = fmaxf(rb, radius[__float_as_int(q.w)]);
        const float w0 = q.x - o0, w1 = q.y - o1, w2 = q.z - o2;
        const float x = sqnorm3(w0, w1, w2);
        if (!(x > 1e-30f && x < never)) {
            bad = 1;
            continue;
        }
        const float inv = __builtin_amdgcn_rsqf(x);
        const float n = x * inv;
        const float u0 = w0 * inv, u1 = w1 * inv, u2 = w2 * inv;
        const float sg = dot3(u0, u1, u2, f0, f1, f2) < 0.0f ? -1.0f : 1.0f;
        s0 = s0 + sg * u0;
        s1 = s1 + sg * u1;
        s2 = s2 + sg * u2;
        rho = fminf(rho, n);
    }
    for (int off = 1; off < G; off <<= 1) {
        s0 = s0 + __shfl_xor(s0, off);
        s1 = s1 + __shfl_xor(s1, off);
        s2 = s2 + __shfl_xor(s2, off);
        rho = fminf(rho, __shfl_xor(rho, off));
        rb = fmaxf(rb, __shfl_xor(rb, off));
        bad |= __shfl_xor(bad, off);
    }
    const float sn = sqrtf(sqnorm3(s0, s1, s2));
    const bool never_culls = bad || !(sn > 0.0f) || !(rb <= never);
    const float a0 = never_culls ? 1.0f : s0 / sn, a1 = never_culls ? 0.0f : s1 / sn,
                a2 = never_culls ? 0.0f : s2 / sn;
    float cmin = 1.0f, s2max = 0.0f;
    if (!never_culls) {
        for (int j = sub; j < e.y; j += G) {
            const float4 q = s_pts[e.x + j - base];
            const float w0 = q.x - o0, w1 = q.y - o1, w2 = q.z - o2;
            const float inv = __builtin_amdgcn_rsqf(sqnorm3(w0, w1, w2));
            const float u0 = w0 * inv, u1 = w1 * inv, u2 = w2 * inv;
            cmin = fminf(cmin, fabsf(dot3(u0, u1, u2, a0, a1, a2)));
            const float x0 = u1 * a2 - u2 * a1, x1 = u2 * a0 - u0 * a2, x2 = u0 * a1 - u1 * a0;
            s2max = fmaxf(s2max, sqnorm3(x0, x1, x2));
        }
    }
    for (int off = 1; off < G; off <<= 1) {
        cmin = fminf(cmin, __shfl_xor(cmin, off));
        s2max = fmaxf(s2max, __shfl_xor(s2max, off));
    }
    if (e.y <= 0 || sub != 0) return;
    if (never_culls) {
        cone_a[p] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
        cone_b[p] = make_float2(0.0f, never);
        return;
    }
    const float smax = sqrtf(s2max);
    const float rl = rho * (1.0f - 2e-6f);
    cone_a[p] = make_float4(a0, a1, a2, rl * (cmin - 2e-6f));
    cone_b[p] = make_float2(rl * (smax + 2e-6f), rb * (1.0f + 1e-6f) + 4e-6f * rho);
}

// fine, then for the same bin: the pixel-ordered copy of its points and the
// cones of its pixels (the samplers' bucket bounds), while they are in cache.
__global__ __launch_bounds__(kBuildBlock) void batch_fine_kernel(const BatchHeader* __restrict__ h,
                                                                 const BatchView* __restrict__ views,
                                                                 const int* __restrict__ prefix,
                                                                 const float* __restrict__ radius, float r0) {
    // placement's pixel counts and cursors (2 * kMaxFineTile ints), then the
    // bin's pixel-ordered points for the cones when they fit
    __shared__ float4 s_buf[kMaxFineTile / 2];
    __shared__ int s_wave[kBuildWaves];
    int* s_mem = reinterpret_cast<int*>(s_buf);
    const int n_items = h->items[kFine], V = h->V;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int v = find_view(prefix, V, item);
        const BatchView& bv = views[v];
        const int b = item - prefix[v];
        fine_body(bv.b, b, s_mem, s_wave);
        const int bs = bv.b.bstart[b], tot = bv.b.btotal[b];
        // fine_body wrote the pixel-ordered point copy from the binned records;
        // the cones read it from LDS (one coalesced copy instead of each
        // thread's serial walks through L2) when the bin fits
        const float4* pts = bv.view.pts;
        const bool staged = tot <= kMaxFineTile / 2;
        if (staged)
            for (int i = threadIdx.x; i < tot; i += kBuildBlock) s_buf[i] = pts[bs + i];
        if (radius) {
            float* prad = const_cast<float*>(bv.view.prad);
            for (int i = bs + (int)threadIdx.x; i < bs + tot; i += kBuildBlock)
                prad[i] = point_threshold(radius[bv.b.out_idx[i]]);
        }
        __syncthreads();
        const int64_t p0 = (int64_t)b << bv.b.ts;
        float4* cone_a = const_cast<float4*>(bv.view.cone_a);
        float2* cone_b = const_cast<float2*>(bv.view.cone_b);
        // bins of fewer pixels than threads give each pixel 2^gs threads (at
        // most 8); gs is uniform over the block, so whole waves take the path
        int gs = 0;
        while (gs < 3 && (kBuildBlock >> (bv.b.ts + gs + 1)) > 0) ++gs;
        if (staged && gs > 0) {
            const int f = (int)threadIdx.x >> gs;
            const int64_t p = p0 + f;
            bucket_cone_group(bv.view.hash, s_buf, bs, radius, r0, bv.view.c2w, p,
                              f < (1 << bv.b.ts) && p < bv.b.P, gs, cone_a, cone_b);
        } else for (int f = threadIdx.x; f < (1 << bv.b.ts); f += kBuildBlock) {
            const int64_t p = p0 + f;
            if (p >= bv.b.P) continue;
            if (staged)
                bucket_cone_body(bv.view.hash, nullptr, nullptr, s_buf, radius, r0, bv.view.c2w, p, cone_a, cone_b,
                                 nullptr, nullptr, bs);
            else
                bucket_cone_body(bv.view.hash, nullptr, nullptr, pts, radius, r0, bv.view.c2w, p, cone_a, cone_b,
                                 nullptr, nullptr);
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace epcq

using namespace epcq;

// Host object behind the opaque epcq_batch handle.
struct epcq_batch {
    epcq_options opt;
    int mode, s_max, bound;
    TilePlan tp;
    int heavy;
    int max_views;
    char* arena;
    size_t arena_bytes;
    size_t data_off;  // first byte after header, descriptors and prefixes
    BatchHeader* d_header;
    BatchView* d_views;
    int* d_prefix;  // [kStages][max_views + 1]
    const float* radius;
    float r_max;
    // current views
    int V;
    int64_t total_rays;
    std::vector<BatchView> views;
    std::vector<int> order;    // caller view of each descriptor slot
    std::vector<int> slot_of;  // descriptor slot of each caller view
    std::vector<int> prefix;  // host copy of the item prefixes
    BatchHeader header;
    int grid_build, grid_sample;
    hipGraph_t graph;
    hipGraphExec_t exec;
    // compact outputs (whole batch)
    SampleOut out;
};

namespace {

size_t header_bytes(int max_views) {
    size_t off = align_up(sizeof(BatchHeader), 256);
    off = align_up(off + (size_t)max_views * sizeof(BatchView), 256);
    off = align_up(off + (size_t)kTableRows * (max_views + 1) * sizeof(int), 256);
    return off;
}

struct ViewLayout {
    BuildPlan bp;
    int64_t P;
    size_t ws, idx, hash, ca, cb, pts, prad;  // byte offsets relative to the view's block
    size_t bytes;
};

ViewLayout layout_view(const epcq_view_desc& d) {
    ViewLayout L{};
    L.bp = plan_build(d.n, d.height, d.width, true, kBatchBinPoints);
    L.P = (int64_t)d.width * d.height;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes, 256);
        return at;
    };
    L.ws = take(L.bp.bytes);
    L.idx = take((size_t)d.n * 4);
    L.hash = take((size_t)L.P * 8);
    L.ca = take((size_t)L.P * 16);
    L.cb = take((size_t)L.P * 8);
    L.pts = take((size_t)d.n * 16);
    L.prad = take((size_t)d.n * 4);  // pixel-ordered radii (used with per-point radius)
    L.bytes = off;
    return L;
}

// Compact outputs of `rays` rays.
size_t output_bytes(int64_t rays, int s_max, int K, int mode) {
    size_t per = 4 + (size_t)s_max * (4 + 4 * K);
    if (mode == EPCQ_MODE_KNN) per += (size_t)s_max * 8 * K;
    return align_up((size_t)rays * per + 8 * 256, 256);  // + the alignment of each array
}

int check_views(const epcq_view_desc* views, int n_views) {
    if (n_views < 1 || !views) return fail(EPCQ_ERR_INVALID, "batch: at least one view");
    for (int i = 0; i < n_views; ++i) {
        const epcq_view_desc& d = views[i];
        if (!d.xyz && d.n > 0) return fail(EPCQ_ERR_INVALID, "batch: view without points");
        if (d.width <= 0 || d.height <= 0) return fail(EPCQ_ERR_INVALID, "batch: image size must be positive");
        if ((int64_t)d.width * d.height >= (int64_t(1) << 30)) return fail(EPCQ_ERR_INVALID, "batch: image too large");
        if (d.n < 0 || d.n >= (int64_t(1) << 31)) return fail(EPCQ_ERR_INVALID, "batch: point count out of range");
        if ((1 << plan_build(d.n, d.height, d.width, true, kBatchBinPoints).ts) > kMaxFineTile)
            return fail(EPCQ_ERR_INVALID, "batch: image too large for the batched build");
    }
    return EPCQ_OK;
}

template <int MODE, int KMAX, int SMAX>
void enqueue_sample(epcq_batch* b, const SampleOpts& o, hipStream_t st) {
    const int* pt = b->d_prefix + kTile * (b->max_views + 1);  // tiles, densest views first
    launch_batch_sample<MODE, KMAX, SMAX>(b->grid_sample, b->tp, b->d_header, b->d_views, pt, o, st);
}

template <int MODE, int KMAX>
void enqueue_sample_smax(epcq_batch* b, const SampleOpts& o, hipStream_t st) {
    if (b->bound <= 2) enqueue_sample<MODE, KMAX, 2>(b, o, st);
    else enqueue_sample<MODE, KMAX, kRegSamples>(b, o, st);
}

// Enqueues the pass; ev (optional, 4 events) brackets the stages
// projection and binning | per-pixel placement with the bucket cones |
// sampling.
int enqueue_all(epcq_batch* b, hipStream_t st, hipEvent_t* ev = nullptr) {
    const int M = b->max_views + 1;
    BatchHeader* h = b->d_header;
    const BatchView* dv = b->d_views;
    int* pf = b->d_prefix;
    const dim3 blk(kBuildBlock), gb(b->grid_build);
    if (ev) (void)hipEventRecord(ev[0], st);
    batch_reset_kernel<<<1, 1, 0, st>>>(h);
    batch_count_kernel<<<gb, blk, 0, st>>>(dv, pf, M);
    batch_colpart_kernel<<<gb, blk, 0, st>>>(h, dv, pf + kCol * M, 0);
    batch_bintotal_kernel<<<gb, blk, 0, st>>>(h, dv, pf + kTot * M);
    batch_binscan_kernel<<<gb, blk, 0, st>>>(h, dv);
    batch_colpart_kernel<<<gb, blk, 0, st>>>(h, dv, pf + kCol * M, 1);
    batch_bin_kernel<<<gb, blk, 0, st>>>(dv, pf, M);
    if (ev) (void)hipEventRecord(ev[1], st);
    batch_fine_kernel<<<gb, blk, 0, st>>>(h, dv, pf + kFine * M, b->radius, b->opt.radius_threshold);
    if (ev) (void)hipEventRecord(ev[2], st);
    SampleOpts o{};
    o.k = b->opt.kernel_size;
    o.ko = (b->opt.kernel_size + 1) / 2;
    o.kdiv = (65536 + b->opt.kernel_size - 1) / b->opt.kernel_size;
    o.L = o.k * o.k;
    o.r0 = b->opt.radius_threshold;
    o.r2lim = squared_threshold(b->opt.radius_threshold);
    o.radius = b->radius;
    o.t_intvl = b->opt.t_intvl;
    o.alpha_p = b->opt.sdf2weight_gaussian_alpha;
    o.gamma = b->opt.sdf2weight_gaussian_gamma;
    o.S = b->opt.num_sample_points_per_ray;
    o.Kc = b->opt.num_point_cloud_per_sp;
    o.xyz = nullptr;  // each view's cloud comes from its descriptor
    o.cull = 1;
    o.heavy = b->heavy;
    if (b->mode == EPCQ_MODE_COMPAT) enqueue_sample_smax<EPCQ_MODE_COMPAT, kCompatBuffer>(b, o, st);
    else if (o.Kc <= 8) enqueue_sample_smax<EPCQ_MODE_KNN, 8>(b, o, st);
    else if (o.Kc <= 16) enqueue_sample_smax<EPCQ_MODE_KNN, 16>(b, o, st);
    else enqueue_sample_smax<EPCQ_MODE_KNN, 32>(b, o, st);
    if (ev) (void)hipEventRecord(ev[3], st);
    return check_launch("batch");
}

int device_grid(int per_cu) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cus * per_cu;
}

}  // namespace

extern "C" size_t epcq_batch_arena_bytes(const epcq_view_desc* views, int32_t n_views, int32_t max_views,
                                         const epcq_options* opt, int32_t mode, int32_t s_max) {
    if (!views || n_views < 1 || !opt || max_views < n_views) return 0;
    size_t total = header_bytes(max_views);
    int64_t rays = 0;
    for (int i = 0; i < n_views; ++i) {
        total += layout_view(views[i]).bytes;
        rays += (int64_t)views[i].width * views[i].height;
    }
    total += output_bytes(rays, s_max, opt->num_point_cloud_per_sp, mode);
    return total;
}

extern "C" size_t epcq_batch_header_bytes(int32_t max_views) {
    return max_views < 1 ? 0 : header_bytes(max_views);
}

extern "C" size_t epcq_batch_view_bytes(const epcq_view_desc* view, const epcq_options* opt, int32_t mode,
                                        int32_t s_max) {
    if (!view || !opt) return 0;
    return layout_view(*view).bytes +
           output_bytes((int64_t)view->width * view->height, s_max, opt->num_point_cloud_per_sp, mode);
}

extern "C" int epcq_batch_create(const epcq_options* opt, int32_t mode, int32_t s_max, int32_t max_views,
                                 const float* radius, float r_max, const epcq_tuning* tuning, void* arena,
                                 size_t arena_bytes, epcq_batch** out) {
    if (!opt || !out) return fail(EPCQ_ERR_INVALID, "batch_create: null argument");
    if (mode != EPCQ_MODE_COMPAT && mode != EPCQ_MODE_KNN) return fail(EPCQ_ERR_INVALID, "batch_create: unknown mode");
    if (opt->kernel_size < 1 || opt->kernel_size > 64) return fail(EPCQ_ERR_INVALID, "batch_create: bad kernel_size");
    if (opt->num_point_cloud_per_sp < 1 || opt->num_point_cloud_per_sp > 32)
        return fail(EPCQ_ERR_INVALID, "batch_create: num_point_cloud_per_sp must be in [1, 32]");
    if (s_max < 1 || max_views < 1) return fail(EPCQ_ERR_INVALID, "batch_create: s_max and max_views must be >= 1");
    if (!arena || arena_bytes < header_bytes(max_views)) return fail(EPCQ_ERR_WORKSPACE, "batch_create: arena too small");
    const float rm = radius ? r_max : opt->radius_threshold;
    int bound = std::min<int>(epcq_max_live_samples(opt, rm), s_max);
    if (radius) bound = s_max;
    if (tuning)
        if (int e = check_tuning(*tuning, "batch_create")) return e;
    epcq_batch* b = new (std::nothrow) epcq_batch();
    if (!b) return fail(EPCQ_ERR_INVALID, "batch_create: out of host memory");
    b->opt = *opt;
    b->mode = mode;
    b->s_max = s_max;
    b->bound = bound;
    epcq_tuning tu{32, 8, 1024, 0, kHeavyCandidates, kListWaveRays};
    if (tuning) tu = *tuning;
    else epcq_get_tuning(&tu);
    b->tp = plan_tiles(opt->kernel_size, tu.tile_w, tu.tile_h, tu.tile_cap, tu.flags, radius != nullptr);
    b->tp.xcd = 0;
    b->heavy = std::max(tu.heavy, 0);
    b->max_views = max_views;
    b->arena = static_cast<char*>(arena);
    b->arena_bytes = arena_bytes;
    b->d_header = reinterpret_cast<BatchHeader*>(b->arena);
    b->d_views = reinterpret_cast<BatchView*>(b->arena + align_up(sizeof(BatchHeader), 256));
    b->d_prefix = reinterpret_cast<int*>(reinterpret_cast<char*>(b->d_views) +
                                         align_up((size_t)max_views * sizeof(BatchView), 256));
    b->data_off = header_bytes(max_views);
    b->radius = radius;
    b->r_max = rm;
    b->V = 0;
    b->total_rays = 0;
    b->grid_build = device_grid(8);
    b->grid_sample = device_grid(4);
    b->graph = nullptr;
    b->exec = nullptr;
    *out = b;
    return EPCQ_OK;
}

extern "C" int epcq_batch_set_views(epcq_batch* b, const epcq_view_desc* views, int32_t n_views, void* stream) {
    if (!b) return fail(EPCQ_ERR_INVALID, "batch_set_views: null batch");
    if (int e = check_views(views, n_views)) return e;
    if (n_views > b->max_views) return fail(EPCQ_ERR_INVALID, "batch_set_views: more views than max_views");
    if (b->radius)
        for (int i = 1; i < n_views; ++i)
            if (views[i].xyz != views[0].xyz)
                return fail(EPCQ_ERR_INVALID, "batch_set_views: per-point radii need one shared point cloud");
    const int V = n_views;
    std::vector<ViewLayout> L(V);
    size_t off = b->data_off;
    int64_t rays = 0;
    for (int i = 0; i < V; ++i) {
        L[i] = layout_view(views[i]);
        rays += L[i].P;
    }
    const int K = b->opt.num_point_cloud_per_sp;
    std::vector<size_t> base(V);
    for (int i = 0; i < V; ++i) {
        base[i] = off;
        off += L[i].bytes;
    }
    // compact outputs of the whole batch
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes, 256);
        return at;
    };
    const size_t o_ns = take((size_t)rays * 4);
    const size_t o_t = take((size_t)rays * b->s_max * 4);
    const size_t o_idx = take((size_t)rays * b->s_max * K * 4);
    size_t o_nbt = 0, o_nbd = 0;
    if (b->mode == EPCQ_MODE_KNN) {
        o_nbt = take((size_t)rays * b->s_max * K * 4);
        o_nbd = take((size_t)rays * b->s_max * K * 4);
    }
    if (off > b->arena_bytes) return fail(EPCQ_ERR_WORKSPACE, "batch_set_views: arena too small for these views");
    char* A = b->arena;
    b->out = SampleOut{};
    b->out.s_max = b->s_max;
    b->out.n_samples = reinterpret_cast<int*>(A + o_ns);
    b->out.sample_t = reinterpret_cast<float*>(A + o_t);
    b->out.nb_idx = reinterpret_cast<int*>(A + o_idx);
    b->out.nb_t = b->mode == EPCQ_MODE_KNN ? reinterpret_cast<float*>(A + o_nbt) : nullptr;
    b->out.nb_dist = b->mode == EPCQ_MODE_KNN ? reinterpret_cast<float*>(A + o_nbd) : nullptr;

    // Descriptors in priority order (denser views -- more points per pixel,
    // heavier tiles -- first, so the work counter hands them out early);
    // outputs in the caller's order.
    b->order.resize(V);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return (double)views[x].n / (double)L[x].P > (double)views[y].n / (double)L[y].P;
    });
    b->slot_of.assign(V, 0);
    for (int r = 0; r < V; ++r) b->slot_of[b->order[r]] = r;
    std::vector<int64_t> ray_off(V, 0);
    for (int i = 1; i < V; ++i) ray_off[i] = ray_off[i - 1] + L[i - 1].P;
    b->views.assign(V, BatchView{});
    for (int r = 0; r < V; ++r) {
        const int i = b->order[r];
        const epcq_view_desc& d = views[i];
        BatchView& bv = b->views[r];
        std::memcpy(bv.c2w, d.c2w, sizeof(bv.c2w));
        std::memcpy(bv.w2c, d.w2c, sizeof(bv.w2c));
        std::memcpy(bv.K, d.K, sizeof(bv.K));
        const char* dptr = reinterpret_cast<const char*>(b->d_views + r);
        const float* d_w2c = reinterpret_cast<const float*>(dptr + offsetof(BatchView, w2c));
        const float* d_K = reinterpret_cast<const float*>(dptr + offsetof(BatchView, K));
        const float* d_c2w = reinterpret_cast<const float*>(dptr + offsetof(BatchView, c2w));
        char* vb = A + base[i];
        int* out_idx = reinterpret_cast<int*>(vb + L[i].idx);
        int* out_hash = reinterpret_cast<int*>(vb + L[i].hash);
        bv.b = make_build_view(L[i].bp, d.xyz, d.n, d_w2c, d_K, d.height, d.width, 0, vb + L[i].ws, out_idx, nullptr,
                               out_hash, reinterpret_cast<float4*>(vb + L[i].pts));
        bv.view = ViewArgs{d_c2w, d.fx, d.fy, d.cx, d.cy, d.width, d.height, reinterpret_cast<const int2*>(out_hash),
                           out_idx, reinterpret_cast<const float4*>(vb + L[i].ca),
                           reinterpret_cast<const float2*>(vb + L[i].cb), reinterpret_cast<const float4*>(vb + L[i].pts),
                           b->radius ? reinterpret_cast<const float*>(vb + L[i].prad) : nullptr, d.xyz};
        bv.tiles_x = (int)ceil_div(d.width, b->tp.tw);
        bv.tiles = bv.tiles_x * (int)ceil_div(d.height, b->tp.th);
        bv.col_groups = (int)ceil_div(L[i].bp.nbins, kBuildBlock);
        bv.ray_off = ray_off[i];
        bv.pix_off = ray_off[i];
    }
    const int M = b->max_views + 1;
    b->prefix.assign((size_t)kTableRows * M, 0);
    int* P = b->prefix.data();
    // each view's chunks go to one XCD: largest first to the least loaded
    {
        std::vector<int> byload(V);
        std::iota(byload.begin(), byload.end(), 0);
        std::stable_sort(byload.begin(), byload.end(),
                         [&](int x, int y) { return b->views[x].b.nblk > b->views[y].b.nblk; });
        int64_t load[kXcds] = {};
        for (int r : byload) {
            const int x = (int)(std::min_element(load, load + kXcds) - load);
            int* xprefix = P + (kStages + x) * M;
            int* xlist = P + (kStages + kXcds + x) * M;
            const int j = xlist[M - 1]++;
            xlist[j] = r;
            xprefix[j + 1] = xprefix[j] + b->views[r].b.nblk;
            load[x] += b->views[r].b.nblk;
        }
    }
    for (int r = 0; r < V; ++r) {
        const BatchView& bv = b->views[r];
        P[kCount * M + r + 1] = P[kCount * M + r] + bv.b.nblk;
        P[kCol * M + r + 1] = P[kCol * M + r] + bv.col_groups * bv.b.nseg;
        P[kScan * M + r + 1] = P[kScan * M + r] + 1;
        P[kFine * M + r + 1] = P[kFine * M + r] + bv.b.nbins;
        P[kTot * M + r + 1] = P[kTot * M + r] + bv.col_groups;
        P[kTile * M + r + 1] = P[kTile * M + r] + bv.tiles;
    }
    b->header = BatchHeader{};
    b->header.V = V;
    for (int k = 0; k < kStages; ++k) b->header.items[k] = P[k * M + V];
    b->header.out = b->out;
    b->V = V;
    b->total_rays = rays;
    hipStream_t st = as_stream(stream);
    if (hipMemcpyAsync(b->d_header, &b->header, sizeof(BatchHeader), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(b->d_views, b->views.data(), sizeof(BatchView) * V, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(b->d_prefix, b->prefix.data(), sizeof(int) * b->prefix.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess)
        return check_launch("batch_set_views");
    // the sources are host vectors of this object: wait until they are read
    return hipStreamSynchronize(st) == hipSuccess ? EPCQ_OK : check_launch("batch_set_views");
}

extern "C" int epcq_batch_view_outputs(const epcq_batch* b, int32_t view, epcq_batch_view_out* out) {
    if (!b || !out || view < 0 || view >= b->V) return fail(EPCQ_ERR_INVALID, "batch_view_outputs: bad view");
    const BatchView& bv = b->views[b->slot_of[view]];
    out->sorted_idx = bv.b.out_idx;
    out->hash = bv.b.out_hash;
    out->rays = bv.b.P;
    out->ray_offset = bv.ray_off;
    out->n_samples = b->out.n_samples + bv.ray_off;
    out->sample_t = b->out.sample_t + bv.ray_off * b->s_max;
    out->nb_idx = b->out.nb_idx + bv.ray_off * b->s_max * b->opt.num_point_cloud_per_sp;
    out->nb_t = b->out.nb_t ? b->out.nb_t + bv.ray_off * b->s_max * b->opt.num_point_cloud_per_sp : nullptr;
    out->nb_dist = b->out.nb_dist ? b->out.nb_dist + bv.ray_off * b->s_max * b->opt.num_point_cloud_per_sp : nullptr;
    return EPCQ_OK;
}

extern "C" int epcq_batch_run(epcq_batch* b, void* stream) {
    if (!b || b->V < 1) return fail(EPCQ_ERR_INVALID, "batch_run: no views");
    return enqueue_all(b, as_stream(stream));
}

extern "C" int epcq_batch_capture(epcq_batch* b, void* stream) {
    if (!b || b->V < 1) return fail(EPCQ_ERR_INVALID, "batch_capture: no views");
    hipStream_t st = as_stream(stream);
    if (!st) return fail(EPCQ_ERR_INVALID, "batch_capture: capture needs a non-default stream");
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return check_launch("batch_capture/begin");
    const int rc = enqueue_all(b, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != EPCQ_OK) return rc;
    if (e != hipSuccess || !g) return fail(EPCQ_ERR_HIP, std::string("batch_capture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(EPCQ_ERR_HIP, std::string("batch_capture/instantiate: ") + hipGetErrorString(e2));
    }
    b->graph = g;
    b->exec = ex;
    return EPCQ_OK;
}

extern "C" int epcq_batch_launch(epcq_batch* b, void* stream) {
    if (!b || !b->exec) return fail(EPCQ_ERR_INVALID, "batch_launch: not captured");
    const hipError_t e = hipGraphLaunch(b->exec, as_stream(stream));
    if (e != hipSuccess) return fail(EPCQ_ERR_HIP, std::string("batch_launch: ") + hipGetErrorString(e));
    return EPCQ_OK;
}

extern "C" int64_t epcq_batch_rays(const epcq_batch* b) { return b ? b->total_rays : 0; }

extern "C" int epcq_batch_profile(epcq_batch* b, void* stream, float* stage_ms) {
    if (!b || b->V < 1 || !stage_ms) return fail(EPCQ_ERR_INVALID, "batch_profile: no views");
    hipStream_t st = as_stream(stream);
    hipEvent_t ev[4];
    for (int i = 0; i < 4; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) return check_launch("batch_profile/event");
    int rc = enqueue_all(b, st, ev);
    if (rc == EPCQ_OK && hipEventSynchronize(ev[3]) != hipSuccess) rc = check_launch("batch_profile/sync");
    for (int i = 0; rc == EPCQ_OK && i < 3; ++i) (void)hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]);
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
    return rc;
}

extern "C" void epcq_batch_destroy(epcq_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    delete b;
}
