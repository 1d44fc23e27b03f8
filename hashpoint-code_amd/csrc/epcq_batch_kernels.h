// Shared layout of the batched pass (descriptors, header, item spaces) and its
// persistent sampling kernel.  launch_batch_sample<MODE, KMAX, SMAX> is
// instantiated in epcq_batch_inst_*.hip, one translation unit per neighbour
// mode, so that they compile in parallel; epcq_batch.hip holds the rest of
// the pass and the C ABI.
#pragma once

#include "epcq_hashbuild.h"
#include "epcq_sampler.h"

namespace epcq {

constexpr int kMaxBins = 1 << kMaxBinBits;      // LDS cursors of count / bin
constexpr int kMaxFineTile = 1 << 11;          // pixels per bin (fine LDS; 16 KB keeps 8 blocks per CU)
// The batch sizes dense views' bins for half the single view's points: more,
// smaller placement items (placement 2.07 -> 1.95 ms, binning 0.89 -> 0.97 ms
// on the 64-view batch).
constexpr int kBatchBinPoints = kBinPoints / 2;
constexpr int kStages = 6;                     // item spaces: count, col, scan, fine, pixels, tiles
enum { kCount = 0, kCol = 1, kScan = 2, kFine = 3, kTot = 4, kTile = 5 };
// The GPU's XCDs (8 on MI355X), each with its own L2; workgroup b of a launch
// runs on XCD b % kXcds.  The projection and binning passes give every view
// to one XCD, so the partial cache lines its binned records are appended to
// fill up in that XCD's L2 instead of being written back from several.
constexpr int kXcds = 8;
// Rows of the item tables: kStages prefix rows over the views, then per XCD
// a prefix row over its views' chunks and a row listing those views (their
// count in the row's last entry).
constexpr int kTableRows = kStages + 2 * kXcds;

struct BatchView {
    BuildView b;
    ViewArgs view;
    float c2w[12];
    float w2c[12];
    float K[9];
    int tiles_x, tiles;
    int col_groups;
    int64_t ray_off;
    int64_t pix_off;
};

struct BatchHeader {
    int V;
    int items[kStages];
    int counter;    // next sample item
    SampleOut out;  // compact outputs of the current views (moves with the view set)
};

// Largest v with prefix[v] <= item (prefix has V + 1 entries, prefix[0] = 0).
__device__ __forceinline__ int find_view(const int* __restrict__ prefix, int V, int item) {
    int lo = 0, hi = V;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

namespace {

template <int MODE, int KMAX, int SMAX, bool PR>
__global__ __launch_bounds__(256) void batch_sample_kernel(BatchHeader* __restrict__ h,
                                                           const BatchView* __restrict__ views,
                                                           const int* __restrict__ prefix, SampleOpts o,
                                                           TilePlan tp) {
    extern __shared__ float4 s_dyn[];
    __shared__ int s_total;
    __shared__ int s_item;
    __shared__ int s_heavy[kMaxTileRays + 1];
    const TileSmem sm = carve_tile_smem(s_dyn, tp, &s_total, s_heavy);
    const int n_items = h->items[kTile], V = h->V;
    const SampleOut out = h->out;
    const int* pf = prefix;
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(&h->counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= n_items) break;
        const int v = find_view(pf, V, item);
        const BatchView& bv = views[v];
        // a tile whose halo region exceeds the LDS reads global memory right
        // away: the densest views come first, so these long tiles overlap
        // with the rest of the batch
        const bool staged = sample_tile_body<MODE, KMAX, SMAX, kLayCompact, PR>(o, out, bv.view, tp,
                                                                                item - pf[v], bv.tiles_x,
                                                                                bv.ray_off, sm);
        if (!staged)
            sample_tile_global<MODE, KMAX, SMAX, kLayCompact, PR>(o, out, bv.view, tp, item - pf[v],
                                                                  bv.tiles_x, bv.ray_off, sm);
    }
}

}  // namespace

template <int MODE, int KMAX, int SMAX>
void launch_batch_sample(int grid, const TilePlan& tp, BatchHeader* h, const BatchView* views, const int* prefix,
                         const SampleOpts& o, hipStream_t st) {
    if (o.radius)
        batch_sample_kernel<MODE, KMAX, SMAX, true><<<dim3(grid), dim3(tp.tw * tp.th), tp.lds_bytes, st>>>(h, views, prefix,
                                                                                                     o, tp);
    else
        batch_sample_kernel<MODE, KMAX, SMAX, false><<<dim3(grid), dim3(tp.tw * tp.th), tp.lds_bytes, st>>>(h, views, prefix,
                                                                                                      o, tp);
}

}  // namespace epcq

#define EPCQ_INSTANTIATE_BATCH(MODE, KMAX)                                                                   \
    template void epcq::launch_batch_sample<MODE, KMAX, 2>(int, const epcq::TilePlan&, epcq::BatchHeader*,   \
                                                           const epcq::BatchView*, const int*,               \
                                                           const epcq::SampleOpts&, hipStream_t);            \
    template void epcq::launch_batch_sample<MODE, KMAX, epcq::kRegSamples>(                                  \
        int, const epcq::TilePlan&, epcq::BatchHeader*, const epcq::BatchView*, const int*,                  \
        const epcq::SampleOpts&, hipStream_t);

// Declarations that keep a unit from instantiating launch_batch_sample itself.
#define EPCQ_EXTERN_BATCH(MODE, KMAX)                                                                        \
    extern template void epcq::launch_batch_sample<MODE, KMAX, 2>(int, const epcq::TilePlan&, epcq::BatchHeader*, \
                                                                  const epcq::BatchView*, const int*,        \
                                                                  const epcq::SampleOpts&, hipStream_t);     \
    extern template void epcq::launch_batch_sample<MODE, KMAX, epcq::kRegSamples>(                           \
        int, const epcq::TilePlan&, epcq::BatchHeader*, const epcq::BatchView*, const int*,                  \
        const epcq::SampleOpts&, hipStream_t);
