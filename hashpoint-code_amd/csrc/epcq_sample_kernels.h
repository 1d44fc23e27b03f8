// Full-image (tiled) and ray-list sampling kernels and their launch, one
// template per (neighbour mode, buffer size, sample bound, lane-set words).
// The launch templates are instantiated in epcq_sample_inst_*.hip (windows of
// at most 64 lanes) and epcq_sample_wide_*.hip (up to 256 lanes), one
// translation unit per (mode, buffer size, sample bound, words), so that they
// compile in parallel; epcq_sample.hip dispatches to them.
#pragma once

#include "epcq_common.h"
#include "epcq_sampler.h"

namespace epcq {

constexpr int kSampleBlock = 256;

namespace {

constexpr int kBlock = kSampleBlock;

template <int MODE, int KMAX, int SMAX, int LAY, bool PR, int NW>
__global__ __launch_bounds__(kBlock) void sample_tile_kernel(SampleOpts o, SampleOut out, ViewArgs view, TilePlan tp,
                                                             int tiles_x) {
    extern __shared__ float4 s_dyn[];
    __shared__ int s_total;
    __shared__ int s_heavy[kMaxTileRays + 1];
    const TileSmem sm = carve_tile_smem(s_dyn, tp, &s_total, s_heavy);
    const int tile = tp.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    // a tile whose halo region exceeds the LDS reads global memory
    if (!sample_tile_body<MODE, KMAX, SMAX, LAY, PR, NW>(o, out, view, tp, tile, tiles_x, 0, sm))
        sample_tile_global<MODE, KMAX, SMAX, LAY, PR, NW>(o, out, view, tp, tile, tiles_x, 0, sm);
}

// One thread per listed ray; rays with many candidates are deferred and then
// run one per wave (the block's waves share the queue).
template <int MODE, int KMAX, int SMAX, int LAY, bool PR, int NW>
__global__ __launch_bounds__(kBlock) void sample_list_kernel(SampleOpts o, SampleOut out, ViewArgs view,
                                                             const int* __restrict__ rays, int64_t R) {
    __shared__ int s_heavy[kMaxTileRays + 1];
    static_assert(kBlock <= kMaxTileRays, "queue holds one entry per thread");
    if (threadIdx.x == 0) s_heavy[kMaxTileRays] = 0;
    __syncthreads();
    const SrcGlobal<PR> src{view.hash, view.sorted_idx, o.xyz, view.pts, view.prad, o.radius, view.cone_a,
                            view.cone_b, view.W};
    const int64_t r0 = (int64_t)blockIdx.x * kBlock;
    const int64_t ri = r0 + threadIdx.x;
    if (ri < R) {
        const int64_t q = rays[ri];
        const bool valid = q >= 0 && q < (int64_t)view.W * view.H;
        const int u = valid ? (int)(q % view.W) : 0, v = valid ? (int)(q / view.W) : 0;
        Ray ray;
        view_ray(view, u, v, ray);
        if (!run_ray<MODE, KMAX, SMAX, LAY, NW>(o, out, src, ray, u, v, view.W, view.H, ri, valid, true))
            defer_ray(s_heavy, (int)threadIdx.x);
    }
    __syncthreads();
    const int n = s_heavy[kMaxTileRays];
    for (int h = (int)threadIdx.x / kWave; h < n; h += kBlock / kWave) {
        const int64_t rj = r0 + s_heavy[h];
        const int64_t q = rays[rj];  // deferred rays are valid
        const int u = (int)(q % view.W), v = (int)(q / view.W);
        Ray ray;
        view_ray(view, u, v, ray);
        run_ray_wave<MODE, KMAX, SMAX, LAY, NW>(o, out, src, ray, u, v, view.W, view.H, rj);
    }
}

// One wave per listed ray (run_ray_wave, the same outputs as run_ray): for
// short lists, such as a training batch, whose one-lane-per-ray form would
// keep a few CUs busy while the rest of the GPU idles.
template <int MODE, int KMAX, int SMAX, int LAY, bool PR, int NW>
__global__ __launch_bounds__(kBlock) void sample_list_wave_kernel(SampleOpts o, SampleOut out, ViewArgs view,
                                                                  const int* __restrict__ rays, int64_t R) {
    const int64_t ri = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (ri >= R) return;  // the whole wave: ri is uniform
    const SrcGlobal<PR> src{view.hash, view.sorted_idx, o.xyz, view.pts, view.prad, o.radius, view.cone_a,
                            view.cone_b, view.W};
    const int64_t q = rays[ri];
    if (!(q >= 0 && q < (int64_t)view.W * view.H)) {
        // a ray off the image: the empty outputs run_ray writes for it
        write_tail_wave<MODE, LAY>(o, out, ri, 0);
        return;
    }
    const int u = (int)(q % view.W), v = (int)(q / view.W);
    Ray ray;
    view_ray(view, u, v, ray);
    run_ray_wave<MODE, KMAX, SMAX, LAY, NW, true>(o, out, src, ray, u, v, view.W, view.H, ri);
}

}  // namespace

template <int MODE, int KMAX, int SMAX, int LAY, bool PR, int NW>
void launch_kernels(const SampleOpts& o, const SampleOut& out, const ViewArgs& view, const int* rays, int64_t R,
                const TilePlan& tp, hipStream_t st) {
    if (rays && R <= o.list_wave) {
        sample_list_wave_kernel<MODE, KMAX, SMAX, LAY, PR, NW>
            <<<dim3((unsigned)ceil_div(R, kBlock / kWave)), dim3(kBlock), 0, st>>>(o, out, view, rays, R);
    } else if (rays) {
        sample_list_kernel<MODE, KMAX, SMAX, LAY, PR, NW><<<dim3((unsigned)ceil_div(R, kBlock)), dim3(kBlock), 0, st>>>(
            o, out, view, rays, R);
    } else {
        const int tiles_x = (int)ceil_div(view.W, tp.tw);
        const int tiles = tiles_x * (int)ceil_div(view.H, tp.th);
        const dim3 block(tp.tw * tp.th);
        sample_tile_kernel<MODE, KMAX, SMAX, LAY, PR, NW><<<dim3(tiles), block, tp.lds_bytes, st>>>(o, out, view, tp,
                                                                                              tiles_x);
    }
}

// Per-point radius is a kernel property (PR); the dense drop-in layout has
// none (quick_sampling takes a uniform radius).
template <int MODE, int KMAX, int SMAX, int LAY, int NW>
void launch_one(const SampleOpts& o, const SampleOut& out, const ViewArgs& view, const int* rays, int64_t R,
                const TilePlan& tp, hipStream_t st) {
    if constexpr (LAY != kLayDense) {
        if (o.radius) {
            launch_kernels<MODE, KMAX, SMAX, LAY, true, NW>(o, out, view, rays, R, tp, st);
            return;
        }
    }
    launch_kernels<MODE, KMAX, SMAX, LAY, false, NW>(o, out, view, rays, R, tp, st);
}

template <int MODE, int KMAX, int SMAX, int NW>
void launch_lay(const SampleOpts& o, const SampleOut& out, const ViewArgs& view, const int* rays, int64_t R,
                const TilePlan& tp, hipStream_t st) {
    if (out.sp_xyz) launch_one<MODE, KMAX, SMAX, kLayDense, NW>(o, out, view, rays, R, tp, st);
    else if (o.feat) launch_one<MODE, KMAX, SMAX, kLayFeat, NW>(o, out, view, rays, R, tp, st);
    else launch_one<MODE, KMAX, SMAX, kLayCompact, NW>(o, out, view, rays, R, tp, st);
}


}  // namespace epcq

// Explicit instantiation of launch_lay<MODE, KMAX, SMAX, NW> (in its own unit).
#define EPCQ_INSTANTIATE_LAUNCH(MODE, KMAX, SMAX, NW)                                                      \
    template void epcq::launch_lay<MODE, KMAX, SMAX, NW>(const epcq::SampleOpts&, const epcq::SampleOut&,  \
                                                          const epcq::ViewArgs&, const int*, int64_t,     \
                                                          const epcq::TilePlan&, hipStream_t);
