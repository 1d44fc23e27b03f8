/*
 * epcq_abi.h -- C ABI of the MI355X (gfx950) epcq searching + sampling path.
 *
 * This is the boundary the Python module `epcq.csrc` binds with ctypes.  It
 * replaces the pybind11 extension `epcq.csrc` of the reference
 * (epcq/csrc/epcq.cpp:5-55, setup.py:37-46).  Every entry point:
 *   - takes plain device pointers and sizes (no torch types);
 *   - never allocates device memory: outputs and workspace are supplied by the
 *     caller, so calls can be captured into a HIP graph;
 *   - enqueues its kernels on `stream` (a hipStream_t; NULL = the legacy
 *     default stream) and returns without synchronising;
 *   - returns EPCQ_OK or an error code; epcq_last_error() gives the message
 *     of the last failure on the calling thread (launch errors included).
 * Integers are int32 and floats IEEE binary32 unless stated otherwise.
 */
#ifndef EPCQ_ABI_H
#define EPCQ_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPCQ_ABI_VERSION 3

enum epcq_status {
    EPCQ_OK = 0,
    EPCQ_ERR_INVALID = 1,   /* bad argument (shape, size, option) */
    EPCQ_ERR_HIP = 2,       /* a HIP runtime / launch error */
    EPCQ_ERR_WORKSPACE = 3  /* workspace too small */
};

/* neighbour selection around each sample point */
enum epcq_mode {
    EPCQ_MODE_COMPAT = 0,  /* the reference's 8-entry buffer (epcq_render_topk.cu:35-79) */
    EPCQ_MODE_KNN = 1      /* K nearest by (distance, index), ascending */
};

/* CameraSpec (epcq/csrc/include/data_spec.hpp:15-31, epcq/epcq.py:15-37).
 * c2w / w2c are device pointers to row-major matrices with >= 3 rows of 4
 * floats (the reference's [>=3,4] CUDA tensors). */
typedef struct epcq_camera {
    const float* c2w;
    const float* w2c;
    float fx, fy, cx, cy;
    int32_t width, height;
} epcq_camera;

/* RenderOptions (epcq/csrc/include/data_spec.hpp:49-69, epcq/epcq.py:81-114). */
typedef struct epcq_options {
    int32_t kernel_size;
    float t_intvl;
    float radius_threshold;
    float sigma_threshold;
    float background_brightness;
    float stop_threshold;
    int32_t num_sample_points_per_ray;
    int32_t num_point_cloud_per_sp;
    int32_t num_point_per_ray;
    float sdf2weight_gaussian_alpha;
    float sdf2weight_gaussian_gamma;
    int32_t topK;
} epcq_options;

int epcq_abi_version(void);
const char* epcq_last_error(void);

/*
 * Launch shape of the full-image samplers (no counterpart in the reference,
 * whose launch shape is fixed).  One block samples a tile_w x tile_h tile of
 * rays (tile_w * tile_h in {64, 128, 256}, tile_w a power of two <= 64);
 * tile_cap = points of the tile's halo region staged in LDS (clamped to the
 * 64 KB block limit).  flags: EPCQ_TUNE_XCD_ORDER gives each XCD a
 * contiguous run of tiles (other bits are rejected).  heavy: a ray whose
 * unculled window buckets hold more than this many points (more than 4x this
 * many when over 16 rays of its wave of 64 exceed it) is sampled by a whole
 * wave (64 lanes share its candidate walks) after the tile's other rays,
 * instead of by one lane; 0 never does.  list_wave: a ray list (the rays
 * argument of the samplers) of at most this many rays is sampled one wave per
 * ray, longer ones one lane per ray; 0 never does.  Results are identical
 * either way.  Process-wide, guarded by a mutex; the autotuner sets it.
 * Defaults: 32, 8, 1024, 0, 128, 65536.
 */
typedef struct epcq_tuning {
    int32_t tile_w;
    int32_t tile_h;
    int32_t tile_cap;
    int32_t flags;
    int32_t heavy;
    int32_t list_wave;
} epcq_tuning;
#define EPCQ_TUNE_XCD_ORDER 1
int epcq_set_tuning(const epcq_tuning* t);
int epcq_get_tuning(epcq_tuning* t);

/* Smallest float s with sqrtf(s) >= r (host helper; sqrtf(e2) < r <=> e2 < s). */
float epcq_squared_threshold(float r);

/*
 * scatter_hashtable -- replaces `scatter_hashtable`
 * (epcq/csrc/epcq.cpp:6, epcq/csrc/rasterization.cu:130-171).
 * Inputs are pixel-sorted (keys ascending, off-image key = H*W) as produced
 * by epcq.py:283-286.  Writes every element of out_idx[n], out_depth[n]
 * (-1 outside pixel runs) and out_hash[H*W*2] ((start, count) or (-1, -1));
 * with sort != 0 each run is ordered by depth ascending, stable (runs longer
 * than 32 points by a block-wide radix sort through the workspace).
 * workspace: device memory of at least
 * epcq_scatter_hashtable_workspace_size(n, sort) bytes (0 when sort == 0).
 */
size_t epcq_scatter_hashtable_workspace_size(int64_t n, int32_t sort);
int epcq_scatter_hashtable(const int32_t* point_idx, const int32_t* keys, const float* depth,
                           int64_t n, int32_t height, int32_t width, int32_t sort,
                           int32_t* out_idx, float* out_depth, int32_t* out_hash,
                           void* workspace, size_t workspace_bytes, void* stream);

/*
 * scatter -- replaces `scatter` (epcq/csrc/epcq.cpp:5,
 * epcq/csrc/rasterization.cu:88-128).  Per pixel the `bins` smallest depths
 * ascending: out_depth / out_idx [H*W, bins], -1 where unused.  Negative
 * depths are skipped.
 */
int epcq_scatter(const int32_t* point_idx, const int32_t* keys, const float* depth, int64_t n,
                 int32_t height, int32_t width, int32_t bins, float* out_depth, int32_t* out_idx,
                 void* stream);

/*
 * build_hashtable -- fused replacement of NeuralPoints.pts2plane
 * (epcq/epcq.py:279-295 = _world2cam, _cam2image, _convetpixel, torch.sort,
 * scatter_hashtable).  xyz [n,3]; w2c >= 3 rows x 4 and K 3x3, row-major
 * device pointers.  Output as epcq_scatter_hashtable, with pixel lists
 * ordered by (pixel, point index), or (pixel, depth, point index) when
 * sort != 0.  workspace: device memory of at least
 * epcq_build_hashtable_workspace_size(n, height, width) bytes.
 */
size_t epcq_build_hashtable_workspace_size(int64_t n, int32_t height, int32_t width);
int epcq_build_hashtable(const float* xyz, int64_t n, const float* w2c, const float* K,
                         int32_t height, int32_t width, int32_t sort,
                         int32_t* out_idx, float* out_depth, int32_t* out_hash,
                         void* workspace, size_t workspace_bytes, void* stream);

/*
 * build_hashtable_noalloc -- no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): epcq_build_hashtable with
 * its workspace in a device array the library holds, EPCQ_NOALLOC_SCRATCH_BYTES
 * on each GPU it runs on, for the life of the process.  Same outputs as
 * epcq_build_hashtable; returns EPCQ_ERR_WORKSPACE when
 * epcq_build_hashtable_workspace_size(n, height, width) exceeds that size.
 * It allocates nothing.  All calls on one GPU share its array: calls that
 * overlap on one GPU (different streams, or threads) race and must not be
 * made.
 */
#define EPCQ_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int epcq_build_hashtable_noalloc(const float* xyz, int64_t n, const float* w2c, const float* K,
                                 int32_t height, int32_t width, int32_t sort,
                                 int32_t* out_idx, float* out_depth, int32_t* out_hash, void* stream);

/*
 * quick_sampling -- replaces `quick_sampling` (epcq/csrc/epcq.cpp:7,
 * epcq/csrc/epcq_render_topk.cu:341-417).  hash [H*W,2]; writes every element
 * of sp_xyz [H,W,S,3], pc_idx [H,W,S,Kc], sp_t [H,W,S] and mask [H,W,S]
 * (bool bytes), S = num_sample_points_per_ray, Kc = num_point_cloud_per_sp.
 * For kernel_size > 5 (where the reference overflows its 32-entry neighbour
 * list) all kernel_size^2 neighbour pixels are used in both stages.
 */
int epcq_quick_sampling(const float* xyz, int64_t n, const epcq_camera* cam,
                        const epcq_options* opt, const int32_t* sorted_idx, const int32_t* hash,
                        int32_t mode, float* sp_xyz, int32_t* pc_idx, float* sp_t, uint8_t* mask,
                        void* workspace, size_t workspace_bytes, void* stream);

/*
 * quick_sampling of a ray list -- the training form of the reference's
 * renderer, which samples the whole image and then keeps the rays listed in
 * training_ray_index (opt/render_image.py:75-88).  rays [R] pixel ids
 * (v * W + u); writes sp_xyz [R,S,3], pc_idx [R,S,Kc], sp_t [R,S] and
 * mask [R,S], row r equal to row rays[r] of epcq_quick_sampling's outputs.
 * Ids outside [0, H*W) give an empty row (zeros / -1 / false).
 */
int epcq_quick_sampling_rays(const float* xyz, int64_t n, const epcq_camera* cam,
                             const epcq_options* opt, const int32_t* sorted_idx, const int32_t* hash,
                             int32_t mode, const int32_t* rays, int64_t R, float* sp_xyz, int32_t* pc_idx,
                             float* sp_t, uint8_t* mask, void* workspace, size_t workspace_bytes,
                             void* stream);

/* Workspace of the samplers for n points and an H x W image: per-pixel
 * bucket cones that let the walks skip neighbour pixels provably outside the
 * radius, and a pixel-ordered copy of the points.  Optional: with a NULL or
 * smaller workspace the samplers skip nothing and gather the points through
 * sorted_idx (same results). */
size_t epcq_sampling_workspace_size(int64_t n, int32_t height, int32_t width);

/*
 * quick_query_for_nearby_point -- replaces `quick_query_for_nearby_point`
 * (epcq/csrc/epcq.cpp:8, epcq/csrc/epcq_render_topk.cu:287-339).
 * rays [R] pixel ids; out [R, num_point_per_ray], hits appended in
 * neighbour-lane order, -1 padded.  Ray ids outside [0, H*W) give a -1 row.
 */
int epcq_quick_query(const float* xyz, int64_t n, const epcq_camera* cam, const epcq_options* opt,
                     const int32_t* sorted_idx, const int32_t* hash, const int32_t* rays,
                     int64_t R, int32_t* out, void* stream);

/*
 * knn_sampling -- compact K-nearest sampling (no counterpart op in the
 * reference; same samples as quick_sampling, epcq_render_topk.cu:82-220).
 *   rays      [R] pixel ids, or NULL for all H*W pixels in order (R ignored);
 *   radius    [n] per-point radius or NULL (= radius_threshold);
 *   s_max     samples stored per ray (>= epcq_max_live_samples(opt, r_max));
 * outputs (all per ray, s = 0..s_max-1, k = 0..Kc-1):
 *   n_samples [R]            number of live samples;
 *   sample_t  [R,s_max]      sample depth t (0 where not live);
 *   nb_idx    [R,s_max,Kc]   neighbour point index (-1 where empty);
 *   nb_t      [R,s_max,Kc]   neighbour ray depth (EPCQ_MODE_KNN, may be NULL);
 *   nb_dist   [R,s_max,Kc]   neighbour point-to-ray distance (KNN, may be NULL);
 * and, when feat != NULL (f32 [n,F] or bf16 [n,F] with feat_bf16 = 1), the
 * inverse-distance weights of epcq/epcq.py:434-437,449-452:
 *   w   [R,s_max,Kc]   weights;   agg [R,s_max,F]   sum_k w_k * feat[idx_k].
 */
int epcq_knn_sampling(const float* xyz, int64_t n, const epcq_camera* cam,
                      const epcq_options* opt, const int32_t* sorted_idx, const int32_t* hash,
                      const float* radius, const int32_t* rays, int64_t R, int32_t mode,
                      int32_t s_max, int32_t* n_samples, float* sample_t, int32_t* nb_idx,
                      float* nb_t, float* nb_dist, const void* feat, int32_t feat_bf16,
                      int32_t F, float* w, float* agg, void* workspace, size_t workspace_bytes,
                      void* stream);

/* Upper bound on live samples per ray for these options and radius bound. */
int32_t epcq_max_live_samples(const epcq_options* opt, float r_max);

/*
 * idw_features -- FeatureQuery gathers + Aggregator.calculate_weight
 * (epcq/epcq.py:345-387, 434-437, 449-452) on dense samples:
 * sp_xyz [M,3], nb_idx [M,K] -> w [M,K], agg [M,F].
 */
int epcq_idw_features(const float* xyz, const void* feat, int32_t feat_bf16, int32_t F,
                      const float* sp_xyz, const int32_t* nb_idx, int64_t M, int32_t K,
                      float* w, float* agg, void* stream);

/*
 * Volume-rendering epilogue: clamp_t and render (epcq/epcq.py:508-530, also
 * utils/utils.py:422-467) as the reference renderer applies them to the
 * samples (opt/render_image.py:101-109), plus render's backward.
 *   clamp_t: delta [R,S] from sample depths t [R,S] (neighbour differences,
 *            vsize for the last and where below 1e-8 or above 2 vsize).
 *   render_rays: out_rgb [R,3] from sigma [R,S] and rgb [R,S,3]; give either
 *            t (clamp_t is then applied on the fly) or delta, not both.
 *            white_bg != 0 adds the remaining transmittance times white.
 *   render_rays_backward: grad_sigma [R,S] and grad_rgb [R,S,3] from
 *            grad_rgb_out [R,3] (same t / delta convention).
 */
int epcq_clamp_t(const float* t, int64_t R, int32_t S, float vsize, float* delta, void* stream);
int epcq_render_rays(const float* t, const float* delta, float vsize, const float* sigma, const float* rgb,
                     int64_t R, int32_t S, int32_t white_bg, float* out_rgb, void* stream);
int epcq_render_rays_backward(const float* t, const float* delta, float vsize, const float* sigma,
                              const float* rgb, const float* grad_rgb_out, int64_t R, int32_t S,
                              int32_t white_bg, float* grad_sigma, float* grad_rgb, void* stream);

/*
 * Batched multi-view sampling (no counterpart in the reference, which renders
 * one view per call).  A batch holds many views -- each its own camera,
 * resolution and point cloud -- and runs the hash-table build, bucket bounds
 * and compact sampling of all of them in one persistent pass, replayable as
 * a HIP graph.  All device memory is the caller's arena.
 */
typedef struct epcq_view_desc {
    const float* xyz;  /* device [n, 3] point cloud of this view */
    int64_t n;
    float c2w[12];     /* row-major 3 x 4 */
    float w2c[12];     /* row-major 3 x 4 */
    float K[9];        /* row-major 3 x 3 */
    float fx, fy, cx, cy;
    int32_t width, height;
} epcq_view_desc;

typedef struct epcq_batch epcq_batch; /* opaque host object */

/* Device pointers of one view's results inside the arena (layout as
 * epcq_knn_sampling's compact outputs; nb_t / nb_dist NULL in compat mode). */
typedef struct epcq_batch_view_out {
    int32_t* sorted_idx;  /* [n] pixel-ordered point lists */
    int32_t* hash;        /* [H*W, 2] */
    int64_t rays;         /* H*W */
    int64_t ray_offset;   /* first row of this view in the batch outputs */
    int32_t* n_samples;   /* [H*W] */
    float* sample_t;      /* [H*W, s_max] */
    int32_t* nb_idx;      /* [H*W, s_max, Kc] */
    float* nb_t;
    float* nb_dist;
} epcq_batch_view_out;

/* Arena bytes for these views (capacity for up to max_views descriptors). */
size_t epcq_batch_arena_bytes(const epcq_view_desc* views, int32_t n_views, int32_t max_views,
                              const epcq_options* opt, int32_t mode, int32_t s_max);
/* Arena bytes of the header part (descriptors for max_views views), and an
 * upper bound of the bytes one view adds (its build workspace, point list,
 * hash table, bucket spheres and outputs): for any view set,
 * arena_bytes <= header_bytes + sum of view_bytes.  Used for admission. */
size_t epcq_batch_header_bytes(int32_t max_views);
size_t epcq_batch_view_bytes(const epcq_view_desc* view, const epcq_options* opt, int32_t mode, int32_t s_max);
/* radius: optional per-point radius of the (single) cloud, r_max its maximum;
 * tuning: launch shape (NULL = the process-wide setting). */
int epcq_batch_create(const epcq_options* opt, int32_t mode, int32_t s_max, int32_t max_views,
                      const float* radius, float r_max, const epcq_tuning* tuning, void* arena,
                      size_t arena_bytes, epcq_batch** out);
/* Lays the views out in the arena and uploads their descriptors (returns
 * after the upload).  A captured graph stays valid for any view set. */
int epcq_batch_set_views(epcq_batch* b, const epcq_view_desc* views, int32_t n_views, void* stream);
int epcq_batch_view_outputs(const epcq_batch* b, int32_t view, epcq_batch_view_out* out);
int64_t epcq_batch_rays(const epcq_batch* b);
int epcq_batch_run(epcq_batch* b, void* stream);      /* enqueue the pass */
int epcq_batch_capture(epcq_batch* b, void* stream);  /* capture the pass into a HIP graph */
int epcq_batch_launch(epcq_batch* b, void* stream);   /* replay the graph */
/* Runs the pass eagerly with HIP events between its stages and waits for it:
 * stage_ms[0..2] = projection and binning, per-pixel placement (hash table,
 * point lists, bucket cones), sampling (milliseconds). */
int epcq_batch_profile(epcq_batch* b, void* stream, float* stage_ms);
void epcq_batch_destroy(epcq_batch* b);

/*
 * aggregator_forward -- replaces Aggregator.forward (epcq/epcq.py:439-507)
 * for inference: block1 and block3 over every valid (sample, neighbour) pair,
 * alpha_branch with softplus(x - 1), the inverse-distance aggregation of the
 * pair alphas and features (calculate_weight, :434-437), and color_branch
 * with sigmoid(x) * 1.002 - 0.001, on bf16 matrix cores with fp32
 * accumulation.  M samples, K in [1, 32] neighbours, F = 32 features:
 * pc_xyz / pc_rgb / pc_dir [M,K,3], pc_feat [M,K,F], sp_xyz [M,3],
 * valid [M,K] (bool bytes), rays_d [M,3], all f32 unless stated ->
 * out_alpha [M], out_rgb [M,3].  params: the module's Linear layers as bf16
 * [out, in] row-major matrices with the inputs zero-padded to a multiple of
 * 16 (w1 256x288, w2 256x256, w3 256x272, w4 256x256, c1 128x288, c2 and c3
 * 128x128) and f32 biases; alpha_branch (wa [256], ba [1]) and the last
 * colour layer (c4 [3,128], cb4 [3]) stay f32.  workspace: 16-byte aligned,
 * at least epcq_aggregator_workspace_size(M) bytes.
 */
typedef struct epcq_aggregator_params {
    const void* w1;  const float* b1;   /* block1.0: 284 -> 256 */
    const void* w2;  const float* b2;   /* block1.2: 256 -> 256 */
    const void* w3;  const float* b3;   /* block3.0: 263 -> 256 */
    const void* w4;  const float* b4;   /* block3.2: 256 -> 256 */
    const float* wa; const float* ba;   /* alpha_branch.0: 256 -> 1 */
    const void* c1;  const float* cb1;  /* color_branch.0: 280 -> 128 */
    const void* c2;  const float* cb2;  /* color_branch.2: 128 -> 128 */
    const void* c3;  const float* cb3;  /* color_branch.4: 128 -> 128 */
    const float* c4; const float* cb4;  /* color_branch.6: 128 -> 3 */
} epcq_aggregator_params;
size_t epcq_aggregator_workspace_size(int64_t M);
int epcq_aggregator_forward(const float* pc_xyz, const float* pc_rgb, const float* pc_dir, const float* pc_feat,
                            const float* sp_xyz, const uint8_t* valid, const float* rays_d, int64_t M, int32_t K,
                            int32_t F, const epcq_aggregator_params* params, float* out_alpha, float* out_rgb,
                            void* workspace, size_t workspace_bytes, void* stream);

/* Test hook: our exp() on the device, out[i] = exp(in[i]). */
int epcq_debug_expf(const float* in, float* out, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EPCQ_ABI_H */
