/*
 * CPU oracle for the epcq searching + sampling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against; nothing in the product (the epcq package, the HIP
 * library) links, loads or calls it.  Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark use it.
 *
 * It restates the reference algorithm sequentially on host memory, following
 * (paths relative to the reference tree):
 *   epcq/epcq.py:251-295                  projection, pixel keys, sort (A4, A5, A8)
 *   epcq/csrc/rasterization.cu:7-171      scatter_hashtable / scatter (A6, A7)
 *   epcq/csrc/epcq_render_topk.cu:10-417  quick_sampling / query (A11-A14)
 *   epcq/csrc/include/cuda_util.cuh       ray, projection, neighbours (A9, A10)
 *   epcq/epcq.py:345-387,434-452          feature gather + IDW weights (A15)
 *
 * Where the reference is order-nondeterministic or relies on NVIDIA's
 * approximate math the oracle fixes one canonical choice:
 *   - pixel lists ordered by (pixel, point index), or (pixel, depth, index)
 *     when sorted by depth (torch.sort tie order is unspecified);
 *   - per-slot distance sums accumulated in (neighbour lane, bucket) order
 *     (the reference uses shared-memory float atomics);
 *   - radius-query output appended in lane order (the reference uses a
 *     global atomic counter);
 *   - exp() is our own Cody-Waite + degree-7 Taylor routine (orc_expf), norms
 *     are sqrtf((x*x + y*y) + z*z) and normalisation multiplies by
 *     1/sqrtf(...), every operation rounded once, no contraction.
 * All arithmetic is IEEE single precision with -ffp-contract=off, except the
 * pixel-centre step which the reference performs in double.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_camera {
    float c2w[12];   /* row-major 3x4 camera-to-world */
    float w2c[12];   /* row-major 3x4 world-to-camera */
    float K[9];      /* row-major 3x3 intrinsics (projection only) */
    float fx, fy, cx, cy;
    int32_t width, height;
} orc_camera;

typedef struct orc_options {
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
} orc_options;

/* neighbour selection around each sample point */
enum { ORC_MODE_COMPAT = 0, ORC_MODE_KNN = 1 };

void orc_set_threads(int n);
float orc_expf(float x);

/* A10: (du, dv) offsets of neighbour lane l for window size k. */
void orc_neighbor_offset(int32_t k, int32_t lane, int32_t* du, int32_t* dv);

/* A9: ray through pixel (u, v): origin[3], normalised dir[3]. */
void orc_ray(const orc_camera* cam, int32_t u, int32_t v, float* origin, float* dir);

/* A9: projection of point p on the ray: t and point-to-ray distance. */
void orc_point_ray(const float* origin, const float* dir, const float* p, float* t, float* dist);

/* A4: pixel key (u + v*W, or H*W off-image) and camera depth per point. */
int orc_project(const float* xyz, int64_t n, const float* w2c12, const float* K9,
                int32_t H, int32_t W, int32_t* key, float* depth);

/* A5: stable sort by key -> (sorted keys, permuted point index, permuted depth). */
int orc_sort_by_key(const int32_t* key, const float* depth, int64_t n,
                    int32_t* key_sorted, int32_t* idx_sorted, float* depth_sorted);

/* A6: scatter_hashtable on pixel-sorted inputs. */
int orc_scatter_hashtable(const int32_t* point_idx, const int32_t* keys, const float* depth,
                          int64_t n, int32_t H, int32_t W, int32_t sort,
                          int32_t* out_idx, float* out_depth, int32_t* out_hash);

/* A8: pts2plane = A4 + A5 + A6. */
int orc_pts2plane(const float* xyz, int64_t n, const float* w2c12, const float* K9,
                  int32_t H, int32_t W, int32_t sort,
                  int32_t* out_idx, float* out_depth, int32_t* out_hash);

/* A7: per-pixel B nearest by depth. out arrays [H*W, bins] pre-filled with -1. */
int orc_scatter(const int32_t* point_idx, const int32_t* keys, const float* depth,
                int64_t n, int32_t H, int32_t W, int32_t bins,
                float* out_depth, int32_t* out_idx);

/*
 * A11 + A12: quick_sampling in the dense drop-in layout.
 *   sp_xyz [P,S,3] f32, pc_idx [P,S,Kc] i32, sp_t [P,S] f32, mask [P,S] u8.
 * radius: optional per-point radius [n] (NULL = options.radius_threshold).
 * mode ORC_MODE_COMPAT reproduces the reference's 8-entry buffer;
 * ORC_MODE_KNN returns the Kc nearest by (distance, index), ascending, and
 * fills nb_t / nb_dist [P,S,Kc] (may be NULL) with each neighbour's ray
 * depth and point-to-ray distance.
 * rays: optional list of R pixel ids to evaluate (NULL = all P pixels in
 * order); outputs are then [R, ...].
 */
int orc_quick_sampling(const float* xyz, int64_t n, const orc_camera* cam, const orc_options* opt,
                       const int32_t* sorted_idx, const int32_t* hash, const float* radius,
                       int32_t mode, const int32_t* rays, int64_t R,
                       float* sp_xyz, int32_t* pc_idx, float* sp_t, uint8_t* mask,
                       float* nb_t, float* nb_dist);

/* A14: radius query for a ray list -> out [R, num_point_per_ray] i32. */
int orc_quick_query(const float* xyz, int64_t n, const orc_camera* cam, const orc_options* opt,
                    const int32_t* sorted_idx, const int32_t* hash,
                    const int32_t* rays, int64_t R, int32_t* out);

/*
 * A15: inverse-distance weights and weighted feature sum for each sample.
 *   sp_xyz [M,3], nb_idx [M,K] (-1 = empty) -> w [M,K], agg [M,F].
 * feat is f32 [n,F] (feat_bf16 == 0) or bf16 bits [n,F] (feat_bf16 == 1).
 */
int orc_idw(const float* xyz, const void* feat, int32_t feat_bf16, int32_t F,
            const float* sp_xyz, const int32_t* nb_idx, int64_t M, int32_t K,
            float* w, float* agg);

/*
 * Volume-rendering epilogue (epcq/epcq.py:508-530): clamp_t delta [R,S] from
 * t [R,S]; render out [R,3] from sigma [R,S], rgb [R,S,3] and either t
 * (clamped on the fly) or delta (the other NULL); its backward grad_sigma
 * [R,S], grad_rgb [R,S,3] from grad_out [R,3].
 */
int orc_clamp_t(const float* t, int64_t R, int32_t S, float vsize, float* delta);
int orc_render(const float* t, const float* delta, float vsize, const float* sigma, const float* rgb, int64_t R,
               int32_t S, int32_t white_bg, float* out);
int orc_render_backward(const float* t, const float* delta, float vsize, const float* sigma, const float* rgb,
                        const float* grad_out, int64_t R, int32_t S, int32_t white_bg, float* grad_sigma,
                        float* grad_rgb);

#ifdef __cplusplus
}
#endif
