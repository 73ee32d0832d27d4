/*
 * recon_mvs_hip.h — C ABI of the MI355X PatchMatch multi-view-stereo library.
 *
 * This is the drop-in boundary for the dense-geometry stage of recon-scanner.
 * The reference implements the stage as C++ free functions in
 * core/include/recon/mvs.h:133-214 (bodies in core/src/mvs/patchmatch_stereo.cpp)
 * and exposes it to apps only through recon_run_reconstruction
 * (core/include/recon/recon_c_api.h:534-542).  This header is the POD-only,
 * exception-free C layer that the C++ entry points of include/recon/mvs.h are
 * re-implemented on, and that a ctypes / JNI / cgo binding calls directly.
 *
 * Conventions (same as recon_c_api.h:3-13):
 *   - every function returns an int32 status whose values equal recon_status
 *     (recon_c_api.h:135-207): 0 = OK, 1 = invalid argument, 2 = out of
 *     memory, 3 = not implemented, 4 = internal, 5 = cancelled,
 *     500 = no depth, 502 = GPU failure;
 *   - all buffers are caller-owned host memory unless a function says
 *     otherwise; images are row-major with stride = width (mvs.h:130-150);
 *   - one context drives one GPU on one HIP stream; a context is not safe
 *     for concurrent use from several threads (recon_c_api.h:10).
 */
#ifndef RECON_MVS_HIP_H_
#define RECON_MVS_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RECON_MVS_API __attribute__((visibility("default")))

/* Status codes: same integers as recon_status (recon_c_api.h:135-176). */
enum {
  RECON_MVS_OK = 0,
  RECON_MVS_ERR_INVALID_ARGUMENT = 1,
  RECON_MVS_ERR_OUT_OF_MEMORY = 2,
  RECON_MVS_ERR_NOT_IMPLEMENTED = 3,
  RECON_MVS_ERR_INTERNAL = 4,
  RECON_MVS_ERR_CANCELLED = 5,
  RECON_MVS_ERR_NO_DEPTH = 500,
  RECON_MVS_ERR_GPU_FAILURE = 502
};

/* Pixel formats of images handed over or stored on the device. */
enum {
  RECON_MVS_PIXEL_U8 = 0,   /* 0..255 gray, the reference's only format (mvs.h:176) */
  RECON_MVS_PIXEL_F32 = 1,
  RECON_MVS_PIXEL_F16 = 2,
  RECON_MVS_PIXEL_BF16 = 3
};

/*
 * Camera = CameraIntrinsics (core/include/recon/camera.h:24-58) + CameraPose
 * (camera.h:61-78) in one POD.  R is the world->camera rotation, row-major;
 * t the world->camera translation in metres.
 */
typedef struct recon_mvs_camera {
  double fx, fy, cx, cy;
  double k1, k2, k3, p1, p2;
  int32_t width, height;
  double R[9];
  double t[3];
} recon_mvs_camera;

/* MvsConfig (core/include/recon/mvs.h:24-67), field for field, same defaults. */
typedef struct recon_mvs_config {
  int32_t num_views;
  int32_t ref_view_index;
  double ideal_angle_min_deg;
  double ideal_angle_max_deg;
  double max_view_scale_mismatch;
  double view_angle_weight;
  double view_texture_weight;
  double view_spread_weight;
  double view_scale_weight;
  double reference_depth_m;
  int32_t patch_size;
  int32_t iterations;
  int32_t random_search_count;
  int32_t propagation_directions;
  double random_search_radius_ratio;
  double restart_ratio;
  int32_t propagation_enabled;
  int32_t num_threads;
  double photo_threshold;
  double depth_regularization;
  double geo_consistency_threshold;
  double weak_texture_variance;
  int32_t subpixel_iterations;
  int32_t subpixel_fit_points;
  double subpixel_window_ratio;
  double min_depth_m;
  double max_depth_m;
  int32_t min_view_consistency;
  int32_t reserved0;
  double fusion_voxel_size_m;
} recon_mvs_config;

/* Fills *cfg with the MvsConfig defaults of mvs.h:24-67. */
RECON_MVS_API void recon_mvs_config_default(recon_mvs_config* cfg);

/* NccResult (mvs.h:70-75). */
typedef struct recon_mvs_ncc_result {
  int32_t valid;
  int32_t weak_texture;
  int32_t valid_samples;
  int32_t reserved0;
  double ncc;
} recon_mvs_ncc_result;

/* CostResult of EvaluateCandidate (patchmatch_stereo.cpp:289-295). */
typedef struct recon_mvs_cost_result {
  double cost;
  double ncc_sum;
  double weight_sum;
  int32_t weak_texture_count;
  int32_t valid_views;
} recon_mvs_cost_result;

/*
 * One reference view and its sources, already resolved through the index
 * contract of mvs.h:151-161 (src_cameras[k] pairs src_pixels[k]).  Pixels are
 * host pointers in `pixel_format`; the reference-mode entry points accept
 * RECON_MVS_PIXEL_U8 only, as the reference does.
 */
typedef struct recon_mvs_problem {
  const void* ref_pixels;
  int32_t ref_w, ref_h;
  int32_t num_src;
  int32_t pixel_format;
  const void* const* src_pixels;
  const int32_t* src_w;
  const int32_t* src_h;
  recon_mvs_camera ref_camera;
  const recon_mvs_camera* src_cameras;
} recon_mvs_problem;

/* Opaque per-GPU context. */
typedef struct recon_mvs_hip_context recon_mvs_hip_context;

/* Creates a context on HIP device `device`; allocates its stream. */
RECON_MVS_API int32_t recon_mvs_hip_create(int32_t device, recon_mvs_hip_context** out);
RECON_MVS_API int32_t recon_mvs_hip_destroy(recon_mvs_hip_context* ctx);
/* Text of the last error on this context ("" if none).  Owned by the context. */
RECON_MVS_API const char* recon_mvs_hip_last_error(const recon_mvs_hip_context* ctx);
/* Blocks until all work queued on the context's stream has finished. */
RECON_MVS_API int32_t recon_mvs_hip_synchronize(recon_mvs_hip_context* ctx);

/* ------------------------------------------------------------------------- */
/* Reference mode: the reference's own algorithm, FP64, reproduced on the GPU */
/* ------------------------------------------------------------------------- */

/*
 * Batched ComputeNcc (mvs.h:168-170; patchmatch_stereo.cpp:214-274).
 * coords holds n rows of (ref_u, ref_v, src_u, src_v); out receives n results.
 */
RECON_MVS_API int32_t recon_mvs_hip_compute_ncc(
    recon_mvs_hip_context* ctx, const uint8_t* ref_img, int32_t ref_w,
    int32_t ref_h, const uint8_t* src_img, int32_t src_w, int32_t src_h,
    int32_t n, const double* coords, int32_t patch_size,
    recon_mvs_ncc_result* out);

/*
 * Batched EvaluateCandidate (patchmatch_stereo.cpp:344-418): n rows of
 * (u, v, inverse_depth) evaluated against all sources of `problem`.
 */
RECON_MVS_API int32_t recon_mvs_hip_evaluate_candidates(
    recon_mvs_hip_context* ctx, const recon_mvs_problem* problem,
    const recon_mvs_config* config, int32_t n, const double* u_v_inv,
    recon_mvs_cost_result* out);

/*
 * PatchMatchStereo (mvs.h:193-206; patchmatch_stereo.cpp:422-764).
 * seed_points: num_seeds rows of (u, v, depth_m), may be NULL.
 * Outputs are ref_w*ref_h arrays (depth NaN where invalid, as DepthMap).
 * *filled receives the reference's return value.  The result equals the
 * reference CPU path run with the same config.num_threads (the reference's
 * trajectory depends on that value through its row-block RNG streams).
 */
RECON_MVS_API int32_t recon_mvs_hip_patchmatch_stereo(
    recon_mvs_hip_context* ctx, const recon_mvs_problem* problem,
    const recon_mvs_config* config, const double* seed_points,
    int32_t num_seeds, float* depth, uint8_t* support, float* confidence,
    uint32_t* filled);

/*
 * RefineSubpixel (mvs.h:208-218; patchmatch_stereo.cpp:770-979), in place on
 * depth / confidence.  *refined receives the reference's return value.
 */
RECON_MVS_API int32_t recon_mvs_hip_refine_subpixel(
    recon_mvs_hip_context* ctx, const recon_mvs_problem* problem,
    const recon_mvs_config* config, float* depth, float* confidence,
    uint32_t* refined);

/* A depth map handed to the fusion entry point (DepthMap, mvs.h:85-109). */
typedef struct recon_mvs_depth_map {
  int32_t width, height;
  const float* depth;
  const uint8_t* support_count;
  const float* confidence;
} recon_mvs_depth_map;

/*
 * FuseDepthMaps (mvs.h:213-218; patchmatch_stereo.cpp:985-1089), step 1:
 * runs the fusion and keeps the cloud on the context; *num_points receives
 * its size.  Points are emitted in ascending voxel-key order (the reference
 * emits them in unordered_map order, which no caller may rely on).
 */
RECON_MVS_API int32_t recon_mvs_hip_fuse_depth_maps(
    recon_mvs_hip_context* ctx, int32_t num_maps,
    const recon_mvs_depth_map* maps, const recon_mvs_camera* cameras,
    const recon_mvs_config* config, int64_t* num_points);

/*
 * Step 2: copies the cloud of the last fusion (PointCloud, mvs.h:112-125):
 * positions / normals 3 doubles per point, colors 3 floats, confidences 1.
 * Any pointer may be NULL.  `capacity` is the number of points the buffers hold.
 */
RECON_MVS_API int32_t recon_mvs_hip_fetch_point_cloud(
    recon_mvs_hip_context* ctx, int64_t capacity, double* positions,
    double* normals, float* colors, float* confidences, int64_t* voxel_keys);

#ifdef __cplusplus
}
#endif

#endif /* RECON_MVS_HIP_H_ */
