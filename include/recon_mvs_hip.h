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
 *   - one context drives one GPU on one HIP stream (plus an upload stream
 *     ordered against it by events); a context is not safe for concurrent
 *     use from several threads (recon_c_api.h:10).
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
  RECON_MVS_PIXEL_BF16 = 3,
  RECON_MVS_PIXEL_BGRA8 = 4 /* byte order B,G,R,A (the app's camera frames): input only,
                               reduced to gray on upload as BgraToGray does
                               (core/src/orchestrator/orchestrator.cpp:53-67) */
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
 * RECON_MVS_PIXEL_U8 only, as the reference does, at most 32 sources, patch
 * sizes 1..63 and images of at most 2^29 pixels (32-bit indexing);
 * anything else returns RECON_MVS_ERR_INVALID_ARGUMENT.
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
/*
 * Status of the last call on ctx that failed (0 if none): e.g. 502 after a
 * GPU failure, which the C++ entry points turn into recon::MvsGpuError.
 */
RECON_MVS_API int32_t recon_mvs_hip_last_status(const recon_mvs_hip_context* ctx);

/*
 * Cooperative cancellation, the library's side of recon_cancel /
 * recon_is_cancelled (recon_c_api.h:601-608; Orchestrator::CheckCancel,
 * orchestrator.cpp:249, checked per keyframe at orchestrator.cpp:989).  GPU
 * work is not preempted: the flag is checked between dispatch batches
 * (docs/01_接口契约.md §3.11) -- per keyframe in recon_mvs_hip_run_mvs_stage,
 * per directional pass in recon_mvs_hip_patchmatch_stereo, at the start of
 * every call of the plane-mode entry points, and by the persistent batch
 * kernel between its work units (one tile), so a running batch drains within
 * milliseconds.  A call that sees the flag returns RECON_MVS_ERR_CANCELLED
 * (5) after the work already queued has finished; its outputs are undefined.
 *
 * recon_mvs_hip_cancel sets the flag; it is thread-safe and may be called
 * from any thread while another thread is inside a call on ctx (the only
 * exception to the one-thread rule).  The flag stays set, as recon_cancel's
 * does, until recon_mvs_hip_clear_cancel.  A callback (may be NULL) is polled
 * at the same check points; a nonzero answer sets the flag -- e.g. a function
 * returning the app context's cancel flag, so recon_cancel reaches running
 * GPU work without another call.
 */
typedef int32_t (*recon_mvs_cancel_fn)(void* user);
RECON_MVS_API int32_t recon_mvs_hip_cancel(recon_mvs_hip_context* ctx);
RECON_MVS_API int32_t recon_mvs_hip_is_cancelled(recon_mvs_hip_context* ctx);
RECON_MVS_API int32_t recon_mvs_hip_clear_cancel(recon_mvs_hip_context* ctx);
RECON_MVS_API int32_t recon_mvs_hip_set_cancel_callback(recon_mvs_hip_context* ctx,
                                                        recon_mvs_cancel_fn fn, void* user);

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

/*
 * SelectSourceViews (mvs.h:133-141; patchmatch_stereo.cpp:133-211), host only:
 * writes up to `capacity` global view indices to out and returns how many
 * the reference would return (or -1 on a null argument).  Ties keep input
 * order.
 */
RECON_MVS_API int32_t recon_mvs_select_source_views(const recon_mvs_camera* ref, int32_t n,
                                                    const recon_mvs_camera* poses,
                                                    const recon_mvs_config* config,
                                                    int32_t* out, int32_t capacity);

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

/*
 * The same cloud as Poisson input (OrientedPointCloud, poisson_mesh.h:25-38),
 * converted as MVS output reaches meshing: PointCloud -> UnifiedPointCloud
 * (orchestrator.cpp:1056-1063) -> UnifiedPointCloud::ToOrientedCloud
 * (orchestrator.cpp:181-192): positions, normals and colors as fetched,
 * weights[i] = max(0, confidence[i]).  Any pointer may be NULL;
 * *num_points (may be NULL) receives the cloud's size, and with the four
 * buffers NULL the call is only that size query.
 */
RECON_MVS_API int32_t recon_mvs_hip_fetch_oriented_cloud(
    recon_mvs_hip_context* ctx, int64_t capacity, double* positions,
    double* normals, float* colors, float* weights, int64_t* num_points);

/*
 * Depth-map spill to the session cache (docs/13_内存与降级.md §3.3.2-3.3.4:
 * `<cache>/mvs_depth/ref_%05d.f32` "with confidence and support counts",
 * written to *.tmp then renamed, recorded in `<cache>/cache_manifest.json`
 * with path, bytes, sha256 and stage).  Host only, no context.
 *
 * File: 32-byte header {"RMVD", uint32 version 1, int32 width, height,
 * ref_index, uint32 planes 3, 2 x uint32 0}, then float32 planes of
 * width x height: depth (NaN = invalid), confidence, support count.
 * Manifest: {"items": [{"path": "mvs_depth/ref_00007.f32", "bytes": B,
 * "sha256": "<hex>", "stage": "mvs"}, ...]}; items of other paths are kept.
 *
 * spill: creates cache_dir and cache_dir/mvs_depth as needed; replaces an
 * earlier map of the same ref_index; sha256_hex (65 bytes, may be NULL)
 * receives the file's digest.  ref_index in [0, 99999].
 */
RECON_MVS_API int32_t recon_mvs_spill_depth_map(const char* cache_dir, int32_t ref_index,
                                                int32_t width, int32_t height, const float* depth,
                                                const uint8_t* support, const float* confidence,
                                                char* sha256_hex);

/*
 * Reads a spilled map back.  The file's sha256 must match its manifest item
 * and the file must be well formed; otherwise the item (file and manifest
 * entry) is deleted and RECON_MVS_ERR_NO_DEPTH returned: the caller
 * recomputes the map (§3.3.4).  A missing file also returns
 * RECON_MVS_ERR_NO_DEPTH.  With depth, support and confidence all NULL it
 * only reports *width / *height; otherwise `capacity` (pixels each non-NULL
 * buffer holds) must be at least width x height.
 */
RECON_MVS_API int32_t recon_mvs_load_depth_map(const char* cache_dir, int32_t ref_index,
                                               int64_t capacity, int32_t* width, int32_t* height,
                                               float* depth, uint8_t* support, float* confidence);

/* ------------------------------------------------------------------------- */
/* Plane mode: slanted-plane PatchMatch with homography warps, FP32 on the    */
/* GPU, views resident in HBM.  It has no CPU counterpart in the reference:  */
/* it implements the GPU design of docs/05_单目管线.md §5.3.3-§5.6 (plane    */
/* hypotheses, red-black propagation, geometric consistency) behind the      */
/* reference's GPU parameter block (shaders/common/gpu_algo.h:217-258).      */
/* ------------------------------------------------------------------------- */

/* One input view (image + camera) handed to recon_mvs_hip_set_views. */
typedef struct recon_mvs_view {
  const void* pixels;   /* host, row-major */
  int32_t pixel_format; /* RECON_MVS_PIXEL_U8, RECON_MVS_PIXEL_F32 or RECON_MVS_PIXEL_BGRA8 */
  int32_t row_stride;   /* BGRA8: bytes per row (0 = 4 x width); U8 / F32: 0 (rows packed) */
  recon_mvs_camera camera;
} recon_mvs_view;

/*
 * Gray conversion of a BGRA8 frame on the GPU, the reference's BgraToGray
 * (core/src/orchestrator/orchestrator.cpp:53-67): gray = (77 R + 150 G + 29 B)
 * >> 8 per pixel, `stride` bytes per input row, `gray` width x height packed.
 * Host buffers; synchronous.  recon_mvs_hip_set_views applies the same
 * conversion on upload to views given as RECON_MVS_PIXEL_BGRA8.
 */
RECON_MVS_API int32_t recon_mvs_hip_bgra_to_gray(recon_mvs_hip_context* ctx, const uint8_t* bgra,
                                                 int32_t width, int32_t height, int32_t stride,
                                                 uint8_t* gray);

/*
 * Plane-mode parameters.  Image values are compared in the units they are
 * given in; value_scale names that unit's full range (255 for 0..255 data,
 * 1 for [0, 1]) and scales the reference's variance floors (weak-texture
 * variance 12 and NCC denominator 1e-9, both in 0..255 units).
 */
typedef struct recon_mvs_plane_params {
  int32_t patch_size;        /* window side, odd (the reference's patch_size) */
  int32_t window_stride;     /* sample stride inside the window (1 or 2) */
  int32_t top_k;             /* views aggregated per hypothesis; 0 = all */
  int32_t iterations;        /* red-black iterations (red + black half-sweep) */
  uint32_t seed;             /* counter-RNG seed */
  int32_t propagation_reach; /* odd steps scanned per direction, 1..16 (1 = 4-neighbour) */
  int32_t refine_iterations; /* golden-section steps of recon_mvs_hip_plane_refine */
  int32_t tile_schedule;     /* workgroup tile: 0 = 64x8, 1 = 32x16, 2 = 128x4 px,
                                -1 = autotuned (cached per size, views, window, format);
                                results do not depend on it */
  double min_depth_m, max_depth_m;
  double depth_perturb;      /* relative inverse-depth radius at iteration 0 */
  double normal_perturb;     /* normal perturbation radius at iteration 0 */
  double photo_threshold;    /* hypotheses with cost > 3 x this are rejected */
  double value_scale;        /* full range of the image values (see above) */
  /* Occupancy of the unconverged phases.  Random init and the first
     iterations evaluate scattered hypotheses whose window gathers miss the
     caches; fewer resident workgroups per CU thrash them less.
     occupancy_cap: 0 = automatic (FP32 images: init + `capped_iterations`
     iterations at 1 workgroup per CU; FP16 / BF16: uncapped; with
     tile_schedule = -1 tuned per phase), -1 = never capped, n > 0 = n
     workgroups per CU in the capped phases.  Results do not depend on it. */
  int32_t occupancy_cap;
  int32_t capped_iterations; /* leading iterations run capped (default 1; < 0 = 1) */
  /* Leading iterations whose half-sweeps run view-phased: one launch per
     source view instead of one for all, each visiting the tiles in strips
     along that view's epipolar lines, so the unconverged hypotheses' window
     gathers touch one source image at a time and stay in the L2 (11x11
     stride-2 window only; other windows ignore it).  -1 = automatic (the
     default: none; the first iteration for FP32 images when binned_phases
     is 0), 0 = none.  Results do not depend on it. */
  int32_t view_phased_iterations;
  /* Leading phases (phase 0 = random init, then the half-sweeps in order)
     whose hypotheses are evaluated in inverse-depth order: each tile's
     (pixel, hypothesis) pairs are sorted into 64 bins of inverse depth in LDS
     and evaluated bin by bin, so a wave's windows sit in one compact region
     of each source instead of 64 scattered ones (11x11 stride-2 window only;
     other windows ignore it).  -1 = automatic (tuned per phase wherever the
     occupancy is tuned, otherwise none), 0 = none.  Results do not depend on
     it. */
  int32_t binned_phases;
} recon_mvs_plane_params;

/* Fills *p with the defaults: 11x11 window, stride 2, top-4, 5 iterations,
   automatic occupancy (init + 1 iteration capped for FP32 images). */
RECON_MVS_API void recon_mvs_plane_params_default(recon_mvs_plane_params* p);

/*
 * Uploads n views into HBM, stored as `storage_format` (F32, F16, BF16 or
 * U8), replacing any previous set.  Depth-map slots are per view index.  U8
 * storage takes 8-bit views only (RECON_MVS_PIXEL_U8 or BGRA8 frames) and
 * holds them exactly, as FP16 and BF16 do: results are bit-identical to F32
 * storage for such views.  F16 / BF16 / U8 storage is compiled for the
 * 11x11 stride-2 window (F16 / BF16 also for 7x7).
 */
RECON_MVS_API int32_t recon_mvs_hip_set_views(recon_mvs_hip_context* ctx, int32_t n,
                                              const recon_mvs_view* views,
                                              int32_t storage_format);

/*
 * Plane-mode PatchMatch of view `ref_view` against `src_views` (the
 * PatchMatchStereo role, patchmatch_stereo.cpp:422-764): random plane init,
 * `iterations` red-black sweeps.  Asynchronous on the context's stream; the
 * depth map stays resident in slot ref_view.
 */
RECON_MVS_API int32_t recon_mvs_hip_plane_patchmatch(recon_mvs_hip_context* ctx,
                                                     int32_t ref_view, int32_t num_src,
                                                     const int32_t* src_views,
                                                     const recon_mvs_plane_params* params);

/*
 * The same with sparse-point seeding, PatchMatchStereo's seed_points
 * (mvs.h:193-206; patchmatch_stereo.cpp:499-533; docs/05_单目管线.md §5.3.2):
 * num_seeds rows of (u, v, depth_m) in ref_view's pixels.  Seeds outside the
 * image or with depth <= 0 are dropped; each 32 x 32 pixel cell holding seeds
 * takes their mean depth (FP64, seed order), and its pixels start on the
 * fronto-parallel plane at that depth (the reference's depth-only
 * hypothesis) instead of a random plane; other pixels start random.  Then
 * `iterations` sweeps as above.  num_seeds = 0 is the unseeded call.
 */
RECON_MVS_API int32_t recon_mvs_hip_plane_patchmatch_seeded(recon_mvs_hip_context* ctx,
                                                            int32_t ref_view, int32_t num_src,
                                                            const int32_t* src_views,
                                                            const recon_mvs_plane_params* params,
                                                            const double* seed_points,
                                                            int32_t num_seeds);

/*
 * Plane-mode sub-pixel refinement of slot ref_view (the RefineSubpixel role,
 * patchmatch_stereo.cpp:770-979): 5-point least-squares parabola in inverse
 * depth along the pixel's ray with the plane's normal fixed, then golden
 * section; the vertex is scaled by the sample step.  Asynchronous.
 */
RECON_MVS_API int32_t recon_mvs_hip_plane_refine(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                 int32_t num_src, const int32_t* src_views,
                                                 const recon_mvs_plane_params* params);

/*
 * Geometric-consistency filter over the resident depth maps of `views`
 * (docs/05 §5.6.2): each valid pixel of each map is projected into every
 * other listed map; the view agrees when |z_proj - z_map| / max(z_proj,
 * z_map) <= threshold.  geo_support = 1 + agreeing views; the pixel is kept
 * when geo_support >= min_views.  FP64, deterministic.  Asynchronous.
 */
RECON_MVS_API int32_t recon_mvs_hip_geo_consistency(recon_mvs_hip_context* ctx, int32_t num_maps,
                                                    const int32_t* views, double threshold,
                                                    int32_t min_views);

/*
 * Copies slot `view` to the host (synchronous).  depth: NaN where invalid
 * (or filtered out when the consistency filter ran and drop_inconsistent);
 * normals: 3 floats per pixel; support: valid views of the best hypothesis;
 * geo_support: see recon_mvs_hip_geo_consistency.  Any pointer may be NULL.
 */
RECON_MVS_API int32_t recon_mvs_hip_get_depth_map(recon_mvs_hip_context* ctx, int32_t view,
                                                  int32_t drop_inconsistent, float* depth,
                                                  float* normals, float* cost, uint8_t* support,
                                                  float* confidence, uint8_t* geo_support);

/* Sets slot `view`'s plane state (depth + normal per pixel) and scores it. */
RECON_MVS_API int32_t recon_mvs_hip_set_plane_state(recon_mvs_hip_context* ctx, int32_t view,
                                                    int32_t num_src, const int32_t* src_views,
                                                    const recon_mvs_plane_params* params,
                                                    const float* depth, const float* normals);

/*
 * Plane-mode cost of n hypotheses (x, y, depth, nx, ny, nz) of view ref_view
 * (for parity tests): aggregated cost, valid views, and per-view NCC
 * (n x num_src; NaN where the view is invalid).  Synchronous.
 */
RECON_MVS_API int32_t recon_mvs_hip_eval_plane_costs(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                     int32_t num_src, const int32_t* src_views,
                                                     const recon_mvs_plane_params* params,
                                                     int32_t n, const float* hyps, float* cost,
                                                     int32_t* valid_views, float* view_ncc);

/*
 * Autotuner: times random init + two half-sweeps of ref_view under every
 * compiled tile schedule, caches and returns the fastest (*best_schedule);
 * ms_per_schedule (3 entries, may be NULL) receives each time (-1 where a
 * schedule is not compiled for this window / format).  Overwrites the slot.
 */
RECON_MVS_API int32_t recon_mvs_hip_autotune(recon_mvs_hip_context* ctx, int32_t ref_view,
                                             int32_t num_src, const int32_t* src_views,
                                             const recon_mvs_plane_params* params,
                                             int32_t* best_schedule, double* ms_per_schedule);

/* One reference view of a batch and its sources (view indices of set_views). */
typedef struct recon_mvs_batch_job {
  int32_t ref_view;
  int32_t num_src;
  int32_t src_views[16];
} recon_mvs_batch_job;

typedef struct recon_mvs_hip_batch recon_mvs_hip_batch;

/*
 * Batched plane-mode PatchMatch of many reference views of any resolution
 * (no counterpart in the reference, whose StageMvs loops over views,
 * orchestrator.cpp:988-1040).  The work is split into (view, tile) units
 * served from a work queue by one persistent launch per phase (random init,
 * then every half-sweep); all phases are captured once into a HIP graph.
 * Each job's depth map lands in slot ref_view, bit-identical to
 * recon_mvs_hip_plane_patchmatch with the same parameters.  tile_schedule
 * -1 autotunes on the job with the most pixel-views.
 */
RECON_MVS_API int32_t recon_mvs_hip_batch_create(recon_mvs_hip_context* ctx, int32_t num_jobs,
                                                 const recon_mvs_batch_job* jobs,
                                                 const recon_mvs_plane_params* params,
                                                 recon_mvs_hip_batch** out);
/* Enqueues one run of the captured graph on the context's stream. */
RECON_MVS_API int32_t recon_mvs_hip_batch_run(recon_mvs_hip_context* ctx, recon_mvs_hip_batch* batch);
/* Tile schedule the batch was built with. */
RECON_MVS_API int32_t recon_mvs_hip_batch_schedule(const recon_mvs_hip_batch* batch);
RECON_MVS_API int32_t recon_mvs_hip_batch_destroy(recon_mvs_hip_context* ctx, recon_mvs_hip_batch* batch);

/*
 * Adjoint of the plane cost (no counterpart in the reference's CPU code: the
 * reverse mode its design documents plan for pose refinement,
 * docs/05_单目管线.md).  For every pixel of the depth-map slot of ref_view,
 * with the planes it holds (from plane_patchmatch or set_plane_state):
 *   cost[P]          the aggregated cost (1e30: invalid), as plane_patchmatch
 *                    scores it, under source poses perturbed by pose_delta;
 *   grad_plane[P*4]  d(sum_p w_p cost_p) / d(depth, nx, ny, nz) per pixel;
 *   grad_pose[6*num_src]  d(sum_p w_p cost_p) / d(w, tau) per source, for the
 *                    left perturbation Y -> Exp(w) Y + tau of the source's
 *                    world->camera transform (FP64 reduction, deterministic).
 * pose_delta (6*num_src doubles: w, tau per source) and weights (P floats)
 * may be NULL (no perturbation / all ones); any output may be NULL.  Invalid
 * pixels get zero gradient.  Windows: (11, 2) in every storage format, (7, 1)
 * and (5, 1) in FP32.  Synchronous.
 */
RECON_MVS_API int32_t recon_mvs_hip_plane_cost_grad(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                    int32_t num_src, const int32_t* src_views,
                                                    const recon_mvs_plane_params* params,
                                                    const double* pose_delta, const float* weights,
                                                    float* cost, float* grad_plane,
                                                    double* grad_pose);

/*
 * recon_mvs_hip_plane_cost_grad without any device allocation, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); it has no counterpart in the reference.
 * Same kernels and results (bit for bit) as recon_mvs_hip_plane_cost_grad,
 * but d_weights (P floats, may be NULL), d_cost (P floats), d_grad_plane
 * (P x 4 floats, 16-byte aligned) and d_grad_pose (6*num_src doubles) are
 * DEVICE pointers (any may be NULL); pose_delta stays a host array (read on
 * the host).  Asynchronous on the context's stream (recon_mvs_hip_synchronize
 * before reading the outputs).  The per-tile pose partials live in a fixed
 * 2 MiB device array per GPU: maps needing more (tiles of 32 x 16 pixels x
 * num_src x 6 doubles, e.g. 1920x1440 with more than 8 sources) are refused
 * with RECON_MVS_ERR_INVALID_ARGUMENT when d_grad_pose is requested.  All
 * calls on one GPU share that array: calls that overlap on one GPU (two
 * contexts, two threads) race.  The views, the plane state of ref_view and
 * the output buffers must exist beforehand.
 */
RECON_MVS_API int32_t recon_mvs_hip_plane_cost_grad_noalloc(
    recon_mvs_hip_context* ctx, int32_t ref_view, int32_t num_src, const int32_t* src_views,
    const recon_mvs_plane_params* params, const double* pose_delta, const float* d_weights,
    float* d_cost, float* d_grad_plane, double* d_grad_pose);

/*
 * Gradient polish of the depth-map slot of ref_view (no counterpart in the
 * reference): `steps` steps (10 is the intended count); even steps move each
 * pixel's inverse depth, odd steps its normal, against the cost gradient by a
 * per-pixel step length, keeping a move only when the pixel's cost drops
 * (costs never increase).  Clears the slot's consistency result.  Async.
 */
RECON_MVS_API int32_t recon_mvs_hip_plane_polish(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                 int32_t num_src, const int32_t* src_views,
                                                 const recon_mvs_plane_params* params, int32_t steps);

/*
 * The MVS stage of a reconstruction over keyframe arrays, the body of
 * Orchestrator::StageMvs (core/src/orchestrator/orchestrator.cpp:942-1071).
 * For every keyframe r with an image (gray[r] != NULL, cams[r].width x height
 * uint8): SelectSourceViews with r's intrinsics for every pose, minus r and
 * keyframes without an image, at most num_views; then
 *   RECON_MVS_STAGE_REFERENCE: PatchMatchStereo + RefineSubpixel per keyframe
 *     (the reference's algorithm, bit-exact for the same num_threads);
 *   RECON_MVS_STAGE_PLANE: every keyframe in one plane-mode batch (`plane`,
 *     NULL = parameters derived from cfg: its window (stride 2 from 9 px),
 *     iterations, depth range and photometric threshold), sub-pixel
 *     refinement and the 10-step gradient polish of every map, then geometric
 *     consistency of each map against its sources (geo_consistency_threshold;
 *     pixels seen consistently by fewer than min_view_consistency maps
 *     dropped); replaces the views resident in ctx;
 * and FuseDepthMaps over the produced maps in keyframe order; fetch the cloud
 * with recon_mvs_hip_fetch_point_cloud (*num_points points).  depth,
 * support, confidence (n pointers each, may be NULL or hold NULLs) receive
 * each keyframe's map (DepthMap fields, mvs.h:85-109; NaN / 0 where invalid
 * or not produced).  *produced = depth maps produced (0: the caller's
 * RECON_ERR_MVS_NO_DEPTH, orchestrator.cpp:1041).
 */
enum { RECON_MVS_STAGE_REFERENCE = 0, RECON_MVS_STAGE_PLANE = 1 };
RECON_MVS_API int32_t recon_mvs_hip_run_mvs_stage(recon_mvs_hip_context* ctx, int32_t n,
                                                  const uint8_t* const* gray,
                                                  const recon_mvs_camera* cams,
                                                  const recon_mvs_config* cfg, int32_t mode,
                                                  const recon_mvs_plane_params* plane,
                                                  float* const* depth, uint8_t* const* support,
                                                  float* const* confidence, int32_t* produced,
                                                  int64_t* num_points);

/*
 * Kernel timers.  enable != 0 records HIP events around every launch of the
 * plane-mode kernels, and of reference mode's pass / initial-cost / refine
 * kernels, on the context's stream; ms / counts (8 entries each, may be NULL)
 * receive the accumulated milliseconds and launch counts per kernel class:
 * 0 sweep (reference mode: one directional pass), 1 init, 2 refine,
 * 3 consistency, 4 score, 5 eval, 6 batch (one graph run), 7 adjoint / polish.
 * reset != 0 clears the accumulators.
 */
RECON_MVS_API int32_t recon_mvs_hip_timing(recon_mvs_hip_context* ctx, int32_t enable,
                                           int32_t reset, double* ms, int64_t* counts);

#ifdef __cplusplus
}
#endif

#endif /* RECON_MVS_HIP_H_ */
