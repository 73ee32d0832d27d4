// ref_shim.cpp — TEST INFRASTRUCTURE ONLY (part of the oracle).
//
// extern "C" shim over the reference's own CPU MVS entry points, so that the
// tests can drive the unmodified reference (compiled in place from
// core/src/mvs/patchmatch_stereo.cpp, core/src/util/camera.cpp and
// core/src/util/linalg.cpp by oracle/Makefile) through ctypes.  Nothing in
// the product links this file; the product path never calls it.
//
// Each function forwards to exactly one reference function:
//   ref_config_default   -> MvsConfig defaults        (mvs.h:24-67)
//   ref_compute_ncc      -> ComputeNcc                (patchmatch_stereo.cpp:214-274)
//   ref_patchmatch       -> PatchMatchStereo          (patchmatch_stereo.cpp:422-764)
//   ref_refine_subpixel  -> RefineSubpixel            (patchmatch_stereo.cpp:770-979)
//   ref_fuse             -> FuseDepthMaps             (patchmatch_stereo.cpp:985-1089)
//   ref_select_views     -> SelectSourceViews         (patchmatch_stereo.cpp:133-211)
//   ref_pixel_to_ray / ref_project_point -> camera.cpp:144-162
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "recon/camera.h"
#include "recon/mvs.h"
#include "recon_mvs_hip.h"

namespace {

recon::CameraIntrinsics ToIntr(const recon_mvs_camera& c) {
  recon::CameraIntrinsics k;
  k.fx = c.fx; k.fy = c.fy; k.cx = c.cx; k.cy = c.cy;
  k.k1 = c.k1; k.k2 = c.k2; k.k3 = c.k3; k.p1 = c.p1; k.p2 = c.p2;
  k.width = c.width; k.height = c.height;
  return k;
}

recon::CameraPose ToPose(const recon_mvs_camera& c) {
  recon::CameraPose p;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) p.R(i, j) = c.R[i * 3 + j];
    p.t(i) = c.t[i];
  }
  return p;
}

recon::MvsConfig ToConfig(const recon_mvs_config& s) {
  recon::MvsConfig c;
  c.num_views = s.num_views;
  c.ref_view_index = s.ref_view_index;
  c.ideal_angle_min_deg = s.ideal_angle_min_deg;
  c.ideal_angle_max_deg = s.ideal_angle_max_deg;
  c.max_view_scale_mismatch = s.max_view_scale_mismatch;
  c.view_angle_weight = s.view_angle_weight;
  c.view_texture_weight = s.view_texture_weight;
  c.view_spread_weight = s.view_spread_weight;
  c.view_scale_weight = s.view_scale_weight;
  c.reference_depth_m = s.reference_depth_m;
  c.patch_size = s.patch_size;
  c.iterations = s.iterations;
  c.random_search_count = s.random_search_count;
  c.random_search_radius_ratio = s.random_search_radius_ratio;
  c.restart_ratio = s.restart_ratio;
  c.propagation_directions = s.propagation_directions;
  c.propagation_enabled = s.propagation_enabled != 0;
  c.num_threads = s.num_threads;
  c.photo_threshold = s.photo_threshold;
  c.depth_regularization = s.depth_regularization;
  c.geo_consistency_threshold = s.geo_consistency_threshold;
  c.weak_texture_variance = s.weak_texture_variance;
  c.subpixel_iterations = s.subpixel_iterations;
  c.subpixel_fit_points = s.subpixel_fit_points;
  c.subpixel_window_ratio = s.subpixel_window_ratio;
  c.min_depth_m = s.min_depth_m;
  c.max_depth_m = s.max_depth_m;
  c.min_view_consistency = s.min_view_consistency;
  c.fusion_voxel_size_m = s.fusion_voxel_size_m;
  return c;
}

struct Views {
  std::vector<const uint8_t*> gray;
  std::vector<int> w, h;
  std::vector<recon::CameraPose> poses;
  std::vector<recon::CameraIntrinsics> intr;
  std::vector<int> indices;
};

Views MakeViews(int num_src, const uint8_t* const* src, const int32_t* sw,
                const int32_t* sh, int num_poses, const recon_mvs_camera* poses,
                int num_intr, const recon_mvs_camera* intr,
                const int32_t* source_indices) {
  Views v;
  for (int k = 0; k < num_src; ++k) {
    v.gray.push_back(src[k]);
    v.w.push_back(sw[k]);
    v.h.push_back(sh[k]);
    v.indices.push_back(source_indices[k]);
  }
  for (int i = 0; i < num_poses; ++i) v.poses.push_back(ToPose(poses[i]));
  for (int i = 0; i < num_intr; ++i) v.intr.push_back(ToIntr(intr[i]));
  return v;
}

void CopyDepth(const recon::DepthMap& d, float* depth, uint8_t* support,
               float* conf) {
  const size_t n = d.depth.size();
  std::memcpy(depth, d.depth.data(), n * sizeof(float));
  std::memcpy(support, d.support_count.data(), n);
  std::memcpy(conf, d.confidence.data(), n * sizeof(float));
}

}  // namespace

extern "C" {

void ref_config_default(recon_mvs_config* out) {
  const recon::MvsConfig c;
  std::memset(out, 0, sizeof(*out));
  out->num_views = c.num_views;
  out->ref_view_index = c.ref_view_index;
  out->ideal_angle_min_deg = c.ideal_angle_min_deg;
  out->ideal_angle_max_deg = c.ideal_angle_max_deg;
  out->max_view_scale_mismatch = c.max_view_scale_mismatch;
  out->view_angle_weight = c.view_angle_weight;
  out->view_texture_weight = c.view_texture_weight;
  out->view_spread_weight = c.view_spread_weight;
  out->view_scale_weight = c.view_scale_weight;
  out->reference_depth_m = c.reference_depth_m;
  out->patch_size = c.patch_size;
  out->iterations = c.iterations;
  out->random_search_count = c.random_search_count;
  out->random_search_radius_ratio = c.random_search_radius_ratio;
  out->restart_ratio = c.restart_ratio;
  out->propagation_directions = c.propagation_directions;
  out->propagation_enabled = c.propagation_enabled ? 1 : 0;
  out->num_threads = c.num_threads;
  out->photo_threshold = c.photo_threshold;
  out->depth_regularization = c.depth_regularization;
  out->geo_consistency_threshold = c.geo_consistency_threshold;
  out->weak_texture_variance = c.weak_texture_variance;
  out->subpixel_iterations = c.subpixel_iterations;
  out->subpixel_fit_points = c.subpixel_fit_points;
  out->subpixel_window_ratio = c.subpixel_window_ratio;
  out->min_depth_m = c.min_depth_m;
  out->max_depth_m = c.max_depth_m;
  out->min_view_consistency = c.min_view_consistency;
  out->fusion_voxel_size_m = c.fusion_voxel_size_m;
}

void ref_compute_ncc(const uint8_t* a, int aw, int ah, const uint8_t* b,
                     int bw, int bh, int n, const double* coords, int patch,
                     recon_mvs_ncc_result* out) {
  for (int i = 0; i < n; ++i) {
    const double* c = coords + 4 * i;
    const recon::NccResult r =
        recon::ComputeNcc(a, aw, ah, b, bw, bh, c[0], c[1], c[2], c[3], patch);
    out[i].valid = r.valid ? 1 : 0;
    out[i].weak_texture = r.weak_texture ? 1 : 0;
    out[i].valid_samples = r.valid_samples;
    out[i].reserved0 = 0;
    out[i].ncc = r.ncc;
  }
}

void ref_pixel_to_ray(const recon_mvs_camera* cam, int n, const double* uv,
                      double* rays) {
  const recon::CameraIntrinsics k = ToIntr(*cam);
  for (int i = 0; i < n; ++i) {
    recon::Vec3d r;
    k.PixelToRay(uv[2 * i], uv[2 * i + 1], &r);
    for (int j = 0; j < 3; ++j) rays[3 * i + j] = r(j);
  }
}

void ref_project_point(const recon_mvs_camera* cam, int n, const double* pts,
                       double* uv, int32_t* ok) {
  const recon::CameraIntrinsics k = ToIntr(*cam);
  for (int i = 0; i < n; ++i) {
    recon::Vec3d p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    double u = 0.0, v = 0.0;
    ok[i] = k.ProjectPoint(p, &u, &v) ? 1 : 0;
    uv[2 * i] = u;
    uv[2 * i + 1] = v;
  }
}

uint32_t ref_patchmatch(const uint8_t* ref, int rw, int rh, int num_src,
                        const uint8_t* const* src, const int32_t* sw,
                        const int32_t* sh, const recon_mvs_camera* ref_cam,
                        int num_poses, const recon_mvs_camera* poses,
                        int num_intr, const recon_mvs_camera* intr,
                        const int32_t* source_indices,
                        const recon_mvs_config* cfg, const double* seeds,
                        int num_seeds, float* depth, uint8_t* support,
                        float* conf) {
  Views v = MakeViews(num_src, src, sw, sh, num_poses, poses, num_intr, intr,
                      source_indices);
  std::vector<std::array<double, 3>> seed_vec;
  for (int i = 0; i < num_seeds; ++i) {
    seed_vec.push_back({seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]});
  }
  recon::DepthMap d;
  const uint32_t filled = recon::PatchMatchStereo(
      ref, rw, rh, v.gray, v.w, v.h, ToPose(*ref_cam), v.poses, ToIntr(*ref_cam),
      v.intr, v.indices, ToConfig(*cfg), seeds ? &seed_vec : nullptr, &d);
  if (d.width == rw && d.height == rh) CopyDepth(d, depth, support, conf);
  return filled;
}

uint32_t ref_refine_subpixel(const uint8_t* ref, int rw, int rh, int num_src,
                             const uint8_t* const* src, const int32_t* sw,
                             const int32_t* sh, const recon_mvs_camera* ref_cam,
                             int num_poses, const recon_mvs_camera* poses,
                             int num_intr, const recon_mvs_camera* intr,
                             const int32_t* source_indices,
                             const recon_mvs_config* cfg, float* depth,
                             uint8_t* support, float* conf) {
  Views v = MakeViews(num_src, src, sw, sh, num_poses, poses, num_intr, intr,
                      source_indices);
  recon::DepthMap d;
  d.width = rw;
  d.height = rh;
  const size_t n = static_cast<size_t>(rw) * rh;
  d.depth.assign(depth, depth + n);
  d.support_count.assign(support, support + n);
  d.confidence.assign(conf, conf + n);
  const uint32_t refined = recon::RefineSubpixel(
      ref, rw, rh, v.gray, v.w, v.h, ToPose(*ref_cam), v.poses, ToIntr(*ref_cam),
      v.intr, v.indices, ToConfig(*cfg), &d);
  CopyDepth(d, depth, support, conf);
  return refined;
}

int64_t ref_fuse(int num_maps, const int32_t* widths, const int32_t* heights,
                 const float* const* depth, const uint8_t* const* support,
                 const float* const* conf, const recon_mvs_camera* cams,
                 const recon_mvs_config* cfg, int64_t capacity, double* pos,
                 double* nrm, float* out_conf, float* colors) {
  std::vector<recon::DepthMap> maps(num_maps);
  std::vector<const recon::DepthMap*> ptrs;
  std::vector<recon::CameraPose> poses;
  std::vector<recon::CameraIntrinsics> intr;
  for (int m = 0; m < num_maps; ++m) {
    const size_t n = static_cast<size_t>(widths[m]) * heights[m];
    maps[m].width = widths[m];
    maps[m].height = heights[m];
    maps[m].depth.assign(depth[m], depth[m] + n);
    maps[m].support_count.assign(support[m], support[m] + n);
    maps[m].confidence.assign(conf[m], conf[m] + n);
    ptrs.push_back(&maps[m]);
    poses.push_back(ToPose(cams[m]));
    intr.push_back(ToIntr(cams[m]));
  }
  const recon::PointCloud cloud =
      recon::FuseDepthMaps(ptrs, poses, intr, ToConfig(*cfg));
  const int64_t count = static_cast<int64_t>(cloud.size());
  const int64_t m = count < capacity ? count : capacity;
  for (int64_t i = 0; i < m; ++i) {
    for (int j = 0; j < 3; ++j) {
      pos[3 * i + j] = cloud.positions[i](j);
      nrm[3 * i + j] = cloud.normals[i](j);
      colors[3 * i + j] = cloud.colors[3 * i + j];
    }
    out_conf[i] = cloud.confidences[i];
  }
  return count;
}

int ref_select_views(const recon_mvs_camera* ref_cam, int num_poses,
                     const recon_mvs_camera* poses, const recon_mvs_config* cfg,
                     int32_t* out, int capacity) {
  std::vector<recon::CameraPose> p;
  for (int i = 0; i < num_poses; ++i) p.push_back(ToPose(poses[i]));
  const std::vector<int> sel = recon::SelectSourceViews(
      ToPose(*ref_cam), p, ToIntr(*ref_cam), ToConfig(*cfg));
  const int n = static_cast<int>(sel.size());
  for (int i = 0; i < n && i < capacity; ++i) out[i] = sel[i];
  return n;
}

}  // extern "C"
