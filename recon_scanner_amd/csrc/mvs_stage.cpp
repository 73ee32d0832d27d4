// mvs_stage.cpp — the MVS stage of the reconstruction over keyframe arrays:
// the body of Orchestrator::StageMvs (core/src/orchestrator/orchestrator.cpp:
// 942-1071) on the library's entry points, in the reference's per-keyframe
// form or as one batched plane-mode run on the GPU.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mvs_common.hpp"
#include "recon_mvs_hip.h"

namespace {

// Source views of keyframe r as StageMvs picks them (orchestrator.cpp:996-1010):
// SelectSourceViews with r's intrinsics for every pose, then r itself and
// keyframes without an image removed, at most num_views kept.
std::vector<int> KeptSources(int n, int r, const uint8_t* const* gray, const recon_mvs_camera* cams,
                             const recon_mvs_config* cfg) {
  std::vector<recon_mvs_camera> pc(cams, cams + n);
  for (int i = 0; i < n; ++i) {
    recon_mvs_camera& c = pc[i];
    const recon_mvs_camera& k = cams[r];
    c.fx = k.fx; c.fy = k.fy; c.cx = k.cx; c.cy = k.cy;
    c.k1 = k.k1; c.k2 = k.k2; c.k3 = k.k3; c.p1 = k.p1; c.p2 = k.p2;
    c.width = k.width; c.height = k.height;
  }
  std::vector<int32_t> sel(n + 1);
  const int m = recon_mvs_select_source_views(&pc[r], n, pc.data(), cfg, sel.data(), n + 1);
  std::vector<int> kept;
  for (int q = 0; q < m; ++q) {
    const int si = sel[q];
    if (si == r || si < 0 || si >= n || gray[si] == nullptr) continue;
    kept.push_back(si);
  }
  if (static_cast<int>(kept.size()) > cfg->num_views) kept.resize(std::max(0, cfg->num_views));
  return kept;
}

// Plane-mode parameters derived from the reference's MvsConfig when the
// caller gives none: its window, iterations, depth range and photometric
// threshold; stride 2 for windows of 9 px and more; uint8 value range.
recon_mvs_plane_params DerivedPlaneParams(const recon_mvs_config* cfg) {
  recon_mvs_plane_params p;
  recon_mvs_plane_params_default(&p);
  p.patch_size = cfg->patch_size;
  p.window_stride = cfg->patch_size >= 9 ? 2 : 1;
  p.iterations = std::max(1, cfg->iterations);
  p.min_depth_m = cfg->min_depth_m;
  p.max_depth_m = cfg->max_depth_m;
  p.photo_threshold = cfg->photo_threshold;
  p.value_scale = 255.0;
  p.tile_schedule = -1;
  return p;
}

}  // namespace

extern "C" RECON_MVS_API int32_t recon_mvs_hip_run_mvs_stage(
    recon_mvs_hip_context* ctx, int32_t n, const uint8_t* const* gray,
    const recon_mvs_camera* cams, const recon_mvs_config* cfg, int32_t mode,
    const recon_mvs_plane_params* plane, float* const* depth, uint8_t* const* support,
    float* const* confidence, int32_t* produced, int64_t* num_points) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (n < 0 || (n > 0 && (!gray || !cams)) || !cfg || !produced || !num_points ||
      (mode != RECON_MVS_STAGE_REFERENCE && mode != RECON_MVS_STAGE_PLANE))
    return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "run_mvs_stage: bad arguments");
  *produced = 0;
  *num_points = 0;
  // per keyframe: depth map storage (kept until fusion)
  std::vector<std::vector<float>> dep(n), conf(n);
  std::vector<std::vector<uint8_t>> sup(n);
  std::vector<int> made;
  auto alloc = [&](int r) {
    const size_t P = static_cast<size_t>(cams[r].width) * cams[r].height;
    dep[r].assign(P, 0.0f);
    conf[r].assign(P, 0.0f);
    sup[r].assign(P, 0);
  };
  int32_t st = RECON_MVS_OK;
  if (rmvs::CancelRequested(ctx)) return rmvs::Cancelled(ctx);
  if (mode == RECON_MVS_STAGE_REFERENCE) {
    for (int r = 0; r < n && st == RECON_MVS_OK; ++r) {
      // the reference's check before every keyframe (orchestrator.cpp:989)
      if (rmvs::CancelRequested(ctx)) return rmvs::Cancelled(ctx);
      if (!gray[r]) continue;
      const std::vector<int> kept = KeptSources(n, r, gray, cams, cfg);
      if (kept.empty()) continue;
      std::vector<const void*> sp;
      std::vector<int32_t> sw, sh;
      std::vector<recon_mvs_camera> sc;
      for (int si : kept) {
        sp.push_back(gray[si]);
        sw.push_back(cams[si].width);
        sh.push_back(cams[si].height);
        sc.push_back(cams[si]);
      }
      recon_mvs_problem p;
      std::memset(&p, 0, sizeof(p));
      p.ref_pixels = gray[r];
      p.ref_w = cams[r].width;
      p.ref_h = cams[r].height;
      p.num_src = static_cast<int32_t>(kept.size());
      p.pixel_format = RECON_MVS_PIXEL_U8;
      p.src_pixels = sp.data();
      p.src_w = sw.data();
      p.src_h = sh.data();
      p.ref_camera = cams[r];
      p.src_cameras = sc.data();
      alloc(r);
      uint32_t filled = 0, refined = 0;
      st = recon_mvs_hip_patchmatch_stereo(ctx, &p, cfg, nullptr, 0, dep[r].data(), sup[r].data(),
                                           conf[r].data(), &filled);
      if (st != RECON_MVS_OK || filled == 0) continue;
      st = recon_mvs_hip_refine_subpixel(ctx, &p, cfg, dep[r].data(), conf[r].data(), &refined);
      if (st == RECON_MVS_OK) made.push_back(r);
    }
  } else {
    // every keyframe with an image becomes a resident view
    std::vector<int> view_of(n, -1), frame_of;
    std::vector<recon_mvs_view> views;
    for (int k = 0; k < n; ++k) {
      if (!gray[k]) continue;
      view_of[k] = static_cast<int>(views.size());
      frame_of.push_back(k);
      recon_mvs_view v;
      std::memset(&v, 0, sizeof(v));
      v.pixels = gray[k];
      v.pixel_format = RECON_MVS_PIXEL_U8;
      v.camera = cams[k];
      views.push_back(v);
    }
    const recon_mvs_plane_params pp = plane ? *plane : DerivedPlaneParams(cfg);
    std::vector<recon_mvs_batch_job> jobs;
    std::vector<std::vector<int>> kept_of(n);
    for (int r = 0; r < n; ++r) {
      if (!gray[r]) continue;
      std::vector<int> kept = KeptSources(n, r, gray, cams, cfg);
      if (kept.size() > 16) kept.resize(16);
      if (kept.empty()) continue;
      kept_of[r] = kept;
      recon_mvs_batch_job j;
      std::memset(&j, 0, sizeof(j));
      j.ref_view = view_of[r];
      j.num_src = static_cast<int32_t>(kept.size());
      for (size_t q = 0; q < kept.size(); ++q) j.src_views[q] = view_of[kept[q]];
      jobs.push_back(j);
    }
    if (!jobs.empty()) {
      // 8-bit keyframes stored as 8 bits: the results equal FP32 storage's
      // bit for bit at a quarter of the bytes.  The (11, 2) window has its
      // sweep, batch and polish kernels in every storage format; other
      // windows polish in FP32 only.
      const int32_t storage = (pp.patch_size == 11 && pp.window_stride == 2) ? RECON_MVS_PIXEL_U8
                                                                             : RECON_MVS_PIXEL_F32;
      st = recon_mvs_hip_set_views(ctx, static_cast<int32_t>(views.size()), views.data(), storage);
      recon_mvs_hip_batch* b = nullptr;
      if (st == RECON_MVS_OK)
        st = recon_mvs_hip_batch_create(ctx, static_cast<int32_t>(jobs.size()), jobs.data(), &pp, &b);
      if (st == RECON_MVS_OK) st = recon_mvs_hip_batch_run(ctx, b);
      if (b) recon_mvs_hip_batch_destroy(ctx, b);
      // a cancel during the batch drained it early: its maps are partial
      if (st == RECON_MVS_OK && rmvs::CancelRequested(ctx)) return rmvs::Cancelled(ctx);
      // the sub-pixel step (RefineSubpixel's place, orchestrator.cpp:1033-1036):
      // refinement along each ray, then the 10-step gradient polish where its
      // window is compiled
      const bool polish = (pp.patch_size == 11 && pp.window_stride == 2) ||
                          (pp.patch_size == 7 && pp.window_stride == 1) ||
                          (pp.patch_size == 5 && pp.window_stride == 1);
      for (const auto& j : jobs) {
        if (st != RECON_MVS_OK) break;
        if (rmvs::CancelRequested(ctx)) return rmvs::Cancelled(ctx);
        st = recon_mvs_hip_plane_refine(ctx, j.ref_view, j.num_src, j.src_views, &pp);
        if (st == RECON_MVS_OK && polish)
          st = recon_mvs_hip_plane_polish(ctx, j.ref_view, j.num_src, j.src_views, &pp, 10);
      }
      // consistency of each map against its sources that have maps, then
      // the map with inconsistent pixels dropped
      std::vector<char> has_map(n, 0);
      for (const auto& j : jobs) has_map[frame_of[j.ref_view]] = 1;
      for (int r = 0; r < n && st == RECON_MVS_OK; ++r) {
        if (!has_map[r]) continue;
        if (rmvs::CancelRequested(ctx)) return rmvs::Cancelled(ctx);
        std::vector<int32_t> list = {view_of[r]};
        for (int si : kept_of[r])
          if (has_map[si] && list.size() < 16) list.push_back(view_of[si]);
        if (list.size() > 1)
          st = recon_mvs_hip_geo_consistency(ctx, static_cast<int32_t>(list.size()), list.data(),
                                             cfg->geo_consistency_threshold, cfg->min_view_consistency);
        if (st != RECON_MVS_OK) break;
        alloc(r);
        st = recon_mvs_hip_get_depth_map(ctx, view_of[r],
                                         list.size() > 1 ? std::max(1, cfg->min_view_consistency) : 0,
                                         dep[r].data(), nullptr, nullptr, sup[r].data(),
                                         conf[r].data(), nullptr);
        if (st == RECON_MVS_OK) made.push_back(r);
      }
    }
  }
  if (st != RECON_MVS_OK) return st;
  // fusion over the produced maps, in keyframe order (orchestrator.cpp:1043-1054)
  *produced = static_cast<int32_t>(made.size());
  if (!made.empty()) {
    std::vector<recon_mvs_depth_map> maps;
    std::vector<recon_mvs_camera> mc;
    for (int r : made) {
      recon_mvs_depth_map m;
      m.width = cams[r].width;
      m.height = cams[r].height;
      m.depth = dep[r].data();
      m.support_count = sup[r].data();
      m.confidence = conf[r].data();
      maps.push_back(m);
      mc.push_back(cams[r]);
    }
    st = recon_mvs_hip_fuse_depth_maps(ctx, static_cast<int32_t>(maps.size()), maps.data(), mc.data(),
                                       cfg, num_points);
    if (st != RECON_MVS_OK) return st;
  }
  for (int r = 0; r < n; ++r) {
    const size_t P = static_cast<size_t>(cams[r].width) * cams[r].height;
    const bool has = std::find(made.begin(), made.end(), r) != made.end();
    if (depth && depth[r]) {
      if (has) std::memcpy(depth[r], dep[r].data(), P * sizeof(float));
      else std::fill(depth[r], depth[r] + P, __builtin_nanf(""));
    }
    if (support && support[r]) {
      if (has) std::memcpy(support[r], sup[r].data(), P);
      else std::memset(support[r], 0, P);
    }
    if (confidence && confidence[r]) {
      if (has) std::memcpy(confidence[r], conf[r].data(), P * sizeof(float));
      else std::fill(confidence[r], confidence[r] + P, 0.0f);
    }
  }
  return RECON_MVS_OK;
}
