// mvs_api.cpp — the C++ entry points of include/recon/mvs.h, implemented on
// the C ABI of this library (include/recon_mvs_hip.h).  Argument checks, the
// source index contract and the return-0 rules follow the reference
// (core/src/mvs/patchmatch_stereo.cpp:133-211, 434-485, 780-803, 985-994);
// the compute runs on HIP device 0.  Also the camera methods of
// include/recon/camera.h (core/src/util/camera.cpp:33-162).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "mvs_common.hpp"
#include "recon/camera.h"
#include "recon/mvs.h"
#include "recon/mvs_cache.h"
#include "recon/mvs_gpu.h"
#include "recon/mvs_stage.h"
#include "recon_mvs_hip.h"

namespace recon {
namespace {

recon_mvs_camera ToCam(const CameraPose& p, const CameraIntrinsics& k) {
  recon_mvs_camera c;
  std::memset(&c, 0, sizeof(c));
  c.fx = k.fx; c.fy = k.fy; c.cx = k.cx; c.cy = k.cy;
  c.k1 = k.k1; c.k2 = k.k2; c.k3 = k.k3; c.p1 = k.p1; c.p2 = k.p2;
  c.width = k.width; c.height = k.height;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) c.R[3 * i + j] = p.R(i, j);
    c.t[i] = p.t(i);
  }
  return c;
}

recon_mvs_camera IntrOnly(const CameraIntrinsics& k) { return ToCam(CameraPose(), k); }

recon_mvs_config ToConfig(const MvsConfig& m) {
  recon_mvs_config c;
  recon_mvs_config_default(&c);
  c.num_views = m.num_views;
  c.ref_view_index = m.ref_view_index;
  c.ideal_angle_min_deg = m.ideal_angle_min_deg;
  c.ideal_angle_max_deg = m.ideal_angle_max_deg;
  c.max_view_scale_mismatch = m.max_view_scale_mismatch;
  c.view_angle_weight = m.view_angle_weight;
  c.view_texture_weight = m.view_texture_weight;
  c.view_spread_weight = m.view_spread_weight;
  c.view_scale_weight = m.view_scale_weight;
  c.reference_depth_m = m.reference_depth_m;
  c.patch_size = m.patch_size;
  c.iterations = m.iterations;
  c.random_search_count = m.random_search_count;
  c.propagation_directions = m.propagation_directions;
  c.random_search_radius_ratio = m.random_search_radius_ratio;
  c.restart_ratio = m.restart_ratio;
  c.propagation_enabled = m.propagation_enabled ? 1 : 0;
  c.num_threads = m.num_threads;
  c.photo_threshold = m.photo_threshold;
  c.depth_regularization = m.depth_regularization;
  c.geo_consistency_threshold = m.geo_consistency_threshold;
  c.weak_texture_variance = m.weak_texture_variance;
  c.subpixel_iterations = m.subpixel_iterations;
  c.subpixel_fit_points = m.subpixel_fit_points;
  c.subpixel_window_ratio = m.subpixel_window_ratio;
  c.min_depth_m = m.min_depth_m;
  c.max_depth_m = m.max_depth_m;
  c.min_view_consistency = m.min_view_consistency;
  c.fusion_voxel_size_m = m.fusion_voxel_size_m;
  return c;
}

std::mutex g_mu;
recon_mvs_hip_context* g_ctx = nullptr;

// Cancellation of the recon:: entry points: CancelMvs (any thread) and the
// caller's callback, polled through the context's own callback at the
// library's check points.
std::atomic<bool> g_cancel{false};
std::atomic<int (*)(void*)> g_cancel_fn{nullptr};
std::atomic<void*> g_cancel_user{nullptr};

int32_t PollCancel(void*) {
  if (g_cancel.load(std::memory_order_acquire)) return 1;
  int (*fn)(void*) = g_cancel_fn.load(std::memory_order_acquire);
  return (fn && fn(g_cancel_user.load(std::memory_order_acquire))) ? 1 : 0;
}

recon_mvs_hip_context* Ctx() {
  if (!g_ctx) {
    const int32_t st = recon_mvs_hip_create(0, &g_ctx);
    if (st != RECON_MVS_OK) {
      g_ctx = nullptr;
      throw MvsGpuError(RECON_MVS_ERR_GPU_FAILURE, "recon_mvs_hip_create failed");
    }
    (void)recon_mvs_hip_set_cancel_callback(g_ctx, PollCancel, nullptr);
  }
  return g_ctx;
}

thread_local int g_last_status = 0;

// True when the call succeeded; false when it was cancelled (the caller
// returns the reference's "nothing produced" value); throws MvsGpuError
// otherwise.  Records the status for LastMvsGpuStatus.
bool Check(int32_t st, const char* what) {
  if (st == RECON_MVS_OK) return true;
  g_last_status = st;
  if (st == RECON_MVS_ERR_CANCELLED) return false;
  const int code = st == RECON_MVS_ERR_OUT_OF_MEMORY ? st : RECON_MVS_ERR_GPU_FAILURE;
  g_last_status = code;
  throw MvsGpuError(code, std::string(what) + ": " + recon_mvs_hip_last_error(g_ctx));
}

// Resolves the index contract (mvs.h:151-161); false where the reference returns 0.
bool Resolve(const std::vector<const uint8_t*>& src_gray, const std::vector<int>& src_w,
             const std::vector<int>& src_h, const std::vector<CameraPose>& src_poses,
             const std::vector<CameraIntrinsics>& src_intrinsics,
             const std::vector<int>& source_indices, std::vector<recon_mvs_camera>* cams) {
  const size_t n = source_indices.size();
  if (src_gray.size() != n || src_w.size() != n || src_h.size() != n) return false;
  const bool global = src_intrinsics.size() == src_poses.size();
  const bool relative = src_intrinsics.size() == n;
  if (!global && !relative) return false;
  if (global)
    for (int si : source_indices)
      if (si < 0 || static_cast<size_t>(si) >= src_poses.size()) return false;
  cams->clear();
  for (size_t k = 0; k < n; ++k) {
    const int si = source_indices[k];
    if (si < 0 || static_cast<size_t>(si) >= src_poses.size()) return false;
    cams->push_back(ToCam(src_poses[si], global ? src_intrinsics[si] : src_intrinsics[k]));
  }
  return true;
}

recon_mvs_problem MakeProblem(const uint8_t* ref, int w, int h,
                              const std::vector<const uint8_t*>& src_gray,
                              const std::vector<int>& src_w, const std::vector<int>& src_h,
                              const recon_mvs_camera& ref_cam,
                              const std::vector<recon_mvs_camera>& cams,
                              std::vector<const void*>* ptrs, std::vector<int32_t>* sw,
                              std::vector<int32_t>* sh) {
  ptrs->assign(src_gray.begin(), src_gray.end());
  sw->assign(src_w.begin(), src_w.end());
  sh->assign(src_h.begin(), src_h.end());
  recon_mvs_problem p;
  std::memset(&p, 0, sizeof(p));
  p.ref_pixels = ref;
  p.ref_w = w;
  p.ref_h = h;
  p.num_src = static_cast<int32_t>(cams.size());
  p.pixel_format = RECON_MVS_PIXEL_U8;
  p.src_pixels = ptrs->data();
  p.src_w = sw->data();
  p.src_h = sh->data();
  p.ref_camera = ref_cam;
  p.src_cameras = cams.data();
  return p;
}

}  // namespace

// ---- camera.h methods (camera.cpp:33-162) ----
void CameraIntrinsics::ApplyDistortion(double x, double y, double* ox, double* oy) const {
  rmvs::ApplyDistortion(IntrOnly(*this), x, y, ox, oy);
}
void CameraIntrinsics::RemoveDistortion(double xd, double yd, double* ox, double* oy) const {
  rmvs::RemoveDistortion(IntrOnly(*this), xd, yd, ox, oy);
}
void CameraIntrinsics::PixelToNormalized(double u, double v, double* x, double* y) const {
  RemoveDistortion((u - cx) / fx, (v - cy) / fy, x, y);
}
void CameraIntrinsics::NormalizedToPixel(double x, double y, double* u, double* v) const {
  *u = fx * x + cx;
  *v = fy * y + cy;
}
void CameraIntrinsics::PixelToRay(double u, double v, Vec3d* ray) const {
  double r[3];
  rmvs::PixelToRay(IntrOnly(*this), u, v, r);
  *ray = Vec3d(r[0], r[1], r[2]);
}
bool CameraIntrinsics::ProjectPoint(const Vec3d& p, double* u, double* v) const {
  const double q[3] = {p(0), p(1), p(2)};
  return rmvs::Project(IntrOnly(*this), q, u, v);
}

// ---- mvs_gpu.h: status and cancellation ----
int LastMvsGpuStatus() { return g_last_status; }
void ClearMvsGpuStatus() { g_last_status = 0; }

void SetMvsCancelCallback(int (*fn)(void* user), void* user) {
  g_cancel_user.store(user, std::memory_order_release);
  g_cancel_fn.store(fn, std::memory_order_release);
}

void CancelMvs() { g_cancel.store(true, std::memory_order_release); }

void ClearMvsCancel() {
  g_cancel.store(false, std::memory_order_release);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ctx) (void)recon_mvs_hip_clear_cancel(g_ctx);
}

bool MvsCancelled() { return PollCancel(nullptr) != 0; }

// ---- mvs.h entry points ----
std::vector<int> SelectSourceViews(const CameraPose& ref_pose, const std::vector<CameraPose>& poses,
                                   const CameraIntrinsics& intrinsics, const MvsConfig& config) {
  std::vector<recon_mvs_camera> cams;
  for (const auto& p : poses) cams.push_back(ToCam(p, intrinsics));
  const recon_mvs_camera ref = ToCam(ref_pose, intrinsics);
  const recon_mvs_config cfg = ToConfig(config);
  std::vector<int32_t> out(poses.size() + 1);
  const int n = recon_mvs_select_source_views(&ref, static_cast<int32_t>(cams.size()), cams.data(),
                                              &cfg, out.data(), static_cast<int32_t>(out.size()));
  return std::vector<int>(out.begin(), out.begin() + std::max(0, n));
}

NccResult ComputeNcc(const uint8_t* ref_img, int ref_w, int ref_h, const uint8_t* src_img,
                     int src_w, int src_h, double ref_u, double ref_v, double src_u, double src_v,
                     int patch_size) {
  std::lock_guard<std::mutex> lock(g_mu);
  const double q[4] = {ref_u, ref_v, src_u, src_v};
  recon_mvs_ncc_result r;
  Check(recon_mvs_hip_compute_ncc(Ctx(), ref_img, ref_w, ref_h, src_img, src_w, src_h, 1, q,
                                  patch_size, &r),
        "ComputeNcc");
  NccResult o;
  o.valid = r.valid != 0;
  o.ncc = r.ncc;
  o.weak_texture = r.weak_texture != 0;
  o.valid_samples = r.valid_samples;
  return o;
}

uint32_t PatchMatchStereo(const uint8_t* ref_gray, int ref_w, int ref_h,
                          const std::vector<const uint8_t*>& src_gray,
                          const std::vector<int>& src_w, const std::vector<int>& src_h,
                          const CameraPose& ref_pose, const std::vector<CameraPose>& src_poses,
                          const CameraIntrinsics& ref_intr,
                          const std::vector<CameraIntrinsics>& src_intrinsics,
                          const std::vector<int>& source_indices, const MvsConfig& config,
                          const std::vector<std::array<double, 3>>* seed_points,
                          DepthMap* out_depth) {
  if (ref_gray == nullptr || out_depth == nullptr) return 0;
  if (source_indices.empty()) return 0;
  std::vector<recon_mvs_camera> cams;
  if (!Resolve(src_gray, src_w, src_h, src_poses, src_intrinsics, source_indices, &cams)) return 0;
  out_depth->Allocate(ref_w, ref_h);
  std::vector<const void*> ptrs;
  std::vector<int32_t> sw, sh;
  const recon_mvs_problem p = MakeProblem(ref_gray, ref_w, ref_h, src_gray, src_w, src_h,
                                          ToCam(ref_pose, ref_intr), cams, &ptrs, &sw, &sh);
  const recon_mvs_config cfg = ToConfig(config);
  std::vector<double> seeds;
  if (seed_points)
    for (const auto& s : *seed_points) seeds.insert(seeds.end(), s.begin(), s.end());
  uint32_t filled = 0;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!Check(recon_mvs_hip_patchmatch_stereo(Ctx(), &p, &cfg, seeds.empty() ? nullptr : seeds.data(),
                                             static_cast<int32_t>(seeds.size() / 3),
                                             out_depth->depth.data(), out_depth->support_count.data(),
                                             out_depth->confidence.data(), &filled),
             "PatchMatchStereo")) {
    out_depth->Allocate(ref_w, ref_h);  // cancelled: nothing produced
    return 0;
  }
  return filled;
}

uint32_t RefineSubpixel(const uint8_t* ref_gray, int ref_w, int ref_h,
                        const std::vector<const uint8_t*>& src_gray,
                        const std::vector<int>& src_w, const std::vector<int>& src_h,
                        const CameraPose& ref_pose, const std::vector<CameraPose>& src_poses,
                        const CameraIntrinsics& ref_intr,
                        const std::vector<CameraIntrinsics>& src_intrinsics,
                        const std::vector<int>& source_indices, const MvsConfig& config,
                        DepthMap* depth) {
  if (ref_gray == nullptr || depth == nullptr) return 0;
  if (source_indices.empty()) return 0;
  std::vector<recon_mvs_camera> cams;
  if (!Resolve(src_gray, src_w, src_h, src_poses, src_intrinsics, source_indices, &cams)) return 0;
  if (depth->width != ref_w || depth->height != ref_h) return 0;
  std::vector<const void*> ptrs;
  std::vector<int32_t> sw, sh;
  const recon_mvs_problem p = MakeProblem(ref_gray, ref_w, ref_h, src_gray, src_w, src_h,
                                          ToCam(ref_pose, ref_intr), cams, &ptrs, &sw, &sh);
  const recon_mvs_config cfg = ToConfig(config);
  uint32_t refined = 0;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!Check(recon_mvs_hip_refine_subpixel(Ctx(), &p, &cfg, depth->depth.data(),
                                           depth->confidence.data(), &refined),
             "RefineSubpixel"))
    return 0;  // cancelled before the refinement ran: the map is unchanged
  return refined;
}

PointCloud FuseDepthMaps(const std::vector<const DepthMap*>& depth_maps,
                         const std::vector<CameraPose>& poses,
                         const std::vector<CameraIntrinsics>& intrinsics, const MvsConfig& config) {
  PointCloud cloud;
  if (depth_maps.empty() || config.fusion_voxel_size_m <= 0.0) return cloud;
  std::vector<recon_mvs_depth_map> maps(depth_maps.size());
  std::vector<recon_mvs_camera> cams(depth_maps.size());
  for (size_t i = 0; i < depth_maps.size(); ++i) {
    std::memset(&maps[i], 0, sizeof(maps[i]));
    cams[i] = ToCam(poses[i], intrinsics[i]);
    const DepthMap* d = depth_maps[i];
    if (!d) continue;
    maps[i].width = d->width;
    maps[i].height = d->height;
    maps[i].depth = d->depth.data();
    maps[i].support_count = d->support_count.data();
    maps[i].confidence = d->confidence.data();
  }
  const recon_mvs_config cfg = ToConfig(config);
  std::lock_guard<std::mutex> lock(g_mu);
  int64_t n = 0;
  Check(recon_mvs_hip_fuse_depth_maps(Ctx(), static_cast<int32_t>(maps.size()), maps.data(),
                                      cams.data(), &cfg, &n),
        "FuseDepthMaps");
  std::vector<double> pos(3 * n), nrm(3 * n);
  cloud.confidences.resize(n);
  cloud.colors.assign(3 * n, 0.0f);
  Check(recon_mvs_hip_fetch_point_cloud(Ctx(), n, pos.data(), nrm.data(), cloud.colors.data(),
                                        cloud.confidences.data(), nullptr),
        "FuseDepthMaps");
  cloud.positions.resize(n);
  cloud.normals.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    cloud.positions[i] = Vec3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    cloud.normals[i] = Vec3d(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
  return cloud;
}

uint32_t RunMvsStage(const std::vector<const uint8_t*>& gray, const std::vector<CameraPose>& poses,
                     const std::vector<CameraIntrinsics>& intrinsics, const MvsConfig& config,
                     MvsStageMode mode, MvsStageResult* out) {
  if (!out) return 0;
  *out = MvsStageResult();
  const size_t n = gray.size();
  if (poses.size() != n || intrinsics.size() != n) return 0;
  std::vector<recon_mvs_camera> cams(n);
  for (size_t k = 0; k < n; ++k) cams[k] = ToCam(poses[k], intrinsics[k]);
  out->depth_maps.resize(n);
  std::vector<float*> dptr(n, nullptr), cptr(n, nullptr);
  std::vector<uint8_t*> sptr(n, nullptr);
  for (size_t k = 0; k < n; ++k) {
    if (!gray[k]) continue;
    out->depth_maps[k].Allocate(intrinsics[k].width, intrinsics[k].height);
    dptr[k] = out->depth_maps[k].depth.data();
    sptr[k] = out->depth_maps[k].support_count.data();
    cptr[k] = out->depth_maps[k].confidence.data();
  }
  const recon_mvs_config cfg = ToConfig(config);
  int32_t produced = 0;
  int64_t npts = 0;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!Check(recon_mvs_hip_run_mvs_stage(Ctx(), static_cast<int32_t>(n), gray.data(), cams.data(),
                                         &cfg, static_cast<int32_t>(mode), nullptr, dptr.data(),
                                         sptr.data(), cptr.data(), &produced, &npts),
             "RunMvsStage")) {
    *out = MvsStageResult();  // cancelled: nothing produced
    return 0;
  }
  for (size_t k = 0; k < n; ++k) {
    DepthMap& d = out->depth_maps[k];
    bool any = false;
    for (float z : d.depth) any = any || std::isfinite(z);
    if (!any) d = DepthMap();
  }
  std::vector<double> pos(3 * npts), nrm(3 * npts);
  out->cloud.confidences.resize(npts);
  out->cloud.colors.assign(3 * npts, 0.0f);
  Check(recon_mvs_hip_fetch_point_cloud(Ctx(), npts, pos.data(), nrm.data(), out->cloud.colors.data(),
                                        out->cloud.confidences.data(), nullptr),
        "RunMvsStage");
  out->cloud.positions.resize(npts);
  out->cloud.normals.resize(npts);
  for (int64_t i = 0; i < npts; ++i) {
    out->cloud.positions[i] = Vec3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    out->cloud.normals[i] = Vec3d(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
  out->produced = static_cast<uint32_t>(produced);
  return out->produced;
}

OrientedPointCloud ToOrientedCloud(const PointCloud& cloud) {
  OrientedPointCloud out;
  out.positions = cloud.positions;
  out.normals = cloud.normals;
  out.colors = cloud.colors;
  out.weights.assign(cloud.positions.size(), 1.0f);
  for (size_t i = 0; i < cloud.positions.size() && i < cloud.confidences.size(); ++i)
    out.weights[i] = std::max(0.0f, cloud.confidences[i]);
  return out;
}

std::string SpillDepthMap(const std::string& cache_dir, int ref_index, const DepthMap& map) {
  const size_t n = static_cast<size_t>(std::max(map.width, 0)) * std::max(map.height, 0);
  if (n == 0 || map.depth.size() != n || map.support_count.size() != n || map.confidence.size() != n)
    return "";
  char hex[65] = {0};
  if (recon_mvs_spill_depth_map(cache_dir.c_str(), ref_index, map.width, map.height, map.depth.data(),
                                map.support_count.data(), map.confidence.data(), hex) != RECON_MVS_OK)
    return "";
  return hex;
}

bool LoadDepthMap(const std::string& cache_dir, int ref_index, DepthMap* map) {
  if (!map) return false;
  int32_t w = 0, h = 0;
  if (recon_mvs_load_depth_map(cache_dir.c_str(), ref_index, 0, &w, &h, nullptr, nullptr, nullptr) !=
      RECON_MVS_OK)
    return false;
  DepthMap m;
  m.Allocate(w, h);
  if (recon_mvs_load_depth_map(cache_dir.c_str(), ref_index, static_cast<int64_t>(w) * h, &w, &h,
                               m.depth.data(), m.support_count.data(), m.confidence.data()) != RECON_MVS_OK)
    return false;
  *map = std::move(m);
  return true;
}

}  // namespace recon

// ---- C ABI: SelectSourceViews (patchmatch_stereo.cpp:133-211), host only ----
extern "C" RECON_MVS_API int32_t recon_mvs_select_source_views(const recon_mvs_camera* ref,
                                                               int32_t n,
                                                               const recon_mvs_camera* poses,
                                                               const recon_mvs_config* cfg,
                                                               int32_t* out, int32_t capacity) {
  if (!ref || !cfg || n < 0 || (n > 0 && !poses)) return -1;
  struct Cand {
    int index;
    double score;
  };
  std::vector<Cand> c;
  double rc[3];
  rmvs::Center(*ref, rc);
  for (int i = 0; i < n; ++i) {
    if (i == cfg->ref_view_index) continue;
    double sc[3];
    rmvs::Center(poses[i], sc);
    const double b[3] = {sc[0] - rc[0], sc[1] - rc[1], sc[2] - rc[2]};
    if (rmvs::Norm3(b) < 1e-6) continue;
    const double z_ref = cfg->reference_depth_m;
    const double z_src = z_ref + b[2];
    if (z_src <= 1e-6) continue;
    if (std::fabs(z_ref / z_src - 1.0) > cfg->max_view_scale_mismatch) continue;
    double rel[3];
    rmvs::MatVec(ref->R, b, rel);
    const double dist = rmvs::Norm3(rel);
    if (dist < 1e-6) continue;
    const double ang = std::acos(std::clamp(rel[2] / dist, -1.0, 1.0)) * 180.0 / recon::kPi;
    double score;
    if (ang < cfg->ideal_angle_min_deg) score = ang / cfg->ideal_angle_min_deg;
    else if (ang > cfg->ideal_angle_max_deg)
      score = std::max(0.0, 1.0 - (ang - cfg->ideal_angle_max_deg) / (cfg->ideal_angle_max_deg * 2.0));
    else score = 1.0;
    c.push_back({i, score * cfg->view_angle_weight});
  }
  std::stable_sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
  int m = 0;
  for (size_t i = 0; i < c.size() && m < cfg->num_views; ++i) {
    if (m < capacity && out) out[m] = c[i].index;
    ++m;
  }
  return m;
}
