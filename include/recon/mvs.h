// recon/mvs.h — the MVS entry points of recon-scanner, served by the MI355X
// library.  Types and signatures are those of the reference's
// core/include/recon/mvs.h:24-214, so callers (Orchestrator::StageMvs,
// orchestrator.cpp:942-1071; tests/unit/test_mvs.cpp) link against this
// library unchanged.  Semantics, return values and the source index contract
// (mvs.h:151-161) are the reference's; PatchMatchStereo / RefineSubpixel /
// FuseDepthMaps run on HIP device 0 in reference mode (bit-exact with the
// reference CPU path for the same num_threads).  A GPU failure throws
// recon::MvsGpuError (status 502 = RECON_ERR_MVS_GPU_FAILURE,
// recon_c_api.h:176); the reference's C ABI guard turns exceptions into
// status codes (c_api.cpp:31-42).
#ifndef RECON_MVS_H_
#define RECON_MVS_H_

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "recon/camera.h"
#include "recon/linalg.h"

namespace recon {

struct MvsConfig {
  int num_views = 7;
  int ref_view_index = -1;
  double ideal_angle_min_deg = 8.0;
  double ideal_angle_max_deg = 35.0;
  double max_view_scale_mismatch = 0.15;
  double view_angle_weight = 1.0;
  double view_texture_weight = 0.8;
  double view_spread_weight = 0.6;
  double view_scale_weight = 0.4;
  double reference_depth_m = 0.40;
  int patch_size = 11;
  int iterations = 4;
  int random_search_count = 5;
  double random_search_radius_ratio = 0.30;
  double restart_ratio = 0.05;
  int propagation_directions = 4;
  bool propagation_enabled = true;
  int num_threads = 0;
  double photo_threshold = 1.0;
  double depth_regularization = 0.02;
  double geo_consistency_threshold = 0.02;
  double weak_texture_variance = 12.0;
  int subpixel_iterations = 4;
  int subpixel_fit_points = 5;
  double subpixel_window_ratio = 0.005;
  double min_depth_m = 0.10;
  double max_depth_m = 1.50;
  int min_view_consistency = 2;
  double fusion_voxel_size_m = 0.001;
};

struct NccResult {
  bool valid = false;
  double ncc = 0.0;
  bool weak_texture = false;
  int valid_samples = 0;
};

struct ViewSelection {
  int ref_index = 0;
  std::vector<int> source_indices;
  std::vector<double> weights;
};

struct DepthMap {
  int width = 0;
  int height = 0;
  std::vector<float> depth;            // metres along the optical axis, NaN = invalid
  std::vector<uint8_t> support_count;  // views supporting the best hypothesis
  std::vector<float> confidence;       // [0, 1]

  void Allocate(int w, int h) {
    width = w;
    height = h;
    const size_t n = static_cast<size_t>(w) * h;
    depth.assign(n, std::nanf(""));
    support_count.assign(n, 0);
    confidence.assign(n, 0.0f);
  }
  bool IsValid(int x, int y) const {
    if (x < 0 || y < 0 || x >= width || y >= height) return false;
    return std::isfinite(depth[static_cast<size_t>(y) * width + x]);
  }
};

struct PointCloud {
  std::vector<Vec3d> positions;
  std::vector<Vec3d> normals;
  std::vector<float> colors;
  std::vector<float> confidences;
  size_t size() const { return positions.size(); }
  void Clear() { positions.clear(); normals.clear(); colors.clear(); confidences.clear(); }
};

// Thrown when the GPU path fails; status() is 502 (RECON_ERR_MVS_GPU_FAILURE).
class MvsGpuError : public std::runtime_error {
 public:
  MvsGpuError(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

std::vector<int> SelectSourceViews(const CameraPose& ref_pose, const std::vector<CameraPose>& poses,
                                   const CameraIntrinsics& intrinsics, const MvsConfig& config);

NccResult ComputeNcc(const uint8_t* ref_img, int ref_w, int ref_h, const uint8_t* src_img,
                     int src_w, int src_h, double ref_u, double ref_v, double src_u, double src_v,
                     int patch_size);

uint32_t PatchMatchStereo(const uint8_t* ref_gray, int ref_w, int ref_h,
                          const std::vector<const uint8_t*>& src_gray,
                          const std::vector<int>& src_w, const std::vector<int>& src_h,
                          const CameraPose& ref_pose, const std::vector<CameraPose>& src_poses,
                          const CameraIntrinsics& ref_intr,
                          const std::vector<CameraIntrinsics>& src_intrinsics,
                          const std::vector<int>& source_indices, const MvsConfig& config,
                          const std::vector<std::array<double, 3>>* seed_points,
                          DepthMap* out_depth);

uint32_t RefineSubpixel(const uint8_t* ref_gray, int ref_w, int ref_h,
                        const std::vector<const uint8_t*>& src_gray,
                        const std::vector<int>& src_w, const std::vector<int>& src_h,
                        const CameraPose& ref_pose, const std::vector<CameraPose>& src_poses,
                        const CameraIntrinsics& ref_intr,
                        const std::vector<CameraIntrinsics>& src_intrinsics,
                        const std::vector<int>& source_indices, const MvsConfig& config,
                        DepthMap* depth);

PointCloud FuseDepthMaps(const std::vector<const DepthMap*>& depth_maps,
                         const std::vector<CameraPose>& poses,
                         const std::vector<CameraIntrinsics>& intrinsics, const MvsConfig& config);

}  // namespace recon

#endif  // RECON_MVS_H_
