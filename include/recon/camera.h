// recon/camera.h — pinhole + Brown-Conrady camera and world->camera pose.
// Same types and members as the reference's core/include/recon/camera.h:24-78;
// the methods are implemented in the MI355X library (mvs_api.cpp) with the
// reference's arithmetic (core/src/util/camera.cpp:33-162).
#ifndef RECON_CAMERA_H_
#define RECON_CAMERA_H_

#include "recon/linalg.h"

namespace recon {

struct CameraIntrinsics {
  double fx = 1.0, fy = 1.0, cx = 0.0, cy = 0.0;
  double k1 = 0.0, k2 = 0.0, k3 = 0.0, p1 = 0.0, p2 = 0.0;
  int width = 0, height = 0;

  void ApplyDistortion(double x, double y, double* out_x, double* out_y) const;
  void RemoveDistortion(double x_d, double y_d, double* out_x, double* out_y) const;
  void PixelToNormalized(double u, double v, double* x, double* y) const;
  void NormalizedToPixel(double x, double y, double* u, double* v) const;
  void PixelToRay(double u, double v, Vec3d* ray) const;
  bool ProjectPoint(const Vec3d& p_cam, double* u, double* v) const;
  double AxialDepth(const Vec3d& p_cam) const { return p_cam(2); }
};

// World -> camera.
struct CameraPose {
  Mat3d R = Mat3d::Identity();
  Vec3d t = Vec3d::Zero();
  Vec3d CameraToWorld(const Vec3d& p_cam) const { return R.transpose() * (p_cam - t); }
  Vec3d WorldToCamera(const Vec3d& p_world) const { return R * p_world + t; }
  Vec3d CenterWorld() const { return -(R.transpose() * t); }
};

inline Vec3d Baseline(const CameraPose& ref, const CameraPose& src) {
  return src.CenterWorld() - ref.CenterWorld();
}

inline Mat3d RelativeRotation(const CameraPose& ref, const CameraPose& src) {
  return src.R * ref.R.transpose();
}

}  // namespace recon

#endif  // RECON_CAMERA_H_
