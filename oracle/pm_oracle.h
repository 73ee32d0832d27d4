/*
 * pm_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * Shared helpers of the CPU oracle: a plain-C restatement of the reference's
 * camera and small-matrix arithmetic, written to perform the same double
 * operations in the same order as the reference (so results agree bit for
 * bit with a build without FMA contraction).  Only tests/, smoke() and
 * bench.py's CPU-baseline leg may load the library built from these files.
 */
#ifndef RECON_PM_ORACLE_H_
#define RECON_PM_ORACLE_H_

#include <math.h>
#include <stdint.h>

#include "recon_mvs_hip.h"

/* Mat * Vec, left-to-right sums starting from 0.0 (linalg.h:242-250). */
static inline void orc_matvec(const double R[9], const double v[3], double o[3]) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    s += R[i * 3 + 0] * v[0];
    s += R[i * 3 + 1] * v[1];
    s += R[i * 3 + 2] * v[2];
    o[i] = s;
  }
}

/* R^T * v, i.e. transpose() then Mat * Vec (linalg.h:252-257). */
static inline void orc_matTvec(const double R[9], const double v[3], double o[3]) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    s += R[0 * 3 + i] * v[0];
    s += R[1 * 3 + i] * v[1];
    s += R[2 * 3 + i] * v[2];
    o[i] = s;
  }
}

static inline double orc_norm3(const double v[3]) {
  double s = 0.0;
  s += v[0] * v[0];
  s += v[1] * v[1];
  s += v[2] * v[2];
  return sqrt(s);
}

/* CameraPose::CameraToWorld = R^T (p - t) (camera.h:64-66). */
static inline void orc_cam_to_world(const recon_mvs_camera* c, const double p[3],
                                    double o[3]) {
  const double d[3] = {p[0] - c->t[0], p[1] - c->t[1], p[2] - c->t[2]};
  orc_matTvec(c->R, d, o);
}

/* CameraPose::WorldToCamera = R p + t (camera.h:68-70). */
static inline void orc_world_to_cam(const recon_mvs_camera* c, const double p[3],
                                    double o[3]) {
  double q[3];
  orc_matvec(c->R, p, q);
  o[0] = q[0] + c->t[0];
  o[1] = q[1] + c->t[1];
  o[2] = q[2] + c->t[2];
}

/* CameraPose::CenterWorld = -(R^T t) (camera.h:72). */
static inline void orc_center(const recon_mvs_camera* c, double o[3]) {
  double q[3];
  orc_matTvec(c->R, c->t, q);
  o[0] = -q[0];
  o[1] = -q[1];
  o[2] = -q[2];
}

/* Brown-Conrady forward model (camera.cpp:33-45). */
static inline void orc_apply_distortion(const recon_mvs_camera* c, double x,
                                        double y, double* ox, double* oy) {
  const double r2 = x * x + y * y;
  const double r4 = r2 * r2;
  const double r6 = r4 * r2;
  const double radial = 1.0 + c->k1 * r2 + c->k2 * r4 + c->k3 * r6;
  const double pxy = c->p1 * x * y + c->p2 * (r2 + 2.0 * x * x);
  const double pxy2 = c->p1 * (r2 + 2.0 * y * y) + c->p2 * x * y;
  *ox = x * radial + pxy;
  *oy = y * radial + pxy2;
}

void orc_remove_distortion(const recon_mvs_camera* c, double xd, double yd,
                           double* ox, double* oy);

/* PixelToRay (camera.cpp:131-150). */
static inline void orc_pixel_to_ray(const recon_mvs_camera* c, double u, double v,
                                    double ray[3]) {
  const double xn = (u - c->cx) / c->fx;
  const double yn = (v - c->cy) / c->fy;
  orc_remove_distortion(c, xn, yn, &ray[0], &ray[1]);
  ray[2] = 1.0;
}

/* ProjectPoint (camera.cpp:152-162). */
static inline int orc_project(const recon_mvs_camera* c, const double p[3],
                              double* u, double* v) {
  if (p[2] <= 1e-9) return 0;
  const double x = p[0] / p[2];
  const double y = p[1] / p[2];
  double xd, yd;
  orc_apply_distortion(c, x, y, &xd, &yd);
  *u = c->fx * xd + c->cx;
  *v = c->fy * yd + c->cy;
  return 1;
}

#endif /* RECON_PM_ORACLE_H_ */
