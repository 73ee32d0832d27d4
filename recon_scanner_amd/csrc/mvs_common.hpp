// mvs_common.hpp — shared pieces of the MI355X MVS library: the per-GPU
// context, error plumbing, and FP64 camera arithmetic usable on host and
// device.
//
// The camera functions perform the same double operations in the same order
// as the reference (core/src/util/camera.cpp:33-162, core/include/recon/
// linalg.h:242-257, camera.h:64-78).  Files that must reproduce the reference
// bit for bit are compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "recon_mvs_hip.h"

#define RMVS_HD __host__ __device__ __forceinline__

namespace rmvs {

// ---------------------------------------------------------------------------
// Device buffers and the context
// ---------------------------------------------------------------------------

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Scratch slots of a context.  Each slot grows on demand and is reused.
enum Slot : int {
  kSlotRefImg = 0,
  kSlotSrcImg,
  kSlotCams,
  kSlotInvA,
  kSlotInvB,
  kSlotCostA,
  kSlotCostB,
  kSlotSupA,
  kSlotSupB,
  kSlotRng0,
  kSlotRng1,
  kSlotOutDepth,
  kSlotOutSupport,
  kSlotOutConf,
  kSlotCounter,
  kSlotQueryIn,
  kSlotQueryOut,
  kSlotQueryImgA,
  kSlotQueryImgB,
  kSlotFuse0,
  kSlotFuse1,
  kSlotFuse2,
  kSlotFuse3,
  kSlotFuse4,
  kSlotFuse5,
  kSlotFuse6,
  kSlotFuse7,
  kSlotFuse8,
  kSlotFuse9,
  kSlotAdjIn,
  kSlotAdjOut,
  kSlotAdjPart,
  kSlotPhase,
  kSlotRestart,
  kSlotCount
};

}  // namespace rmvs

struct ResidentView;      // plane mode: one uploaded view (defined in plane_mode.hip)
struct PlaneState;        // plane mode: one depth-map slot

struct recon_mvs_hip_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  int32_t last_status = 0;  // status of the last failed call (0: none)
  // cooperative cancellation: a flag in pinned host memory mapped into the
  // device's address space (the persistent batch kernel polls it), set by
  // recon_mvs_hip_cancel from any thread, or latched from the caller's
  // callback at the library's check points
  int32_t* cancel_host = nullptr;
  int32_t* cancel_dev = nullptr;
  recon_mvs_cancel_fn cancel_fn = nullptr;
  void* cancel_user = nullptr;
  rmvs::DevBuf slots[rmvs::kSlotCount];
  void* pinned[2] = {nullptr, nullptr};
  size_t pinned_bytes[2] = {0, 0};
  hipEvent_t pinned_ready[2] = {nullptr, nullptr};
  // uploads that overlap the kernels on `stream` (reference mode's per-pass
  // random tables); tab_free[i]: the kernel reading device table i is done
  hipStream_t copy_stream = nullptr;
  hipEvent_t tab_free[2] = {nullptr, nullptr};
  // fusion result kept for recon_mvs_hip_fetch_point_cloud
  std::vector<double> cloud_pos, cloud_nrm;
  std::vector<float> cloud_conf;
  std::vector<int64_t> cloud_keys;
  // plane mode
  std::vector<ResidentView*> views;
  std::vector<PlaneState*> states;
  int storage_format = RECON_MVS_PIXEL_F32;
  // kernel timing (HIP events around the dominant kernel launches)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double timed_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t timed_count[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // autotuned tile schedule per (size, views, window, format)
  std::map<std::string, int> tuned_schedule;
  // autotuned per-phase occupancy caps (plane mode, tile_schedule = -1)
  std::map<std::string, std::vector<int>> tuned_caps;
  // tile orders of view-phased half-sweeps, per (reference, sources, tile)
  std::map<std::string, int*> phase_orders;
};

namespace rmvs {

inline int32_t Fail(recon_mvs_hip_context* ctx, int32_t code, const std::string& msg) {
  if (ctx) {
    ctx->error = msg;
    ctx->last_status = code;
  }
  return code;
}

// Whether cancellation has been requested: the context's flag, or the
// caller's callback (a yes is latched into the flag, so the device sees it
// too).  Checked between dispatch batches: per keyframe, per directional
// pass, per batch launch (docs/01_接口契约.md §3.11: GPU work is not
// preempted; the flag is checked between batches).
inline bool CancelRequested(recon_mvs_hip_context* ctx) {
  if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) return true;
  if (ctx->cancel_fn && ctx->cancel_fn(ctx->cancel_user)) {
    __atomic_store_n(ctx->cancel_host, 1, __ATOMIC_RELEASE);
    return true;
  }
  return false;
}

// Waits for the work already queued on the context's streams, then reports
// the cancellation.
inline int32_t Cancelled(recon_mvs_hip_context* ctx) {
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  return Fail(ctx, RECON_MVS_ERR_CANCELLED, "cancelled");
}

#define RMVS_HIP(ctx, expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      return ::rmvs::Fail((ctx), RECON_MVS_ERR_GPU_FAILURE,                    \
                          std::string(#expr) + ": " + hipGetErrorString(_e));  \
    }                                                                          \
  } while (0)

#define RMVS_LAUNCH_CHECK(ctx)                                                 \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) {                                                    \
      return ::rmvs::Fail((ctx), RECON_MVS_ERR_GPU_FAILURE,                    \
                          std::string("kernel launch: ") + hipGetErrorString(_e)); \
    }                                                                          \
  } while (0)

// Grows slot s of ctx to at least `bytes`; returns nullptr on failure.
void* Ensure(recon_mvs_hip_context* ctx, Slot s, size_t bytes);
// Pinned host staging buffer i (0/1), grown to `bytes`.
void* EnsurePinned(recon_mvs_hip_context* ctx, int i, size_t bytes);

// Timing of kernel class `k` (0..7): records events around a launch when
// ctx->timing is on.  Collect() folds finished pairs into timed_ms.
void TimerBegin(recon_mvs_hip_context* ctx, int k, hipEvent_t* b, hipEvent_t* e);
void TimerEnd(recon_mvs_hip_context* ctx, int k, hipEvent_t b, hipEvent_t e);
void TimerCollect(recon_mvs_hip_context* ctx);

// ---------------------------------------------------------------------------
// FP64 camera arithmetic, reference operation order
// ---------------------------------------------------------------------------

RMVS_HD void MatVec(const double* R, const double* v, double* o) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    s += R[i * 3 + 0] * v[0];
    s += R[i * 3 + 1] * v[1];
    s += R[i * 3 + 2] * v[2];
    o[i] = s;
  }
}

RMVS_HD void MatTVec(const double* R, const double* v, double* o) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    s += R[0 * 3 + i] * v[0];
    s += R[1 * 3 + i] * v[1];
    s += R[2 * 3 + i] * v[2];
    o[i] = s;
  }
}

RMVS_HD double Norm3(const double* v) {
  double s = 0.0;
  s += v[0] * v[0];
  s += v[1] * v[1];
  s += v[2] * v[2];
  return sqrt(s);
}

RMVS_HD void CamToWorld(const recon_mvs_camera& c, const double* p, double* o) {
  const double d[3] = {p[0] - c.t[0], p[1] - c.t[1], p[2] - c.t[2]};
  MatTVec(c.R, d, o);
}

RMVS_HD void WorldToCam(const recon_mvs_camera& c, const double* p, double* o) {
  double q[3];
  MatVec(c.R, p, q);
  o[0] = q[0] + c.t[0];
  o[1] = q[1] + c.t[1];
  o[2] = q[2] + c.t[2];
}

RMVS_HD void Center(const recon_mvs_camera& c, double* o) {
  double q[3];
  MatTVec(c.R, c.t, q);
  o[0] = -q[0];
  o[1] = -q[1];
  o[2] = -q[2];
}

RMVS_HD void ApplyDistortion(const recon_mvs_camera& c, double x, double y, double* ox,
                             double* oy) {
  const double r2 = x * x + y * y;
  const double r4 = r2 * r2;
  const double r6 = r4 * r2;
  const double radial = 1.0 + c.k1 * r2 + c.k2 * r4 + c.k3 * r6;
  const double pxy = c.p1 * x * y + c.p2 * (r2 + 2.0 * x * x);
  const double pxy2 = c.p1 * (r2 + 2.0 * y * y) + c.p2 * x * y;
  *ox = x * radial + pxy;
  *oy = y * radial + pxy2;
}

RMVS_HD bool HasDistortion(const recon_mvs_camera& c) {
  return !(c.k1 == 0.0 && c.k2 == 0.0 && c.k3 == 0.0 && c.p1 == 0.0 && c.p2 == 0.0);
}

// Damped Newton inversion of the distortion (camera.cpp:47-129).
RMVS_HD void RemoveDistortion(const recon_mvs_camera& c, double x_d, double y_d, double* ox,
                              double* oy) {
  if (!HasDistortion(c)) {
    *ox = x_d;
    *oy = y_d;
    return;
  }
  double x = x_d, y = y_d;
  for (int iter = 0; iter < 50; ++iter) {
    double xr, yr;
    ApplyDistortion(c, x, y, &xr, &yr);
    const double ex = x_d - xr;
    const double ey = y_d - yr;
    const double err = sqrt(ex * ex + ey * ey);
    if (err < 1e-15) break;
    const double r2 = x * x + y * y;
    const double r4 = r2 * r2;
    const double radial = 1.0 + c.k1 * r2 + c.k2 * r4 + c.k3 * r4 * r2;
    const double dr = c.k1 + 2.0 * c.k2 * r2 + 3.0 * c.k3 * r4;
    const double j00 = radial + 2.0 * x * x * dr + 2.0 * c.p1 * y + c.p2 * (3.0 * x * x + y * y);
    const double j01 = 2.0 * x * y * dr + c.p1 * (x * x + 3.0 * y * y) + 2.0 * c.p2 * x * y;
    const double j10 = 2.0 * x * y * dr + 2.0 * c.p1 * x + c.p2 * (x * x + 3.0 * y * y);
    const double j11 = radial + 2.0 * y * y * dr + c.p1 * (3.0 * x * x + y * y) + 2.0 * c.p2 * x * y;
    const double det = j00 * j11 - j01 * j10;
    if (fabs(det) < 1e-12) break;
    const double dx = (j11 * ex - j01 * ey) / det;
    const double dy = (-j10 * ex + j00 * ey) / det;
    double lambda = 1.0;
    bool accepted = false;
    for (int ls = 0; ls < 20; ++ls) {
      const double xn = x + lambda * dx;
      const double yn = y + lambda * dy;
      double xrn, yrn;
      ApplyDistortion(c, xn, yn, &xrn, &yrn);
      const double rn = sqrt((x_d - xrn) * (x_d - xrn) + (y_d - yrn) * (y_d - yrn));
      if (rn < err) {
        x = xn;
        y = yn;
        accepted = true;
        break;
      }
      lambda *= 0.5;
    }
    if (!accepted) break;
  }
  *ox = x;
  *oy = y;
}

RMVS_HD void PixelToRay(const recon_mvs_camera& c, double u, double v, double* ray) {
  const double xn = (u - c.cx) / c.fx;
  const double yn = (v - c.cy) / c.fy;
  RemoveDistortion(c, xn, yn, &ray[0], &ray[1]);
  ray[2] = 1.0;
}

RMVS_HD bool Project(const recon_mvs_camera& c, const double* p, double* u, double* v) {
  if (p[2] <= 1e-9) return false;
  const double x = p[0] / p[2];
  const double y = p[1] / p[2];
  double xd, yd;
  ApplyDistortion(c, x, y, &xd, &yd);
  *u = c.fx * xd + c.cx;
  *v = c.fy * yd + c.cy;
  return true;
}

RMVS_HD double DepthToInv(double d) { return d < 1e-9 ? 1e9 : 1.0 / d; }
RMVS_HD double InvToDepth(double x) { return x <= 1e-9 ? 1e9 : 1.0 / x; }

// Mean source baseline of a problem (patchmatch_stereo.cpp:472-479).
inline double MeanBaseline(const recon_mvs_camera& ref, const recon_mvs_camera* src, int n) {
  double rc[3];
  Center(ref, rc);
  double sum = 0.0;
  for (int k = 0; k < n; ++k) {
    double sc[3];
    Center(src[k], sc);
    const double d[3] = {sc[0] - rc[0], sc[1] - rc[1], sc[2] - rc[2]};
    sum += Norm3(d);
  }
  return sum / static_cast<double>(n);
}

}  // namespace rmvs
