// context.hip — per-GPU context of the MVS library: creation, scratch
// buffers, pinned staging, error text, kernel timers.
#include <cstring>

#include "mvs_common.hpp"

namespace rmvs {

void* Ensure(recon_mvs_hip_context* ctx, Slot s, size_t bytes) {
  DevBuf& b = ctx->slots[s];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    // the stream may still read the old buffer
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  want += want / 8;  // headroom for slightly larger later calls
  if (hipMalloc(&b.ptr, want) != hipSuccess) {
    b.ptr = nullptr;
    return nullptr;
  }
  b.bytes = want;
  return b.ptr;
}

void* EnsurePinned(recon_mvs_hip_context* ctx, int i, size_t bytes) {
  if (ctx->pinned_bytes[i] >= bytes && ctx->pinned[i]) return ctx->pinned[i];
  if (ctx->pinned[i]) {
    if (hipEventSynchronize(ctx->pinned_ready[i]) != hipSuccess) return nullptr;
    (void)hipHostFree(ctx->pinned[i]);
    ctx->pinned[i] = nullptr;
  }
  if (hipHostMalloc(&ctx->pinned[i], bytes, hipHostMallocDefault) != hipSuccess) {
    ctx->pinned[i] = nullptr;
    ctx->pinned_bytes[i] = 0;
    return nullptr;
  }
  ctx->pinned_bytes[i] = bytes;
  return ctx->pinned[i];
}

void TimerBegin(recon_mvs_hip_context* ctx, int k, hipEvent_t* b, hipEvent_t* e) {
  *b = nullptr;
  *e = nullptr;
  if (!ctx->timing || k < 0 || k >= 8) return;
  for (hipEvent_t* p : {b, e}) {
    if (!ctx->ev_pool.empty()) {
      *p = ctx->ev_pool.back();
      ctx->ev_pool.pop_back();
    } else if (hipEventCreate(p) != hipSuccess) {
      *p = nullptr;
    }
  }
  if (*b) (void)hipEventRecord(*b, ctx->stream);
}

void TimerEnd(recon_mvs_hip_context* ctx, int k, hipEvent_t b, hipEvent_t e) {
  if (!b || !e) return;
  (void)hipEventRecord(e, ctx->stream);
  ctx->ev_pending.push_back({k, {b, e}});
}

void TimerCollect(recon_mvs_hip_context* ctx) {
  for (auto& pe : ctx->ev_pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(pe.second.second) == hipSuccess &&
        hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      ctx->timed_ms[pe.first] += ms;
      ctx->timed_count[pe.first] += 1;
    }
    ctx->ev_pool.push_back(pe.second.first);
    ctx->ev_pool.push_back(pe.second.second);
  }
  ctx->ev_pending.clear();
}

}  // namespace rmvs

using namespace rmvs;

void DestroyPlaneState(recon_mvs_hip_context* ctx);  // plane_mode.hip
namespace rmvs {
int32_t AdjointScratchInit(int device);  // plane_adjoint.hip
}

extern "C" {

RECON_MVS_API void recon_mvs_config_default(recon_mvs_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->num_views = 7;
  c->ref_view_index = -1;
  c->ideal_angle_min_deg = 8.0;
  c->ideal_angle_max_deg = 35.0;
  c->max_view_scale_mismatch = 0.15;
  c->view_angle_weight = 1.0;
  c->view_texture_weight = 0.8;
  c->view_spread_weight = 0.6;
  c->view_scale_weight = 0.4;
  c->reference_depth_m = 0.40;
  c->patch_size = 11;
  c->iterations = 4;
  c->random_search_count = 5;
  c->propagation_directions = 4;
  c->random_search_radius_ratio = 0.30;
  c->restart_ratio = 0.05;
  c->propagation_enabled = 1;
  c->num_threads = 0;
  c->photo_threshold = 1.0;
  c->depth_regularization = 0.02;
  c->geo_consistency_threshold = 0.02;
  c->weak_texture_variance = 12.0;
  c->subpixel_iterations = 4;
  c->subpixel_fit_points = 5;
  c->subpixel_window_ratio = 0.005;
  c->min_depth_m = 0.10;
  c->max_depth_m = 1.50;
  c->min_view_consistency = 2;
  c->fusion_voxel_size_m = 0.001;
}

RECON_MVS_API int32_t recon_mvs_hip_create(int32_t device, recon_mvs_hip_context** out) {
  if (!out) return RECON_MVS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RECON_MVS_ERR_GPU_FAILURE;
  if (device < 0 || device >= n) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return RECON_MVS_ERR_GPU_FAILURE;
  auto* ctx = new (std::nothrow) recon_mvs_hip_context();
  if (!ctx) return RECON_MVS_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  // the device address of the allocation-free adjoint's scratch, once per device
  if (rmvs::AdjointScratchInit(device) != RECON_MVS_OK) {
    delete ctx;
    return RECON_MVS_ERR_GPU_FAILURE;
  }
  // both streams at the default priority
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return RECON_MVS_ERR_GPU_FAILURE;
  }
  if (hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    ctx->copy_stream = nullptr;
    (void)recon_mvs_hip_destroy(ctx);
    return RECON_MVS_ERR_GPU_FAILURE;
  }
  // the cancel flag: coherent pinned host memory, read by kernels through its
  // device mapping
  if (hipHostMalloc(reinterpret_cast<void**>(&ctx->cancel_host), 64,
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    ctx->cancel_host = nullptr;
    (void)recon_mvs_hip_destroy(ctx);
    return RECON_MVS_ERR_GPU_FAILURE;
  }
  *ctx->cancel_host = 0;
  if (hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->cancel_dev), ctx->cancel_host, 0) !=
      hipSuccess) {
    (void)recon_mvs_hip_destroy(ctx);
    return RECON_MVS_ERR_GPU_FAILURE;
  }
  for (int i = 0; i < 2; ++i) {
    if (hipEventCreateWithFlags(&ctx->pinned_ready[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->tab_free[i], hipEventDisableTiming) != hipSuccess) {
      (void)recon_mvs_hip_destroy(ctx);
      return RECON_MVS_ERR_GPU_FAILURE;
    }
    (void)hipEventRecord(ctx->pinned_ready[i], ctx->stream);
    (void)hipEventRecord(ctx->tab_free[i], ctx->stream);
  }
  *out = ctx;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_destroy(recon_mvs_hip_context* ctx) {
  if (!ctx) return RECON_MVS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  DestroyPlaneState(ctx);
  for (auto& b : ctx->slots)
    if (b.ptr) (void)hipFree(b.ptr);
  for (int i = 0; i < 2; ++i) {
    if (ctx->pinned[i]) (void)hipHostFree(ctx->pinned[i]);
    if (ctx->pinned_ready[i]) (void)hipEventDestroy(ctx->pinned_ready[i]);
    if (ctx->tab_free[i]) (void)hipEventDestroy(ctx->tab_free[i]);
  }
  TimerCollect(ctx);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  (void)hipStreamDestroy(ctx->stream);
  if (ctx->cancel_host) (void)hipHostFree(ctx->cancel_host);
  delete ctx;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_cancel(recon_mvs_hip_context* ctx) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  __atomic_store_n(ctx->cancel_host, 1, __ATOMIC_RELEASE);
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_is_cancelled(recon_mvs_hip_context* ctx) {
  if (!ctx) return 0;
  return CancelRequested(ctx) ? 1 : 0;
}

RECON_MVS_API int32_t recon_mvs_hip_clear_cancel(recon_mvs_hip_context* ctx) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  __atomic_store_n(ctx->cancel_host, 0, __ATOMIC_RELEASE);
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_set_cancel_callback(recon_mvs_hip_context* ctx,
                                                        recon_mvs_cancel_fn fn, void* user) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  ctx->cancel_fn = fn;
  ctx->cancel_user = user;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_last_status(const recon_mvs_hip_context* ctx) {
  return ctx ? ctx->last_status : RECON_MVS_ERR_INVALID_ARGUMENT;
}

RECON_MVS_API const char* recon_mvs_hip_last_error(const recon_mvs_hip_context* ctx) {
  return ctx ? ctx->error.c_str() : "null context";
}

RECON_MVS_API int32_t recon_mvs_hip_synchronize(recon_mvs_hip_context* ctx) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

}  // extern "C"
