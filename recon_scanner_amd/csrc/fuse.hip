// fuse.hip — FuseDepthMaps (core/src/mvs/patchmatch_stereo.cpp:985-1089) on
// the GPU, bit-exact with the reference (compiled with -ffp-contract=off).
//
// The reference walks every map pixel in (map, y, x) order, hashes its voxel
// into an int64 key ((vx*73856093) ^ (vy*19349663) ^ (vz*83492791), line
// 1045) and accumulates position, normal and confidence per key in an
// unordered_map.  Here:
//   1. k_fuse_points: one thread per map pixel computes the filter, the world
//      point, the outward normal and the key (HBM-bound streaming);
//   2. a stable flagged select compacts the surviving pixel indices;
//   3. a stable radix sort orders them by key, so the members of each voxel
//      stay in (map, y, x) order;
//   4. k_fuse_reduce: one thread per voxel sums its members sequentially in
//      that order, which reproduces the reference's double sums bit for bit.
// Points come out in ascending key order (the reference's order is the hash
// map's iteration order, which callers cannot rely on).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cstring>

#include "mvs_common.hpp"

namespace rmvs {

struct FuseMap {
  int width, height;
  uint64_t offset;  // first pixel of this map in the concatenated arrays
};

__global__ void k_fuse_points(const FuseMap* __restrict__ maps, const float* __restrict__ depth,
                              const uint8_t* __restrict__ support, const float* __restrict__ conf,
                              const recon_mvs_camera* __restrict__ cams, int min_view,
                              double photo_threshold, double voxel, int64_t* __restrict__ keys,
                              double* __restrict__ pw, double* __restrict__ nw,
                              uint8_t* __restrict__ flags) {
  const int m = blockIdx.y;
  const FuseMap fm = maps[m];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fm.width * fm.height) return;
  const uint64_t g = fm.offset + i;
  const float d = depth[g];
  bool keep = isfinite(d) && support[g] >= min_view;
  if (keep && photo_threshold < 1.0) {
    const float photo = 1.0f - conf[g];
    if (static_cast<double>(photo) > photo_threshold) keep = false;
  }
  flags[g] = keep ? 1 : 0;
  if (!keep) return;
  const recon_mvs_camera& c = cams[m];
  const int x = i % fm.width, y = i / fm.width;
  double ray[3];
  PixelToRay(c, x + 0.5, y + 0.5, ray);
  const double dd = static_cast<double>(d);
  const double pc[3] = {ray[0] * dd, ray[1] * dd, ray[2] * dd};
  double p[3];
  CamToWorld(c, pc, p);
  double rt[3];
  MatTVec(c.R, ray, rt);
  const double nn = Norm3(rt);
  double n[3];
  for (int j = 0; j < 3; ++j) n[j] = nn < 1e-15 ? 0.0 : rt[j] / nn;
  const int64_t vx = static_cast<int64_t>(floor(p[0] / voxel));
  const int64_t vy = static_cast<int64_t>(floor(p[1] / voxel));
  const int64_t vz = static_cast<int64_t>(floor(p[2] / voxel));
  keys[g] = static_cast<int64_t>((static_cast<uint64_t>(vx) * 73856093ULL) ^
                                 (static_cast<uint64_t>(vy) * 19349663ULL) ^
                                 (static_cast<uint64_t>(vz) * 83492791ULL));
  for (int j = 0; j < 3; ++j) {
    pw[3 * g + j] = p[j];
    nw[3 * g + j] = -n[j];
  }
}

__global__ void k_gather_keys(uint32_t n, const uint32_t* __restrict__ idx,
                              const int64_t* __restrict__ keys, int64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = keys[idx[i]];
}

__global__ void k_boundaries(uint32_t n, const int64_t* __restrict__ k, uint8_t* __restrict__ f) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

__global__ void k_fuse_reduce(uint32_t nseg, uint32_t n, const uint32_t* __restrict__ seg,
                              const uint32_t* __restrict__ idx, const int64_t* __restrict__ keys,
                              const double* __restrict__ pw, const double* __restrict__ nw,
                              const float* __restrict__ conf, double* __restrict__ out_pos,
                              double* __restrict__ out_nrm, float* __restrict__ out_conf,
                              int64_t* __restrict__ out_key) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const uint32_t b = seg[s];
  const uint32_t e = (s + 1 < nseg) ? seg[s + 1] : n;
  double sp[3] = {0.0, 0.0, 0.0}, sn[3] = {0.0, 0.0, 0.0}, sc = 0.0;
  uint32_t cnt = 0;
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t g = idx[i];
    for (int j = 0; j < 3; ++j) {
      sp[j] += pw[3 * static_cast<uint64_t>(g) + j];
      sn[j] += nw[3 * static_cast<uint64_t>(g) + j];
    }
    sc += conf[g];
    ++cnt;
  }
  const double nf = static_cast<double>(cnt);
  double nrm[3] = {sn[0], sn[1], sn[2]};
  if (Norm3(nrm) > 1e-12) {
    const double l = Norm3(nrm);
    if (!(l < 1e-15))
      for (int j = 0; j < 3; ++j) nrm[j] /= l;
  } else {
    nrm[0] = 0.0;
    nrm[1] = 0.0;
    nrm[2] = 1.0;
  }
  const double cf = fmin(1.0, (sc / nf) * sqrt(nf));
  for (int j = 0; j < 3; ++j) {
    out_pos[3 * s + j] = sp[j] / nf;
    out_nrm[3 * s + j] = nrm[j];
  }
  out_conf[s] = static_cast<float>(cf);
  out_key[s] = keys[b];
}

inline unsigned Blk(size_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace rmvs

using namespace rmvs;

extern "C" {

RECON_MVS_API int32_t recon_mvs_hip_fuse_depth_maps(recon_mvs_hip_context* ctx, int32_t num_maps,
                                                    const recon_mvs_depth_map* maps,
                                                    const recon_mvs_camera* cameras,
                                                    const recon_mvs_config* config,
                                                    int64_t* num_points) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!num_points || !config || num_maps < 0 || (num_maps > 0 && (!maps || !cameras)))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "fuse: bad arguments");
  *num_points = 0;
  ctx->cloud_pos.clear();
  ctx->cloud_nrm.clear();
  ctx->cloud_conf.clear();
  ctx->cloud_keys.clear();
  const double voxel = config->fusion_voxel_size_m;
  if (num_maps == 0 || voxel <= 0.0) return RECON_MVS_OK;  // lines 990-993
  std::vector<FuseMap> fm(num_maps);
  uint64_t total = 0;
  int max_px = 0;
  for (int m = 0; m < num_maps; ++m) {
    const recon_mvs_depth_map& d = maps[m];
    fm[m].width = d.depth ? d.width : 0;  // a null map is skipped (line 1009)
    fm[m].height = d.depth ? d.height : 0;
    if (d.depth && (d.width <= 0 || d.height <= 0 || !d.support_count || !d.confidence))
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "fuse: bad depth map");
    fm[m].offset = total;
    total += static_cast<uint64_t>(fm[m].width) * fm[m].height;
    max_px = std::max(max_px, fm[m].width * fm[m].height);
  }
  if (total == 0) return RECON_MVS_OK;
  if (total >= (1ull << 31)) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "fuse: too many pixels");
  const uint32_t T = static_cast<uint32_t>(total);
  auto* d_depth = static_cast<float*>(Ensure(ctx, kSlotFuse0, total * 4));
  auto* d_sup = static_cast<uint8_t*>(Ensure(ctx, kSlotFuse1, total));
  auto* d_conf = static_cast<float*>(Ensure(ctx, kSlotFuse2, total * 4));
  auto* d_keys = static_cast<int64_t*>(Ensure(ctx, kSlotFuse3, total * 8 * 2));
  auto* d_pw = static_cast<double*>(Ensure(ctx, kSlotFuse4, total * 24));
  auto* d_nw = static_cast<double*>(Ensure(ctx, kSlotFuse5, total * 24));
  auto* d_flags = static_cast<uint8_t*>(Ensure(ctx, kSlotFuse6, total * 2));
  auto* d_idx = static_cast<uint32_t*>(Ensure(ctx, kSlotFuse7, total * 4 * 3 + 64));
  auto* d_meta = static_cast<uint8_t*>(
      Ensure(ctx, kSlotFuse8, sizeof(FuseMap) * num_maps + sizeof(recon_mvs_camera) * num_maps + 64));
  if (!d_depth || !d_sup || !d_conf || !d_keys || !d_pw || !d_nw || !d_flags || !d_idx || !d_meta)
    return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "fuse: device allocation failed");
  auto* d_maps = reinterpret_cast<FuseMap*>(d_meta);
  auto* d_cams = reinterpret_cast<recon_mvs_camera*>(d_meta + ((sizeof(FuseMap) * num_maps + 63) & ~size_t(63)));
  for (int m = 0; m < num_maps; ++m) {
    const size_t n = static_cast<size_t>(fm[m].width) * fm[m].height;
    if (!n) continue;
    RMVS_HIP(ctx, hipMemcpyAsync(d_depth + fm[m].offset, maps[m].depth, n * 4, hipMemcpyHostToDevice, ctx->stream));
    RMVS_HIP(ctx, hipMemcpyAsync(d_sup + fm[m].offset, maps[m].support_count, n, hipMemcpyHostToDevice, ctx->stream));
    RMVS_HIP(ctx, hipMemcpyAsync(d_conf + fm[m].offset, maps[m].confidence, n * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  RMVS_HIP(ctx, hipMemcpyAsync(d_maps, fm.data(), sizeof(FuseMap) * num_maps, hipMemcpyHostToDevice, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(d_cams, cameras, sizeof(recon_mvs_camera) * num_maps, hipMemcpyHostToDevice, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // fm must outlive its copy

  hipLaunchKernelGGL(k_fuse_points, dim3(Blk(max_px, 256), num_maps), dim3(256), 0, ctx->stream,
                     d_maps, d_depth, d_sup, d_conf, d_cams, config->min_view_consistency,
                     config->photo_threshold, voxel, d_keys, d_pw, d_nw, d_flags);
  RMVS_LAUNCH_CHECK(ctx);

  uint32_t* idx_c = d_idx;            // compacted pixel indices
  uint32_t* idx_s = d_idx + T;        // sorted by key
  uint32_t* seg = d_idx + 2 * T;      // segment starts
  uint32_t* d_count = d_idx + 3 * T;  // 2 counters
  int64_t* keys_c = d_keys + T;
  // temp storage sized for the largest of the three primitives
  size_t t1 = 0, t2 = 0, t3 = 0;
  rocprim::counting_iterator<uint32_t> it(0);
  RMVS_HIP(ctx, rocprim::select(nullptr, t1, it, d_flags, idx_c, d_count, total, ctx->stream));
  RMVS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, t2, keys_c, keys_c, idx_c, idx_s, T, 0, 64, ctx->stream));
  RMVS_HIP(ctx, rocprim::select(nullptr, t3, it, d_flags, seg, d_count + 1, total, ctx->stream));
  const size_t tmp_bytes = std::max(t1, std::max(t2, t3));
  // sorted keys need their own buffer: slot 9 holds [sorted keys | temp storage]
  auto* d_tmp = static_cast<uint8_t*>(Ensure(ctx, kSlotFuse9, total * 8 + tmp_bytes + 256));
  if (!d_tmp) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "fuse: temp allocation failed");
  int64_t* keys_sorted = reinterpret_cast<int64_t*>(d_tmp);
  void* temp = d_tmp + ((total * 8 + 255) & ~uint64_t(255));

  size_t tb = tmp_bytes;
  RMVS_HIP(ctx, rocprim::select(temp, tb, it, d_flags, idx_c, d_count, total, ctx->stream));
  uint32_t n_valid = 0;
  RMVS_HIP(ctx, hipMemcpyAsync(&n_valid, d_count, 4, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (n_valid == 0) return RECON_MVS_OK;
  hipLaunchKernelGGL(k_gather_keys, dim3(Blk(n_valid, 256)), dim3(256), 0, ctx->stream, n_valid,
                     idx_c, d_keys, keys_c);
  RMVS_LAUNCH_CHECK(ctx);
  tb = tmp_bytes;
  RMVS_HIP(ctx, rocprim::radix_sort_pairs(temp, tb, keys_c, keys_sorted, idx_c, idx_s, n_valid, 0, 64,
                                          ctx->stream));
  uint8_t* bflags = d_flags + total;
  hipLaunchKernelGGL(k_boundaries, dim3(Blk(n_valid, 256)), dim3(256), 0, ctx->stream, n_valid,
                     keys_sorted, bflags);
  RMVS_LAUNCH_CHECK(ctx);
  tb = tmp_bytes;
  RMVS_HIP(ctx, rocprim::select(temp, tb, it, bflags, seg, d_count + 1, static_cast<size_t>(n_valid),
                                ctx->stream));
  uint32_t nseg = 0;
  RMVS_HIP(ctx, hipMemcpyAsync(&nseg, d_count + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // outputs reuse the per-pixel scratch (nseg <= n_valid <= total)
  double* o_pos = static_cast<double*>(Ensure(ctx, kSlotQueryIn, static_cast<size_t>(nseg) * 48 + 64));
  auto* o_conf = static_cast<float*>(Ensure(ctx, kSlotQueryOut, static_cast<size_t>(nseg) * 12 + 64));
  if (!o_pos || !o_conf) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "fuse: output allocation failed");
  double* o_nrm = o_pos + 3 * static_cast<size_t>(nseg);
  int64_t* o_key = reinterpret_cast<int64_t*>(o_conf + ((nseg + 1) & ~1u));
  hipLaunchKernelGGL(k_fuse_reduce, dim3(Blk(nseg, 128)), dim3(128), 0, ctx->stream, nseg, n_valid,
                     seg, idx_s, keys_sorted, d_pw, d_nw, d_conf, o_pos, o_nrm, o_conf, o_key);
  RMVS_LAUNCH_CHECK(ctx);
  ctx->cloud_pos.resize(3 * static_cast<size_t>(nseg));
  ctx->cloud_nrm.resize(3 * static_cast<size_t>(nseg));
  ctx->cloud_conf.resize(nseg);
  ctx->cloud_keys.resize(nseg);
  RMVS_HIP(ctx, hipMemcpyAsync(ctx->cloud_pos.data(), o_pos, 24ull * nseg, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(ctx->cloud_nrm.data(), o_nrm, 24ull * nseg, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(ctx->cloud_conf.data(), o_conf, 4ull * nseg, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(ctx->cloud_keys.data(), o_key, 8ull * nseg, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *num_points = nseg;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_fetch_point_cloud(recon_mvs_hip_context* ctx, int64_t capacity,
                                                      double* positions, double* normals,
                                                      float* colors, float* confidences,
                                                      int64_t* voxel_keys) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  const int64_t n = static_cast<int64_t>(ctx->cloud_conf.size());
  if (capacity < n) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "fetch: capacity below cloud size");
  if (positions) std::memcpy(positions, ctx->cloud_pos.data(), 24 * n);
  if (normals) std::memcpy(normals, ctx->cloud_nrm.data(), 24 * n);
  if (confidences) std::memcpy(confidences, ctx->cloud_conf.data(), 4 * n);
  if (voxel_keys) std::memcpy(voxel_keys, ctx->cloud_keys.data(), 8 * n);
  if (colors) std::memset(colors, 0, 12 * n);  // colors are 0 (lines 1083-1085)
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_fetch_oriented_cloud(recon_mvs_hip_context* ctx, int64_t capacity,
                                                         double* positions, double* normals,
                                                         float* colors, float* weights,
                                                         int64_t* num_points) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  const int64_t n = static_cast<int64_t>(ctx->cloud_conf.size());
  if (num_points) *num_points = n;
  if (!positions && !normals && !colors && !weights) return RECON_MVS_OK;  // size query
  if (capacity < n) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "fetch: capacity below cloud size");
  if (positions) std::memcpy(positions, ctx->cloud_pos.data(), 24 * n);
  if (normals) std::memcpy(normals, ctx->cloud_nrm.data(), 24 * n);
  if (colors) std::memset(colors, 0, 12 * n);
  // the Poisson point weight is the confidence clamped at 0
  // (UnifiedPointCloud::ToOrientedCloud, orchestrator.cpp:181-192)
  if (weights)
    for (int64_t i = 0; i < n; ++i) weights[i] = std::max(0.0f, ctx->cloud_conf[i]);
  return RECON_MVS_OK;
}

}  // extern "C"
