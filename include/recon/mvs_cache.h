// mvs_cache.h — what the MVS stage hands downstream, for callers of the
// recon:: API: the fused cloud in the Poisson input form (OrientedPointCloud,
// core/include/recon/poisson_mesh.h:25-38, converted as
// UnifiedPointCloud::ToOrientedCloud does, core/src/orchestrator/
// orchestrator.cpp:181-192), and depth maps spilled to / restored from the
// session cache (docs/13_内存与降级.md §3.3.2-3.3.4).  Implemented on the C
// ABI (include/recon_mvs_hip.h: recon_mvs_spill_depth_map,
// recon_mvs_load_depth_map).
#ifndef RECON_MVS_CACHE_H_
#define RECON_MVS_CACHE_H_

#include <string>
#include <vector>

#include "recon/mvs.h"

namespace recon {

struct OrientedPointCloud {
  std::vector<Vec3d> positions;
  std::vector<Vec3d> normals;  // outward
  std::vector<float> colors;   // RGB 0-255, 3 per point
  std::vector<float> weights;  // >= 0

  size_t size() const { return positions.size(); }
  void Clear() { positions.clear(); normals.clear(); colors.clear(); weights.clear(); }
};

// positions / normals / colors copied; weights[i] = max(0, confidences[i]),
// 1 where the cloud has no confidence for point i.
OrientedPointCloud ToOrientedCloud(const PointCloud& cloud);

// Writes `map` as <cache_dir>/mvs_depth/ref_%05d.f32 and records it in
// <cache_dir>/cache_manifest.json.  Returns the file's sha256 (hex), "" on
// failure (bad argument, I/O error).
std::string SpillDepthMap(const std::string& cache_dir, int ref_index, const DepthMap& map);

// Reads the map back into *map.  false when it is missing or fails
// verification (a failed item is deleted): the caller recomputes it.
bool LoadDepthMap(const std::string& cache_dir, int ref_index, DepthMap* map);

}  // namespace recon

#endif  // RECON_MVS_CACHE_H_
