// mvs_stage.h — the MVS stage of a reconstruction as one call, for callers
// of the recon:: API (Orchestrator::StageMvs, core/src/orchestrator/
// orchestrator.cpp:942-1071, is a member over the orchestrator's own state;
// this is the same body over keyframe arrays).  Implemented on the C ABI
// recon_mvs_hip_run_mvs_stage (include/recon_mvs_hip.h).
#ifndef RECON_MVS_STAGE_H_
#define RECON_MVS_STAGE_H_

#include <cstdint>
#include <vector>

#include "recon/mvs.h"

namespace recon {

enum class MvsStageMode {
  kReference = 0,     // per keyframe PatchMatchStereo + RefineSubpixel (the reference's algorithm)
  kPlaneBatched = 1,  // all keyframes in one GPU plane-mode batch + geometric consistency
};

struct MvsStageResult {
  std::vector<DepthMap> depth_maps;  // per keyframe; width 0 = not produced
  PointCloud cloud;                  // FuseDepthMaps over the produced maps
  uint32_t produced = 0;             // 0: RECON_ERR_MVS_NO_DEPTH for the caller
};

// gray[k]: keyframe k's uint8 image (intrinsics[k].width x height), nullptr
// = no image; poses / intrinsics: every keyframe.  Throws MvsGpuError on a
// GPU failure.
uint32_t RunMvsStage(const std::vector<const uint8_t*>& gray, const std::vector<CameraPose>& poses,
                     const std::vector<CameraIntrinsics>& intrinsics, const MvsConfig& config,
                     MvsStageMode mode, MvsStageResult* out);

}  // namespace recon

#endif  // RECON_MVS_STAGE_H_
