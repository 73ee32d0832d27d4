// mvs_gpu.h — GPU-side controls of the recon:: MVS entry points that the
// reference's CPU path has no use for: the status behind a thrown
// recon::MvsGpuError, and cooperative cancellation of running GPU work.
//
// Neither has a counterpart in the reference's mvs.h.  They exist for the
// two places where the reference's own callers meet the GPU library:
//   - errors: the reference's C ABI wraps every call in Guard
//     (core/src/util/c_api.cpp:31-42), which maps any std::exception to
//     RECON_ERR_INTERNAL; RECON_ERR_MVS_GPU_FAILURE (502,
//     recon_c_api.h:176) reaches the app only where a caller catches
//     recon::MvsGpuError (status()) or reads LastMvsGpuStatus();
//   - cancellation: recon_cancel (recon_c_api.h:601-608) sets the app
//     context's flag, which Orchestrator::CheckCancel reads between
//     keyframes (orchestrator.cpp:989); SetMvsCancelCallback hands that
//     flag to the GPU library so a running depth map stops at its next
//     dispatch batch (docs/01_接口契约.md §3.11).
#ifndef RECON_MVS_GPU_H_
#define RECON_MVS_GPU_H_

namespace recon {

// Status (recon_status integers) of the last failure of a recon:: MVS entry
// point on this thread: 502 for a GPU failure, 2 out of memory, 5
// cancelled; 0 when none since ClearMvsGpuStatus.
int LastMvsGpuStatus();
void ClearMvsGpuStatus();

// Cancellation of the GPU work behind the recon:: MVS entry points.  fn
// (may be null) is polled between dispatch batches; a nonzero answer
// cancels.  A cancelled PatchMatchStereo / RefineSubpixel returns 0 (the
// reference's "no depth, skip" value, orchestrator.cpp:1030), FuseDepthMaps
// an empty cloud, RunMvsStage 0 produced; LastMvsGpuStatus() is then 5.
// CancelMvs is thread-safe (callable while another thread is inside an entry
// point); the state stays set until ClearMvsCancel.
void SetMvsCancelCallback(int (*fn)(void* user), void* user);
void CancelMvs();
void ClearMvsCancel();
bool MvsCancelled();

}  // namespace recon

#endif  // RECON_MVS_GPU_H_
