// plane_mode.hip — placeholder, replaced by the plane-mode kernels.
#include "mvs_common.hpp"
void DestroyPlaneState(recon_mvs_hip_context* ctx) { (void)ctx; }
