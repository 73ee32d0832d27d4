/*
 * pm_oracle_plane.c — TEST INFRASTRUCTURE ONLY (the oracle).
 * FP64 restatement of the plane-mode (slanted plane + homography) path.
 */
#include "pm_oracle.h"

/* lowbias32 (shaders/common/gpu_algo.h:150-154; shaders/metal/gpu_common.h:239-244). */
uint32_t orc_lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
