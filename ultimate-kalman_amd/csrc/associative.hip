/*
 * associative.hip -- gfx950 kernels of the associative-scan smoother
 * (c/kalman_associative_smoother.c:309-1240).
 */
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ukhip_internal.h"

struct ukhip_batch;
extern "C" void uk_set_error(const char* fmt, ...);

extern "C" int uk_assoc_smooth(struct ukhip_batch* b, uint32_t options) {
  (void)b;
  (void)options;
  uk_set_error("kalman_smooth_associative: the associative kernels are not built yet");
  return -1;
}
