// Internal definition of the residual context, shared by the translation
// units of the library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/fluid_gfx950.h"

struct fsg_context {
  fsg_params params;
  std::vector<fsg_block_desc> blocks;
  std::vector<fsg_group_desc> groups;
  const int32_t* support;
  void* face_out;
  // Optional per-launch timing: an event pair around every launch, on the
  // launch stream; kind = group index, or ngroups + block index.  1 times
  // every launch, 2 only the launches of unmasked residual passes.
  int timing = 0;
  std::vector<hipEvent_t> pool;
  std::vector<int> kinds;
  size_t used = 0;
  int element_mfma = -1;   // FSG_OPT_ELEMENT_MFMA; -1: default (vector units)
  int hex_sum_fact = -1;  // FSG_OPT_HEX_SUM_FACTORISATION; -1: default (on)
};

// Records a formatted error for fsg_last_error() and returns code.
int fsg_fail(int code, const char* fmt, ...);
