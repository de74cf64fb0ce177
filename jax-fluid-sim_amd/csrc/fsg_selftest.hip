// Device check of the matrix-core lane layout assumed by the batched element
// kernel (no counterpart in the reference): C = A B for one 16x16x16 product
// through the same 16x16x4 tiles.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fluid_gfx950.h"
#include "fsg_context.h"
#include "fsg_mfma.h"

namespace {

template <typename real>
__global__ __launch_bounds__(64) void selftest_kernel(const real* A, const real* B, real* C) {
  const int l = threadIdx.x;
  typename fsg::Acc<real>::v4 acc = fsg::zero4<real>();
  for (int k0 = 0; k0 < 16; k0 += 4) {
    const real a = A[(l % 16) * 16 + k0 + l / 16];
    const real b = B[(k0 + l / 16) * 16 + l % 16];
    acc = fsg::mma(a, b, acc);
  }
  for (int j = 0; j < 4; ++j) C[fsg::crow<real>(l, j) * 16 + l % 16] = acc[j];
}

}  // namespace

extern "C" int fsg_mfma_selftest(int32_t dtype, const void* A, const void* B, void* C,
                                 void* stream) {
  if (!A || !B || !C) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == FSG_F64)
    hipLaunchKernelGGL(selftest_kernel<double>, dim3(1), dim3(64), 0, s,
                       static_cast<const double*>(A), static_cast<const double*>(B),
                       static_cast<double*>(C));
  else
    hipLaunchKernelGGL(selftest_kernel<float>, dim3(1), dim3(64), 0, s,
                       static_cast<const float*>(A), static_cast<const float*>(B),
                       static_cast<float*>(C));
  return hipGetLastError() == hipSuccess ? 0 : fsg_fail(-1, "self-test launch failed");
}
