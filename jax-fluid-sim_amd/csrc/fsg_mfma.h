// 16x16x4 matrix-core tiles (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64)
// for the batched element operators.
//
// Lane layout of one wavefront (64 lanes) for C[16x16] += A[16x4] B[4x16]:
//   A operand: lane l holds A[l % 16][l / 16]
//   B operand: lane l holds B[l / 16][l % 16]
//   C / D:     lane l, element j holds C[crow(l, j)][l % 16], j = 0..3, with
//              crow = 4 (l / 16) + j for f32 and (l / 16) + 4 j for f64
// (measured on the device with fsg_mfma_selftest).
#pragma once
#include <hip/hip_runtime.h>

namespace fsg {

template <typename real> struct Acc;
template <> struct Acc<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
};
template <> struct Acc<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
};

__device__ __forceinline__ Acc<float>::v4 mma(float a, float b, Acc<float>::v4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Acc<double>::v4 mma(double a, double b, Acc<double>::v4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row of the 16x16 tile held by element j of lane l's accumulator.
template <typename real>
__device__ __forceinline__ int crow(int l, int j) {
  return sizeof(real) == 4 ? 4 * (l >> 4) + j : (l >> 4) + 4 * j;
}

template <typename real>
__device__ __forceinline__ typename Acc<real>::v4 zero4() {
  typename Acc<real>::v4 z = {real(0), real(0), real(0), real(0)};
  return z;
}

}  // namespace fsg
