// Lane exchanges for butterfly sums inside a wavefront, on DPP (data
// parallel primitives: a VALU operand read from another lane of the same
// row of 16) instead of ds_bpermute (an LDS-unit round trip with an address
// computation and a wait per exchange).
//
// butterfly<O>(x) returns the value of x on the partner lane of step O of an
// all-reduce over aligned groups of 2 O lanes:
//   O = 1, 2: the lane l ^ O (quad permutations [1,0,3,2] / [2,3,0,1]);
//   O = 4:    lane 7 - (l & 7) of the same half-row (row_half_mirror);
//   O = 8:    lane 15 - (l & 15) of the same row (row_mirror);
//   O >= 16:  the lane l ^ O, by ds_bpermute.
// For O = 4 and 8 the mirror partner lies in the other O-lane half of the
// group, so where x is already uniform over aligned groups of O lanes (the
// earlier steps of the same butterfly have run) the result equals x on lane
// l ^ O: a sum or extremum built from these steps is the one the xor
// butterfly gives, with the same operands in the same order, bit for bit.
// Every lane of the wavefront must be active (a DPP read of an inactive
// lane gives 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg {

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, true);
}

template <int O>
constexpr int butterfly_ctrl() {
  return O == 1 ? 0xB1 : O == 2 ? 0x4E : O == 4 ? 0x141 : 0x140;
}

template <int O>
__device__ __forceinline__ float butterfly(float x) {
  if constexpr (O >= 16) {
    return __shfl_xor(x, O, 64);
  } else {
    return __builtin_bit_cast(float, dpp_i32<butterfly_ctrl<O>()>(__builtin_bit_cast(int, x)));
  }
}

template <int O>
__device__ __forceinline__ double butterfly(double x) {
  if constexpr (O >= 16) {
    return __shfl_xor(x, O, 64);
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)dpp_i32<butterfly_ctrl<O>()>((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)dpp_i32<butterfly_ctrl<O>()>((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
  }
}

// x summed over aligned groups of G lanes (G a power of two <= 64), every
// lane of a group holding the group's sum: the xor butterfly 1, 2, ..., G/2.
template <int G, typename real>
__device__ __forceinline__ real group_sum(real x) {
  if constexpr (G >= 2) x += butterfly<1>(x);
  if constexpr (G >= 4) x += butterfly<2>(x);
  if constexpr (G >= 8) x += butterfly<4>(x);
  if constexpr (G >= 16) x += butterfly<8>(x);
  if constexpr (G >= 32) x += butterfly<16>(x);
  if constexpr (G >= 64) x += butterfly<32>(x);
  return x;
}

}  // namespace fsg
