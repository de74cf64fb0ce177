// Launch dispatch from runtime (element type, face type, degree, nvars,
// precision) to the compile-time kernel instantiations.
#pragma once
#include "fsg_kernels.h"
#include "fsg_elem_mfma.h"
#include "fsg_hex_sf.h"

namespace fsg {

// Threads per element in the vector element kernel: one wave when the
// element has at most 64 volume points (hexahedra up to p = 2, where a
// second wave would idle), else two (the points are the widest loop).
constexpr int elem_threads(int T, int P) { return n_vol_q(T, P) <= 64 ? 64 : 128; }

template <int P, int V, typename real>
int launch_element(int type, int64_t count, const ElemArgs<real>& a, const Phys<real>& ph,
                   hipStream_t s, int mfma);

template <int P, int V, typename real>
int launch_face(int face_type, int owner_type, int neighbor_type, int sub, int64_t count,
                const FaceArgs<real>& a, const Phys<real>& ph, hipStream_t s);

#define FSG_DEFINE_LAUNCHERS(P, V, real)                                                        \
  template <>                                                                                   \
  int launch_element<P, V, real>(int type, int64_t count, const ElemArgs<real>& a,              \
                                 const Phys<real>& ph, hipStream_t s, int mfma) {              \
    if (count <= 0) return 0;                                                                   \
    if (mfma) {                                                                                 \
      switch (type) {                                                                           \
        FSG_MFMA_CASE(HEX, P, V, real)                                                          \
        FSG_MFMA_CASE(WEDGE, P, V, real)                                                        \
        FSG_MFMA_CASE(TET, P, V, real)                                                          \
        default: return -2;                                                                     \
      }                                                                                         \
      return hipGetLastError() == hipSuccess ? 0 : -1;                                          \
    }                                                                                           \
    if constexpr (P >= 2) {                                                                     \
      if (type == HEX && a.sf1d) {                                                              \
        constexpr int NTS = P == 3 ? 256 : 64;                                                  \
        const dim3 gs((unsigned)pass_grid(count, a.mask != nullptr));                           \
        if (a.mask)                                                                             \
          hipLaunchKernelGGL((hex_sf_kernel<P, V, real, NTS, true>), gs, dim3(NTS), 0, s, a, ph); \
        else                                                                                    \
          hipLaunchKernelGGL((hex_sf_kernel<P, V, real, NTS, false>), gs, dim3(NTS), 0, s, a, ph); \
        return hipGetLastError() == hipSuccess ? 0 : -1;                                        \
      }                                                                                         \
    }                                                                                           \
    constexpr int NTH = elem_threads(HEX, P), NTW = elem_threads(WEDGE, P);                    \
    constexpr int NTT = elem_threads(TET, P);                                                   \
    dim3 grid((unsigned)count);                                                                 \
    switch (type) {                                                                             \
      case HEX: hipLaunchKernelGGL((element_kernel<HEX, P, V, real, NTH>), grid, dim3(NTH), 0, s, a, ph); break; \
      case WEDGE: hipLaunchKernelGGL((element_kernel<WEDGE, P, V, real, NTW>), grid, dim3(NTW), 0, s, a, ph); break; \
      case TET: hipLaunchKernelGGL((element_kernel<TET, P, V, real, NTT>), grid, dim3(NTT), 0, s, a, ph); break; \
      default: return -2;                                                                       \
    }                                                                                           \
    return hipGetLastError() == hipSuccess ? 0 : -1;                                            \
  }                                                                                             \
  template <>                                                                                   \
  int launch_face<P, V, real>(int ft, int to, int tn, int sub, int64_t count,                   \
                              const FaceArgs<real>& a, const Phys<real>& ph, hipStream_t s) {   \
    if (count <= 0) return 0;                                                                   \
    dim3 grid((unsigned)pass_grid(count, a.mask_o != nullptr)), block(64);                      \
    const int key = ((ft * 2 + sub) * 4 + to) * 4 + (tn + 1);                                   \
    switch (key) {                                                                              \
      FSG_FACE_CASE(QUAD, 0, HEX, HEX, P, V, real)                                              \
      FSG_FACE_CASE(QUAD, 0, HEX, WEDGE, P, V, real)                                            \
      FSG_FACE_CASE(QUAD, 0, HEX, -1, P, V, real)                                               \
      FSG_FACE_CASE(QUAD, 0, WEDGE, HEX, P, V, real)                                            \
      FSG_FACE_CASE(QUAD, 0, WEDGE, WEDGE, P, V, real)                                          \
      FSG_FACE_CASE(QUAD, 0, WEDGE, -1, P, V, real)                                             \
      FSG_FACE_CASE(TRI, 0, WEDGE, WEDGE, P, V, real)                                           \
      FSG_FACE_CASE(TRI, 0, WEDGE, TET, P, V, real)                                             \
      FSG_FACE_CASE(TRI, 0, WEDGE, -1, P, V, real)                                              \
      FSG_FACE_CASE(TRI, 0, TET, WEDGE, P, V, real)                                             \
      FSG_FACE_CASE(TRI, 0, TET, TET, P, V, real)                                               \
      FSG_FACE_CASE(TRI, 0, TET, -1, P, V, real)                                                \
      FSG_FACE_CASE(TRI, 1, HEX, WEDGE, P, V, real)                                             \
      FSG_FACE_CASE(TRI, 1, HEX, TET, P, V, real)                                               \
      FSG_FACE_CASE(TRI, 1, WEDGE, WEDGE, P, V, real)                                           \
      FSG_FACE_CASE(TRI, 1, WEDGE, TET, P, V, real)                                             \
      default: return -3;                                                                       \
    }                                                                                           \
    return hipGetLastError() == hipSuccess ? 0 : -1;                                            \
  }

#define FSG_MFMA_CASE(T, P, V, real)                                                           \
  case T: {                                                                                     \
    constexpr int EB = MfmaShape<T, P, V, real>::EB;                                            \
    dim3 grid((unsigned)((count + EB - 1) / EB)), block(256);                                   \
    hipLaunchKernelGGL((element_mfma_kernel<T, P, V, real>), grid, block, 0, s, a, ph, count);  \
    break;                                                                                      \
  }

// WALE and the restricted pass are separate instantiations (the plain full
// pass keeps its registers).
#define FSG_FACE_CASE(FT, SUB, TO, TN, P, V, real)                                             \
  case ((FT * 2 + SUB) * 4 + TO) * 4 + (TN + 1):                                               \
    if (a.mask_o) {                                                                             \
      if (ph.wale)                                                                              \
        hipLaunchKernelGGL((face_kernel<FT, TO, TN, SUB, P, V, real, true, true>), grid, block, \
                           0, s, a, ph);                                                        \
      else                                                                                      \
        hipLaunchKernelGGL((face_kernel<FT, TO, TN, SUB, P, V, real, false, true>), grid, block, \
                           0, s, a, ph);                                                        \
    } else if (ph.wale) {                                                                       \
      hipLaunchKernelGGL((face_kernel<FT, TO, TN, SUB, P, V, real, true, false>), grid, block, 0, \
                         s, a, ph);                                                             \
    } else {                                                                                    \
      hipLaunchKernelGGL((face_kernel<FT, TO, TN, SUB, P, V, real, false, false>), grid, block, \
                         0, s, a, ph);                                                          \
    }                                                                                           \
    break;

}  // namespace fsg
