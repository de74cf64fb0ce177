// MLP surrogate of the stiff air-11 chemistry substep (no counterpart in the
// reference solver).
//
// Training (FP32): tiled SGEMM forward with fused bias + SiLU epilogue,
// backward (activation derivative, weight/bias gradients, input gradient) and
// Adam.  Inference: one fused kernel per tile of cells on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): features -> 4 layers ->
// log-ratio softmax -> exact element/charge/mass projection (FP64 Newton) ->
// hull and convergence checks that flag cells for the plain integrator.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string.h>

#include <string>

#include "../../include/fluid_gfx950.h"

namespace fsgmlp {

// ---------------------------------------------------------------------------
// FP32 tiled GEMM: C[M,N] (+)= sum_k A(m,k) B(k,n), A(m,k) = TA ? A[k*lda+m] :
// A[m*lda+k], B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  64x64 tile, 256 threads,
// 4x4 outputs per thread, K step 16 through LDS.
// Epilogue: 0 store, 1 store + bias[n] (pre-activation Z) and H = silu(Z),
// 2 accumulate (C += result).
constexpr int TM = 64, TN = 64, TK = 16;

__device__ __forceinline__ double dmax_(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb,
                                                   float* __restrict__ C, int ldc,
                                                   const float* __restrict__ bias,
                                                   float* __restrict__ H, int M, int N, int K,
                                                   int silu_on, int kchunk, int64_t c_zstride) {
  __shared__ float sA[TK][TM + 4];
  __shared__ float sB[TK][TN + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  // blockIdx.z takes the k range [kbeg, kend) and writes its partial product
  // to C + z * c_zstride (split-k; one z and kchunk = K otherwise)
  const int kbeg = blockIdx.z * kchunk;
  const int kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  C += (int64_t)blockIdx.z * c_zstride;
  float acc[4][4] = {};
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    for (int e = threadIdx.x; e < TK * TM; e += 256) {
      int kk = e / TM, mm = e % TM;
      int m = m0 + mm, k = k0 + kk;
      float v = 0.f;
      if (m < M && k < kend) v = TA ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
      sA[kk][mm] = v;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      int kk = e / TN, nn = e % TN;
      int n = n0 + nn, k = k0 + kk;
      float v = 0.f;
      if (n < N && k < kend) v = TB ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n];
      sB[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sB[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int m = m0 + ty * 4 + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int n = n0 + tx * 4 + j;
      if (n >= N) continue;
      float* c = C + (int64_t)m * ldc + n;
      if (EPI == 0) {
        *c = acc[i][j];
      } else if (EPI == 1) {
        float z = acc[i][j] + bias[n];
        *c = z;
        H[(int64_t)m * ldc + n] = silu_on ? silu(z) : z;
      } else {
        *c += acc[i][j];
      }
    }
  }
}

// dZ = dH * silu'(Z) in place on dH.
__global__ void dsilu_kernel(float* __restrict__ g, const float* __restrict__ Z, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] *= dsilu(Z[i]);
}

// Partial column sums of dZ over row chunk z: part[z * N + n] = sum over
// rows [z * chunk, min(M, (z + 1) * chunk)) of dZ[m, n]; 256 threads stride
// the columns (N <= 256).
__global__ __launch_bounds__(256) void colsum_partial_kernel(const float* __restrict__ g,
                                                             float* __restrict__ part, int M,
                                                             int N, int chunk) {
  const int n = threadIdx.x;
  if (n >= N) return;
  const int mbeg = blockIdx.x * chunk;
  const int mend = mbeg + chunk < M ? mbeg + chunk : M;
  float s = 0.f;
  for (int m = mbeg; m < mend; ++m) s += g[(int64_t)m * N + n];
  part[(int64_t)blockIdx.x * N + n] = s;
}

// out[i] = sum over z = 0..S-1 of part[z * n + i], in order of z (the
// split-k reduction; deterministic).
__global__ void reduce_splits_kernel(const float* __restrict__ part, float* __restrict__ out,
                                     int64_t n, int S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < S; ++z) s += part[(int64_t)z * n + i];
  out[i] = s;
}

// Split of the batch dimension for the weight / bias gradients: chunks of
// at least 64 rows, at most 64 chunks.
inline int backward_splits(int M) {
  int S = (M + 63) / 64;
  return S < 1 ? 1 : (S > 64 ? 64 : S);
}

// dO = 2 (O - T) / (M N)
__global__ void mse_grad_kernel(const float* __restrict__ O, const float* __restrict__ T,
                                float* __restrict__ g, int64_t n, float scale) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = scale * (O[i] - T[i]);
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Inference on BF16 MFMA.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int HID = 128;
constexpr int NIN = 14;
constexpr int NOUT = 11;
constexpr int PROJ_ITERS = 32;   // damped dual Newton iterations, at most
constexpr int PROJ_ITERS_F32 = 8;  // of which first in FP32, at most
// Packed bf16 weights (A-operand order, see pack_weights in the host code):
//   W1: 8 out-tiles x 1 k-step, W2/W3: 8 x 4, W4: 1 x 4 ; each (tile, step)
//   block holds 64 lanes x 8 bf16.
constexpr int BLK = 64 * 8;
constexpr int W1_BLK = 8, W2_BLK = 32, W3_BLK = 32, W4_BLK = 4;
constexpr int W_TOTAL = (W1_BLK + W2_BLK + W3_BLK + W4_BLK) * BLK;

// k index held by lane l, slot j (K=32 step): the C-layout rows of two
// consecutive 16-row tiles, so a layer's accumulators feed the next layer's
// B operand without a transpose.
__host__ __device__ constexpr int perm_k(int l, int j) {
  return j < 4 ? (l / 16) * 4 + j : 16 + (l / 16) * 4 + (j - 4);
}

struct InferConsts {
  float mean[NIN], inv_std[NIN];
  float out_mean[16], out_std[16];
  float hull_lo[NIN], hull_hi[NIN];
  float eps_y;         // floor inside the log of the mass fractions
  float hull_margin;   // allowed excursion, in standard deviations
  double proj_tol;     // relative constraint residual accepted
  double molar_mass[NOUT];
  double charge[NOUT];
  double n_atoms[NOUT];
  double o_atoms[NOUT];
};

// Constraint rows of the projection per unit mass (N atoms, O atoms,
// charge), built once on the host: A[i][s] = (atoms or charge)_s / M_s.
struct ProjRows {
  double A[3][NOUT];
};

__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// SiLU of the inference path: the hardware reciprocal (1 ulp) in place of
// the IEEE division; the result is rounded to BF16 next, so the two agree
// after rounding but for rare ties, at a fraction of the VALU work (the
// activations are the kernel's largest VALU cost).
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// Converts two accumulator tiles (rows 0-15 and 16-31 of a K=32 step) to the
// B operand of the next layer after bias + SiLU.
__device__ __forceinline__ bf16x8 to_b(f32x4 lo, f32x4 hi, const float* blo, const float* bhi,
                                       int l) {
  bf16x8 r;
  const int base = (l / 16) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = (__bf16)silu_fast(lo[j] + blo[base + j]);
    r[4 + j] = (__bf16)silu_fast(hi[j] + bhi[base + j]);
  }
  return r;
}

// One wavefront = 16 cells.  Block = IW wavefronts sharing the weights of
// all layers, staged once per block in LDS (78 KB: one block per CU, so the
// block is wide to keep four waves per SIMD); each wave runs its own
// grid-stride loop over 16-cell tiles.
constexpr int IW = 16;
__global__ __launch_bounds__(IW * 64) void infer_kernel(const __bf16* __restrict__ Wp,
                                                    const float* __restrict__ bias,
                                                    const InferConsts cstv,
                                                    const double* __restrict__ rho_in,
                                                    const double* __restrict__ T_in,
                                                    const double* __restrict__ dt_cell,
                                                    double dt, double* __restrict__ rho_out,
                                                    int32_t* __restrict__ flag, int64_t n) {
  __shared__ __bf16 sW[W_TOTAL];
  __shared__ float sb[3 * HID + 16];
  __shared__ float sO[IW][16][17];
  for (int k = threadIdx.x; k < W_TOTAL / 8; k += IW * 64)
    reinterpret_cast<bf16x8*>(sW)[k] = reinterpret_cast<const bf16x8*>(Wp)[k];
  for (int k = threadIdx.x; k < 3 * HID + 16; k += IW * 64) sb[k] = bias[k];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cellw = lane & 15;
  const __bf16* W1 = sW;
  const __bf16* W2 = W1 + W1_BLK * BLK;
  const __bf16* W3 = W2 + W2_BLK * BLK;
  const __bf16* W4 = W3 + W3_BLK * BLK;
  const InferConsts* cst = &cstv;
  const int64_t ngroups = (n + 63) / 64;
  // no block barriers below: each wave's loop is its own.  A wave takes 64
  // cells: lane l computes the features of cell l once, then the four
  // 16-cell tiles run through the network in turn, each lane fetching its
  // B-operand slots from the lane that owns the tile's cell.
  for (int64_t grp = (int64_t)blockIdx.x * IW + wave; grp < ngroups;
       grp += (int64_t)gridDim.x * IW) {
    float feat[NIN];
    {
      const int64_t c = grp * 64 + lane;
      const bool valid = c < n;
      double r[NOUT], rt = 0.0;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) {
        r[s] = valid ? rho_in[c * NOUT + s] : 1.0;
        rt += r[s];
      }
      double T = valid ? T_in[c] : 1000.0;
      double h = dt_cell ? (valid ? dt_cell[c] : 1.0) : dt;
      float f[NIN];
      f[0] = (float)log(T);
      f[1] = (float)log(rt);
#pragma unroll
      for (int s = 0; s < NOUT; ++s) f[2 + s] = (float)(r[s] / rt);
      f[13] = (float)log(h);
#pragma unroll
      for (int k = 0; k < NIN; ++k) feat[k] = (f[k] - cst->mean[k]) * cst->inv_std[k];
    }
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
    const int64_t c = grp * 64 + t * 16 + cellw;
    const bool valid = c < n;
    // B operand: slot j < 4 of lane l holds feature perm_k(l, j) = 4 (l / 16)
    // + j of cell 16 t + l % 16 (slots 4-7 index the zero padding 16-31)
    bf16x8 x;
    {
      const int gq = lane >> 4, src = t * 16 + cellw;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float v = 0.f;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          if (4 * g + j < NIN) {
            const float w = __shfl(feat[4 * g + j], src, 64);
            v = gq == g ? w : v;
          }
        }
        x[j] = (__bf16)v;
        x[4 + j] = (__bf16)0.f;
      }
    }
    // Layer 1: 8 output tiles, one K step.
    f32x4 h1[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) h1[o] = mfma(ld8(W1 + (o * 64 + lane) * 8), x, f32x4{0, 0, 0, 0});
    bf16x8 b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) b[s] = to_b(h1[2 * s], h1[2 * s + 1], sb + 32 * s, sb + 32 * s + 16, lane);
    // Layers 2 and 3.
    f32x4 h2[8];
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
      const __bf16* W = layer == 0 ? W2 : W3;
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(ld8(W + ((o * 4 + s) * 64 + lane) * 8), b[s], acc);
        h2[o] = acc;
      }
      const float* bb = sb + HID * (layer + 1);
#pragma unroll
      for (int s = 0; s < 4; ++s) b[s] = to_b(h2[2 * s], h2[2 * s + 1], bb + 32 * s, bb + 32 * s + 16, lane);
    }
    // Layer 4: one 16-row output tile.
    f32x4 out = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) out = mfma(ld8(W4 + (s * 64 + lane) * 8), b[s], out);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int row = (lane / 16) * 4 + j;
      sO[wave][cellw][row] = out[j] + sb[3 * HID + row];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // The network output of each cell goes to its own rho_out slot as NOUT
    // floats (the slot holds NOUT doubles); project_kernel reads them there.
    if (lane < 16 && valid) {
      float* o = reinterpret_cast<float*>(rho_out + c * NOUT);
#pragma unroll
      for (int s = 0; s < NOUT; ++s) o[s] = sO[wave][lane][s];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
}

// Log-ratio softmax, then the minimum-KL projection onto the element,
// charge and mass constraints of the input (FP64), one cell per thread; the
// network output is read from the cell's rho_out slot (NOUT floats written by
// infer_kernel) before the slot is overwritten with the result.
__global__ __launch_bounds__(64) void project_kernel(const InferConsts cstv, const ProjRows rows,
                                                     const double* __restrict__ rho_in,
                                                     const double* __restrict__ T_in,
                                                     const double* __restrict__ dt_cell,
                                                     double dt, double* __restrict__ rho_out,
                                                     int32_t* __restrict__ flag, int64_t n) {
  const InferConsts* cst = &cstv;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float onet[NOUT];
  {
    const float* o = reinterpret_cast<const float*>(rho_out + c * NOUT);
#pragma unroll
    for (int s = 0; s < NOUT; ++s) onet[s] = o[s];
  }
  {
      // Log-ratio softmax, then the minimum-KL projection onto the element,
      // charge and mass constraints of the input (FP64).
      double r[NOUT], rt = 0.0;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) {
        r[s] = rho_in[c * NOUT + s];
        rt += r[s];
      }
      double Y0[NOUT], Yp[NOUT], mx = -1e300;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) {
        Y0[s] = r[s] / rt;
        double o = (double)onet[s] * cst->out_std[s] + cst->out_mean[s];
        Yp[s] = log(dmax_(Y0[s], 0.0) + (double)cst->eps_y) + o;
        mx = Yp[s] > mx ? Yp[s] : mx;
      }
      double z = 0.0;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) {
        Yp[s] = exp(Yp[s] - mx);
        z += Yp[s];
      }
#pragma unroll
      for (int s = 0; s < NOUT; ++s) Yp[s] /= z;
      // Rows: N atoms, O atoms, charge (per unit mass).  Total mass is not a
      // fourth row: the species table gives ions the neutral molar mass, so
      // sum(Y) = M_N * (N row) + M_O * (O row) + Y_e identically and the four
      // would freeze Y_e.  After the projection the total mass is restored by
      // the integrator's own final rescale (hybrid/chemistry/integrator.py:102),
      // which keeps the charge exact and moves N and O by the change of Y_e,
      // as the integrator itself does.
      const double (*A)[NOUT] = rows.A;   // uniform: held in scalar registers
      double cvec[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cvec[i] = 0.0;
#pragma unroll
        for (int s = 0; s < NOUT; ++s) cvec[i] += A[i][s] * Y0[s];
      }
      // Newton on the dual with the step limited to a change of 2 in any
      // species' log (the electron column of A is ~1e6 times the others, so
      // a full step from lam = 0 can overflow exp); stops when converged.
      // The first iterations run in FP32 (hardware exponentials) until the
      // step falls below 1e-4; the FP64 iterations then start from there and
      // converge to the same stop (a step of at most 1e-14), in one or two
      // iterations instead of four or five from lam = 0.  A non-finite FP32
      // result is discarded (the FP64 iterations then start from 0).
      double lam[3] = {0, 0, 0};
      {
        float lf[3] = {0.f, 0.f, 0.f};
        for (int it = 0; it < PROJ_ITERS_F32; ++it) {
          float Yq[NOUT];
#pragma unroll
          for (int s = 0; s < NOUT; ++s)
            Yq[s] = (float)Yp[s] * __expf((float)A[0][s] * lf[0] + (float)A[1][s] * lf[1] +
                                          (float)A[2][s] * lf[2]);
          float rr[3], J[3][3];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            float acc = 0.f;
#pragma unroll
            for (int s = 0; s < NOUT; ++s) acc += (float)A[i][s] * Yq[s];
            rr[i] = acc - (float)cvec[i];
#pragma unroll
            for (int k = i; k < 3; ++k) {
              float jk = 0.f;
#pragma unroll
              for (int s = 0; s < NOUT; ++s) jk += (float)A[i][s] * Yq[s] * (float)A[k][s];
              J[i][k] = jk;
              J[k][i] = jk;
            }
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int i = k + 1; i < 3; ++i) {
              const float f = J[i][k] / J[k][k];
#pragma unroll
              for (int j = k; j < 3; ++j) J[i][j] -= f * J[k][j];
              rr[i] -= f * rr[k];
            }
          }
          float d[3];
#pragma unroll
          for (int k = 2; k >= 0; --k) {
            float acc = rr[k];
#pragma unroll
            for (int j = k + 1; j < 3; ++j) acc -= J[k][j] * d[j];
            d[k] = acc / J[k][k];
          }
          float step = 0.f;
#pragma unroll
          for (int s = 0; s < NOUT; ++s) {
            const float e = fabsf((float)A[0][s] * d[0] + (float)A[1][s] * d[1] +
                                  (float)A[2][s] * d[2]);
            step = e > step ? e : step;
          }
          const float damp = step > 2.f ? 2.f / step : 1.f;
#pragma unroll
          for (int k = 0; k < 3; ++k) lf[k] -= damp * d[k];
          if (!(step > 1e-4f)) break;
        }
        const bool fin = (lf[0] - lf[0] == 0.f) && (lf[1] - lf[1] == 0.f) && (lf[2] - lf[2] == 0.f);
#pragma unroll
        for (int k = 0; k < 3; ++k) lam[k] = fin ? (double)lf[k] : 0.0;
      }
      double res = 0.0;
      for (int it = 0; it < PROJ_ITERS; ++it) {
        double Yq[NOUT];
#pragma unroll
        for (int s = 0; s < NOUT; ++s)
          Yq[s] = Yp[s] * exp(A[0][s] * lam[0] + A[1][s] * lam[1] + A[2][s] * lam[2]);
        double rr[3], J[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          rr[i] = -cvec[i];
#pragma unroll
          for (int s = 0; s < NOUT; ++s) rr[i] += A[i][s] * Yq[s];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int s = 0; s < NOUT; ++s) acc += A[i][s] * Yq[s] * A[k][s];
            J[i][k] = acc;
          }
        }
        // Gaussian elimination on the 3x3 SPD system (diagonal pivots).
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
          for (int i = k + 1; i < 3; ++i) {
            double f = J[i][k] / J[k][k];
#pragma unroll
            for (int j = k; j < 3; ++j) J[i][j] -= f * J[k][j];
            rr[i] -= f * rr[k];
          }
        }
        double d[3];
#pragma unroll
        for (int k = 2; k >= 0; --k) {
          double acc = rr[k];
#pragma unroll
          for (int j = k + 1; j < 3; ++j) acc -= J[k][j] * d[j];
          d[k] = acc / J[k][k];
        }
        double step = 0.0;
#pragma unroll
        for (int s = 0; s < NOUT; ++s) {
          double e = fabs(A[0][s] * d[0] + A[1][s] * d[1] + A[2][s] * d[2]);
          step = e > step ? e : step;
        }
        const double damp = step > 2.0 ? 2.0 / step : 1.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) lam[k] -= damp * d[k];
        if (!(step > 1e-14)) break;   // converged (or NaN: caught by the checks below)
      }
      // the input is read again rather than held in registers through the
      // iterations (the same values: r and Y0 = r / rt bit for bit)
#pragma unroll
      for (int s = 0; s < NOUT; ++s) {
        r[s] = __builtin_nontemporal_load(rho_in + c * NOUT + s);
        Y0[s] = r[s] / rt;
      }
      double Yf[NOUT], zf = 0.0;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) {
        Yf[s] = Yp[s] * exp(A[0][s] * lam[0] + A[1][s] * lam[1] + A[2][s] * lam[2]);
        zf += Yf[s];
      }
      // Residual of each row relative to the size of the terms it sums, before
      // and after the step (a row whose species are all absent initially, such
      // as the charge of neutral air, has cvec = 0 and is measured against the
      // predicted traces).
      res = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double acc = -cvec[i], scale = 0.0;
#pragma unroll
        for (int s = 0; s < NOUT; ++s) {
          acc += A[i][s] * Yf[s];
          scale += fabs(A[i][s]) * (Y0[s] + Yf[s]);
        }
        double rel = fabs(acc) / (scale + 1e-300);
        res = rel > res ? rel : res;
      }
#pragma unroll
      for (int s = 0; s < NOUT; ++s) Yf[s] /= zf;
      // Out-of-hull check on the raw features.
      bool inside = true;
      {
        double T = T_in[c];
        double h = dt_cell ? dt_cell[c] : dt;
        double f[NIN];
        f[0] = log(T);
        f[1] = log(rt);
#pragma unroll
        for (int s = 0; s < NOUT; ++s) f[2 + s] = Y0[s];
        f[13] = log(h);
#pragma unroll
        for (int k = 0; k < NIN; ++k) {
          // margin in standard deviations, plus the FP32 rounding of the bounds
          double span = (cst->inv_std[k] > 0.f ? cst->hull_margin / cst->inv_std[k] : 0.0) +
                        1e-6 * (1.0 + fabs(f[k]));
          inside = inside && f[k] >= cst->hull_lo[k] - span && f[k] <= cst->hull_hi[k] + span;
        }
      }
      bool ok = inside && res <= cst->proj_tol;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) ok = ok && (Yf[s] - Yf[s] == 0.0) && Yf[s] >= 0.0;
      flag[c] = ok ? 0 : 1;
#pragma unroll
      for (int s = 0; s < NOUT; ++s) rho_out[c * NOUT + s] = ok ? rt * Yf[s] : r[s];
  }
}

}  // namespace fsgmlp

static thread_local std::string g_mlp_error;

static int mlp_fail(int code, const char* msg) {
  g_mlp_error = msg;
  return code;
}

static int launched() {
  return hipGetLastError() == hipSuccess ? 0 : mlp_fail(-1, "launch failed");
}

extern "C" {

int fsg_mlp_linear_forward(const float* X, const float* W, const float* b, float* Z, float* H,
                           int32_t M, int32_t N, int32_t K, int32_t silu_on, void* stream) {
  if (!X || !W || !b || !Z || !H) return mlp_fail(-10, "null argument");
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  dim3 grid((N + fsgmlp::TN - 1) / fsgmlp::TN, (M + fsgmlp::TM - 1) / fsgmlp::TM);
  hipLaunchKernelGGL((fsgmlp::gemm_kernel<false, true, 1>), grid, dim3(256), 0,
                     static_cast<hipStream_t>(stream), X, K, W, K, Z, N, b, H, M, N, K, silu_on,
                     K, (int64_t)0);
  return launched();
}

int64_t fsg_mlp_backward_work_size(int32_t M, int32_t N, int32_t K) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  return (int64_t)fsgmlp::backward_splits(M) * ((int64_t)N * K + N);
}

int fsg_mlp_linear_backward(float* dH, const float* Z, const float* X, const float* W, float* dW,
                            float* db, float* dX, float* work, int32_t M, int32_t N, int32_t K,
                            int32_t silu_on, void* stream) {
  if (!dH || !Z || !X || !W || !dW || !db || !work) return mlp_fail(-10, "null argument");
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  if (N > 256) return mlp_fail(-12, "layer width above 256");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t n = (int64_t)M * N;
  if (silu_on)
    hipLaunchKernelGGL(fsgmlp::dsilu_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       dH, Z, n);
  // dW[N, K] = dZ^T X : A(m = out n, k = batch) = dZ[k * N + n] (transposed),
  // split over S batch chunks into work, then reduced in chunk order
  const int S = fsgmlp::backward_splits(M);
  const int chunk = (M + S - 1) / S;
  const int64_t nk = (int64_t)N * K;
  float* part_w = work;
  float* part_b = work + (int64_t)S * nk;
  dim3 gw((K + fsgmlp::TN - 1) / fsgmlp::TN, (N + fsgmlp::TM - 1) / fsgmlp::TM, S);
  hipLaunchKernelGGL((fsgmlp::gemm_kernel<true, false, 0>), gw, dim3(256), 0, s, dH, N, X, K,
                     part_w, K, nullptr, nullptr, N, K, M, 0, chunk, nk);
  hipLaunchKernelGGL(fsgmlp::reduce_splits_kernel, dim3((unsigned)((nk + 255) / 256)), dim3(256),
                     0, s, part_w, dW, nk, S);
  hipLaunchKernelGGL(fsgmlp::colsum_partial_kernel, dim3(S), dim3(256), 0, s, dH, part_b, M, N,
                     chunk);
  hipLaunchKernelGGL(fsgmlp::reduce_splits_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256),
                     0, s, part_b, db, (int64_t)N, S);
  if (dX) {
    dim3 gx((K + fsgmlp::TN - 1) / fsgmlp::TN, (M + fsgmlp::TM - 1) / fsgmlp::TM);
    hipLaunchKernelGGL((fsgmlp::gemm_kernel<false, false, 0>), gx, dim3(256), 0, s, dH, N, W, K,
                       dX, K, nullptr, nullptr, M, K, N, 0, N, (int64_t)0);
  }
  return launched();
}

int fsg_mlp_mse_grad(const float* O, const float* T, float* g, int64_t n, float scale,
                     void* stream) {
  if (!O || !T || !g) return mlp_fail(-10, "null argument");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fsgmlp::mse_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), O, T, g, n, scale);
  return launched();
}

int fsg_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int32_t step, void* stream) {
  if (!p || !g || !m || !v) return mlp_fail(-10, "null argument");
  if (n <= 0) return 0;
  if (step < 1) return mlp_fail(-11, "Adam step count starts at 1");
  float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(fsgmlp::adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return launched();
}

int32_t fsg_mlp_packed_weight_count(void) { return fsgmlp::W_TOTAL; }
int32_t fsg_mlp_consts_size(void) { return (int32_t)sizeof(fsgmlp::InferConsts); }

int fsg_mlp_infer(const void* packed_weights, const float* biases, const void* consts,
                  const double* rho_in, const double* T_in, const double* dt_cell, double dt,
                  double* rho_out, int32_t* fallback, int64_t n, int32_t blocks, void* stream) {
  if (!packed_weights || !biases || !consts || !rho_in || !T_in || !rho_out || !fallback)
    return mlp_fail(-10, "null argument");
  if (n <= 0) return 0;
  // the network's raw outputs are staged in rho_out before the projection
  // reads rho_in again: the two must not overlap
  {
    const char* a = reinterpret_cast<const char*>(rho_in);
    const char* b = reinterpret_cast<const char*>(rho_out);
    const int64_t bytes = n * fsgmlp::NOUT * (int64_t)sizeof(double);
    if (a < b + bytes && b < a + bytes) return mlp_fail(-11, "rho_out overlaps rho_in");
  }
  fsgmlp::InferConsts c;
  memcpy(&c, consts, sizeof c);
  int64_t groups = (n + 63) / 64;
  int64_t want = (groups + fsgmlp::IW - 1) / fsgmlp::IW;
  unsigned grid = (unsigned)(blocks > 0 && want > blocks ? blocks : want);
  hipLaunchKernelGGL(fsgmlp::infer_kernel, dim3(grid), dim3(fsgmlp::IW * 64), 0,
                     static_cast<hipStream_t>(stream),
                     static_cast<const __bf16*>(packed_weights), biases, c, rho_in, T_in,
                     dt_cell, dt, rho_out, fallback, n);
  if (int rc = launched()) return rc;
  fsgmlp::ProjRows rows;
  for (int s = 0; s < fsgmlp::NOUT; ++s) {
    rows.A[0][s] = c.n_atoms[s] / c.molar_mass[s];
    rows.A[1][s] = c.o_atoms[s] / c.molar_mass[s];
    rows.A[2][s] = c.charge[s] / c.molar_mass[s];
  }
  hipLaunchKernelGGL(fsgmlp::project_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0,
                     static_cast<hipStream_t>(stream), c, rows, rho_in, T_in, dt_cell, dt, rho_out,
                     fallback, n);
  return launched();
}

const char* fsg_mlp_last_error(void) { return g_mlp_error.c_str(); }

}  // extern "C"
