// Batched dense products on the gfx950 matrix cores.
//
// C = alpha * diag(dl) * op(A) * diag(dmid) * op(B) * diag(dr) + beta * C + diag(dadd)
//
// Every slice-matrix multiply of the DQMC stack (reference stack.jl:319-367: eT^2 eV M,
// M eT^2 eV, M eV^-1 eT^-2, eV^-1 eT^-2 M, eV' eT^2' M), the Green's-function wraps
// (stack.jl:594-603), the products of calculate_greens_AVX! (stack.jl:442-496) and the
// physical-greens transform (greens.jl:114-125) are one or two launches of this kernel:
// the diagonal interaction matrices enter as row / inner / column scalings generated
// straight from the Hirsch field column, so they never exist as separate arrays.
//
// Tile: 64 x 64 x 16 per 256-thread workgroup.  Each wave64 owns a 32 x 32 quadrant as
// 2 x 2 16x16 accumulators of v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32
// (FP32).  Operand layout of the 16x16x4 instructions (wave64, lane l):
//   A (16 x 4): lane l supplies A[l % 16][l / 16]
//   B (4 x 16): lane l supplies B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[row][l % 16], r = 0..3, with row = l / 16 + 4 r (FP64)
//                or 4 (l / 16) + r (FP32); see mfma_row in common.h
// Global tiles are staged in LDS with the next tile's loads in flight during the MFMAs.
#include "common.h"

namespace dqmc {

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

template <typename T> struct AccT;
template <> struct AccT<double> { typedef dx4 type; };
template <> struct AccT<float> { typedef fx4 type; };

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4 mfma16(float a, float b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int kBM = 64, kBN = 64, kBK = 16, kPad = 1;

template <typename T, int TA, int TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmBatch<T> batch) {
  typedef typename AccT<T>::type acc_t;
  const GemmItem<T>& it = batch.it[blockIdx.z];
  const int M = it.M, N = it.N, K = it.K;
  const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
  if (m0 >= M || n0 >= N) return;

  __shared__ T As[kBK][kBM + kPad];
  __shared__ T Bs[kBK][kBN + kPad];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const T* __restrict__ A = it.A;
  const T* __restrict__ B = it.B;
  const int lda = it.lda, ldb = it.ldb;
  const bool has_mid = diag_active(it.dmid);

  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int mm, kk;
      if (TA == 0) { mm = tid & 63; kk = (tid >> 6) + 4 * q; }
      else { kk = tid & 15; mm = (tid >> 4) + 16 * q; }
      const int gm = m0 + mm, gk = k0 + kk;
      T v = T(0);
      if (gm < M && gk < K) {
        v = TA == 0 ? A[(size_t)gm + (size_t)gk * lda] : A[(size_t)gk + (size_t)gm * lda];
        if (has_mid) v = T(v * diag_at(it.dmid, gk));
      }
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int nn, kk;
      if (TB == 0) { kk = tid & 15; nn = (tid >> 4) + 16 * q; }
      else { nn = tid & 63; kk = (tid >> 6) + 4 * q; }
      const int gn = n0 + nn, gk = k0 + kk;
      T v = T(0);
      if (gn < N && gk < K)
        v = TB == 0 ? B[(size_t)gk + (size_t)gn * ldb] : B[(size_t)gn + (size_t)gk * ldb];
      rb[q] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (TA == 0) As[(tid >> 6) + 4 * q][tid & 63] = ra[q];
      else As[tid & 15][(tid >> 4) + 16 * q] = ra[q];
      if (TB == 0) Bs[tid & 15][(tid >> 4) + 16 * q] = rb[q];
      else Bs[(tid >> 6) + 4 * q][tid & 63] = rb[q];
    }
  };

  load(0);
  for (int k0 = 0; k0 < K; k0 += kBK) {
    store();
    __syncthreads();
    if (k0 + kBK < K) load(k0 + kBK);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      const T a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
      const T b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
      acc[0][0] = mfma16(a0, b0, acc[0][0]);
      acc[0][1] = mfma16(a0, b1, acc[0][1]);
      acc[1][0] = mfma16(a1, b0, acc[1][0]);
      acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }

  const T alpha = T(batch.alpha), beta = T(batch.beta);
  const bool has_l = diag_active(it.dl), has_r = diag_active(it.dr), has_add = diag_active(it.dadd);
  T* __restrict__ Cm = it.C;
  const int ldc = it.ldc;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int col = n0 + wn + 16 * j + (lane & 15);
      if (col >= N) continue;
      const T sr = has_r ? T(diag_at(it.dr, col)) : T(1);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = m0 + wm + 16 * i + mfma_row<T>(lane, r);
        if (row >= M) continue;
        T v = alpha * acc[i][j][r];
        if (has_l) v *= T(diag_at(it.dl, row));
        if (has_r) v *= sr;
        T* dst = Cm + (size_t)row + (size_t)col * ldc;
        if (beta != T(0)) v += beta * (*dst);
        if (has_add && row == col) v += T(diag_at(it.dadd, row));
        *dst = v;
      }
    }
}

template <typename T>
static hipError_t launch_gemm_t(const GemmBatch<T>& b, hipStream_t s) {
  if (b.count <= 0) return hipSuccess;
  if (b.count > kMaxGemmItems) return hipErrorInvalidValue;
  int Mx = 0, Nx = 0;
  for (int i = 0; i < b.count; i++) {
    if (b.it[i].M > Mx) Mx = b.it[i].M;
    if (b.it[i].N > Nx) Nx = b.it[i].N;
  }
  dim3 grid((Mx + kBM - 1) / kBM, (Nx + kBN - 1) / kBN, b.count);
  dim3 block(256);
  const int sel = b.transA * 2 + b.transB;
  switch (sel) {
    case 0: hipLaunchKernelGGL((gemm_kernel<T, 0, 0>), grid, block, 0, s, b); break;
    case 1: hipLaunchKernelGGL((gemm_kernel<T, 0, 1>), grid, block, 0, s, b); break;
    case 2: hipLaunchKernelGGL((gemm_kernel<T, 1, 0>), grid, block, 0, s, b); break;
    default: hipLaunchKernelGGL((gemm_kernel<T, 1, 1>), grid, block, 0, s, b); break;
  }
  return hipGetLastError();
}

hipError_t launch_gemm_f64(const GemmBatch<double>& b, hipStream_t s) { return launch_gemm_t(b, s); }
hipError_t launch_gemm_f32(const GemmBatch<float>& b, hipStream_t s) { return launch_gemm_t(b, s); }

}  // namespace dqmc
