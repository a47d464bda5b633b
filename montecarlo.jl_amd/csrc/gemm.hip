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
// Tile: 32 x 32 per 256-thread workgroup (many workgroups even at n = 256); each wave64
// owns one 16x16 output tile as a v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32
// (FP32) accumulator.  Operand layout of the 16x16x4 instructions (wave64, lane l):
//   A (16 x 4): lane l supplies A[l % 16][l / 16]
//   B (4 x 16): lane l supplies B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[row][l % 16], r = 0..3, with row = l / 16 + 4 r (FP64)
//                or 4 (l / 16) + r (FP32); see mfma_row in common.h
// K is staged through LDS in chunks of 128 with two buffers: a chunk's global loads are
// all issued at once and the next chunk's loads are in flight during the current chunk's
// products, so a product of K <= 256 costs about one memory round trip.
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

constexpr int kTM = 32, kTN = 32, kKC = 128, kLD = 33;
constexpr int kLoadsPerThread = (kTM + kTN) * kKC / 256;   // 32

template <typename T>
constexpr size_t gemm_lds_bytes() { return sizeof(T) * 2 * 2 * kKC * kLD; }

template <typename T, int TA, int TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmBatch<T> batch) {
  typedef typename AccT<T>::type acc_t;
  const GemmItem<T>& it = batch.items ? batch.items[blockIdx.z] : batch.it[blockIdx.z];
  const int M = it.M, N = it.N, K = it.K;
  const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
  if (m0 >= M || n0 >= N) return;  // uniform per workgroup, before any barrier

  extern __shared__ __align__(16) unsigned char smem[];
  T* As = reinterpret_cast<T*>(smem);              // [2][kKC][kLD]  As[k][m]
  T* Bs = As + 2 * kKC * kLD;                      // [2][kKC][kLD]  Bs[k][n]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 16, wn = (wave >> 1) * 16;
  const T* __restrict__ A = it.A;
  const T* __restrict__ B = it.B;
  const int lda = it.lda, ldb = it.ldb;
  const bool has_mid = diag_active(it.dmid);
  const int nch = (K + kKC - 1) / kKC;

  // element q of this thread in a chunk: q < 16 -> A, else B; index e = tid + 256 (q % 16)
  auto coord = [&](int q, int& mn, int& kk) {
    const int e = tid + 256 * (q & 15);
    const bool isA = q < 16;
    const bool kfast = isA ? TA == 1 : TB == 0;    // consecutive threads along k in memory
    if (kfast) { kk = e % kKC; mn = e / kKC; }
    else { mn = e & 31; kk = e >> 5; }
  };
  // two register sets (chunks of even / odd index), so that two chunks can be in flight
  T ra[kLoadsPerThread], rb[kLoadsPerThread];
  // A chunk's loads are all issued before any is used: in-range (clamped) addresses, then
  // the out-of-range elements zeroed; the middle diagonal (a uniform choice per launch) is
  // gathered for the chunk's k values in a second batch of loads.
  auto load = [&](int c, T* reg) {
    const int k0 = c * kKC;
#pragma unroll
    for (int q = 0; q < kLoadsPerThread; q++) {
      int mn, kk;
      coord(q, mn, kk);
      const int gk = k0 + kk;
      const int gkc = gk < K ? gk : K - 1;
      if (q < 16) {
        const int gm = m0 + mn, gmc = gm < M ? gm : M - 1;
        reg[q] = TA == 0 ? A[(size_t)gmc + (size_t)gkc * lda] : A[(size_t)gkc + (size_t)gmc * lda];
      } else {
        const int gn = n0 + mn, gnc = gn < N ? gn : N - 1;
        reg[q] = TB == 0 ? B[(size_t)gkc + (size_t)gnc * ldb] : B[(size_t)gnc + (size_t)gkc * ldb];
      }
    }
    if (has_mid) {
      double f[16];
      if (it.dmid.vec) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
          int mn, kk;
          coord(q, mn, kk);
          const int gk = k0 + kk;
          f[q] = it.dmid.vec[gk < K ? gk : K - 1];
        }
        if (it.dmid.invert) {
#pragma unroll
          for (int q = 0; q < 16; q++) f[q] = 1.0 / f[q];
        }
      } else {
        int8_t cf[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          int mn, kk;
          coord(q, mn, kk);
          const int gk = k0 + kk;
          cf[q] = it.dmid.conf[gk < K ? gk : K - 1];
        }
#pragma unroll
        for (int q = 0; q < 16; q++) f[q] = cf[q] > 0 ? it.dmid.pos : it.dmid.neg;
      }
#pragma unroll
      for (int q = 0; q < 16; q++) reg[q] = T(reg[q] * f[q]);
    }
#pragma unroll
    for (int q = 0; q < kLoadsPerThread; q++) {
      int mn, kk;
      coord(q, mn, kk);
      const int gk = k0 + kk;
      const bool ok = q < 16 ? (m0 + mn < M && gk < K) : (n0 + mn < N && gk < K);
      reg[q] = ok ? reg[q] : T(0);
    }
  };
  auto store = [&](int buf, const T* reg) {
    T* Ab = As + buf * kKC * kLD;
    T* Bb = Bs + buf * kKC * kLD;
#pragma unroll
    for (int q = 0; q < kLoadsPerThread; q++) {
      int mn, kk;
      coord(q, mn, kk);
      if (q < 16) Ab[kk * kLD + mn] = reg[q];
      else Bb[kk * kLD + mn] = reg[q];
    }
  };
  acc_t acc = acc_t{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  auto compute = [&](int c, int buf) {
    const T* Ab = As + buf * kKC * kLD;
    const T* Bb = Bs + buf * kKC * kLD;
    const int kend = K - c * kKC < kKC ? K - c * kKC : kKC;
#pragma unroll 8
    for (int kk = 0; kk < kKC; kk += 4) {
      if (kk < kend) {
        const T av = Ab[(kk + lk) * kLD + wm + li];
        const T bv = Bb[(kk + lk) * kLD + wn + li];
        acc = mfma16(av, bv, acc);
      }
    }
  };

  // chunk c lives in buffer c % 2 and arrives in register set ra (even c) or rb (odd c)
  load(0, ra);
  if (nch > 1) load(1, rb);
  store(0, ra);
  __syncthreads();
  for (int c = 0; c < nch; c += 2) {
    if (c + 2 < nch) load(c + 2, ra);
    compute(c, 0);
    if (c + 1 < nch) store(1, rb);
    __syncthreads();
    if (c + 1 < nch) {
      if (c + 3 < nch) load(c + 3, rb);
      compute(c + 1, 1);
      if (c + 2 < nch) store(0, ra);
      __syncthreads();
    }
  }

  const T alpha = T(batch.alpha), beta = T(batch.beta);
  const bool has_l = diag_active(it.dl), has_r = diag_active(it.dr), has_add = diag_active(it.dadd);
  T* __restrict__ Cm = it.C;
  const int ldc = it.ldc;
  const int col = n0 + wn + (lane & 15);
  if (col < N) {
    const T sr = has_r ? T(diag_at(it.dr, col)) : T(1);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = m0 + wm + mfma_row<T>(lane, r);
      if (row >= M) continue;
      T v = alpha * acc[r];
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
  int Mx = 0, Nx = 0;
  if (b.items) {
    // item table in device memory (checked by the host code that packed it)
    if (b.max_m < 1 || b.max_n < 1 || b.count > 65535) return hipErrorInvalidValue;
    Mx = b.max_m;
    Nx = b.max_n;
  } else {
    if (b.count > kMaxGemmItems) return hipErrorInvalidValue;
    for (int i = 0; i < b.count; i++) {
      if (b.it[i].K < 1) return hipErrorInvalidValue;
      if (b.it[i].M > Mx) Mx = b.it[i].M;
      if (b.it[i].N > Nx) Nx = b.it[i].N;
    }
  }
  dim3 grid((Mx + kTM - 1) / kTM, (Nx + kTN - 1) / kTN, b.count);
  dim3 block(256);
  const size_t lds = gemm_lds_bytes<T>();
  const int sel = b.transA * 2 + b.transB;
  switch (sel) {
    case 0: hipLaunchKernelGGL((gemm_kernel<T, 0, 0>), grid, block, lds, s, b); break;
    case 1: hipLaunchKernelGGL((gemm_kernel<T, 0, 1>), grid, block, lds, s, b); break;
    case 2: hipLaunchKernelGGL((gemm_kernel<T, 1, 0>), grid, block, lds, s, b); break;
    default: hipLaunchKernelGGL((gemm_kernel<T, 1, 1>), grid, block, lds, s, b); break;
  }
  return hipGetLastError();
}

hipError_t gemm_kernels_init() {
  const void* fns[] = {reinterpret_cast<const void*>(gemm_kernel<double, 0, 0>),
                       reinterpret_cast<const void*>(gemm_kernel<double, 0, 1>),
                       reinterpret_cast<const void*>(gemm_kernel<double, 1, 0>),
                       reinterpret_cast<const void*>(gemm_kernel<double, 1, 1>),
                       reinterpret_cast<const void*>(gemm_kernel<float, 0, 0>),
                       reinterpret_cast<const void*>(gemm_kernel<float, 0, 1>),
                       reinterpret_cast<const void*>(gemm_kernel<float, 1, 0>),
                       reinterpret_cast<const void*>(gemm_kernel<float, 1, 1>)};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_gemm_f64(const GemmBatch<double>& b, hipStream_t s) { return launch_gemm_t(b, s); }
hipError_t launch_gemm_f32(const GemmBatch<float>& b, hipStream_t s) { return launch_gemm_t(b, s); }

}  // namespace dqmc
