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

// Tile geometry: TILE x TILE outputs per workgroup, K staged in chunks of KC.  TILE = 32
// (each wave one 16x16 accumulator, KC = 128) for launches with few tiles; TILE = 64 (each
// wave a 32x32 quadrant as 2 x 2 accumulators, KC = 64) for launches with many, where it
// halves the operand traffic per flop.  Every output element accumulates its k-steps of 4
// in the same order with either geometry, so the two give identical results.
template <int TILE> struct TileCfg {
  static constexpr int KC = TILE == 32 ? 128 : 64;
  static constexpr int LD = TILE + 1;
  static constexpr int F = TILE / 32;                     // 16x16 fragments per wave and side
  static constexpr int LPT = 2 * TILE * KC / 256;         // loads per thread per chunk (32)
};

template <typename T, int TILE>
constexpr size_t gemm_lds_bytes() { return sizeof(T) * 2 * 2 * TileCfg<TILE>::KC * TileCfg<TILE>::LD; }

// XCD-aware tile order.  Workgroups go to the 8 XCDs round-robin by linear id, and each XCD
// has its own L2: with the plain order, XCD k gets the tiles of row block k, so every XCD
// fetches every column panel of B.  When the grid splits evenly, the tiles are renumbered
// so that XCD k holds a contiguous run of 4 x 4 super-tiles (of one item), whose 4 row and
// 4 column panels then come from its own L2.  A permutation of the tiles: results are equal.
constexpr int kXcds = 8, kSuper = 4;
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
  const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  bx = blockIdx.x; by = blockIdx.y; bz = blockIdx.z;
  const int per_item = gx * gy;
  const int total = per_item * gz;
  if (gx % kSuper || gy % kSuper || total % kXcds) return;
  const int hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int logical = (hw % kXcds) * (total / kXcds) + hw / kXcds;
  bz = logical / per_item;
  const int t = logical % per_item;
  const int sb = t / (kSuper * kSuper), w = t % (kSuper * kSuper);
  const int sbx = sb % (gx / kSuper), sby = sb / (gx / kSuper);
  bx = sbx * kSuper + w % kSuper;
  by = sby * kSuper + w / kSuper;
}

// TAB: the items come from the device table batch.items (many-chain launches), else from
// the kernel arguments (a separate instantiation, so the one-chain path reads its items
// straight from the argument segment)
template <typename T, int TA, int TB, bool TAB, int TILE>
__global__ __launch_bounds__(256) void gemm_kernel(GemmBatch<T> batch) {
  typedef typename AccT<T>::type acc_t;
  constexpr int kTM = TILE, kTN = TILE, kKC = TileCfg<TILE>::KC, kLD = TileCfg<TILE>::LD;
  constexpr int kF = TileCfg<TILE>::F, kLoadsPerThread = TileCfg<TILE>::LPT;
  constexpr int kHalf = kLoadsPerThread / 2;
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const GemmItem<T>& it = TAB ? batch.items[bz] : batch.it[bz];
  const int M = it.M, N = it.N, K = it.K;
  const int m0 = bx * kTM, n0 = by * kTN;
  if (m0 >= M || n0 >= N) return;  // uniform per workgroup, before any barrier

  extern __shared__ __align__(16) unsigned char smem[];
  T* As = reinterpret_cast<T*>(smem);              // [2][kKC][kLD]  As[k][m]
  T* Bs = As + 2 * kKC * kLD;                      // [2][kKC][kLD]  Bs[k][n]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (16 * kF), wn = (wave >> 1) * (16 * kF);
  const T* __restrict__ A = it.A;
  const T* __restrict__ B = it.B;
  const int lda = it.lda, ldb = it.ldb;
  const bool has_mid = diag_active(it.dmid);
  const int nch = (K + kKC - 1) / kKC;

  // element q of this thread in a chunk: q < 16 -> A, else B; index e = tid + 256 (q % 16)
  auto coord = [&](int q, int& mn, int& kk) {
    const int e = tid + 256 * (q % kHalf);
    const bool isA = q < kHalf;
    const bool kfast = isA ? TA == 1 : TB == 0;    // consecutive threads along k in memory
    if (kfast) { kk = e % kKC; mn = e / kKC; }
    else { mn = e % kTM; kk = e / kTM; }
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
      if (q < kHalf) {
        const int gm = m0 + mn, gmc = gm < M ? gm : M - 1;
        reg[q] = TA == 0 ? A[(size_t)gmc + (size_t)gkc * lda] : A[(size_t)gkc + (size_t)gmc * lda];
      } else {
        const int gn = n0 + mn, gnc = gn < N ? gn : N - 1;
        reg[q] = TB == 0 ? B[(size_t)gkc + (size_t)gnc * ldb] : B[(size_t)gnc + (size_t)gkc * ldb];
      }
    }
    if (has_mid) {
      double f[kHalf];
      if (it.dmid.vec) {
#pragma unroll
        for (int q = 0; q < kHalf; q++) {
          int mn, kk;
          coord(q, mn, kk);
          const int gk = k0 + kk;
          f[q] = it.dmid.vec[gk < K ? gk : K - 1];
        }
        if (it.dmid.mode == kDiagInv) {
#pragma unroll
          for (int q = 0; q < kHalf; q++) f[q] = 1.0 / f[q];
        } else if (it.dmid.mode != kDiagPlain) {
#pragma unroll
          for (int q = 0; q < kHalf; q++) f[q] = diag_map(f[q], it.dmid.mode);
        }
      } else {
        int8_t cf[kHalf];
#pragma unroll
        for (int q = 0; q < kHalf; q++) {
          int mn, kk;
          coord(q, mn, kk);
          const int gk = k0 + kk;
          cf[q] = it.dmid.conf[gk < K ? gk : K - 1];
        }
#pragma unroll
        for (int q = 0; q < kHalf; q++) f[q] = cf[q] > 0 ? it.dmid.pos : it.dmid.neg;
      }
#pragma unroll
      for (int q = 0; q < kHalf; q++) reg[q] = T(reg[q] * f[q]);
    }
#pragma unroll
    for (int q = 0; q < kLoadsPerThread; q++) {
      int mn, kk;
      coord(q, mn, kk);
      const int gk = k0 + kk;
      const bool ok = q < kHalf ? (m0 + mn < M && gk < K) : (n0 + mn < N && gk < K);
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
      if (q < kHalf) Ab[kk * kLD + mn] = reg[q];
      else Bb[kk * kLD + mn] = reg[q];
    }
  };
  acc_t acc[kF][kF];
#pragma unroll
  for (int i = 0; i < kF; i++)
#pragma unroll
    for (int j = 0; j < kF; j++) acc[i][j] = acc_t{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  auto step = [&](const T* Ab, const T* Bb, int kk) {
    T av[kF], bv[kF];
#pragma unroll
    for (int i = 0; i < kF; i++) av[i] = Ab[(kk + lk) * kLD + wm + 16 * i + li];
#pragma unroll
    for (int j = 0; j < kF; j++) bv[j] = Bb[(kk + lk) * kLD + wn + 16 * j + li];
#pragma unroll
    for (int i = 0; i < kF; i++)
#pragma unroll
      for (int j = 0; j < kF; j++) acc[i][j] = mfma16(av[i], bv[j], acc[i][j]);
  };
  auto compute = [&](int c, int buf) {
    const T* Ab = As + buf * kKC * kLD;
    const T* Bb = Bs + buf * kKC * kLD;
    const int kend = K - c * kKC < kKC ? K - c * kKC : kKC;
    if (kend == kKC) {
      // a whole chunk: straight-line k-steps, so the LDS reads run ahead of the products
#pragma unroll
      for (int kk = 0; kk < kKC; kk += 4) step(Ab, Bb, kk);
    } else {
#pragma unroll 8
      for (int kk = 0; kk < kKC; kk += 4)
        if (kk < kend) step(Ab, Bb, kk);
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
#pragma unroll
  for (int j = 0; j < kF; j++) {
    const int col = n0 + wn + 16 * j + (lane & 15);
    if (col >= N) continue;
    const T sr = has_r ? T(diag_at(it.dr, col)) : T(1);
#pragma unroll
    for (int i = 0; i < kF; i++) {
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
}

// 64-wide tiles once a launch has at least 8 workgroups of 32 x 32 per CU to spread
constexpr int64_t kWideTileMin = 8 * 256;

template <typename T, int TILE>
static void launch_gemm_tile(const GemmBatch<T>& b, int Mx, int Nx, hipStream_t s) {
  dim3 grid((Mx + TILE - 1) / TILE, (Nx + TILE - 1) / TILE, b.count);
  dim3 block(256);
  const size_t lds = gemm_lds_bytes<T, TILE>();
  const int sel = b.transA * 2 + b.transB;
  if (b.items) {
    switch (sel) {
      case 0: hipLaunchKernelGGL((gemm_kernel<T, 0, 0, true, TILE>), grid, block, lds, s, b); break;
      case 1: hipLaunchKernelGGL((gemm_kernel<T, 0, 1, true, TILE>), grid, block, lds, s, b); break;
      case 2: hipLaunchKernelGGL((gemm_kernel<T, 1, 0, true, TILE>), grid, block, lds, s, b); break;
      default: hipLaunchKernelGGL((gemm_kernel<T, 1, 1, true, TILE>), grid, block, lds, s, b); break;
    }
  } else {
    switch (sel) {
      case 0: hipLaunchKernelGGL((gemm_kernel<T, 0, 0, false, TILE>), grid, block, lds, s, b); break;
      case 1: hipLaunchKernelGGL((gemm_kernel<T, 0, 1, false, TILE>), grid, block, lds, s, b); break;
      case 2: hipLaunchKernelGGL((gemm_kernel<T, 1, 0, false, TILE>), grid, block, lds, s, b); break;
      default: hipLaunchKernelGGL((gemm_kernel<T, 1, 1, false, TILE>), grid, block, lds, s, b); break;
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
  const int64_t tiles32 = (int64_t)((Mx + 31) / 32) * ((Nx + 31) / 32) * b.count;
  if (Mx >= 64 && Nx >= 64 && tiles32 >= kWideTileMin) launch_gemm_tile<T, 64>(b, Mx, Nx, s);
  else launch_gemm_tile<T, 32>(b, Mx, Nx, s);
  return hipGetLastError();
}

template <typename T, bool TAB, int TILE>
static hipError_t set_gemm_lds() {
  const void* fns[] = {reinterpret_cast<const void*>(gemm_kernel<T, 0, 0, TAB, TILE>),
                       reinterpret_cast<const void*>(gemm_kernel<T, 0, 1, TAB, TILE>),
                       reinterpret_cast<const void*>(gemm_kernel<T, 1, 0, TAB, TILE>),
                       reinterpret_cast<const void*>(gemm_kernel<T, 1, 1, TAB, TILE>)};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t gemm_kernels_init() {
  hipError_t e = set_gemm_lds<double, false, 32>();
  if (e == hipSuccess) e = set_gemm_lds<double, true, 32>();
  if (e == hipSuccess) e = set_gemm_lds<double, false, 64>();
  if (e == hipSuccess) e = set_gemm_lds<double, true, 64>();
  if (e == hipSuccess) e = set_gemm_lds<float, false, 32>();
  if (e == hipSuccess) e = set_gemm_lds<float, true, 32>();
  if (e == hipSuccess) e = set_gemm_lds<float, false, 64>();
  if (e == hipSuccess) e = set_gemm_lds<float, true, 64>();
  return e;
}

hipError_t launch_gemm_f64(const GemmBatch<double>& b, hipStream_t s) { return launch_gemm_t(b, s); }
hipError_t launch_gemm_f32(const GemmBatch<float>& b, hipStream_t s) { return launch_gemm_t(b, s); }

}  // namespace dqmc
