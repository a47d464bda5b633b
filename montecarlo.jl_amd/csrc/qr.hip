// UDT decomposition (column-pivoted Householder QR) and rdivp on gfx950.
//
// Reference: udt_AVX_pivot! (linalg/UDT.jl:216-334) and rdivp! (linalg/real.jl:198-226).
//
// The reference picks, at every column step, the trailing column of largest 2-norm
// (recomputed exactly, strict ">" so the first maximum wins), applies one Householder
// reflector at a time (BLAS-2), then accumulates Q backwards, sets D = |diag R| (0 -> 1)
// and T = D^-1 R, column-permuted (apply_pivot) or left "dirty upper" for rdivp.
//
// GPU form (a short sequence of launches per batch of matrices):
//  * Blocked Householder with panels of w <= 16 columns factorised in LDS by one
//    workgroup, and the trailing matrix updated with the compact-WY block reflector
//    I - V T V^T by one wave per 16-column strip on the matrix cores, so the trailing
//    matrix is streamed once per panel instead of once per column.
//  * Pivoting: at every panel start the trailing columns' residual norms are recomputed
//    exactly and the w largest (first index wins ties) are moved to the front with the
//    reference's sequential swap semantics.  Inside a panel the order is kept.  This is
//    a column-pivoted QR whose pivot choices can differ from the reference's only where
//    the residual-norm order changes inside one panel; U D T reconstructs the input and
//    the Green's functions built from it agree with the reference's to ~1e-11 relative.
//  * Reflector convention as the reference: nu = copysign(||x||, x1), R_jj = -nu,
//    tau = (x1 + nu) / nu, v = x / (x1 + nu) with v_j = 1.
#include "common.h"

#include <vector>

namespace dqmc {

constexpr int kQrThreads = 256;
constexpr int kQrMaxW = 16;
constexpr int kPanelDoubles = 12 * 1024;  // 96 KB of panel storage

// Wave64 sum of a double, result in every lane: butterflies inside each 16-lane row
// with DPP (quad_perm xor 1, xor 2, row_ror 4, row_ror 8), then the four row sums are
// combined through v_readlane.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wsum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x124>(v);  // row_ror 4
  v += dpp_mov<0x128>(v);  // row_ror 8
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

typedef double dx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Entry (i, a) of the unit lower-trapezoidal reflector block V (panel-local row i),
// zero outside the w reflectors of the panel.
__device__ __forceinline__ double vget(const double* P, int rows, int w, int i, int a) {
  if (i >= rows || a >= w) return 0.0;
  return i == a ? 1.0 : (i > a ? P[i + (size_t)rows * a] : 0.0);
}

// X[p0:n, cb:cb+16] -= V op(T) V^T X[p0:n, cb:cb+16] with op(T) = T^T (trans) or T, on
// the matrix cores, by one wave.  W = V^T X_strip and Y = op(T) W stay in registers: the
// f64 16x16x4 accumulator element r of lane l is row l/16 + 4r, which is exactly the
// B-operand element of k-chunk r, so W and Y feed the next products directly.  Loads are
// issued kChunk k-steps at a time so that many are in flight.  When nrm is given, the
// squared norms of the updated columns over rows >= p0 + w are written to it (the
// residual norms the next panel pivots on).
constexpr int kChunk = 32;

__device__ void reflect_strip(double* __restrict__ X, int n, int p0, int rows, int cb, const double* P, int w,
                              const double* Tf, bool trans, int lane, double* nrm) {
  const int li = lane & 15, lk = lane >> 4;
  const int col = cb + li;
  const bool colok = col < n;
  const double* xc = X + (size_t)(colok ? col : n - 1) * n + p0;
  dx4 W = dx4{0, 0, 0, 0};
  for (int kb = 0; kb < rows; kb += 4 * kChunk) {
    double xb[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; q++) {
      const int i = kb + 4 * q + lk;
      xb[q] = (i < rows && colok) ? xc[i < rows ? i : rows - 1] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kChunk; q++) {
      const int i = kb + 4 * q + lk;
      if (kb + 4 * q < rows) W = mfma16(vget(P, rows, w, i, li), xb[q], W);
    }
  }
  dx4 Y = dx4{0, 0, 0, 0};
#pragma unroll
  for (int kc = 0; kc < 4; kc++) {
    const int kk = 4 * kc + lk;
    const double ta = trans ? Tf[kk * 16 + li] : Tf[li * 16 + kk];
    Y = mfma16(ta, W[kc], Y);
  }
  double* xw = X + (size_t)(colok ? col : n - 1) * n + p0;
  double ss = 0.0;
  constexpr int kRB = 4;  // 16-row blocks per batch
  for (int rb = 0; rb < rows; rb += 16 * kRB) {
    dx4 acc[kRB];
#pragma unroll
    for (int q = 0; q < kRB; q++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = rb + 16 * q + mfma_row<double>(lane, r);
        acc[q][r] = (row < rows && colok) ? xw[row < rows ? row : rows - 1] : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < kRB; q++) {
      const int r0 = rb + 16 * q;
      if (r0 >= rows) break;
#pragma unroll
      for (int kc = 0; kc < 4; kc++) acc[q] = mfma16(-vget(P, rows, w, r0 + li, 4 * kc + lk), Y[kc], acc[q]);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = r0 + mfma_row<double>(lane, r);
        if (row < rows && colok) {
          xw[row] = acc[q][r];
          if (row >= w) ss += acc[q][r] * acc[q][r];
        }
      }
    }
  }
  if (nrm) {
    ss += __shfl_xor(ss, 16, 64);
    ss += __shfl_xor(ss, 32, 64);
    if (lk == 0 && colok) nrm[col] = ss;
  }
}

// The UDT runs as a sequence of launches per batch (one stream, so each sees the previous
// one's results): column norms; per panel a single-workgroup panel kernel (pivot choice,
// column moves, Householder factorisation of the panel in LDS, compact-WY factor T) and a
// many-workgroup trailing update; Q by backward accumulation, one many-workgroup launch
// per panel; finally D and T.  Per-item state between launches: it.tau holds the residual
// norms, it.pivot the working permutation, it.tfac the T factor of each panel.

// Panel width for `rows` remaining rows (the panel must fit kPanelDoubles).
static int udt_panel_width(int rows) {
  int w = kPanelDoubles / rows;
  if (w > kQrMaxW) w = kQrMaxW;
  if (w > rows) w = rows;
  if (w < 1) w = 1;
  return w;
}

__global__ __launch_bounds__(256) void udt_norms_kernel(UdtBatch batch) {
  const UdtItem it = batch.it[blockIdx.y];
  const int n = batch.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = blockIdx.x * 4 + wave;
  if (col >= n) return;  // per-wave work only: no workgroup barrier in this kernel
  const double* a = it.A + (size_t)col * n;
  double s = 0.0;
  for (int i0 = 0; i0 < n; i0 += 64 * 8) {
    double x[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int i = i0 + 64 * q + lane;
      x[q] = i < n ? a[i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) s += x[q] * x[q];
  }
  s = wsum(s);
  if (lane == 0) {
    it.tau[col] = s;
    it.pivot[col] = col;
  }
}

static size_t udt_panel_lds_bytes(int n) {
  size_t b = 0;
  b += (size_t)n * 8;                 // nrm
  b += (size_t)kPanelDoubles * 8;     // panel
  b += 256 * 8 * 2;                   // Tf, S
  b += 32 * 8;                        // scalars, taus
  b += (size_t)n * 4;                 // perm
  b += 3 * 32 * 4 + 16;               // moves, selection, misc
  return b;
}

__global__ __launch_bounds__(kQrThreads) void udt_panel_kernel(UdtBatch batch, int p0, int w, int last) {
  const UdtItem it = batch.it[blockIdx.x];
  const int n = batch.n;
  const int rows = n - p0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = kQrThreads / 64;
  extern __shared__ __align__(16) unsigned char smem[];
  double* nrm = reinterpret_cast<double*>(smem);
  double* P = nrm + n;
  double* Tf = P + kPanelDoubles;
  double* S = Tf + 256;
  double* sc = S + 256;          // sc[0..15] taus of the panel, sc[16..] scalars
  int* perm = reinterpret_cast<int*>(sc + 32);
  int* mv_dst = perm + n;
  int* mv_src = mv_dst + 32;
  int* selr = mv_src + 32;       // 16
  int* misc = selr + 32;         // [0] moves
  double* __restrict__ A = it.A;

  for (int k = p0 + tid; k < n; k += kQrThreads) {
    nrm[k] = it.tau[k];
    perm[k] = it.pivot[k];
  }
  if (tid < kQrMaxW) selr[tid] = -1;
  __syncthreads();

  // the w largest residual norms, ties to the lower index (reference: strict ">")
  for (int k = p0 + tid; k < n; k += kQrThreads) {
    const double v = nrm[k];
    int cnt = 0;
#pragma unroll 8
    for (int k2 = p0; k2 < n; k2++) {
      const double u = nrm[k2];
      cnt += (u > v) || (u == v && k2 < k);
    }
    if (cnt < w) selr[cnt] = k;
  }
  __syncthreads();

  // sequential-swap semantics -> list of column moves (thread 0, O(w^2))
  if (tid == 0) {
    bool ok = true;
    for (int t = 0; t < w; t++) ok = ok && selr[t] >= p0 && selr[t] < n;
    if (!ok)
      for (int t = 0; t < w; t++) selr[t] = p0 + t;  // non-finite norms: keep the order
    int dcol[kQrMaxW + 1], dslot[kQrMaxW + 1], nd = 0;
    for (int t = 0; t < w; t++) {
      const int c = selr[t];
      const int target = p0 + t;
      int slot_c = c, dc = -1;
      for (int d = 0; d < nd; d++)
        if (dcol[d] == c) { slot_c = dslot[d]; dc = d; }
      if (slot_c == target) continue;
      int content = target, dt = -1;
      for (int d = 0; d < nd; d++)
        if (dslot[d] == target) { content = dcol[d]; dt = d; }
      if (dt >= 0) dslot[dt] = slot_c;
      else { dcol[nd] = content; dslot[nd] = slot_c; nd++; }
      if (dc >= 0) { dcol[dc] = dcol[nd - 1]; dslot[dc] = dslot[nd - 1]; nd--; }
      const int tp = perm[target]; perm[target] = perm[slot_c]; perm[slot_c] = tp;
      const double tn = nrm[target]; nrm[target] = nrm[slot_c]; nrm[slot_c] = tn;
    }
    int nm = 0;
    for (int t = 0; t < w; t++)
      if (selr[t] != p0 + t) { mv_dst[nm] = p0 + t; mv_src[nm] = selr[t]; nm++; }
    for (int d = 0; d < nd; d++)
      if (dcol[d] != dslot[d]) { mv_dst[nm] = dslot[d]; mv_src[nm] = dcol[d]; nm++; }
    misc[0] = nm;
  }
  __syncthreads();

  // physical column moves (whole columns), 256 rows at a time: every source element is
  // read into registers before any destination is written
  const int nm = misc[0];
  if (nm > 0) {
    for (int r0 = 0; r0 < n; r0 += kQrThreads) {
      const int r = r0 + tid;
      const int rc = r < n ? r : n - 1;
      double v[2 * kQrMaxW];
#pragma unroll
      for (int q = 0; q < 2 * kQrMaxW; q++) {
        const int src = q < nm ? mv_src[q] : 0;
        v[q] = A[(size_t)rc + (size_t)src * n];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2 * kQrMaxW; q++)
        if (q < nm && r < n) A[(size_t)r + (size_t)mv_dst[q] * n] = v[q];
      __syncthreads();
    }
  }
  for (int k = p0 + tid; k < n; k += kQrThreads) it.pivot[k] = perm[k];

  // panel factorisation in LDS
  for (int idx = tid; idx < rows * w; idx += kQrThreads) {
    const int jj = idx / rows, i = idx - jj * rows;
    P[idx] = A[(size_t)(p0 + i) + (size_t)(p0 + jj) * n];
  }
  if (tid == 0) sc[16] = nrm[p0];
  __syncthreads();
  for (int jj = 0; jj < w; jj++) {
    const int rj = jj;
    const double normu = sc[16];
    const double xi1 = P[rj + (size_t)rows * jj];
    double tau = 0.0, nu = 0.0;
    if (normu != 0.0) {
      nu = copysign(sqrt(normu), xi1);
      const double xi = xi1 + nu;
      tau = xi / nu;
      for (int i = rj + 1 + tid; i < rows; i += kQrThreads) P[i + (size_t)rows * jj] /= xi;
    }
    __syncthreads();
    if (tid == 0) {
      if (normu != 0.0) P[rj + (size_t)rows * jj] = -nu;
      sc[jj] = tau;
    }
    for (int j2 = jj + 1 + wave; j2 < w; j2 += NW) {
      double s = 0.0;
      for (int i = rj + 1 + lane; i < rows; i += 64) s += P[i + (size_t)rows * jj] * P[i + (size_t)rows * j2];
      s = wsum(s);
      const double vA = tau * (P[rj + (size_t)rows * j2] + s);
      for (int i = rj + 1 + lane; i < rows; i += 64) P[i + (size_t)rows * j2] -= P[i + (size_t)rows * jj] * vA;
      if (lane == 0) P[rj + (size_t)rows * j2] -= vA;
      if (j2 == jj + 1) {
        double t = 0.0;
        for (int i = rj + 1 + lane; i < rows; i += 64) {
          const double x = P[i + (size_t)rows * j2];
          t += x * x;
        }
        t = wsum(t);
        if (lane == 0) sc[16] = t;
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < rows * w; idx += kQrThreads) {
    const int jj = idx / rows, i = idx - jj * rows;
    A[(size_t)(p0 + i) + (size_t)(p0 + jj) * n] = P[idx];
  }

  // compact-WY factor T (upper triangular, 16 x 16, zero beyond w):
  // H_1..H_w = I - V T V^T.  S = V^T V on the matrix cores (wave 0), then lane i of
  // wave 0 builds row i of T from its own earlier entries:
  // T[i][j] = -tau_j sum_{k=i}^{j-1} T[i][k] S[k][j].
  __syncthreads();
  if (wave == 0) {
    dx4 Sacc = dx4{0, 0, 0, 0};
    for (int k0 = 0; k0 < rows; k0 += 4) {
      const double v = vget(P, rows, w, k0 + (lane >> 4), lane & 15);
      Sacc = mfma16(v, v, Sacc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) S[mfma_row<double>(lane, r) * 16 + (lane & 15)] = Sacc[r];
  }
  __syncthreads();
  if (tid < 16) {
    const int i = tid;
    double row[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      double t = 0.0;
      if (i < w && j < w) {
        if (j == i) t = sc[j];
        else if (j > i) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 16; k++)
            if (k >= i && k < j) s += row[k] * S[k * 16 + j];
          t = -sc[j] * s;
        }
      }
      row[j] = t;
    }
#pragma unroll
    for (int j = 0; j < 16; j++) it.tfac[(size_t)p0 * 16 + i * 16 + j] = row[j];
  }
  if (last) {
    // D = |diag R| (0 -> 1); the diagonal of the earlier panels is final, this panel's
    // diagonal is read from LDS
    for (int i = tid; i < n; i += kQrThreads) {
      const double r = i >= p0 ? P[(i - p0) + (size_t)rows * (i - p0)] : A[(size_t)i + (size_t)i * n];
      const double x = fabs(r);
      it.D[i] = x == 0.0 ? 1.0 : x;
    }
  }
}

// One wave per 16-column strip: mode 0 = trailing update of A with T^T (and the residual
// norms of the updated columns), mode 1 = Q accumulation into U with T.
static size_t udt_reflect_lds_bytes(int rows, int w) { return ((size_t)rows * w + 256) * 8; }

__global__ __launch_bounds__(64) void udt_reflect_kernel(UdtBatch batch, int p0, int w, int c0, int mode) {
  const UdtItem it = batch.it[blockIdx.y];
  const int n = batch.n;
  const int rows = n - p0;
  const int cb = c0 + 16 * blockIdx.x;
  const int lane = threadIdx.x;
  extern __shared__ __align__(16) unsigned char smem[];
  double* P = reinterpret_cast<double*>(smem);
  double* Tf = P + (size_t)rows * w;
  const double* __restrict__ A = it.A;
  const int total = rows * w;
  for (int base = 0; base < total; base += 64 * 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      int idx = base + 64 * q + lane;
      idx = idx < total ? idx : total - 1;
      const int jj = idx / rows, i = idx - jj * rows;
      v[q] = A[(size_t)(p0 + i) + (size_t)(p0 + jj) * n];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = base + 64 * q + lane;
      if (idx < total) P[idx] = v[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) Tf[64 * q + lane] = it.tfac[(size_t)p0 * 16 + 64 * q + lane];
  __syncthreads();
  if (mode == 0) reflect_strip(it.A, n, p0, rows, cb, P, w, Tf, true, lane, it.tau);
  else reflect_strip(it.U, n, p0, rows, cb, P, w, Tf, false, lane, nullptr);
}

// T = D^-1 R (pivoted: column j of R goes to column pivot[j]; dirty: the strict lower part
// is copied), one wave per column.
__global__ __launch_bounds__(256) void udt_finish_kernel(UdtBatch batch) {
  const UdtItem it = batch.it[blockIdx.y];
  const int n = batch.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + wave;
  if (j >= n) return;
  const double* __restrict__ A = it.A;
  double* To = it.T;
  const size_t cj = (size_t)j * n;
  if (batch.apply_pivot) {
    const size_t co = (size_t)it.pivot[j] * n;
    for (int i = lane; i < n; i += 64) {
      const double d = 1.0 / it.D[i];
      To[co + i] = j >= i ? d * A[cj + i] : 0.0;
    }
  } else {
    for (int i = lane; i < n; i += 64) {
      if (j >= i) {
        const double d = 1.0 / it.D[i];
        To[cj + i] = d * A[cj + i];
      } else if (To != A) {
        To[cj + i] = A[cj + i];
      }
    }
  }
}

hipError_t qr_kernels_init() {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(udt_panel_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(udt_reflect_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

void udt_plan(int n, std::vector<int>& starts, std::vector<int>& widths) {
  starts.clear();
  widths.clear();
  for (int p0 = 0; p0 < n;) {
    const int w = udt_panel_width(n - p0);
    starts.push_back(p0);
    widths.push_back(w);
    p0 += w;
  }
}

hipError_t udt_check(const UdtBatch& b) {
  if (b.count < 0 || b.count > kMaxUdtItems || b.n < 1 || b.n > 2048) return hipErrorInvalidValue;
  if (b.apply_pivot)
    for (int k = 0; k < b.count; k++)
      if (b.it[k].T == b.it[k].A) return hipErrorInvalidValue;  // pivoted T is written out of place
  return hipSuccess;
}

hipError_t launch_udt_norms(const UdtBatch& b, hipStream_t s) {
  hipLaunchKernelGGL(udt_norms_kernel, dim3((b.n + 3) / 4, b.count), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_udt_panel(const UdtBatch& b, int p0, int w, int last, hipStream_t s) {
  hipLaunchKernelGGL(udt_panel_kernel, dim3(b.count), dim3(kQrThreads), udt_panel_lds_bytes(b.n), s, b, p0, w,
                     last);
  return hipGetLastError();
}

hipError_t launch_udt_reflect(const UdtBatch& b, int p0, int w, int c0, int mode, hipStream_t s) {
  if (c0 >= b.n) return hipSuccess;
  hipLaunchKernelGGL(udt_reflect_kernel, dim3((b.n - c0 + 15) / 16, b.count), dim3(64),
                     udt_reflect_lds_bytes(b.n - p0, w), s, b, p0, w, c0, mode);
  return hipGetLastError();
}

hipError_t launch_udt_finish(const UdtBatch& b, hipStream_t s) {
  hipLaunchKernelGGL(udt_finish_kernel, dim3((b.n + 3) / 4, b.count), dim3(256), 0, s, b);
  return hipGetLastError();
}

// The whole sequence on one stream (the chain's own path times the steps separately).
hipError_t launch_udt(const UdtBatch& b, hipStream_t s) {
  if (b.count == 0) return hipSuccess;
  hipError_t e = udt_check(b);
  if (e != hipSuccess) return e;
  std::vector<int> st, wd;
  udt_plan(b.n, st, wd);
  const int np = (int)st.size();
  if ((e = launch_udt_norms(b, s)) != hipSuccess) return e;
  for (int pi = 0; pi < np; pi++) {
    if ((e = launch_udt_panel(b, st[pi], wd[pi], pi == np - 1, s)) != hipSuccess) return e;
    if ((e = launch_udt_reflect(b, st[pi], wd[pi], st[pi] + wd[pi], 0, s)) != hipSuccess) return e;
  }
  for (int k = 0; k < b.count; k++)
    if ((e = launch_set_identity(b.it[k].U, b.n, 1, s)) != hipSuccess) return e;
  for (int pi = np - 1; pi >= 0; pi--)
    if ((e = launch_udt_reflect(b, st[pi], wd[pi], st[pi], 1, s)) != hipSuccess) return e;
  return launch_udt_finish(b, s);
}

// rdivp: A <- A[:, pivot] * triu(T)^-1 (only the upper triangle of T is read).
// Rows are independent.  Each wave owns 16 rows; for every 16-column panel J the
// coupling to the already solved columns, X_{<J} T_{<J,J}, is a 16 x 16 x j0 product on
// the matrix cores, then the 16 x 16 triangular block is solved by 16 lanes (one row each)
// in the reference's order: x_j = (o_j - sum_{k<j} x_k T_kj) / T_jj.
__global__ __launch_bounds__(256) void rdivp_kernel(RdivpBatch batch) {
  const RdivpItem it = batch.it[blockIdx.y];
  const int n = batch.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * 64 + wave * 16;
  __shared__ double Ysh[4][16][17];
  __shared__ double Tsh[4][16][17];
  if (r0 >= n) return;  // per-wave work only: no workgroup barrier below
  double* __restrict__ A = it.A;
  double* __restrict__ O = it.O;
  const double* __restrict__ T = it.T;
  const int* __restrict__ piv = it.pivot;
  {
    const int r = r0 + li;
    if (r < n)
      for (int j = lk; j < n; j += 4) O[(size_t)r + (size_t)j * n] = A[(size_t)r + (size_t)piv[j] * n];
  }
  __threadfence_block();
  for (int j0 = 0; j0 < n; j0 += 16) {
    const int w = n - j0 < 16 ? n - j0 : 16;
    dx4 acc = dx4{0, 0, 0, 0};
    const int ra = r0 + li;
    const int cb = j0 + li;
#pragma unroll 4
    for (int k0 = 0; k0 < j0; k0 += 4) {
      const int k = k0 + lk;
      const double av = ra < n ? A[(size_t)ra + (size_t)k * n] : 0.0;
      const double bv = cb < n ? T[(size_t)k + (size_t)cb * n] : 0.0;
      acc = mfma16(av, bv, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = mfma_row<double>(lane, r);
      const int gr = r0 + row, gc = j0 + li;
      Ysh[wave][row][li] = (gr < n && gc < n) ? O[(size_t)gr + (size_t)gc * n] - acc[r] : 0.0;
      const int tr = j0 + row;
      Tsh[wave][row][li] = (tr < n && gc < n) ? T[(size_t)tr + (size_t)gc * n] : 1.0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    if (lane < 16 && r0 + lane < n) {
      double x[16];
#pragma unroll
      for (int c = 0; c < 16; c++) {
        double v = Ysh[wave][lane][c];
#pragma unroll
        for (int cp = 0; cp < 16; cp++)
          if (cp < c) v -= x[cp] * Tsh[wave][cp][c];
        x[c] = v / Tsh[wave][c][c];
      }
#pragma unroll
      for (int c = 0; c < 16; c++)
        if (c < w) A[(size_t)(r0 + lane) + (size_t)(j0 + c) * n] = x[c];
    }
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
  }
}

hipError_t launch_rdivp(const RdivpBatch& b, hipStream_t s) {
  if (b.count <= 0) return hipSuccess;
  if (b.count > kMaxUdtItems) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rdivp_kernel, dim3((b.n + 63) / 64, b.count), dim3(256), 0, s, b);
  return hipGetLastError();
}

}  // namespace dqmc
