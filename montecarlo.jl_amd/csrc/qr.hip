// UDT decomposition (column-pivoted Householder QR) and rdivp on gfx950.
//
// Reference: udt_AVX_pivot! (linalg/UDT.jl:216-334) and rdivp! (linalg/real.jl:198-226).
//
// The reference picks, at every column step, the trailing column of largest 2-norm
// (recomputed exactly, strict ">" so the first maximum wins), applies one Householder
// reflector at a time (BLAS-2), then accumulates Q backwards, sets D = |diag R| (0 -> 1)
// and T = D^-1 R, column-permuted (apply_pivot) or left "dirty upper" for rdivp.
//
// GPU form (one workgroup per matrix, matrix resident in L2):
//  * Blocked Householder with panels of w <= 16 columns factorised in LDS and the
//    trailing matrix updated with the compact-WY block reflector I - V T V^T, so the
//    trailing matrix is streamed once per panel instead of once per column.
//  * Pivoting: at every panel start the trailing columns' residual norms are recomputed
//    exactly and the w largest (first index wins ties) are moved to the front with the
//    reference's sequential swap semantics.  Inside a panel the order is kept.  This is
//    a column-pivoted QR whose pivot choices can differ from the reference's only where
//    the residual-norm order changes inside one panel; U D T reconstructs the input and
//    the Green's functions built from it agree with the reference's to ~1e-11 relative.
//  * Reflector convention as the reference: nu = copysign(||x||, x1), R_jj = -nu,
//    tau = (x1 + nu) / nu, v = x / (x1 + nu) with v_j = 1.
#include "common.h"

namespace dqmc {

constexpr int kQrThreads = 256;
constexpr int kQrMaxW = 16;
constexpr int kPanelDoubles = 12 * 1024;  // 96 KB of panel storage

__device__ __forceinline__ double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

static size_t udt_lds_bytes(int n) {
  size_t b = 0;
  b += (size_t)n * 8;                 // nrm / D
  b += (size_t)kPanelDoubles * 8;     // panel
  b += 256 * 8 * 2;                   // Tf, S
  b += 32 * 8;                        // scalars, taus
  b += (size_t)n * 4;                 // perm
  b += (size_t)(n + 1) * 4;           // panel starts
  b += 3 * 32 * 4 + 16;               // moves, selection, misc
  return b;
}

__global__ __launch_bounds__(kQrThreads) void udt_kernel(UdtBatch batch) {
  const UdtItem it = batch.it[blockIdx.x];
  const int n = batch.n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = kQrThreads / 64;
  extern __shared__ __align__(16) unsigned char smem[];
  double* nrm = reinterpret_cast<double*>(smem);
  double* P = nrm + n;
  double* Tf = P + kPanelDoubles;
  double* S = Tf + 256;
  double* sc = S + 256;          // sc[0..15] taus of the current panel, sc[16..] scalars
  int* perm = reinterpret_cast<int*>(sc + 32);
  int* pst = perm + n;
  int* mv_dst = pst + n + 1;
  int* mv_src = mv_dst + 32;
  int* selr = mv_src + 32;       // 16
  int* misc = selr + 32;         // [0] moves, [1] ok

  double* __restrict__ A = it.A;

  for (int k = tid; k < n; k += kQrThreads) perm[k] = k;
  __syncthreads();

  int npan = 0;
  for (int p0 = 0; p0 < n;) {
    const int rows = n - p0;
    int w = kPanelDoubles / rows;
    if (w > kQrMaxW) w = kQrMaxW;
    if (w > rows) w = rows;
    if (w < 1) w = 1;
    if (tid == 0) pst[npan] = p0;
    npan++;

    // 1. exact residual norms of the trailing columns
    for (int k = p0 + wave; k < n; k += NW) {
      double s = 0.0;
      for (int i = p0 + lane; i < n; i += 64) {
        const double x = A[(size_t)i + (size_t)k * n];
        s += x * x;
      }
      s = wsum(s);
      if (lane == 0) nrm[k] = s;
    }
    if (tid < kQrMaxW) selr[tid] = -1;
    __syncthreads();

    // 2. the w largest residual norms, ties to the lower index (reference: strict ">")
    for (int k = p0 + tid; k < n; k += kQrThreads) {
      const double v = nrm[k];
      int cnt = 0;
      for (int k2 = p0; k2 < n && cnt < w; k2++) {
        const double u = nrm[k2];
        cnt += (u > v) || (u == v && k2 < k);
      }
      if (cnt < w) selr[cnt] = k;
    }
    __syncthreads();

    // 3. sequential-swap semantics -> list of column moves (thread 0, O(w^2))
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
        if (dc >= 0) { dcol[dc] = dcol[nd - 1]; dslot[dc] = dslot[nd - 1]; nd--;
          // the moved entry may have been dt; its slot was updated above
        }
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

    // 4. physical column moves through the scratch buffer
    const int nm = misc[0];
    for (int idx = tid; idx < nm * n; idx += kQrThreads) {
      const int q = idx / n, r = idx - q * n;
      it.swap[(size_t)q * n + r] = A[(size_t)r + (size_t)mv_src[q] * n];
    }
    __syncthreads();
    for (int idx = tid; idx < nm * n; idx += kQrThreads) {
      const int q = idx / n, r = idx - q * n;
      A[(size_t)r + (size_t)mv_dst[q] * n] = it.swap[(size_t)q * n + r];
    }
    __syncthreads();

    // 5. panel factorisation in LDS
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

    // 6. compact-WY factor T (upper triangular, w x w): H_1..H_w = I - V T V^T
    const int npairs = w * (w - 1) / 2;
    for (int pr = wave; pr < npairs; pr += NW) {
      int a = 0, rem = pr;
      while (rem >= w - 1 - a) { rem -= w - 1 - a; a++; }
      const int b = a + 1 + rem;
      double s = 0.0;
      for (int i = b + 1 + lane; i < rows; i += 64) s += P[i + (size_t)rows * a] * P[i + (size_t)rows * b];
      s = wsum(s);
      if (lane == 0) S[a * 16 + b] = s + P[b + (size_t)rows * a];
    }
    __syncthreads();
    if (tid == 0) {
      for (int j = 0; j < w; j++) {
        const double tj = sc[j];
        for (int i = 0; i < j; i++) {
          double s = 0.0;
          for (int k = i; k < j; k++) s += Tf[i * 16 + k] * S[k * 16 + j];
          Tf[i * 16 + j] = -tj * s;
        }
        Tf[j * 16 + j] = tj;
        for (int i = j + 1; i < 16; i++) Tf[i * 16 + j] = 0.0;
      }
      for (int j = w; j < 16; j++)
        for (int i = 0; i < 16; i++) Tf[i * 16 + j] = 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < 256; idx += kQrThreads) it.tfac[(size_t)p0 * 16 + idx] = Tf[idx];

    // 7. trailing update A <- (I - V T^T V^T) A for the columns right of the panel
    for (int k = p0 + w + wave; k < n; k += NW) {
      double y[kQrMaxW];
#pragma unroll
      for (int a = 0; a < kQrMaxW; a++) y[a] = 0.0;
      double* col = A + (size_t)k * n + p0;
      for (int i = lane; i < rows; i += 64) {
        const double x = col[i];
#pragma unroll
        for (int a = 0; a < kQrMaxW; a++) {
          if (a >= w) break;
          const double v = i == a ? 1.0 : (i > a ? P[i + (size_t)rows * a] : 0.0);
          y[a] += v * x;
        }
      }
#pragma unroll
      for (int a = 0; a < kQrMaxW; a++) {
        if (a >= w) break;
        y[a] = wsum(y[a]);
      }
      double z[kQrMaxW];
#pragma unroll
      for (int b = 0; b < kQrMaxW; b++) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < kQrMaxW; a++)
          if (a <= b) s += Tf[a * 16 + b] * y[a];
        z[b] = s;
      }
      for (int i = lane; i < rows; i += 64) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < kQrMaxW; a++) {
          if (a >= w) break;
          const double v = i == a ? 1.0 : (i > a ? P[i + (size_t)rows * a] : 0.0);
          s += v * z[a];
        }
        col[i] -= s;
      }
    }
    __syncthreads();
    p0 += w;
  }

  // 8. explicit Q by backward accumulation of the block reflectors
  double* __restrict__ U = it.U;
  for (int idx = tid; idx < n * n; idx += kQrThreads) {
    const int r = idx % n, c = idx / n;
    U[idx] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int pi = npan - 1; pi >= 0; pi--) {
    const int p0 = pst[pi];
    const int w = (pi + 1 < npan ? pst[pi + 1] : n) - p0;
    const int rows = n - p0;
    for (int idx = tid; idx < rows * w; idx += kQrThreads) {
      const int jj = idx / rows, i = idx - jj * rows;
      P[idx] = A[(size_t)(p0 + i) + (size_t)(p0 + jj) * n];
    }
    for (int idx = tid; idx < 256; idx += kQrThreads) Tf[idx] = it.tfac[(size_t)p0 * 16 + idx];
    __syncthreads();
    for (int k = p0 + wave; k < n; k += NW) {
      double y[kQrMaxW];
#pragma unroll
      for (int a = 0; a < kQrMaxW; a++) y[a] = 0.0;
      double* col = U + (size_t)k * n + p0;
      for (int i = lane; i < rows; i += 64) {
        const double x = col[i];
#pragma unroll
        for (int a = 0; a < kQrMaxW; a++) {
          if (a >= w) break;
          const double v = i == a ? 1.0 : (i > a ? P[i + (size_t)rows * a] : 0.0);
          y[a] += v * x;
        }
      }
#pragma unroll
      for (int a = 0; a < kQrMaxW; a++) {
        if (a >= w) break;
        y[a] = wsum(y[a]);
      }
      double z[kQrMaxW];
#pragma unroll
      for (int a = 0; a < kQrMaxW; a++) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < kQrMaxW; b++)
          if (b >= a && b < w) s += Tf[a * 16 + b] * y[b];
        z[a] = s;
      }
      for (int i = lane; i < rows; i += 64) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < kQrMaxW; a++) {
          if (a >= w) break;
          const double v = i == a ? 1.0 : (i > a ? P[i + (size_t)rows * a] : 0.0);
          s += v * z[a];
        }
        col[i] -= s;
      }
    }
    __syncthreads();
  }

  // 9. D = |diag R| (0 -> 1); T = D^-1 R (pivoted or dirty upper); pivot
  double* dd = nrm;
  for (int i = tid; i < n; i += kQrThreads) {
    const double x = fabs(A[(size_t)i + (size_t)i * n]);
    const double d = x == 0.0 ? 1.0 : x;
    dd[i] = d;
    it.D[i] = d;
    it.pivot[i] = perm[i];
  }
  __syncthreads();
  double* __restrict__ To = it.T;
  if (batch.apply_pivot) {
    for (int idx = tid; idx < n * n; idx += kQrThreads) {
      const int i = idx % n, j = idx / n;
      const double d = 1.0 / dd[i];
      To[(size_t)i + (size_t)perm[j] * n] = j >= i ? d * A[idx] : 0.0;
    }
  } else {
    for (int idx = tid; idx < n * n; idx += kQrThreads) {
      const int i = idx % n, j = idx / n;
      if (j >= i) {
        const double d = 1.0 / dd[i];
        To[idx] = d * A[idx];
      } else if (To != A) {
        To[idx] = A[idx];
      }
    }
  }
}

hipError_t qr_kernels_init() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(udt_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

hipError_t launch_udt(const UdtBatch& b, hipStream_t s) {
  if (b.count <= 0) return hipSuccess;
  if (b.count > kMaxUdtItems || b.n < 1 || b.n > 2048) return hipErrorInvalidValue;
  const size_t lds = udt_lds_bytes(b.n);
  hipLaunchKernelGGL(udt_kernel, dim3(b.count), dim3(kQrThreads), lds, s, b);
  return hipGetLastError();
}

// rdivp: A <- A[:, pivot] * triu(T)^-1.  Rows are independent: one lane per row,
// forward substitution over columns in blocks of 16 (T entries are wave-uniform).
__global__ __launch_bounds__(64) void rdivp_kernel(RdivpBatch batch) {
  const RdivpItem it = batch.it[blockIdx.y];
  const int n = batch.n;
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  double* __restrict__ A = it.A;
  double* __restrict__ O = it.O;
  const double* __restrict__ T = it.T;
  for (int j = 0; j < n; j++) O[(size_t)i + (size_t)j * n] = A[(size_t)i + (size_t)it.pivot[j] * n];
  for (int j0 = 0; j0 < n; j0 += 16) {
    const int w = n - j0 < 16 ? n - j0 : 16;
    double acc[16];
#pragma unroll
    for (int c = 0; c < 16; c++) acc[c] = c < w ? O[(size_t)i + (size_t)(j0 + c) * n] : 0.0;
    for (int k = 0; k < j0; k++) {
      const double x = A[(size_t)i + (size_t)k * n];
#pragma unroll
      for (int c = 0; c < 16; c++)
        if (c < w) acc[c] -= x * T[(size_t)k + (size_t)(j0 + c) * n];
    }
#pragma unroll
    for (int c = 0; c < 16; c++) {
      if (c >= w) break;
      double x = acc[c];
#pragma unroll
      for (int cp = 0; cp < 16; cp++)
        if (cp < c) x -= acc[cp] * T[(size_t)(j0 + cp) + (size_t)(j0 + c) * n];
      acc[c] = x / T[(size_t)(j0 + c) + (size_t)(j0 + c) * n];
    }
#pragma unroll
    for (int c = 0; c < 16; c++)
      if (c < w) A[(size_t)i + (size_t)(j0 + c) * n] = acc[c];
  }
}

hipError_t launch_rdivp(const RdivpBatch& b, hipStream_t s) {
  if (b.count <= 0) return hipSuccess;
  if (b.count > kMaxUdtItems) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rdivp_kernel, dim3((b.n + 63) / 64, b.count), dim3(64), 0, s, b);
  return hipGetLastError();
}

}  // namespace dqmc
