// Local Hubbard-Stratonovich Metropolis sweep over the N sites of one time slice.
//
// Reference: sweep_spatial (updates/local_updates.jl:23-60), propose_local for the
// Hirsch fields (fields.jl:388-393, 440-449), calculate_detratio! (fields.jl:63-84),
// accept_local! / update_greens! (fields.jl:271-286, 340-344) with the rank-1 helpers
// of linalg/updates.jl:7-131.
//
// Design (one workgroup per chain, all flavour blocks):
//  * Decisions are evaluated redundantly by every thread from the running diagonal
//    d_p = G'_pp kept in LDS, so a rejected proposal costs no barrier.
//  * Accepted flips are not applied to G immediately ("delayed update"): with m pending
//    flips G' = G - U V^T, where column j of U is (e_i - G'[:, i]) and column j of V is
//    (Delta/R) G'[i, :]^T of the j-th accepted flip.  The column/row of G' needed by a
//    new flip cost O(n m); the diagonal is updated in O(n).
//  * After kmax accepted flips (or at the end of the slice) G <- G - U V^T is applied
//    as a rank-kmax product on the matrix cores (16x16x4 MFMA tiles).
// This is algebraically the reference's sequence of rank-1 updates; only the rounding
// differs, so accept decisions agree except at ties |u - p| ~ 1e-12, which are counted.
#include "common.h"

namespace dqmc {

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4 mfma16(float a, float b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct V4;
template <> struct V4<double> { typedef dx4 type; };
template <> struct V4<float> { typedef fx4 type; };

__device__ __forceinline__ double counter_uniform(unsigned long long seed, unsigned long long k) {
  unsigned long long z = seed + (k + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * 0x1.0p-53;
}

struct Proposal {
  double p, R0, R1, D0, D1;
};

// Same operation order as the reference's Julia expressions, without FMA contraction:
//   R = 1.0 + Delta * (1.0 - G_ii);  detratio = R*R (density) or R1*R2 (magnetic)
//   p = exp(-dE_boson) * detratio
__device__ __forceinline__ Proposal propose(int flv, double g0, double g1, double df, double dsec,
                                            double bw) {
#pragma clang fp contract(off)
  Proposal q;
  if (flv == 1) {
    const double d = df;
    const double r = 1.0 + d * (1.0 - g0);
    q.R0 = r; q.D0 = d; q.R1 = 1.0; q.D1 = 0.0;
    const double detratio = r * r;
    q.p = bw * detratio;
  } else {
    const double d1 = df, d2 = dsec;
    const double r1 = 1.0 + d1 * (1.0 - g0);
    const double r2 = 1.0 + d2 * (1.0 - g1);
    q.R0 = r1; q.R1 = r2; q.D0 = d1; q.D1 = d2;
    const double detratio = r1 * r2;
    q.p = 1.0 * detratio;
  }
  return q;
}

__device__ __forceinline__ double gamma_of(double D, double R) {
#pragma clang fp contract(off)
  return D / R;
}

// G_b -= U_b V_b^T for every block b, U_b / V_b are n x m4 (column j contiguous).
template <typename GT, int NT>
__device__ void flush(GT* __restrict__ G, const GT* U, const GT* V, int n, int flv, int m4,
                      int ldu) {
  typedef typename V4<GT>::type acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles1 = (n + 15) >> 4;
  const int ntiles = tiles1 * tiles1 * flv;
  for (int t = wave; t < ntiles; t += NT / 64) {
    const int b = t / (tiles1 * tiles1);
    const int tt = t - b * tiles1 * tiles1;
    const int ti = tt % tiles1, tj = tt / tiles1;
    GT* Gb = G + (size_t)b * n * n;
    const int col = tj * 16 + (lane & 15);
    const int row0 = ti * 16;
    acc_t acc;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = row0 + mfma_row<GT>(lane, r);
      acc[r] = (row < n && col < n) ? Gb[(size_t)row + (size_t)col * n] : GT(0);
    }
    const int pa = ti * 16 + (lane & 15);  // row of U for the A operand
    const int qb = tj * 16 + (lane & 15);  // row of V for the B operand
    for (int k0 = 0; k0 < m4; k0 += 4) {
      const int k = k0 + (lane >> 4);
      const GT av = pa < n ? -U[(size_t)b * ldu * n + (size_t)k * n + pa] : GT(0);
      const GT bv = qb < n ? V[(size_t)b * ldu * n + (size_t)k * n + qb] : GT(0);
      acc = mfma16(av, bv, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = row0 + mfma_row<GT>(lane, r);
      if (row < n && col < n) Gb[(size_t)row + (size_t)col * n] = acc[r];
    }
  }
}

template <typename GT, int NT>
__global__ __launch_bounds__(NT) void local_sweep_kernel(LocalArgs<GT> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = a.n, flv = a.flv, kmax = a.kmax;
  const int tid = threadIdx.x;
  // LDS layout: diag (flv*n double) | uniforms (n double) | conf (n int8, padded) | U | V
  double* dg = reinterpret_cast<double*>(smem);
  double* us = dg + (size_t)flv * n;
  int8_t* cf = reinterpret_cast<int8_t*>(us + n);
  GT* Ubase = reinterpret_cast<GT*>(smem + (((size_t)(flv * n + n) * 8 + n + 15) & ~(size_t)15));
  GT* U = a.lds_uv ? Ubase : a.ubuf;
  GT* V = a.lds_uv ? Ubase + (size_t)flv * kmax * n : a.vbuf;
  GT* __restrict__ G = a.G;

  const long long pos0 = a.ctr->u_pos;
  for (int idx = tid; idx < flv * n; idx += NT) {
    const int b = idx / n, p = idx - b * n;
    dg[idx] = (double)G[(size_t)b * n * n + (size_t)p + (size_t)p * n];
  }
  for (int p = tid; p < n; p += NT) {
    cf[p] = a.conf[p];
    if (a.uniforms) {
      const long long k = pos0 + p;
      us[p] = k < a.n_uniforms ? a.uniforms[k] : 1.0;
    }
  }
  __syncthreads();

  const double df0 = a.d_first[0], df1 = a.d_first[1];
  const double ds0 = a.d_second[0], ds1 = a.d_second[1];
  const double bw0 = a.bweight[0], bw1 = a.bweight[1];
  long long pos = pos0;
  long long loglen = a.ctr->log_len;
  long long acc_count = 0, neg = 0, ties = 0, exhausted = 0, flushes = 0;
  int m = 0;

  for (int i = 0; i < n; i++) {
    const int x = cf[i];
    const int ix = x > 0 ? 1 : 0;
    const Proposal q = propose(flv, dg[i], flv > 1 ? dg[n + i] : 0.0, ix ? df1 : df0,
                               ix ? ds1 : ds0, ix ? bw1 : bw0);
    const double p = q.p;
    bool accept;
    if (p < 0.0) neg++;
    if (p > 1.0) {
      accept = true;
      if (p - 1.0 < 1e-12) ties++;
    } else {
      double u;
      if (a.uniforms) {
        if (pos >= a.n_uniforms) { exhausted = 1; u = 1.0; }
        else u = us[pos - pos0];
      } else {
        u = counter_uniform(a.seed, (unsigned long long)pos);
      }
      pos++;
      accept = u < p;
      if (fabs(u - p) < 1e-12) ties++;
    }
    if (tid == 0 && a.dec_log && loglen < a.log_cap) {
      a.dec_log[loglen] = accept ? 1 : 0;
      if (a.p_log) a.p_log[loglen] = p;
    }
    loglen++;
    if (!accept) continue;

    const double g0 = gamma_of(q.D0, q.R0);
    const double g1 = flv > 1 ? gamma_of(q.D1, q.R1) : 0.0;
    // Every wave must have evaluated proposal i (read dg[i], cf[i]) before any of them
    // modifies the diagonal or the field: rejected proposals run without barriers, so a
    // wave may still be a few sites behind.
    __syncthreads();
    for (int idx = tid; idx < flv * n; idx += NT) {
      const int b = idx / n, r = idx - b * n;
      const GT* Gb = G + (size_t)b * n * n;
      const GT* Ub = U + (size_t)b * kmax * n;
      const GT* Vb = V + (size_t)b * kmax * n;
      GT c = Gb[(size_t)r + (size_t)i * n];   // G'[r, i]
      GT w = Gb[(size_t)i + (size_t)r * n];   // G'[i, r]
      for (int j = 0; j < m; j++) {
        c -= Ub[(size_t)j * n + r] * Vb[(size_t)j * n + i];
        w -= Ub[(size_t)j * n + i] * Vb[(size_t)j * n + r];
      }
      const GT unew = GT(r == i ? 1 : 0) - c;
      const GT vnew = GT(b == 0 ? g0 : g1) * w;
      GT* Uw = U + (size_t)b * kmax * n + (size_t)m * n;
      GT* Vw = V + (size_t)b * kmax * n + (size_t)m * n;
      Uw[r] = unew;
      Vw[r] = vnew;
      dg[idx] -= (double)(unew * vnew);
    }
    if (tid == 0) {
      cf[i] = (int8_t)(-x);
      a.conf[i] = (int8_t)(-x);
    }
    acc_count++;
    m++;
    __syncthreads();
    if (m == kmax) {
      flush<GT, NT>(G, U, V, n, flv, kmax, kmax);
      flushes++;
      m = 0;
      __syncthreads();
      for (int idx = tid; idx < flv * n; idx += NT) {
        const int b = idx / n, r = idx - b * n;
        dg[idx] = (double)G[(size_t)b * n * n + (size_t)r + (size_t)r * n];
      }
      __syncthreads();
    }
  }
  if (m > 0) {
    const int m4 = (m + 3) & ~3;
    for (int j = m; j < m4; j++)
      for (int idx = tid; idx < flv * n; idx += NT) {
        const int b = idx / n, r = idx - b * n;
        U[(size_t)b * kmax * n + (size_t)j * n + r] = GT(0);
        V[(size_t)b * kmax * n + (size_t)j * n + r] = GT(0);
      }
    __syncthreads();
    flush<GT, NT>(G, U, V, n, flv, m4, kmax);
    flushes++;
  }
  if (tid == 0) {
    ChainCounters* c = a.ctr;
    c->u_pos = pos;
    c->accepted += acc_count;
    c->proposed += n;
    c->negative += neg;
    c->ties += ties;
    c->exhausted |= exhausted;
    c->log_len = loglen;
    c->flushes += flushes;
  }
}

static size_t local_lds_fixed(int n, int flv) {
  return (((size_t)(flv * n + n) * 8 + n + 15) & ~(size_t)15);
}

constexpr size_t kLocalLdsBudget = 150 * 1024;

int local_lds_kmax(int n, int flv, int elem_bytes) {
  const size_t fixed = local_lds_fixed(n, flv);
  if (fixed >= kLocalLdsBudget) return 0;
  size_t k = (kLocalLdsBudget - fixed) / ((size_t)2 * flv * n * elem_bytes);
  if (k > 32) k = 32;
  return (int)(k & ~(size_t)3);
}

template <typename GT>
static hipError_t launch_local_t(const LocalArgs<GT>& a, hipStream_t s) {
  constexpr int NT = 256;
  size_t lds = local_lds_fixed(a.n, a.flv);
  if (a.lds_uv) lds += (size_t)2 * a.flv * a.kmax * a.n * sizeof(GT);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL((local_sweep_kernel<GT, NT>), dim3(1), dim3(NT), lds, s, a);
  return hipGetLastError();
}

hipError_t local_kernels_init() {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(local_sweep_kernel<double, 256>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(local_sweep_kernel<float, 256>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

hipError_t launch_local_f64(const LocalArgs<double>& a, hipStream_t s) { return launch_local_t(a, s); }
hipError_t launch_local_f32(const LocalArgs<float>& a, hipStream_t s) { return launch_local_t(a, s); }

}  // namespace dqmc
