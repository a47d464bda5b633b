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

// ---------------------------------------------------------------------------------------
// Windowed form: the decisions of W consecutive sites need only the W x W block of G.
// A window kernel (one workgroup per chain) runs the Metropolis decisions and the rank-1
// updates on that block in LDS, and accumulates, by bordering, the m x m factors of
//     G_new = G - C Y R,   C = (I - G)[:, S],  R = G[S, :],
//     Y = (I - N)^-1 (I + Gamma L)^-1 Gamma,
// with S the accepted sites, N[l][j] = v_l(s_j), L[j][l] = u_l(s_j) (l < j), u_l and v_l
// the rank-1 vectors of flip l.  In exact arithmetic this is the reference's sequence of
// rank-1 updates.  The O(n^2 m) application is two MFMA GEMM launches over many CUs.
// ---------------------------------------------------------------------------------------
template <typename GT, int NT>
__global__ __launch_bounds__(NT) void window_kernel(WindowArgs<GT> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = a.n, flv = a.flv, W = a.W, w0 = a.w0;
  const int Wn = n - w0 < W ? n - w0 : W;
  const int LD = W + 1;
  const int tid = threadIdx.x;
  double* Gw = reinterpret_cast<double*>(smem);           // [b][row][col], flv*W*LD
  double* Uw = Gw + (size_t)flv * W * LD;                 // [b][site][flip]
  double* Vw = Uw + (size_t)flv * W * LD;
  double* Kf = Vw + (size_t)flv * W * LD;                 // [b][l][l'] lower triangular
  double* Pf = Kf + (size_t)flv * W * LD;                 // [b][l][l'] upper triangular
  double* tu = Pf + (size_t)flv * W * LD;                 // [b][W] new u
  double* tv = tu + (size_t)flv * W;                      // [b][W] new v
  double* tk = tv + (size_t)flv * W;                      // [b][W] new row of K
  double* tp = tk + (size_t)flv * W;                      // [b][W] new column of Pinv
  double* us = tp + (size_t)flv * W;                      // staged uniforms [W]
  int* sites = reinterpret_cast<int*>(us + W);            // [W]
  int8_t* cf = reinterpret_cast<int8_t*>(sites + W);      // [W]
  GT* __restrict__ G = a.G;

  const long long pos0 = a.ctr->u_pos;
  for (int idx = tid; idx < flv * W * LD; idx += NT) {
    Kf[idx] = 0.0;
    Pf[idx] = 0.0;
  }
  // window block of G, rows fastest (coalesced), several loads in flight per thread
  {
    const int total = flv * W * W;
#pragma unroll 4
    for (int idx = tid; idx < total; idx += NT) {
      const int b = idx / (W * W), rem = idx - b * W * W, c = rem / W, r = rem - c * W;
      Gw[(size_t)b * W * LD + (size_t)r * LD + c] =
          (r < Wn && c < Wn) ? (double)G[(size_t)b * n * n + (size_t)(w0 + r) + (size_t)(w0 + c) * n] : 0.0;
    }
  }
  for (int t = tid; t < Wn; t += NT) {
    cf[t] = a.conf[w0 + t];
    if (a.uniforms) {
      const long long k = pos0 + t;
      us[t] = k < a.n_uniforms ? a.uniforms[k] : 1.0;
    }
  }
  __syncthreads();

  const double df0 = a.d_first[0], df1 = a.d_first[1];
  const double ds0 = a.d_second[0], ds1 = a.d_second[1];
  const double bw0 = a.bweight[0], bw1 = a.bweight[1];
  long long pos = pos0;
  long long loglen = a.ctr->log_len;
  long long acc_count = 0, neg = 0, ties = 0, exhausted = 0;
  int m = 0;

  for (int t = 0; t < Wn; t++) {
    const int x = cf[t];
    const int ix = x > 0 ? 1 : 0;
    const double g0d = Gw[(size_t)t * LD + t];
    const double g1d = flv > 1 ? Gw[(size_t)W * LD + (size_t)t * LD + t] : 0.0;
    const Proposal q = propose(flv, g0d, g1d, ix ? df1 : df0, ix ? ds1 : ds0, ix ? bw1 : bw0);
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

    const double gam0 = gamma_of(q.D0, q.R0);
    const double gam1 = flv > 1 ? gamma_of(q.D1, q.R1) : 0.0;
    // phase 1: new rank-1 vectors on the window and the bordering of K and Pinv
    // (reads only: a wave that is behind still reads the diagonal of sites <= t).
    for (int idx = tid; idx < flv * Wn; idx += NT) {
      const int b = idx / Wn, r = idx - b * Wn;
      const double* Gb = Gw + (size_t)b * W * LD;
      tu[b * W + r] = (r == t ? 1.0 : 0.0) - Gb[(size_t)r * LD + t];
      tv[b * W + r] = (b == 0 ? gam0 : gam1) * Gb[(size_t)t * LD + r];
    }
    for (int idx = tid; idx < flv * m; idx += NT) {
      const int b = idx / m, l = idx - b * m;
      const double* Ub = Uw + (size_t)b * W * LD + (size_t)t * LD;  // u_l'(s_t), l' < m
      const double* Vb = Vw + (size_t)b * W * LD + (size_t)t * LD;  // v_l'(s_t)
      const double* Kb = Kf + (size_t)b * W * LD;
      const double* Pb = Pf + (size_t)b * W * LD;
      double sk = 0.0, sp = 0.0;
      for (int lp = l; lp < m; lp++) {
        sk += Ub[lp] * Kb[(size_t)lp * LD + l];
        sp += Pb[(size_t)l * LD + lp] * Vb[lp];
      }
      tk[b * W + l] = -(b == 0 ? gam0 : gam1) * sk;
      tp[b * W + l] = sp;
    }
    __syncthreads();
    // phase 2: rank-1 update of the window, record u, v and the bordered factors
    for (int idx = tid; idx < flv * Wn * Wn; idx += NT) {
      const int b = idx / (Wn * Wn), rem = idx - b * Wn * Wn, r = rem % Wn, c = rem / Wn;
      Gw[(size_t)b * W * LD + (size_t)r * LD + c] -= tu[b * W + r] * tv[b * W + c];
    }
    for (int idx = tid; idx < flv * Wn; idx += NT) {
      const int b = idx / Wn, r = idx - b * Wn;
      Uw[(size_t)b * W * LD + (size_t)r * LD + m] = tu[b * W + r];
      Vw[(size_t)b * W * LD + (size_t)r * LD + m] = tv[b * W + r];
    }
    for (int idx = tid; idx < flv * (m + 1); idx += NT) {
      const int b = idx / (m + 1), l = idx - b * (m + 1);
      double* Kb = Kf + (size_t)b * W * LD;
      double* Pb = Pf + (size_t)b * W * LD;
      Kb[(size_t)m * LD + l] = l < m ? tk[b * W + l] : (b == 0 ? gam0 : gam1);
      Pb[(size_t)l * LD + m] = l < m ? tp[b * W + l] : 1.0;
    }
    if (tid == 0) {
      sites[m] = w0 + t;
      a.conf[w0 + t] = (int8_t)(-x);
    }
    acc_count++;
    m++;
    __syncthreads();
  }
  __syncthreads();

  // Y = Pinv K (m x m, zero-padded to W x W), C = (I - G)[:, S], R = G[S, :] (zero-padded)
  for (int idx = tid; idx < flv * W * W; idx += NT) {
    const int b = idx / (W * W), rem = idx - b * W * W, r = rem % W, c = rem / W;
    double y = 0.0;
    if (r < m && c < m) {
      const double* Kb = Kf + (size_t)b * W * LD;
      const double* Pb = Pf + (size_t)b * W * LD;
      for (int l = (r > c ? r : c); l < m; l++) y += Pb[(size_t)r * LD + l] * Kb[(size_t)l * LD + c];
    }
    a.Y[(size_t)b * W * W + (size_t)r + (size_t)c * W] = (GT)y;
  }
  // C = (I - G)[:, S] (n x W) and R = G[S, :] (W x n), zero-padded beyond m
  // Each thread first issues a batch of kSnap independent loads, then the stores.
  constexpr int kSnap = 8;
  for (int b = 0; b < flv; b++) {
    const GT* __restrict__ Gb = G + (size_t)b * n * n;
    GT* __restrict__ Cb = a.C + (size_t)b * n * W;
    GT* __restrict__ Rb = a.R + (size_t)b * W * n;
    for (int pr = tid; pr < n; pr += NT) {
      for (int k0 = 0; k0 < W; k0 += kSnap) {
        GT gc[kSnap], gr[kSnap];
#pragma unroll
        for (int q = 0; q < kSnap; q++) {
          const int k = k0 + q;
          const int sk = k < m ? sites[k] : -1;
          gc[q] = sk >= 0 ? Gb[(size_t)pr + (size_t)sk * n] : GT(0);
          gr[q] = sk >= 0 ? Gb[(size_t)sk + (size_t)pr * n] : GT(0);
        }
#pragma unroll
        for (int q = 0; q < kSnap; q++) {
          const int k = k0 + q;
          if (k >= W) break;
          const int sk = k < m ? sites[k] : -1;
          Cb[(size_t)pr + (size_t)k * n] = sk >= 0 ? (pr == sk ? GT(1) : GT(0)) - gc[q] : GT(0);
          Rb[(size_t)k + (size_t)pr * W] = gr[q];
        }
      }
    }
  }
  if (tid == 0) {
    ChainCounters* c = a.ctr;
    c->u_pos = pos;
    c->accepted += acc_count;
    c->proposed += Wn;
    c->negative += neg;
    c->ties += ties;
    c->exhausted |= exhausted;
    c->log_len = loglen;
    c->flushes += 1;
  }
}

// G_b -= C_b (Y_b R_b) for one 64 x 64 tile of G per workgroup.  The tile's slice of
// Z = Y R (W x 64) is formed in LDS first; the rank-W product is computed transposed,
// tile^T = Z^T C^T, so accumulator element r of lane l is G[r0 + l%16][c0 + row(l, r)]:
// quarter-waves touch 16 consecutive rows of a column (128-byte coalesced).
template <typename GT>
__global__ __launch_bounds__(256) void window_apply_kernel(GT* G, const GT* C, const GT* R, const GT* Y,
                                                           int n, int W) {
  typedef typename V4<GT>::type acc_t;
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.z;
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int W4 = (W + 3) & ~3;
  GT* Ys = reinterpret_cast<GT*>(smem);          // W4 x W4   [k][l]
  GT* Rs = Ys + W4 * W4;                          // W4 x 64   [l][col]
  GT* Zs = Rs + W4 * 64;                          // W4 x 65   [k][col]
  GT* Cs = Zs + W4 * 65;                          // 64 x W4+1 [row][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const GT* Yb = Y + (size_t)b * W * W;
  const GT* Rb = R + (size_t)b * W * n;
  const GT* Cb = C + (size_t)b * n * W;
  for (int idx = tid; idx < W4 * W4; idx += 256) {
    const int k = idx / W4, l = idx - k * W4;
    Ys[idx] = (k < W && l < W) ? Yb[(size_t)k + (size_t)l * W] : GT(0);
  }
  for (int idx = tid; idx < W4 * 64; idx += 256) {
    const int cc = idx / W4, l = idx - cc * W4;
    const int gc = c0 + cc;
    Rs[l * 64 + cc] = (l < W && gc < n) ? Rb[(size_t)l + (size_t)gc * W] : GT(0);
  }
  for (int idx = tid; idx < W4 * 64; idx += 256) {
    const int k = idx / 64, rr = idx - k * 64;
    const int gr = r0 + rr;
    Cs[rr * (W4 + 1) + k] = (k < W && gr < n) ? Cb[(size_t)gr + (size_t)k * n] : GT(0);
  }
  __syncthreads();
  for (int idx = tid; idx < W4 * 64; idx += 256) {
    const int k = idx / 64, cc = idx - k * 64;
    GT z = GT(0);
    for (int l = 0; l < W4; l++) z += Ys[k * W4 + l] * Rs[l * 64 + cc];
    Zs[k * 65 + cc] = z;
  }
  __syncthreads();
  // each wave: 16-row x 64-column strip of the tile, as four 16x16 transposed MFMA tiles
  const int li = lane & 15, lk = lane >> 4;
  GT* Gb = G + (size_t)b * n * n;
  const int rr0 = wave * 16;
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int grow = r0 + rr0 + li;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int gcol = c0 + 16 * j + mfma_row<GT>(lane, r);
      acc[j][r] = (grow < n && gcol < n) ? Gb[(size_t)grow + (size_t)gcol * n] : GT(0);
    }
  }
  for (int k0 = 0; k0 < W4; k0 += 4) {
    const int k = k0 + lk;
    const GT bv = Cs[(rr0 + li) * (W4 + 1) + k];        // C^T[k][row]
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const GT av = -Zs[k * 65 + 16 * j + li];          // Z^T[col][k]
      acc[j] = mfma16(av, bv, acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int grow = r0 + rr0 + li;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int gcol = c0 + 16 * j + mfma_row<GT>(lane, r);
      if (grow < n && gcol < n) Gb[(size_t)grow + (size_t)gcol * n] = acc[j][r];
    }
  }
}

template <typename GT>
static hipError_t launch_apply_t(GT* G, const GT* C, const GT* R, const GT* Y, int n, int W, int flv,
                                 hipStream_t s) {
  const int W4 = (W + 3) & ~3;
  const size_t lds = sizeof(GT) * ((size_t)W4 * W4 + W4 * 64 + W4 * 65 + 64 * (W4 + 1));
  dim3 grid((n + 63) / 64, (n + 63) / 64, flv);
  hipLaunchKernelGGL(window_apply_kernel<GT>, grid, dim3(256), lds, s, G, C, R, Y, n, W);
  return hipGetLastError();
}
hipError_t launch_window_apply_f64(double* G, const double* C, const double* R, const double* Y, int n,
                                   int W, int flv, hipStream_t s) {
  return launch_apply_t(G, C, R, Y, n, W, flv, s);
}
hipError_t launch_window_apply_f32(float* G, const float* C, const float* R, const float* Y, int n, int W,
                                   int flv, hipStream_t s) {
  return launch_apply_t(G, C, R, Y, n, W, flv, s);
}

size_t window_lds_bytes(int W, int flv) {
  const size_t LD = W + 1;
  return (size_t)flv * W * LD * 8 * 5 + (size_t)flv * W * 8 * 4 + (size_t)W * 8 + (size_t)W * 4 + W + 16;
}

int window_size(int n, int flv) {
  int W = 48;
  while (W > 8 && window_lds_bytes(W, flv) > 150 * 1024) W -= 8;
  return W;
}

template <typename GT>
static hipError_t launch_window_t(const WindowArgs<GT>& a, hipStream_t s) {
  const size_t lds = window_lds_bytes(a.W, a.flv);
  if (lds > 160 * 1024 || a.W < 1 || a.W > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL((window_kernel<GT, 256>), dim3(1), dim3(256), lds, s, a);
  return hipGetLastError();
}
hipError_t launch_window_f64(const WindowArgs<double>& a, hipStream_t s) { return launch_window_t(a, s); }
hipError_t launch_window_f32(const WindowArgs<float>& a, hipStream_t s) { return launch_window_t(a, s); }

hipError_t local_kernels_init() {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(window_kernel<double, 256>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(window_kernel<float, 256>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

}  // namespace dqmc
