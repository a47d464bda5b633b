// Local Hubbard-Stratonovich Metropolis sweep over the N sites of one time slice.
//
// Reference: sweep_spatial (updates/local_updates.jl:23-60), propose_local for the
// Hirsch fields (fields.jl:388-393, 440-449), calculate_detratio! (fields.jl:63-84),
// accept_local! / update_greens! (fields.jl:271-286, 340-344) with the rank-1 helpers
// of linalg/updates.jl:7-131.
//
// Design: the sites of a slice are visited in windows of W <= 64 consecutive sites.
//  * window_kernel (one workgroup per chain): Metropolis decisions and rank-1 updates on
//    the W x W block of G held in registers, and the factors S, Y of the accumulated
//    update G <- G - C Y R with C = (I - G)[:, S], R = G[S, :].
//  * window_apply_kernel (one 64 x 64 tile per workgroup): the rank-m update of all of G
//    on the matrix cores, from one copy of G into the other.
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

__device__ __forceinline__ double readlane_f64(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---------------------------------------------------------------------------------------
// Windowed form: the decisions of W <= 64 consecutive sites need only the W x W block of G.
// A window kernel (one workgroup per chain) runs the Metropolis decisions and the rank-1
// updates on that block, and forms the factors of
//     G_new = G - C Y R,   C = (I - G)[:, S],  R = G[S, :],
//     Y = (I - N)^-1 (I + Gamma L)^-1 Gamma,
// with S the accepted sites, u_l = e_{s_l} - G_l[:, s_l], v_l = gamma_l G_l[s_l, :] the
// rank-1 vectors of flip l (G_l: after l flips), L[j][l] = u_l(s_j) and N[l][j] = v_l(s_j)
// (l < j).  In exact arithmetic this is the reference's sequence of rank-1 updates.
//
// Register layout: lane = row of the window block; wave w keeps columns w, w+4, ... of
// every flavour block in registers, and every wave keeps its own copy of the diagonal
// (lane p holds G_pp), so a decision needs no memory access and a rejected proposal no
// barrier.  An accepted flip at site t: the owner of column t publishes u, lane t of
// every wave publishes v on its columns (LDS), one barrier, then every wave applies the
// rank-1 update to its columns and to its diagonal copy with the same fma, so the copies
// stay bitwise equal to the block.  The published u, v columns are kept for the solves.
// ---------------------------------------------------------------------------------------
template <typename GT>
__global__ __launch_bounds__(256) void window_kernel(WindowArgs<GT> a) {
  constexpr int NW = 4, kCPW = 16, kQ = 32;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = a.n, flv = a.flv, W = a.W, w0 = a.w0;
  const int Wn = n - w0 < W ? n - w0 : W;
  const int LD = W + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // LDS, per flavour block b: Uw[b][l][r] = u_l(r), Vw[b][l][c] = v_l(c)
  double* Uw = reinterpret_cast<double*>(smem);
  double* Vw = Uw + (size_t)flv * W * LD;
  const GT* __restrict__ G = a.G;
  const size_t nn = (size_t)n * n;

  const long long pos0 = a.ctr->u_pos;
  long long loglen = a.ctr->log_len;
  const bool rowok = lane < Wn;
  double g0[kCPW], g1[kCPW];
#pragma unroll
  for (int i = 0; i < kCPW; i++) {
    const int c = wave + NW * i;
    const bool ok = rowok && c < Wn;
    const size_t o = (size_t)(w0 + lane) + (size_t)(w0 + c) * n;
    g0[i] = ok ? (double)G[o] : 0.0;
    g1[i] = (ok && flv > 1) ? (double)G[nn + o] : 0.0;
  }
  const size_t od = (size_t)(w0 + lane) * (n + 1);
  double d0 = rowok ? (double)G[od] : 0.0;
  double d1 = (rowok && flv > 1) ? (double)G[nn + od] : 0.0;
  const int cfl = rowok ? (int)a.conf[w0 + lane] : 0;
  double ul = 1.0;
  if (a.uniforms && pos0 + lane < a.n_uniforms) ul = a.uniforms[pos0 + lane];

  const double df0 = a.d_first[0], df1 = a.d_first[1];
  const double ds0 = a.d_second[0], ds1 = a.d_second[1];
  const double bw0 = a.bweight[0], bw1 = a.bweight[1];
  long long pos = pos0;
  long long neg = 0, ties = 0, exhausted = 0;
  int m = 0;
  int site_r = 0;                  // lane j: window index of flip j
  double gam0_r = 0.0, gam1_r = 0.0;

  for (int t = 0; t < Wn; t++) {
    const int x = __builtin_amdgcn_readlane(cfl, t);
    const int ix = x > 0 ? 1 : 0;
    const double g0d = readlane_f64(d0, t);
    const double g1d = flv > 1 ? readlane_f64(d1, t) : 0.0;
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
        else u = readlane_f64(ul, (int)(pos - pos0));
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
    if (lane == m) { site_r = t; gam0_r = gam0; gam1_r = gam1; }
    double* U0 = Uw + (size_t)m * LD;
    double* U1 = Uw + (size_t)(W + m) * LD;
    double* V0 = Vw + (size_t)m * LD;
    double* V1 = Vw + (size_t)(W + m) * LD;
    if (wave == (t & (NW - 1))) {           // owner of column t: u = e_t - G[:, t]
      const int it = t / NW;
      double c0 = 0.0, c1 = 0.0;
#pragma unroll
      for (int i = 0; i < kCPW; i++) {
        c0 = i == it ? g0[i] : c0;
        c1 = i == it ? g1[i] : c1;
      }
      if (lane < W) {
        U0[lane] = (lane == t ? 1.0 : 0.0) - c0;
        if (flv > 1) U1[lane] = (lane == t ? 1.0 : 0.0) - c1;
      }
    }
    if (lane == t) {                        // v = gamma G[t, :] on this wave's columns
#pragma unroll
      for (int i = 0; i < kCPW; i++) {
        const int c = wave + NW * i;
        if (c < W) {
          V0[c] = gam0 * g0[i];
          if (flv > 1) V1[c] = gam1 * g1[i];
        }
      }
    }
    if (tid == 0) a.conf[w0 + t] = (int8_t)(-x);
    __syncthreads();
    const double u0 = lane < W ? U0[lane] : 0.0;
    const double u1 = (lane < W && flv > 1) ? U1[lane] : 0.0;
#pragma unroll
    for (int i = 0; i < kCPW; i++) {
      const int c = wave + NW * i;
      if (c < W) {
        g0[i] = fma(-u0, V0[c], g0[i]);
        if (flv > 1) g1[i] = fma(-u1, V1[c], g1[i]);
      }
    }
    if (lane < W) {
      d0 = fma(-u0, V0[lane], d0);
      if (flv > 1) d1 = fma(-u1, V1[lane], d1);
    }
    m++;
  }

  // Y = (I - N)^-1 (I + Gamma L)^-1 Gamma by two triangular solves in registers: lane =
  // row j, the wave's columns c = wsub + wpb q of block b.  The u, v columns were published
  // before a barrier, so every wave sees them.
  {
    const int wpb = NW / flv;
    const int b = wave / wpb, wsub = wave - b * wpb;
    const double gj = b == 0 ? gam0_r : gam1_r;
    const double* Ub = Uw + (size_t)b * W * LD;
    const double* Vb = Vw + (size_t)b * W * LD;
    double xq[kQ];
#pragma unroll
    for (int q = 0; q < kQ; q++) {
      const int c = wsub + wpb * q;
      xq[q] = (c < m && lane == c) ? gj : 0.0;
    }
    // forward: K = (I + Gamma L)^-1 Gamma, column-oriented: x_j -= gamma_j u_l(s_j) x_l
    for (int l = 0; l + 1 < m; l++) {
      const double al = (lane > l && lane < m) ? gj * Ub[(size_t)l * LD + site_r] : 0.0;
#pragma unroll
      for (int q = 0; q < kQ; q++) {
        const int c = wsub + wpb * q;
        if (c <= l) xq[q] = fma(-al, readlane_f64(xq[q], l), xq[q]);
      }
    }
    // back: Y = (I - N)^-1 K: y_j += v_j(s_l) y_l for j < l
    for (int l = m - 1; l >= 1; l--) {
      const int sl = __builtin_amdgcn_readlane(site_r, l);
      const int jl = lane < l ? lane : 0;   // keeps the address inside the block
      const double nl = lane < l ? Vb[(size_t)jl * LD + sl] : 0.0;
#pragma unroll
      for (int q = 0; q < kQ; q++) {
        const int c = wsub + wpb * q;
        if (c < m) xq[q] = fma(nl, readlane_f64(xq[q], l), xq[q]);
      }
    }
    GT* Yb = a.Y + (size_t)b * W * W;
#pragma unroll
    for (int q = 0; q < kQ; q++) {
      const int c = wsub + wpb * q;
      if (c < W && lane < W) Yb[(size_t)lane + (size_t)c * W] = (GT)xq[q];
    }
  }
  if (wave == 0) {
    if (lane < m) a.S[1 + lane] = w0 + site_r;
    if (lane == 0) a.S[0] = m;
  }
  if (tid == 0) {
    ChainCounters* c = a.ctr;
    c->u_pos = pos;
    c->accepted += m;
    c->proposed += Wn;
    c->negative += neg;
    c->ties += ties;
    c->exhausted |= exhausted;
    c->log_len = loglen;
    c->flushes += 1;
  }
}

// Load K elements per thread with every load in flight before the first LDS store.
template <int K, typename T, typename Ld, typename St>
__device__ __forceinline__ void stage(int total, int tid, Ld ld, St st) {
  T v[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int idx = tid + 256 * q;
    v[q] = idx < total ? ld(idx) : T(0);
  }
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int idx = tid + 256 * q;
    if (idx < total) st(idx, v[q]);
  }
}

// Gd_b = Gs_b - C_b (Y_b R_b) for one 64 x 64 tile per workgroup, with C = (I - Gs)[:, S]
// and R = Gs[S, :] gathered from the source copy (the update writes the other copy, so
// no tile can change a column or row that another workgroup still gathers).  The tile's
// slice of Z = Y R (W16 x 64) is formed in LDS on the matrix cores; the rank-W16 product
// is computed transposed, tile^T -= Z^T C^T, so accumulator element r of lane l is
// G[r0 + l%16][c0 + row(l, r)]: quarter-waves touch 16 consecutive rows of a column.
template <typename GT, int W16>
__global__ __launch_bounds__(256) void window_apply_kernel(GT* __restrict__ Gd, const GT* __restrict__ Gs,
                                                           const int* __restrict__ S,
                                                           const GT* __restrict__ Y, int n, int W) {
  typedef typename V4<GT>::type acc_t;
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.z;
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const size_t off = (size_t)b * n * n;
  const GT* Gb = Gs + off;
  GT* Gw = Gd + off;
  const int rr0 = wave * 16;
  const int grow = r0 + rr0 + li;
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int gcol = c0 + 16 * j + mfma_row<GT>(lane, r);
      acc[j][r] = (grow < n && gcol < n) ? Gb[(size_t)grow + (size_t)gcol * n] : GT(0);
    }
  const int m = S[0];
  if (m > 0) {
    GT* Ys = reinterpret_cast<GT*>(smem);   // W16 x W16 column-major: Ys[l * W16 + k] = Y[k][l]
    GT* Rs = Ys + W16 * W16;                 // W16 x 64  [k][col]
    GT* Zs = Rs + W16 * 64;                  // W16 x 65  [k][col]
    GT* Cs = Zs + W16 * 65;                  // 64 x W16+1 [row][k]
    const GT* Yb = Y + (size_t)b * W * W;
    stage<(W16 * W16 + 255) / 256, GT>(W16 * W16, tid,
        [&](int idx) { const int l = idx / W16, k = idx - l * W16;
                       return (k < W && l < W) ? Yb[(size_t)k + (size_t)l * W] : GT(0); },
        [&](int idx, GT v) { Ys[idx] = v; });
    stage<(W16 * 64 + 255) / 256, GT>(W16 * 64, tid,
        [&](int idx) { const int cc = idx / W16, k = idx - cc * W16, gc = c0 + cc;
                       return (k < m && gc < n) ? Gb[(size_t)S[1 + k] + (size_t)gc * n] : GT(0); },
        [&](int idx, GT v) { const int cc = idx / W16, k = idx - cc * W16; Rs[k * 64 + cc] = v; });
    stage<(W16 * 64 + 255) / 256, GT>(W16 * 64, tid,
        [&](int idx) { const int k = idx / 64, rr = idx - k * 64, gr = r0 + rr;
                       if (k >= m || gr >= n) return GT(0);
                       const int sk = S[1 + k];
                       return (gr == sk ? GT(1) : GT(0)) - Gb[(size_t)gr + (size_t)sk * n]; },
        [&](int idx, GT v) { const int k = idx / 64, rr = idx - k * 64; Cs[rr * (W16 + 1) + k] = v; });
    __syncthreads();
    // Z[:, 16 wave : 16 wave + 16] = Y R[:, same columns]
#pragma unroll
    for (int i0 = 0; i0 < W16; i0 += 16) {
      acc_t z = {0, 0, 0, 0};
#pragma unroll
      for (int k0 = 0; k0 < W16; k0 += 4) {
        const GT av = Ys[(k0 + lk) * W16 + i0 + li];       // Y[i0 + li][k0 + lk]
        const GT bv = Rs[(k0 + lk) * 64 + 16 * wave + li];  // R[k0 + lk][16 wave + li]
        z = mfma16(av, bv, z);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) Zs[(i0 + mfma_row<GT>(lane, r)) * 65 + 16 * wave + li] = z[r];
    }
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < W16; k0 += 4) {
      const int k = k0 + lk;
      const GT bv = Cs[(rr0 + li) * (W16 + 1) + k];       // C^T[k][row]
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const GT av = -Zs[k * 65 + 16 * j + li];          // Z^T[col][k]
        acc[j] = mfma16(av, bv, acc[j]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int gcol = c0 + 16 * j + mfma_row<GT>(lane, r);
      if (grow < n && gcol < n) Gw[(size_t)grow + (size_t)gcol * n] = acc[j][r];
    }
}

template <typename GT>
static size_t apply_lds_bytes(int W16) {
  return sizeof(GT) * ((size_t)W16 * W16 + W16 * 64 + W16 * 65 + 64 * (W16 + 1));
}

template <typename GT, int W16>
static void launch_apply_w(GT* Gd, const GT* Gs, const int* S, const GT* Y, int n, int W, int flv,
                           hipStream_t s) {
  dim3 grid((n + 63) / 64, (n + 63) / 64, flv);
  hipLaunchKernelGGL((window_apply_kernel<GT, W16>), grid, dim3(256), apply_lds_bytes<GT>(W16), s, Gd, Gs, S,
                     Y, n, W);
}

template <typename GT>
static hipError_t launch_apply_t(GT* Gd, const GT* Gs, const int* S, const GT* Y, int n, int W, int flv,
                                 hipStream_t s) {
  if (W < 1 || W > 64) return hipErrorInvalidValue;
  const int W16 = (W + 15) & ~15;
  switch (W16) {
    case 16: launch_apply_w<GT, 16>(Gd, Gs, S, Y, n, W, flv, s); break;
    case 32: launch_apply_w<GT, 32>(Gd, Gs, S, Y, n, W, flv, s); break;
    case 48: launch_apply_w<GT, 48>(Gd, Gs, S, Y, n, W, flv, s); break;
    default: launch_apply_w<GT, 64>(Gd, Gs, S, Y, n, W, flv, s); break;
  }
  return hipGetLastError();
}
hipError_t launch_window_apply_f64(double* Gd, const double* Gs, const int* S, const double* Y, int n, int W,
                                   int flv, hipStream_t s) {
  return launch_apply_t(Gd, Gs, S, Y, n, W, flv, s);
}
hipError_t launch_window_apply_f32(float* Gd, const float* Gs, const int* S, const float* Y, int n, int W,
                                   int flv, hipStream_t s) {
  return launch_apply_t(Gd, Gs, S, Y, n, W, flv, s);
}

size_t window_lds_bytes(int W, int flv) { return (size_t)2 * flv * W * (W + 1) * 8; }

int window_size(int n, int flv) {
  int W = 64;
  while (W > 16 && W - 16 >= n) W -= 16;   // no wider than the lattice, rounded up to 16
  while (W > 8 && window_lds_bytes(W, flv) > 150 * 1024) W -= 8;
  return W;
}

template <typename GT>
static hipError_t launch_window_t(const WindowArgs<GT>& a, hipStream_t s) {
  const size_t lds = window_lds_bytes(a.W, a.flv);
  if (lds > 160 * 1024 || a.W < 1 || a.W > 64 || a.flv < 1 || a.flv > 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL((window_kernel<GT>), dim3(1), dim3(256), lds, s, a);
  return hipGetLastError();
}
hipError_t launch_window_f64(const WindowArgs<double>& a, hipStream_t s) { return launch_window_t(a, s); }
hipError_t launch_window_f32(const WindowArgs<float>& a, hipStream_t s) { return launch_window_t(a, s); }

hipError_t local_kernels_init() {
  const void* fns[] = {reinterpret_cast<const void*>(window_kernel<double>),
                       reinterpret_cast<const void*>(window_kernel<float>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 16>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 32>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 48>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 64>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 16>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 32>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 48>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 64>)};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dqmc
