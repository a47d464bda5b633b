// Local Hubbard-Stratonovich Metropolis sweep over the N sites of one time slice.
//
// Reference: sweep_spatial (updates/local_updates.jl:23-60), propose_local for the
// Hirsch fields (fields.jl:388-393, 440-449), calculate_detratio! (fields.jl:63-84),
// accept_local! / update_greens! (fields.jl:271-286, 340-344) with the rank-1 helpers
// of linalg/updates.jl:7-131.
//
// Design: the sites of a slice are visited in windows of W in {16, 32, 64} consecutive
// sites, one launch per window:
//  * workgroup 0 forms the window's W x W block of G (applying the previous window's
//    factors to it), runs the Metropolis decisions and rank-1 updates on the block held in
//    registers, and writes the dense factor Y of the accumulated update
//    G <- G - C Y R with C = (I - G)[:, w], R = G[w, :] over the window's positions w;
//  * workgroups 1.. apply the previous window's update to all of G on the matrix cores,
//    one 64 x 64 tile each, from one copy of G into the other;
//  * window_apply_kernel applies the last window's update.
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
// Windowed form: the decisions of W consecutive sites need only the W x W block of G.
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
// v is stored with the columns of one wave contiguous: position pc(c) = (c%4) W/4 + c/4.
// ---------------------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ int vpos(int c) { return (c & 3) * (W / 4) + (c >> 2); }

// LDS of the window workgroup: the decision loop's Uw, Vw, gamma, sites; before the loop,
// the block Bn (over Uw) and, for W <= 32, the staged operands and FP32 Z scratch after it.
template <typename GT, int FLV, int W>
constexpr size_t window_lds() {
  const size_t loop = (size_t)2 * FLV * W * (W + 1) * 8 + (size_t)FLV * W * 8 + (size_t)W * 4 + 16 +
                      (W <= 32 ? (size_t)FLV * W * (W + 1) * 8 : 0);   // + u^T (W <= 32)
  const size_t block = W > 32 ? 0
                              : (size_t)FLV * W * (W + 1) * 8 + (size_t)4 * FLV * W * (W + 1) * sizeof(GT) +
                                    (sizeof(GT) == 4 ? (size_t)4 * W * 17 * 4 : 0);
  return (loop > block ? loop : block) + 16;     // + the two progress flags (last 16 bytes)
}

// Pipelining of the one-launch-per-slice form (workgroup 0, W <= 32): the fourth wave, idle
// during the decisions, prefetches the next window's three block operands into LDS; the
// window's factor is also kept in LDS for the next window's block; and the fourth wave
// publishes the previous window's factor once the current window's block is formed.
template <typename GT>
struct SliceCtl {
  GT* pf;                              // [3 FLV][W][LD]: G[w',w'], G[w',w], G[w,w'] of the next
                                       // window w' (the staging layout of mats 0-2)
  GT* yl;                              // [FLV][W][LD]: the window's factor (staging layout of mat 3)
  bool pf_in;                          // this window's operands: pf and yl (else global memory)
  const GT* gnext;                     // copy that holds the next window's operands (nullptr: none)
  const unsigned long long* wait_p;    // gnext is complete once *wait_p >= wait_v (nullptr: already)
  unsigned long long wait_v;
  unsigned long long* pub_p;           // previous window: *pub_p = pub_v once its stores are done
  unsigned long long pub_v;
  unsigned long long* errc;            // device-error count of a wait that gives up
  long long limit;
};

template <typename GT, int FLV, int W>
constexpr size_t slice_extra_lds() { return (size_t)4 * FLV * W * (W + 1) * sizeof(GT); }

// The window's block of G after the previous window's flips, from the source copy:
//   B = G[w, w] - C[w, :] Y R[:, w] = G[w, w] + G[w, wp] (Y G[wp, w])
// (C = (I - G)[:, wp] has no identity part in the rows of this window), on the matrix
// cores with the operands, order and roles of the application's tiles, so that B equals
// the block the application writes.  Wave unit (b, ci) forms the 16-column tile ci of
// flavour block b; the result goes to LDS as Bn[b][c][r] (leading dimension W + 1).
//
// W <= 32: the four W x W operands of every block (G[w, w], G[w, wp], G[wp, w] and the
// previous factor) are first staged in LDS with coalesced loads (lane = row of a column),
// all in flight together.  W = 64 (no room to stage): operands are loaded per unit
// straight into the matrix-core registers.
template <typename GT, int FLV, int W>
__device__ __forceinline__ void window_block_direct(const WindowArgs<GT>& a, double* Bn, GT* Zs, int Wn,
                                                    int lane, int wave) {
  typedef typename V4<GT>::type acc_t;
  constexpr int CT = W / 16, KS = W / 4, LD = W + 1;
  const int n = a.n, w0 = a.w0, wp = a.wp;
  const size_t nn = (size_t)n * n;
  const int li = lane & 15, lk = lane >> 4;
  const bool corr = a.Yp != nullptr;
  const int mp = corr ? a.Sp[0] : 0;
  for (int unit = wave; unit < FLV * CT; unit += 4) {
    const int b = unit / CT, ci = unit - b * CT;
    const GT* Gb = a.G + (size_t)b * nn;
    const int cc = ci * 16 + li;
    const size_t colo = (size_t)(w0 + (cc < Wn ? cc : Wn - 1)) * n;
    acc_t o[CT];
#pragma unroll
    for (int t = 0; t < CT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = t * 16 + mfma_row<GT>(lane, r);
        o[t][r] = Gb[(size_t)(w0 + (row < Wn ? row : Wn - 1)) + colo];
      }
    if (corr) {
      const GT* Yb = a.Yp + (size_t)b * W * W;
      GT qb[KS], ya[CT][KS], pa[CT][KS];
#pragma unroll
      for (int s = 0; s < KS; s++) {
        const int k = wp + 4 * s + lk;
        qb[s] = Gb[(size_t)(k < n ? k : n - 1) + colo];
      }
#pragma unroll
      for (int t = 0; t < CT; t++) {
        const int rr = t * 16 + li;
        const size_t rowo = (size_t)(w0 + (rr < Wn ? rr : Wn - 1));
#pragma unroll
        for (int s = 0; s < KS; s++) {
          const int k = wp + 4 * s + lk;
          ya[t][s] = Yb[(size_t)(t * 16 + li) * W + 4 * s + lk];
          pa[t][s] = Gb[rowo + (size_t)(k < n ? k : n - 1) * n];
        }
      }
      if (mp > 0) {
        acc_t z[CT];
#pragma unroll
        for (int t = 0; t < CT; t++) {
          const acc_t zero = {0, 0, 0, 0};
          z[t] = zero;
#pragma unroll
          for (int s = 0; s < KS; s++) z[t] = mfma16(ya[t][s], qb[s], z[t]);   // Z = Y G[wp, w]
        }
        if constexpr (sizeof(GT) == 8) {
          // element r of Z tile t is Z[16 t + lk + 4 r][cc]: the B operand of k-step 4 t + r
#pragma unroll
          for (int t = 0; t < CT; t++)
#pragma unroll
            for (int s = 0; s < KS; s++) o[t] = mfma16(pa[t][s], z[s >> 2][s & 3], o[t]);
        } else {
          // FP32 accumulator rows are 4 (lane / 16) + r: regroup Z through LDS (wave-local)
          GT* Zb = Zs + (size_t)b * W * LD;
#pragma unroll
          for (int t = 0; t < CT; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) Zb[(t * 16 + mfma_row<GT>(lane, r)) * LD + cc] = z[t][r];
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
          for (int s = 0; s < KS; s++) {
            const GT bz = Zb[(4 * s + lk) * LD + cc];
#pragma unroll
            for (int t = 0; t < CT; t++) o[t] = mfma16(pa[t][s], bz, o[t]);
          }
        }
      }
    }
    double* Bb = Bn + (size_t)b * W * LD;
#pragma unroll
    for (int t = 0; t < CT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) Bb[cc * LD + t * 16 + mfma_row<GT>(lane, r)] = (double)o[t][r];
  }
}

template <typename GT, int FLV, int W>
__device__ __forceinline__ void window_block_staged(const WindowArgs<GT>& a, double* Bn, GT* St, GT* Zw, int Wn,
                                                    int tid, int lane, int wave,
                                                    const SliceCtl<GT>* sc = nullptr) {
  typedef typename V4<GT>::type acc_t;
  constexpr int CT = W / 16, KS = W / 4, LD = W + 1, WW = W * W, QM = WW / 256;
  static_assert(WW % 256 == 0, "staging assumes whole passes of 256 threads per matrix");
  const int n = a.n, w0 = a.w0, wp = a.wp;
  const size_t nn = (size_t)n * n;
  const int li = lane & 15, lk = lane >> 4;
  const bool corr = a.Yp != nullptr;
  const int mp = corr ? a.Sp[0] : 0;
  // St[(4 b + mat) W LD + c LD + r]: mat 0 = G[w, w], 1 = G[w, wp], 2 = G[wp, w]
  // (column-major, [c][r]); mat 3 = previous factor (row-major, [p][q])
  const bool lds_in = sc != nullptr && sc->pf_in;
  if (!lds_in) {
  GT v[FLV][4][QM];
#pragma unroll
  for (int b = 0; b < FLV; b++) {
    const GT* Gb = a.G + (size_t)b * nn;
#pragma unroll
    for (int q = 0; q < QM; q++) {
      const int idx = tid + 256 * q;
      const int r = idx % W, c = idx / W;
      const int rw = w0 + (r < Wn ? r : Wn - 1), cw = w0 + (c < Wn ? c : Wn - 1);
      const int rp = wp + r < n ? wp + r : n - 1, cp = wp + c < n ? wp + c : n - 1;
      v[b][0][q] = Gb[(size_t)rw + (size_t)cw * n];
      if (corr) {
        v[b][1][q] = Gb[(size_t)rw + (size_t)cp * n];
        v[b][2][q] = Gb[(size_t)rp + (size_t)cw * n];
        v[b][3][q] = a.Yp[(size_t)b * WW + idx];
      }
    }
  }
#pragma unroll
  for (int b = 0; b < FLV; b++)
#pragma unroll
    for (int q = 0; q < QM; q++) {
      const int idx = tid + 256 * q;
      const int r = idx % W, c = idx / W;
      GT* Sb = St + (size_t)(4 * b) * W * LD + c * LD + r;
      Sb[0] = v[b][0][q];
      if (corr) {
        Sb[W * LD] = v[b][1][q];
        Sb[2 * W * LD] = v[b][2][q];
        Sb[3 * W * LD] = v[b][3][q];
      }
    }
  __syncthreads();
  }
  for (int unit = wave; unit < FLV * CT; unit += 4) {
    const int b = unit / CT, ci = unit - b * CT;
    const GT* Bs = lds_in ? sc->pf + (size_t)(3 * b) * W * LD : St + (size_t)(4 * b) * W * LD;
    const GT* Ps = Bs + W * LD;
    const GT* Qs = Bs + 2 * W * LD;
    const GT* Ys = lds_in ? sc->yl + (size_t)b * W * LD : Bs + 3 * W * LD;
    const int cc = ci * 16 + li;
    acc_t o[CT];
#pragma unroll
    for (int t = 0; t < CT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) o[t][r] = Bs[cc * LD + t * 16 + mfma_row<GT>(lane, r)];
    if (corr && mp > 0) {
      acc_t z[CT];
#pragma unroll
      for (int t = 0; t < CT; t++) {
        const acc_t zero = {0, 0, 0, 0};
        z[t] = zero;
#pragma unroll
        for (int s = 0; s < KS; s++)                                   // Z = Y G[wp, w]
          z[t] = mfma16(Ys[(t * 16 + li) * LD + 4 * s + lk], Qs[cc * LD + 4 * s + lk], z[t]);
      }
      if constexpr (sizeof(GT) == 8) {
        // element r of Z tile t is Z[16 t + lk + 4 r][cc]: the B operand of k-step 4 t + r
#pragma unroll
        for (int t = 0; t < CT; t++)
#pragma unroll
          for (int s = 0; s < KS; s++) o[t] = mfma16(Ps[(4 * s + lk) * LD + t * 16 + li], z[s >> 2][s & 3], o[t]);
      } else {
        // FP32 accumulator rows are 4 (lane / 16) + r: regroup Z through this wave's scratch
        GT* Zb = Zw + wave * W * 17;
#pragma unroll
        for (int t = 0; t < CT; t++)
#pragma unroll
          for (int r = 0; r < 4; r++) Zb[(t * 16 + mfma_row<GT>(lane, r)) * 17 + li] = z[t][r];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int s = 0; s < KS; s++) {
          const GT bz = Zb[(4 * s + lk) * 17 + li];
#pragma unroll
          for (int t = 0; t < CT; t++) o[t] = mfma16(Ps[(4 * s + lk) * LD + t * 16 + li], bz, o[t]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    double* Bb = Bn + (size_t)b * W * LD;
#pragma unroll
    for (int t = 0; t < CT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) Bb[cc * LD + t * 16 + mfma_row<GT>(lane, r)] = (double)o[t][r];
  }
}

template <typename GT, int FLV, int W>
__device__ __forceinline__ void window_role(const WindowArgs<GT>& a, unsigned char* smem,
                                            const SliceCtl<GT>* sc = nullptr) {
  constexpr int NW = 4, CPW = W / NW, LD = W + 1;
  typedef double colvec __attribute__((ext_vector_type(CPW)));
  // LDS, per flavour block b: Uw[b][l][r] = u_l(r);  Vw[b][l][vpos(c)] = G_l[s_l][c] (the
  // raw row; v_l = gamma_l times it)
  double* Uw = reinterpret_cast<double*>(smem);
  double* Vw = Uw + FLV * W * LD;
  double* gam = Vw + FLV * W * LD;                        // [b][l]
  int* sites = reinterpret_cast<int*>(gam + FLV * W);     // [l], window index of flip l
  unsigned long long* shm = reinterpret_cast<unsigned long long*>(sites + W);   // m, accepted mask
  double* UwT = reinterpret_cast<double*>(shm + 2);       // W <= 32: u^T[b][r][flip], stride W + 1
  const int n = a.n, w0 = a.w0;
  const int Wn = n - w0 < W ? n - w0 : W;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // progress of the decision wave for the waves that build the factor beside it:
  // flags[0] = flips published, flags[1] = final flip count (-1 while deciding)
  int* flags = reinterpret_cast<int*>(smem + window_lds<GT, FLV, W>() - 16);
  if (tid == 0) {
    flags[0] = 0;
    flags[1] = -1;
  }
  const long long pos0 = a.ctr->u_pos;
  const long long log0 = a.ctr->log_len;
  const int lr = lane < Wn ? lane : Wn - 1;
  const bool rowok = lane < Wn;
  // field and this window's uniforms (lane k: u_{pos0 + k})
  const int cfl = (int)a.conf[w0 + lr];
  double ul = 1.0;
  if (a.uniforms) {
    if (pos0 + lane < a.n_uniforms) ul = a.uniforms[pos0 + lane];
  } else {
    ul = counter_uniform(a.seed, (unsigned long long)(pos0 + lane));
  }
  // window block after the previous window's flips (LDS, aliasing Uw), then into registers:
  // lane = row, wave w keeps columns w, w + 4, ...; every wave keeps the diagonal (lane p: G_pp)
  double* Bn = Uw;
  if constexpr (W <= 32) {
    GT* St = reinterpret_cast<GT*>(Uw + FLV * W * LD);
    window_block_staged<GT, FLV, W>(a, Bn, St, St + 4 * FLV * W * LD, Wn, tid, lane, wave, sc);
  } else {
    window_block_direct<GT, FLV, W>(a, Bn, reinterpret_cast<GT*>(Vw), Wn, lane, wave);
  }
  __syncthreads();
  const double df0 = a.d_first[0], df1 = a.d_first[1];
  const double ds0 = a.d_second[0], ds1 = a.d_second[1];
  const double bw0 = a.bweight[0], bw1 = a.bweight[1];
  int used = 0;                    // uniforms consumed in this window
  int neg = 0, ties = 0;
  int m = 0;                       // flips accepted (every wave, after the loop)
  // per lane: decision and p of site `lane`; site and gammas of flip `lane`
  int acc_r = 0, site_r = 0;
  double p_r = 0.0, gam0_r = 0.0, gam1_r = 0.0;

  if constexpr (W <= 32) {
    // rows and columns not yet published read as exact zeros in the factor waves
    for (int e = tid; e < FLV * W * LD; e += 256) {
      Vw[e] = 0.0;
      UwT[e] = 0.0;
    }
    // the previous window's factor, flip count and field went out before this window
    // started: every wave's stores are complete at the barrier, and the fourth wave
    // publishes them while the decisions run
    if (sc && sc->pub_p) __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (sc && sc->pub_p && tid == 192) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(sc->pub_p, sc->pub_v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  // factor waves (W <= 32): wave 1 column r of X, wave 2 row r of (I - N)^-1, and gammas
  typedef double fvec __attribute__((ext_vector_type(16)));
  fvec fx_lo = 0.0, fx_hi = 0.0, gc_lo = 0.0, gc_hi = 0.0;
  if constexpr (W <= 32) {
    // Single-wave decisions: wave 0 lane 32 f + r holds row r of flavour block f, all W
    // columns in registers; u of an accepted flip at site t is each lane's own G[r][t],
    // and row t reaches the other lanes through LDS inside the wave (no workgroup barrier).
    // Waves 1-3 wait at the barrier after the loop.
    typedef double rowvec __attribute__((ext_vector_type(16)));
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int f = lane >> 5, r = lane & 31;
    const bool fok = f < FLV;
    const int fb = fok ? f : 0;          // lanes of an absent second block read block 0, unused
    constexpr int UT = LD;               // row stride of u^T (fits the Uw storage)
    if (wave == 0) {
      const bool rok = fok && r < Wn;
      const int rc = r < Wn ? r : Wn - 1;
      const double* Bf = Bn + (size_t)fb * W * LD;
      // every read in flight at once: in-range (clamped) addresses, masked by an exact
      // multiplication with 0 or 1 (the block is finite), no per-column branch
      // every read in flight at once: in-range (clamped) addresses, masked by an exact
      // multiplication with 0 or 1 (the block is finite), no per-column branch
      rowvec lo, hi;
#pragma unroll
      for (int c = 0; c < 16; c++) {
        const int cl = c < Wn ? c : Wn - 1;
        lo[c] = Bf[cl * LD + rc] * ((rok && c < Wn) ? 1.0 : 0.0);
        if constexpr (W > 16) {
          const int ch = c + 16 < Wn ? c + 16 : Wn - 1;
          hi[c] = Bf[ch * LD + rc] * ((rok && c + 16 < Wn) ? 1.0 : 0.0);
        } else {
          hi[c] = 0.0;
        }
      }
      double dg = Bf[rc * LD + rc] * (rok ? 1.0 : 0.0);
      // the other flavour block's diagonal entry of this row, updated with the same
      // operands as its owner's copy, so that every lane holds both of its site's entries
      const int fo = FLV > 1 ? 1 - fb : fb;
      double dgo = 0.0;
      if constexpr (FLV > 1) dgo = Bn[(size_t)fo * W * LD + rc * LD + rc] * (rok ? 1.0 : 0.0);
      // site r's field value and couplings (lanes r and 32 + r)
      const int cfr = __shfl(cfl, r, 64);
      const double dfr = cfr > 0 ? df1 : df0, dsr = cfr > 0 ? ds1 : ds0, bwr = cfr > 0 ? bw1 : bw0;
      // order the reads of Bn before the first flip's stores over it
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_s_waitcnt(0xC07F);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // Rounds of decisions: every undecided site's proposal is formed at once (lane r: site
      // r) from the current block; sites before the first accepted one are rejected with
      // exactly those values (no flip has changed G since their turn), the first accepted
      // one is flipped, and the next round starts after it.  Uniforms go to the proposals
      // with p <= 1 in site order, so site r's is the next unused one plus the count of
      // such sites between the round's start and r.
      const unsigned int below = (1u << r) - 1u;
      unsigned int und = Wn >= 32 ? 0xffffffffu : ((1u << Wn) - 1u);
      double u_r = 0.0;                // uniform drawn for site r (p <= 1)
      while (und) {
        const double g0 = f ? dgo : dg, g1 = f ? dg : dgo;
        const Proposal q = propose(FLV, g0, g1, dfr, dsr, bwr);
        const bool big = q.p > 1.0;
        const unsigned int needu = (unsigned int)__ballot(!big) & und;
        const int idx = used + __popc(needu & below);
        const double u = __shfl(ul, idx < 64 ? idx : 63, 64);
        // the flip's division for every site of the round, lane 32 f + r: block f's gamma
        // of site r, while the uniforms are on their way
        double gq = gamma_of(f ? q.D1 : q.D0, f ? q.R1 : q.R0);
        asm volatile("" : "+v"(gq));
        const bool accept = big || u < q.p;
        const unsigned int accm = (unsigned int)__ballot(accept) & und;
        const int ta = accm ? __builtin_ctz(accm) : 32;
        const unsigned int dec = accm ? (und & ((2u << ta) - 1u)) : und;   // decided this round
        const bool mine = f == 0 && ((dec >> r) & 1u);
        p_r = mine ? q.p : p_r;
        u_r = mine ? u : u_r;
        acc_r = mine ? (r == ta ? 1 : 0) : acc_r;
        used += __popc(needu & dec);
        und &= ~dec;
        if (!accm) break;

        const int t = ta;
        const double gam0t = readlane_f64(gq, t);
        const double gam1t = FLV > 1 ? readlane_f64(gq, 32 + t) : 0.0;
        const double gamL = f ? gam1t : gam0t;                    // this lane's block
        const double gt = (W <= 16 || t < 16) ? lo[t & 15] : hi[t & 15];   // G_f[r][t]
        double* Vr = Vw + (size_t)(fb * W + m) * W;                        // raw row t
        const double uraw = (r == t ? 1.0 : 0.0) - gt;                     // u_t(r), unscaled
        if (fok && r < W) UwT[(size_t)(f * W + r) * UT + m] = uraw;         // u^T: [f][r][flip]
        if (fok && r == t) {
#pragma unroll
          for (int c = 0; c < 16; c += 2) {
            *reinterpret_cast<d2*>(Vr + c) = d2{lo[c], lo[c + 1]};
            if constexpr (W > 16) *reinterpret_cast<d2*>(Vr + 16 + c) = d2{hi[c], hi[c + 1]};
          }
        }
        site_r = lane == m ? t : site_r;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (fok && r == m) gam[f * W + m] = gamL;
        const int rw = r < W ? r : W - 1;
        const double ug = fok ? gamL * uraw : 0.0;
        double vo = 0.0, uo = 0.0;
        if constexpr (FLV > 1) {
          vo = Vw[(size_t)(fo * W + m) * W + rw];                // G_{1-f}[t][r]
          uo = UwT[(size_t)(fo * W + rw) * UT + m];             // u_t(r) of the other block
        }
#pragma unroll
        for (int c = 0; c < 16; c += 2) {
          const d2 vl = *reinterpret_cast<const d2*>(Vr + c);
          lo[c] = fma(-ug, vl[0], lo[c]);
          lo[c + 1] = fma(-ug, vl[1], lo[c + 1]);
          if constexpr (W > 16) {
            const d2 vh = *reinterpret_cast<const d2*>(Vr + 16 + c);
            hi[c] = fma(-ug, vh[0], hi[c]);
            hi[c + 1] = fma(-ug, vh[1], hi[c + 1]);
          }
        }
        dg = fma(-ug, Vr[rw], dg);
        if constexpr (FLV > 1) {
          const double ugo = (f ? gam0t : gam1t) * uo;           // the owner's ug, bit for bit
          dgo = fma(-ugo, vo, dgo);
        }
        if (lane == 0) sites[m] = t;
        m++;
        // flip m - 1 (u, row, gamma, site; written by several lanes) is complete: publish it
        // to the factor waves
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(flags, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) __hip_atomic_store(flags + 1, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      {
        // negative probabilities and ties of the window's decided sites
        const unsigned int win = Wn >= 32 ? 0xffffffffu : ((1u << Wn) - 1u);
        const bool tie = p_r > 1.0 ? (p_r - 1.0 < 1e-12) : (fabs(u_r - p_r) < 1e-12);
        neg = __popc((unsigned int)__ballot(f == 0 && p_r < 0.0) & win);
        ties = __popc((unsigned int)__ballot(f == 0 && tie) & win);
      }
      const unsigned long long am = __ballot(acc_r != 0);    // whole wave active here
      if (lane == 0) {
        shm[0] = (unsigned long long)m;
        shm[1] = am;
      }
      if (lane < m) sites[lane] = site_r;
      else if (lane < W) {
        sites[lane] = 0;
        gam[lane] = 0.0;                  // rows beyond m come out exactly zero
        if (FLV > 1) gam[W + lane] = 0.0;
      }
      if (lane < Wn) {
        if (a.dec_log && log0 + lane < a.log_cap) {
          a.dec_log[log0 + lane] = (unsigned char)acc_r;
          if (a.p_log) a.p_log[log0 + lane] = p_r;
        }
        if (acc_r) a.conf[w0 + lane] = (int8_t)(-cfl);
      }
    } else if (wave == 1 || wave == 2) {
      // Beside the decisions, flip by flip as wave 0 publishes them (lane 32 f + r, block f):
      //   wave 1: row k of X = (I + Gamma L)^-1 Gamma,
      //           x_k = gamma_k (e_k - sum_{l<k} u_l(s_k) x_l); lane r holds column r of X;
      //   wave 2: column k of (I - N)^-1, N[i][k] = gamma_i G_i[s_i][s_k] (i < k), by bordering:
      //           Binv[:, k] = Binv_(k) N[:, k], Binv[k][k] = 1; lane r holds row r.
      // Flip k needs only flips < k and its own site and gamma, all published before it.
      // The wait is bounded, so the wave always leaves the loop.
      // Giving up leaves Y wrong, so it is counted as a device error that the host reports.
      int k = 0;
      long long spins = 0;
      for (;;) {
        int avail = __hip_atomic_load(flags, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (avail <= k) {
          const int fin = __hip_atomic_load(flags + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (fin < 0) {
            if (++spins > a.wait_limit) {
              if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&a.ctr->device_errors), 1ull);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          avail = __hip_atomic_load(flags, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (avail <= k) break;
        }
        const int sk = sites[k];
        const double gk = gam[fb * W + k];
        const int kk = k & 15;
        // Terms l >= k carry an unset (zero) x_l or gamma_l and leave the sums unchanged, so
        // only the 8-wide chunks that start below k are read and summed.
        double cf[W];
        if (wave == 1) {
          // L[k][l] = u_l(s_k): row s_k of u^T, contiguous
          const double* ur = UwT + (size_t)(fb * W + sk) * UT;
#pragma unroll
          for (int c8 = 0; c8 < W; c8 += 8)
            if (c8 < k) {
#pragma unroll
              for (int l = c8; l < c8 + 8; l++) cf[l] = ur[l];
            }
          double s4[4] = {r == k ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int c8 = 0; c8 < W; c8 += 8)
            if (c8 < k) {
#pragma unroll
              for (int l = c8; l < c8 + 8; l++) {
                const double xl = l < 16 ? fx_lo[l & 15] : fx_hi[l & 15];
                s4[l & 3] = fma(-cf[l], xl, s4[l & 3]);
              }
            }
          const double xk = gk * ((s4[0] + s4[1]) + (s4[2] + s4[3]));
          const double olo = fx_lo[kk], ohi = fx_hi[kk];
          fx_lo[kk] = k < 16 ? xk : olo;
          fx_hi[kk] = k < 16 ? ohi : xk;
        } else {
          // N[i][k] / gamma_i = G_i[s_i][s_k]: column s_k of the published rows
#pragma unroll
          for (int c8 = 0; c8 < W; c8 += 8)
            if (c8 < k) {
#pragma unroll
              for (int i = c8; i < c8 + 8; i++) cf[i] = Vw[(size_t)(fb * W + i) * W + sk];
            }
          const double glo = gc_lo[kk], ghi = gc_hi[kk];
          gc_lo[kk] = k < 16 ? gk : glo;
          gc_hi[kk] = k < 16 ? ghi : gk;
          double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int c8 = 0; c8 < W; c8 += 8)
            if (c8 < k) {
#pragma unroll
              for (int i = c8; i < c8 + 8; i++) {
                const double gi = i < 16 ? gc_lo[i & 15] : gc_hi[i & 15];
                const double bi = i < 16 ? fx_lo[i & 15] : fx_hi[i & 15];
                s4[i & 3] = fma(bi, gi * cf[i], s4[i & 3]);
              }
            }
          const double yk = r == k ? 1.0 : (s4[0] + s4[1]) + (s4[2] + s4[3]);
          const double olo = fx_lo[kk], ohi = fx_hi[kk];
          fx_lo[kk] = k < 16 ? yk : olo;
          fx_hi[kk] = k < 16 ? ohi : yk;
        }
        k++;
      }
    } else if (sc && sc->gnext) {
      // wave 3: the next window's operands G[w', w'], G[w', w], G[w, w'] from the copy
      // that the updates of the window before this one write, once they are all done
      if (sc->wait_p) {
        bool ok = true;
        long long polls = 0;
        while (__hip_atomic_load(sc->wait_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sc->wait_v) {
          if (++polls > sc->limit) { ok = false; break; }
          __builtin_amdgcn_s_sleep(2);
        }
        if (!ok && lane == 0) atomicAdd(sc->errc, 1ull);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      const int w0n = w0 + W, wpn = w0;
      const int Wnn = n - w0n < W ? n - w0n : W;
      constexpr int QL = W * W / 64;
#pragma unroll
      for (int mat = 0; mat < 3; mat++) {
        GT v[FLV][QL];
#pragma unroll
        for (int b = 0; b < FLV; b++) {
          const GT* Gb = sc->gnext + (size_t)b * n * n;
#pragma unroll
          for (int q = 0; q < QL; q++) {
            const int idx = lane + 64 * q;
            const int rr = idx % W, cc = idx / W;
            const int rw = w0n + (rr < Wnn ? rr : Wnn - 1), cw = w0n + (cc < Wnn ? cc : Wnn - 1);
            const int row = mat == 2 ? wpn + rr : rw;
            const int col = mat == 1 ? wpn + cc : cw;
            v[b][q] = Gb[(size_t)row + (size_t)col * n];
          }
        }
#pragma unroll
        for (int b = 0; b < FLV; b++)
#pragma unroll
          for (int q = 0; q < QL; q++) {
            const int idx = lane + 64 * q;
            const int rr = idx % W, cc = idx / W;
            sc->pf[(size_t)(3 * b + mat) * W * LD + cc * LD + rr] = v[b][q];
          }
      }
    }
    __syncthreads();
    m = (int)shm[0];
  } else {
    colvec g0, g1;
#pragma unroll
    for (int i = 0; i < CPW; i++) {
      const int c = wave + NW * i;
      const int cc = c < Wn ? c : Wn - 1;
      const double x0 = Bn[cc * LD + lr];
      const double x1 = FLV > 1 ? Bn[(W + cc) * LD + lr] : 0.0;
      const bool ok = rowok && c < Wn;
      g0[i] = ok ? x0 : 0.0;
      g1[i] = ok ? x1 : 0.0;
    }
    double d0 = rowok ? Bn[lr * LD + lr] : 0.0;
    double d1 = (FLV > 1 && rowok) ? Bn[(W + lr) * LD + lr] : 0.0;
    __syncthreads();                  // Bn is overwritten by the first flip's u, v

    for (int t = 0; t < Wn; t++) {
      const int x = __builtin_amdgcn_readlane(cfl, t);
      const int ix = x > 0 ? 1 : 0;
      const double g0d = readlane_f64(d0, t);
      const double g1d = FLV > 1 ? readlane_f64(d1, t) : 0.0;
      const Proposal q = propose(FLV, g0d, g1d, ix ? df1 : df0, ix ? ds1 : ds0, ix ? bw1 : bw0);
      const double p = q.p;
      bool accept;
      neg += p < 0.0;
      if (p > 1.0) {
        accept = true;
        ties += (p - 1.0 < 1e-12);
      } else {
        const double u = readlane_f64(ul, used);
        used++;
        accept = u < p;
        ties += (fabs(u - p) < 1e-12);
      }
      p_r = lane == t ? p : p_r;
      acc_r = lane == t ? (int)accept : acc_r;
      if (!accept) continue;

      double* U0 = Uw + m * LD;
      double* U1 = Uw + (W + m) * LD;
      double* V0 = Vw + m * LD;
      double* V1 = Vw + (W + m) * LD;
      if (wave == (t & (NW - 1))) {           // owner of column t: u = e_t - G[:, t]
        const int it = t >> 2;
        const double c0 = g0[it];
        const double c1 = FLV > 1 ? g1[it] : 0.0;
        if (lane < W) {
          const double e = lane == t ? 1.0 : 0.0;
          U0[lane] = e - c0;
          if (FLV > 1) U1[lane] = e - c1;
        }
      }
      if (lane == t) {                        // row t on this wave's columns
#pragma unroll
        for (int i = 0; i < CPW; i++) {
          V0[wave * CPW + i] = g0[i];
          if (FLV > 1) V1[wave * CPW + i] = g1[i];
        }
      }
      __syncthreads();
      // G -= (gamma u) row_t on this wave's columns and on the diagonal copy; the divisions
      // overlap the reads of u
      const int lw = lane < W ? lane : W - 1;
      const double U0v = U0[lw];
      const double U1v = FLV > 1 ? U1[lw] : 0.0;
      const double gam0 = gamma_of(q.D0, q.R0);
      const double gam1 = FLV > 1 ? gamma_of(q.D1, q.R1) : 0.0;
      site_r = lane == m ? t : site_r;
      gam0_r = lane == m ? gam0 : gam0_r;
      gam1_r = lane == m ? gam1 : gam1_r;
      const double u0 = lane < W ? gam0 * U0v : 0.0;
      const double u1 = (FLV > 1 && lane < W) ? gam1 * U1v : 0.0;
#pragma unroll
      for (int i = 0; i < CPW; i++) {
        g0[i] = fma(-u0, V0[wave * CPW + i], g0[i]);
        if (FLV > 1) g1[i] = fma(-u1, V1[wave * CPW + i], g1[i]);
      }
      d0 = fma(-u0, V0[vpos<W>(lw)], d0);
      if (FLV > 1) d1 = fma(-u1, V1[vpos<W>(lw)], d1);
      m++;
    }

    if (wave == 0) {
      if (lane < m) {
        sites[lane] = site_r;
        gam[lane] = gam0_r;
        if (FLV > 1) gam[W + lane] = gam1_r;
      } else if (lane < W) {
        sites[lane] = 0;
        gam[lane] = 0.0;                  // rows beyond m come out exactly zero
        if (FLV > 1) gam[W + lane] = 0.0;
      }
      if (lane < Wn) {
        if (a.dec_log && log0 + lane < a.log_cap) {
          a.dec_log[log0 + lane] = (unsigned char)acc_r;
          if (a.p_log) a.p_log[log0 + lane] = p_r;
        }
        if (acc_r) a.conf[w0 + lane] = (int8_t)(-cfl);
      }
    }
    if (wave == 0) {
      const unsigned long long am = __ballot(acc_r != 0);
      if (lane == 0) shm[1] = am;
    }
    __syncthreads();
  }
  const unsigned long long accm = shm[1];    // accepted window positions

  if constexpr (W <= 32) {
    // Y = (I - N)^-1 X on the matrix cores: X (wave 1) and (I - N)^-1 (wave 2) through LDS
    // over the free Uw / Vw storage, then scattered to the window positions.
    if (m > 0) {
      double* Xs = Uw;                  // [b][j][c], leading dimension LD
      double* Bs = Vw;                  // [b][i][j] = Binv[j][i] (transposed), leading dimension LD
      const int f = lane >> 5, r = lane & 31;
      if ((wave == 1 || wave == 2) && f < FLV && r < W) {
        double* dst = (wave == 1 ? Xs : Bs) + (size_t)f * W * LD;
#pragma unroll
        for (int j = 0; j < W; j++) dst[j * LD + r] = j < 16 ? fx_lo[j & 15] : fx_hi[j & 15];
      }
      __syncthreads();
      constexpr int CT = W / 16, KS = W / 4;
      const int li = lane & 15, lk = lane >> 4;
      for (int unit = wave; unit < FLV * CT * CT; unit += 4) {
        const int b = unit / (CT * CT), t2 = unit - b * CT * CT;
        const int ti = t2 / CT, ci = t2 - ti * CT;
        const double* Bb = Bs + (size_t)b * W * LD;
        const double* Xb = Xs + (size_t)b * W * LD;
        dx4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < KS; s4++)
          acc = mfma16(Bb[(4 * s4 + lk) * LD + ti * 16 + li], Xb[(4 * s4 + lk) * LD + ci * 16 + li], acc);
        GT* Yb = a.Y + (size_t)b * W * W;
        const int c = ci * 16 + li;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int j = ti * 16 + mfma_row<double>(lane, q);
          if (j < m && c < m) {
            Yb[(size_t)sites[j] * W + sites[c]] = (GT)acc[q];
            if (sc) sc->yl[(size_t)b * W * LD + sites[j] * LD + sites[c]] = (GT)acc[q];
          }
        }
      }
    }
  } else {
    // Gather the m x m coefficient blocks, transposed, over the (now free) Uw/Vw storage:
    //   Lt[b][l][j] = u_l(s_j) (l < j),  Nt[b][l][j] = gamma_j G_j[s_j][s_l] (j < l),
    // zero elsewhere: column l of L and of Gamma N is a contiguous row.
    double* Lc = Uw;                 // Lt[b] at 2 b W^2, Nt[b] at (2 b + 1) W^2
    if (m > 0) {
      constexpr int E = 2 * FLV * W * W / 256;
      double gv[E];
  #pragma unroll
      for (int q = 0; q < E; q++) {
        const int e = tid + 256 * q;
        const int mat = e / (W * W), rem = e - mat * W * W;
        const int l = rem / W, j = rem - l * W;
        const int b = mat >> 1;
        // unconditional in-range loads (sites[] is 0 beyond m), then the mask
        const double vl = Uw[(b * W + l) * LD + sites[j]];
        const double vn = W <= 32 ? Vw[(b * W + j) * W + sites[l]] : Vw[(b * W + j) * LD + vpos<W>(sites[l])];
        const double gj = gam[b * W + j];
        const bool isl = (mat & 1) == 0;
        const bool keep = isl ? (l < j && j < m) : (j < l && l < m);
        gv[q] = keep ? (isl ? vl : gj * vn) : 0.0;
      }
      __syncthreads();
  #pragma unroll
      for (int q = 0; q < E; q++) Lc[tid + 256 * q] = gv[q];
      __syncthreads();
    }

    // Y = (I - N)^-1 (I + Gamma L)^-1 Gamma, wave b for flavour block b, lane = column c,
    // column c of Y in registers.  Column-oriented substitutions: once x_j (y_l) is final it
    // is subtracted from (added to) the remaining rows, so each step's own chain is one
    // multiply and one fma; the next column of coefficients is loaded during the step.
    if (wave < FLV && m > 0) {
      const int b = wave;
      const double* Lt = Lc + 2 * b * W * W;
      const double* Nt = Lc + (2 * b + 1) * W * W;
      const double* gb = gam + b * W;
      double xr[W], cc[W], cn[W] = {};
  #pragma unroll
      for (int k = 0; k < W; k++) xr[k] = lane == k ? 1.0 : 0.0;
  #pragma unroll
      for (int k = 1; k < W; k++) cc[k] = Lt[k];
      double gc = gb[0], gn = 0.0;
      // forward: x_j = gamma_j s_j, then s_k -= L[k][j] x_j for k > j
  #pragma unroll
      for (int j = 0; j < W; j++) {
        if (j + 1 < W) {                   // unconditional: columns beyond m are zero
  #pragma unroll
          for (int k = j + 2; k < W; k++) cn[k] = Lt[(j + 1) * W + k];
          gn = gb[j + 1];
        }
        if (j < m) {
          const double xj = gc * xr[j];
          xr[j] = xj;
  #pragma unroll
          for (int kb = 0; kb < W; kb += 16) {
            if (kb + 16 > j + 1 && kb < m) {
  #pragma unroll
              for (int k = (kb > j + 1 ? kb : j + 1); k < kb + 16; k++) xr[k] = fma(-cc[k], xj, xr[k]);
            }
          }
        } else {
          xr[j] = 0.0;
        }
  #pragma unroll
        for (int k = j + 2; k < W; k++) cc[k] = cn[k];
        gc = gn;
      }
      // back: y_l final in descending l, then y_k += gamma_k N[k][l] y_l for k < l
  #pragma unroll
      for (int k = 0; k < W - 1; k++) cc[k] = Nt[(W - 1) * W + k];
  #pragma unroll
      for (int l = W - 1; l >= 1; l--) {
  #pragma unroll
        for (int k = 0; k < l - 1; k++) cn[k] = Nt[(l - 1) * W + k];
        if (l < m) {
          const double yl = xr[l];
  #pragma unroll
          for (int kb = 0; kb < W; kb += 16) {
            if (kb < l && kb < m) {
  #pragma unroll
              for (int k = kb; k < (kb + 16 < l ? kb + 16 : l); k++) xr[k] = fma(cc[k], yl, xr[k]);
            }
          }
        }
  #pragma unroll
        for (int k = 0; k < l - 1; k++) cc[k] = cn[k];
      }
      // dense factor over window positions: Y[b][s_j][s_l] = y_j(l) (lane l holds column l)
      GT* Yb = a.Y + (size_t)b * W * W;
      if (lane < m) {
  #pragma unroll
        for (int j = 0; j < W; j++)
          if (j < m) Yb[(size_t)sites[j] * W + sites[lane]] = (GT)xr[j];
      }
    }
  }
  if (m > 0) {
    // zeros in the rows and columns of rejected sites (disjoint from the entries above)
#pragma unroll
    for (int q = 0; q < FLV * W * W / 256; q++) {
      const int e = tid + 256 * q;
      const int rem = e % (W * W);
      const int pr = rem / W, pc = rem - pr * W;
      if (!(((accm >> pr) & 1ull) && ((accm >> pc) & 1ull))) {
        a.Y[e] = GT(0);
        if constexpr (W <= 32) {
          if (sc) sc->yl[(size_t)(e / (W * W)) * W * LD + pr * LD + pc] = GT(0);
        }
      }
    }
  }
  if (tid == 0) a.S[0] = m;
  if (tid == 0) {
    ChainCounters* c = a.ctr;
    c->u_pos = pos0 + used;
    c->accepted += m;
    c->proposed += Wn;
    c->negative += neg;
    c->ties += ties;
    if (a.uniforms && pos0 + used > a.n_uniforms) c->exhausted = 1;
    c->log_len = log0 + Wn;
    c->flushes += 1;
  }
}

// Gd_b = Gs_b - C_b (Y_b R_b) on one 64 x 64 tile, for the previous window's flips with
// its dense factor Y over positions wp .. wp + W - 1: C = (I - Gs)[:, wp:wp+W] and
// R = Gs[wp:wp+W, :] are contiguous strips of the source copy (the update writes the other
// copy, so no tile changes what another still reads).  The tile's slice of Z = Y R
// (W x 64) is formed in LDS on the matrix cores; the rank-W product is computed
// transposed, tile^T -= Z^T C^T, so accumulator element r of lane l is
// G[r0 + l%16][c0 + row(l, r)]: quarter-waves touch 16 consecutive rows of a column.
template <typename GT, int W>
__device__ __forceinline__ void apply_tile(GT* __restrict__ Gd, const GT* __restrict__ Gs,
                                          const int* __restrict__ Sp, const GT* __restrict__ Y, int n,
                                          int wp, int tile, unsigned char* smem) {
  typedef typename V4<GT>::type acc_t;
  const int nt = (n + 63) >> 6;
  const int b = tile / (nt * nt);
  const int rem = tile - b * nt * nt;
  const int r0 = (rem % nt) * 64, c0 = (rem / nt) * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const size_t off = (size_t)b * n * n;
  const GT* Gb = Gs + off;
  GT* Gw = Gd + off;
  const int rr0 = wave * 16;
  const int grow = r0 + rr0 + li;
  const int growc = grow < n ? grow : n - 1;
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int gcol = c0 + 16 * j + mfma_row<GT>(lane, r);
      acc[j][r] = Gb[(size_t)growc + (size_t)(gcol < n ? gcol : n - 1) * n];
    }
  const int m = Sp[0];
  if (m > 0) {
    GT* Ys = reinterpret_cast<GT*>(smem);   // W x W column-major: Ys[l * W + k] = Y[k][l]
    GT* Rs = Ys + W * W;                     // W x 64  [k][col]
    GT* Zs = Rs + W * 64;                    // W x 65  [k][col]
    GT* Cs = Zs + W * 65;                    // 64 x W+1 [row][k]
    {
      // R[k][c0 + cc] = Gs[wp + k][c0 + cc]: lane = k, cc = wave + 4 q
      const int k = wp + (lane < W ? lane : W - 1);
      const size_t ko = (size_t)(k < n ? k : n - 1);
      GT v[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int gc = c0 + wave + 4 * q;
        v[q] = Gb[ko + (size_t)(gc < n ? gc : n - 1) * n];
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int gc = c0 + wave + 4 * q;
        if (lane < W) Rs[lane * 64 + wave + 4 * q] = gc < n ? v[q] : GT(0);
      }
    }
    {
      // C[r0 + rr][k] = delta - Gs[r0 + rr][wp + k]: lane = rr, k = wave + 4 q
      constexpr int QC = W / 4;
      const int gr = r0 + lane;
      const int grc = gr < n ? gr : n - 1;
      GT v[QC];
#pragma unroll
      for (int q = 0; q < QC; q++) {
        const int k = wp + wave + 4 * q;
        v[q] = Gb[(size_t)grc + (size_t)(k < n ? k : n - 1) * n];
      }
#pragma unroll
      for (int q = 0; q < QC; q++) {
        const int k = wave + 4 * q;
        Cs[lane * (W + 1) + k] = gr < n ? (gr == wp + k ? GT(1) : GT(0)) - v[q] : GT(0);
      }
    }
    {
      // Y (row-major W x W in global): Ys[l * W + k] = Y[k][l]; lane = l, k = wave + 4 q
      constexpr int QY = W / 4;
      const int lc = lane < W ? lane : W - 1;
      GT v[QY];
#pragma unroll
      for (int q = 0; q < QY; q++) v[q] = Y[(size_t)b * W * W + (size_t)(wave + 4 * q) * W + lc];
#pragma unroll
      for (int q = 0; q < QY; q++)
        if (lane < W) Ys[lane * W + wave + 4 * q] = v[q];
    }
    __syncthreads();
    // Z[:, 16 wave : 16 wave + 16] = Y R[:, same columns]
#pragma unroll
    for (int i0 = 0; i0 < W; i0 += 16) {
      acc_t z = {0, 0, 0, 0};
#pragma unroll
      for (int k0 = 0; k0 < W; k0 += 4) {
        const GT av = Ys[(k0 + lk) * W + i0 + li];        // Y[i0 + li][k0 + lk]
        const GT bv = Rs[(k0 + lk) * 64 + 16 * wave + li];  // R[k0 + lk][16 wave + li]
        z = mfma16(av, bv, z);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) Zs[(i0 + mfma_row<GT>(lane, r)) * 65 + 16 * wave + li] = z[r];
    }
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < W; k0 += 4) {
      const int k = k0 + lk;
      const GT bv = Cs[(rr0 + li) * (W + 1) + k];          // C^T[k][row]
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const GT av = -Zs[k * 65 + 16 * j + li];           // Z^T[col][k]
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

// One launch per window: workgroup 0 runs the window (decisions on the block formed from
// the previous window's factors), workgroups 1.. apply the previous window's flips to all
// of G, tile by tile, into the other copy.
template <typename GT, int FLV, int W>
__global__ __launch_bounds__(256) void window_kernel(WindowArgs<GT> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (blockIdx.x == 0) {
    window_role<GT, FLV, W>(a, smem);
    return;
  }
  apply_tile<GT, W>(a.Gd, a.G, a.Sp, a.Yp, a.n, a.wp, (int)blockIdx.x - 1, smem);
}

// The last window's application alone.
template <typename GT, int W>
__global__ __launch_bounds__(256) void window_apply_kernel(WindowArgs<GT> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  apply_tile<GT, W>(a.Gd, a.G, a.Sp, a.Yp, a.n, a.wp, (int)blockIdx.x, smem);
}

// Bounded wait until *p >= v (device scope); false when the bound is reached.
__device__ __forceinline__ bool wait_at_least(const unsigned long long* p, unsigned long long v, long long limit) {
  long long polls = 0;
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) {
    if (++polls > limit) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

// One launch per slice.  Workgroup 0 runs windows 0 .. nw - 1 back to back, as in
// window_kernel, pipelined (SliceCtl): window k's block operands were prefetched during window
// k - 1 once every application of window k - 2 was done (window k reads their output copy),
// and window k - 1's factor is published once window k's block is formed.  Workgroup 1 + i applies window j to tiles
// i, i + napply, ... once window j is published and every application of window j - 1 is
// done (whose source copy window j's application overwrites).  Every wait is bounded; a wait
// that gives up counts a device error and the launch still ends.  All workgroups are resident
// at once (one per CU, 1 + napply <= 129), and a workgroup waits only for work that needs no
// wait of its own to finish first, so the launch always drains.
template <typename GT, int FLV, int W>
__global__ __launch_bounds__(256) void window_slice_kernel(WindowSlice<GT> s) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  unsigned long long* sync = s.sync;
  const unsigned long long e = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long done = (e + 1) * (unsigned long long)s.napply;
  unsigned long long* ready = sync + 1;
  unsigned long long* applied = sync + 1 + s.nw;
  unsigned long long* errc = reinterpret_cast<unsigned long long*>(&s.a.ctr->device_errors);
  const size_t ysz = (size_t)FLV * W * W;
  const long long limit = s.a.wait_limit;
  if (blockIdx.x == 0) {
    // window k's block operands: window 0's from global memory, every later window's
    // prefetched during the previous window (from the copy that window k - 2's updates
    // write, after they are done); window k - 1 is published once window k's block is formed
    GT* pf = reinterpret_cast<GT*>(smem + window_lds<GT, FLV, W>());
    SliceCtl<GT> sc;
    sc.pf = pf;
    sc.yl = pf + (size_t)3 * FLV * W * (W + 1);
    sc.errc = errc;
    sc.limit = limit;
    for (int k = 0; k < s.nw; k++) {
      __syncthreads();
      WindowArgs<GT> a = s.a;
      a.w0 = k * W;
      a.wp = a.w0 - W;
      a.do_window = 1;
      a.G = k ? s.B[(k - 1) & 1] : s.B[0];
      a.Gd = k ? s.B[k & 1] : s.B[1];
      a.Sp = k ? s.S + ((k - 1) & 1) * (W + 1) : nullptr;
      a.Yp = k ? s.Y + ((k - 1) & 1) * ysz : nullptr;
      a.S = s.S + (k & 1) * (W + 1);
      a.Y = s.Y + (k & 1) * ysz;
      sc.pf_in = k > 0;
      sc.gnext = k + 1 < s.nw ? s.B[k & 1] : nullptr;         // window k + 1 reads B[k % 2]
      sc.wait_p = k >= 1 ? applied + (k - 1) : nullptr;
      sc.wait_v = done;
      // Where window k - 1 is published: short windows (W = 16) publish each window as soon
      // as its stores have completed (the fourth wave issues the release while the next window
      // starts), so that its updates are done before the fourth wave's next prefetch waits for
      // them; longer windows publish once the next window's block is formed, which keeps the
      // wait for the stores off their path (measured: 16 sites 71.5 -> 70.5 ms per sweep at
      // n = 256, 32 sites 73.7 -> 74.8 ms with the early form).
      constexpr bool kEarly = W <= 16;
      sc.pub_p = !kEarly && k >= 1 ? ready + (k - 1) : nullptr;
      sc.pub_v = e + 1;
      window_role<GT, FLV, W>(a, smem, &sc);
      if (kEarly || k + 1 == s.nw) {
        __builtin_amdgcn_s_waitcnt(0x0F70);     // this window's global stores have completed
        __syncthreads();
        if (tid == 192) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          __hip_atomic_store(ready + k, e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    if (tid == 0) {
      // every workgroup has read the epoch once the last window's applications are done
      if (!wait_at_least(applied + (s.nw - 1), done, limit)) atomicAdd(errc, 1ull);
      __hip_atomic_store(sync, e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  const int nt = (s.a.n + 63) >> 6;
  const int tiles = FLV * nt * nt;
  for (int j = 0; j < s.nw; j++) {
    if (tid == 0) {
      bool ok = wait_at_least(ready + j, e + 1, limit);
      if (ok && j >= 1) ok = wait_at_least(applied + (j - 1), done, limit);
      if (!ok) atomicAdd(errc, 1ull);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    for (int t = (int)blockIdx.x - 1; t < tiles; t += s.napply) {
      apply_tile<GT, W>(s.B[(j + 1) & 1], s.B[j & 1], s.S + (j & 1) * (W + 1), s.Y + (j & 1) * ysz, s.a.n,
                        j * W, t, smem);
      __syncthreads();
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (tid == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(applied + j, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <typename GT, int W>
constexpr size_t apply_lds() {
  return sizeof(GT) * ((size_t)W * W + W * 64 + W * 65 + 64 * (W + 1));
}

// Sites per window (delay_depth selects 16, 32 or 64 explicitly).  Per-flip work and the
// m x m solves grow with W, the per-window block formation and factor build with 1 / W.
// Measured with one launch per slice: 16 is faster up to n = 256 (config B 71.5 against
// 73.7 ms per sweep; the 64-chain config-D batch 323 against 313 sweeps/s), 32 beyond
// (n = 576: 326 against 335 ms; n = 1024: 373 against 406 ms).
int window_size(int n, int flv) {
  (void)flv;
  return n <= 256 ? 16 : 32;
}

template <typename GT, int FLV, int W>
static void launch_window_w(const WindowArgs<GT>& a, hipStream_t s) {
  const int nt = (a.n + 63) / 64;
  const int tiles = a.Yp ? FLV * nt * nt : 0;
  if (a.do_window) {
    const size_t lds = tiles && apply_lds<GT, W>() > window_lds<GT, FLV, W>() ? apply_lds<GT, W>()
                                                                           : window_lds<GT, FLV, W>();
    hipLaunchKernelGGL((window_kernel<GT, FLV, W>), dim3(1 + tiles), dim3(256), lds, s, a);
  } else if (tiles) {
    hipLaunchKernelGGL((window_apply_kernel<GT, W>), dim3(tiles), dim3(256), (apply_lds<GT, W>()), s, a);
  }
}

// Many chains per launch: workgroup (0, c) runs chain c's window (when it has one in this
// launch), workgroups (1 + t, c) apply chain c's previous window to tile t.  Each chain's
// work is exactly that of its own launch of window_kernel / window_apply_kernel, with the
// flavour count of that chain (one and two flavours share the launch).
template <typename GT, int W>
__global__ __launch_bounds__(256) void window_multi_kernel(const WindowArgs<GT>* __restrict__ items) {
  extern __shared__ __align__(16) unsigned char smem[];
  const WindowArgs<GT>& a = items[blockIdx.y];
  if (blockIdx.x == 0) {
    if (a.do_window) {
      if (a.flv == 1) window_role<GT, 1, W>(a, smem);
      else window_role<GT, 2, W>(a, smem);
    }
    return;
  }
  if (a.Yp == nullptr) return;
  const int nt = (a.n + 63) / 64;
  const int tile = (int)blockIdx.x - 1;
  if (tile >= a.flv * nt * nt) return;
  apply_tile<GT, W>(a.Gd, a.G, a.Sp, a.Yp, a.n, a.wp, tile, smem);
}

template <int W>
constexpr size_t window_multi_lds() {
  constexpr size_t w1 = window_lds<double, 1, W>(), w2 = window_lds<double, 2, W>(), ap = apply_lds<double, W>();
  return (w1 > w2 ? (w1 > ap ? w1 : ap) : (w2 > ap ? w2 : ap));
}

hipError_t launch_window_multi_f64(const WindowArgs<double>* items, int count, int max_flv, int W, int max_n,
                                   hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count > 65535 || max_n < 1 || max_n > 2048 || max_flv < 1 || max_flv > 2) return hipErrorInvalidValue;
  const int nt = (max_n + 63) / 64;
  const dim3 grid(1 + max_flv * nt * nt, count);
  if (W == 16) hipLaunchKernelGGL((window_multi_kernel<double, 16>), grid, dim3(256), window_multi_lds<16>(), s, items);
  else if (W == 32) hipLaunchKernelGGL((window_multi_kernel<double, 32>), grid, dim3(256), window_multi_lds<32>(), s, items);
  else if (W == 64) hipLaunchKernelGGL((window_multi_kernel<double, 64>), grid, dim3(256), window_multi_lds<64>(), s, items);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename GT>
static hipError_t launch_window_t(const WindowArgs<GT>& a, hipStream_t s) {
  if (a.flv < 1 || a.flv > 2 || (a.W != 16 && a.W != 32 && a.W != 64)) return hipErrorInvalidValue;
  if (a.Yp && (a.Sp == nullptr || a.Gd == nullptr || a.Gd == a.G || a.wp < 0 || a.wp >= a.n))
    return hipErrorInvalidValue;
  if (a.do_window && (a.w0 < 0 || a.w0 >= a.n || (a.Yp && a.wp + a.W != a.w0))) return hipErrorInvalidValue;
  if (a.flv == 1) {
    if (a.W == 16) launch_window_w<GT, 1, 16>(a, s);
    else if (a.W == 32) launch_window_w<GT, 1, 32>(a, s);
    else launch_window_w<GT, 1, 64>(a, s);
  } else {
    if (a.W == 16) launch_window_w<GT, 2, 16>(a, s);
    else if (a.W == 32) launch_window_w<GT, 2, 32>(a, s);
    else launch_window_w<GT, 2, 64>(a, s);
  }
  return hipGetLastError();
}
hipError_t launch_window_f64(const WindowArgs<double>& a, hipStream_t s) { return launch_window_t(a, s); }

template <int FLV, int W>
static void launch_slice_w(const WindowSlice<double>& s, hipStream_t st) {
  const size_t wl = window_lds<double, FLV, W>() + slice_extra_lds<double, FLV, W>(), al = apply_lds<double, W>();
  static_assert(window_lds<double, FLV, W>() % 16 == 0, "prefetch area alignment");
  static_assert(window_lds<double, FLV, W>() + slice_extra_lds<double, FLV, W>() <= 160 * 1024, "LDS");
  hipLaunchKernelGGL((window_slice_kernel<double, FLV, W>), dim3(1 + s.napply), dim3(256), wl > al ? wl : al, st, s);
}

hipError_t launch_window_slice_f64(const WindowSlice<double>& s, hipStream_t st) {
  const WindowArgs<double>& a = s.a;
  if (a.flv < 1 || a.flv > 2 || (a.W != 16 && a.W != 32) || a.n < 1) return hipErrorInvalidValue;
  const int nt = (a.n + 63) / 64;
  if (s.nw != (a.n + a.W - 1) / a.W || s.napply < 1 || s.napply > 128 || s.napply > a.flv * nt * nt)
    return hipErrorInvalidValue;
  if (!s.B[0] || !s.B[1] || s.B[0] == s.B[1] || !s.Y || !s.S || !s.sync || !a.ctr || !a.conf)
    return hipErrorInvalidValue;
  if (a.flv == 1) {
    if (a.W == 16) launch_slice_w<1, 16>(s, st);
    else launch_slice_w<1, 32>(s, st);
  } else {
    if (a.W == 16) launch_slice_w<2, 16>(s, st);
    else launch_slice_w<2, 32>(s, st);
  }
  return hipGetLastError();
}
hipError_t launch_window_f32(const WindowArgs<float>& a, hipStream_t s) { return launch_window_t(a, s); }

hipError_t local_kernels_init() {
  const void* fns[] = {reinterpret_cast<const void*>(window_kernel<double, 1, 16>),
                       reinterpret_cast<const void*>(window_kernel<double, 1, 32>),
                       reinterpret_cast<const void*>(window_kernel<double, 1, 64>),
                       reinterpret_cast<const void*>(window_kernel<double, 2, 16>),
                       reinterpret_cast<const void*>(window_kernel<double, 2, 32>),
                       reinterpret_cast<const void*>(window_kernel<double, 2, 64>),
                       reinterpret_cast<const void*>(window_kernel<float, 1, 16>),
                       reinterpret_cast<const void*>(window_kernel<float, 1, 32>),
                       reinterpret_cast<const void*>(window_kernel<float, 1, 64>),
                       reinterpret_cast<const void*>(window_kernel<float, 2, 16>),
                       reinterpret_cast<const void*>(window_kernel<float, 2, 32>),
                       reinterpret_cast<const void*>(window_kernel<float, 2, 64>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 16>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 32>),
                       reinterpret_cast<const void*>(window_apply_kernel<double, 64>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 16>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 32>),
                       reinterpret_cast<const void*>(window_apply_kernel<float, 64>),
                       reinterpret_cast<const void*>(window_multi_kernel<double, 16>),
                       reinterpret_cast<const void*>(window_multi_kernel<double, 32>),
                       reinterpret_cast<const void*>(window_multi_kernel<double, 64>),
                       reinterpret_cast<const void*>(window_slice_kernel<double, 1, 16>),
                       reinterpret_cast<const void*>(window_slice_kernel<double, 1, 32>),
                       reinterpret_cast<const void*>(window_slice_kernel<double, 2, 16>),
                       reinterpret_cast<const void*>(window_slice_kernel<double, 2, 32>)};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dqmc
