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
//    matrix is streamed once per panel instead of once per column.  The same launch
//    accumulates Q = H_1 ... H_P forward on 16-row strips of U (U <- U (I - V T V^T),
//    the first panel's launch starting from the identity without reading U), so Q needs
//    no launches of its own.
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
constexpr int kPanelDoubles = 4 * 64 * 17;  // the column steps' reduction transposes

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

// The UDT runs as a sequence of launches per batch (one stream, so each sees the previous
// one's results): column norms; per panel a single-workgroup panel kernel (pivot choice,
// column moves, Householder factorisation of the panel) and a many-workgroup launch that
// forms the panel's compact-WY factor T, applies the block reflector to the trailing matrix
// and accumulates Q forward on U's rows (U <- U H); finally D and T.  Per-item state
// between launches: it.tau holds the residual norms, it.pivot the working permutation,
// it.tfac the taus of each panel.

// Rows of the panel per thread of the panel kernel (256 threads): 1, 2, 4 or 8.

__host__ __device__ static int udt_rows_per_thread(int rows) {
  int R = 1;
  while (R * 256 < rows) R *= 2;
  return R;
}

// Panel width for `rows` remaining rows: the panel lives in registers (R rows x w columns
// per thread, at most 64 values).
__host__ __device__ static int udt_panel_width(int rows) {
  int w = 64 / udt_rows_per_thread(rows);
  if (w > kQrMaxW) w = kQrMaxW;
  if (w > rows) w = rows;
  if (w < 1) w = 1;
  return w;
}

// TAB (all UDT and rdivp kernels): items from the device table batch.items with their own n
// and panel (many-chain launches), else from the kernel arguments (a separate instantiation
// for the one-chain path)
template <bool TAB>
__global__ __launch_bounds__(256) void udt_norms_kernel(UdtBatch batch) {
  constexpr bool tab = TAB;
  const UdtItem it = tab ? batch.items[blockIdx.y] : batch.it[blockIdx.y];
  const int n = tab ? it.n : batch.n;
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

// (a, ia) comes before (b, ib): larger norm first, the lower column on ties.
__device__ __forceinline__ bool pivot_before(double a, int ia, double b, int ib) {
  return a > b || (a == b && ia < ib);
}

// Value of lane (lane ^ S).  Inside a 16-lane row with DPP: xor 1 and xor 2 are quad
// permutations, xor 4 a half-row mirror (i <-> 7 - i) followed by a quad reversal, xor 8
// a row rotation by 8.  Across rows with the gfx950 swaps of two registers (here two
// copies of x): permlane16_swap exchanges the odd rows of the first with the even rows of
// the second, so lanes of odd rows find lane - 16 in the first result and lanes of even
// rows lane + 16 in the second; permlane32_swap does the same with the two 32-lane halves.
template <int S>
__device__ __forceinline__ int lane_xor(int x) {
  if constexpr (S == 1) {
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
  } else if constexpr (S == 2) {
    return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
  } else if constexpr (S == 4) {
    const int m = __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return __builtin_amdgcn_update_dpp(0, m, 0x1B, 0xF, 0xF, false);    // quad_perm [3,2,1,0]
  } else if constexpr (S == 8) {
    return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);   // row_ror 8
  } else if constexpr (S == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (__lane_id() & 16) ? (int)r[0] : (int)r[1];
  } else {
    static_assert(S == 32, "lane_xor: strides 1 .. 32");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (__lane_id() & 32) ? (int)r[0] : (int)r[1];
  }
}

template <int S>
__device__ __forceinline__ double lane_xor_d(double x) {
  const int lo = lane_xor<S>(__double2loint(x));
  const int hi = lane_xor<S>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

template <int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_stage(double& kv, int& ki, int lane) {
  const int lo = lane_xor<STRIDE>(__double2loint(kv));
  const int hi = lane_xor<STRIDE>(__double2hiint(kv));
  const double pv = __hiloint2double(hi, lo);
  const int pi = lane_xor<STRIDE>(ki);
  const bool up = (lane & SIZE) == 0 || SIZE == 64;
  const bool lower = (lane & STRIDE) == 0;
  const bool pb = pivot_before(pv, pi, kv, ki);
  const bool take = lower == up ? pb : !pb;
  kv = take ? pv : kv;
  ki = take ? pi : ki;
  if constexpr (STRIDE > 1) bitonic_stage<SIZE, STRIDE / 2>(kv, ki, lane);
}

// Bitonic sort of one (norm, column) pair per lane of a wave into pivot order.
__device__ __forceinline__ void bitonic64(double& kv, int& ki, int lane) {
  bitonic_stage<2, 1>(kv, ki, lane);
  bitonic_stage<4, 2>(kv, ki, lane);
  bitonic_stage<8, 4>(kv, ki, lane);
  bitonic_stage<16, 8>(kv, ki, lane);
  bitonic_stage<32, 16>(kv, ki, lane);
  bitonic_stage<64, 32>(kv, ki, lane);
}

static size_t udt_panel_lds_bytes(int n) {
  size_t b = 0;
  b += (size_t)n * 8;                 // nrm
  b += (size_t)kPanelDoubles * 8;     // the factorisation's reduction transposes
  b += 32 * 8;                        // taus, scalars
  b += 2 * (4 * 16 + 16) * 8;         // wave totals, row j (two buffers)
  b += 4 * 16 * 16;                   // top-16 candidates of each wave
  b += (size_t)n * 4;                 // perm
  b += 3 * 32 * 4 + 16;               // moves, selection, misc
  return b;
}

// Column steps J.. of the single-wave panel factorisation (rows <= 256: lane l holds rows
// l + 64 k; rows at or beyond `rows` hold zeros, so they add nothing and stay zero).  A
// template recursion, so every register index is a compile-time constant, and branch-free
// inside a step: only row block k = 0 holds rows <= j, selected per lane.
// Per step: the partial sums of ||x||^2 and x^T a_c (c > j) are transposed through LDS
// (lane 4 v + g adds value v of 16 lanes, quads combined with DPP) together with row j,
// so lane 4 c holds column c's sum and a_jc and computes its w_c; the w_c reach every
// lane by readlane.
template <int J, int RT, int WM>
__device__ __forceinline__ void panel_steps_wave(double (&a)[RT][WM], double (&taus)[WM], double* rw, int lane,
                                                 int w) {
  if constexpr (J < WM) {
    constexpr int j = J;
    taus[j] = 0.0;
    if (j < w) {
      double* rj = rw + 64 * 17;          // row j of the panel
      double part[WM];
      {
        const double x0 = a[0][j];
        const double xe = lane >= j ? x0 : 0.0;
        const double xg = lane > j ? x0 : 0.0;
        part[j] = xe * xe;
#pragma unroll
        for (int c = j + 1; c < WM; c++) part[c] = xg * a[0][c];
      }
#pragma unroll
      for (int k = 1; k < RT; k++) {
        const double x = a[k][j];
        part[j] += x * x;
#pragma unroll
        for (int c = j + 1; c < WM; c++) part[c] += x * a[k][c];
      }
#pragma unroll
      for (int c = j; c < WM; c++) rw[lane * 17 + c] = part[c];
      if (lane == j) {
#pragma unroll
        for (int c = j; c < WM; c++) rj[c] = a[0][c];
      }
      __builtin_amdgcn_wave_barrier();
      const int v = lane >> 2, g = lane & 3;
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 16; q++) s4[q & 3] += rw[(g * 16 + q) * 17 + v];
      const double ajv = rj[v];
      double sv = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      __builtin_amdgcn_wave_barrier();
      sv += dpp_mov<0xB1>(sv);             // quad_perm [1,0,3,2]
      sv += dpp_mov<0x4E>(sv);             // quad_perm [2,3,0,1]
      const double normu = readlane_d(sv, 4 * j);
      const double xi1 = readlane_d(ajv, 4 * j);
      if (normu != 0.0) {                  // wave-uniform
        const double nu = copysign(sqrt(normu), xi1);
        const double xi = xi1 + nu;
        const double tau = xi / nu;
        const double rxi = 1.0 / xi;
        const double wl = tau * (ajv + sv * rxi);   // w_v in lanes 4 v .. 4 v + 3
        double wc[WM];
#pragma unroll
        for (int c = j + 1; c < WM; c++) wc[c] = readlane_d(wl, 4 * c);
        {
          const double x0 = a[0][j];
          const double vv = x0 * rxi;
          const double coef = lane > j ? vv : (lane == j ? 1.0 : 0.0);
          a[0][j] = lane > j ? vv : (lane == j ? -nu : x0);
#pragma unroll
          for (int c = j + 1; c < WM; c++) a[0][c] -= coef * wc[c];
        }
#pragma unroll
        for (int k = 1; k < RT; k++) {
          const double vv = a[k][j] * rxi;
          a[k][j] = vv;
#pragma unroll
          for (int c = j + 1; c < WM; c++) a[k][c] -= vv * wc[c];
        }
        taus[j] = tau;
      }
    }
    panel_steps_wave<J + 1, RT, WM>(a, taus, rw, lane, w);
  }
}

// Column steps J.. of the four-wave panel factorisation (thread t holds rows t + 256 k).
// A template recursion, so every register index is a compile-time constant.  The row-j
// and total buffers alternate with the step parity, so one barrier per step suffices.
template <int J, int R, int WM>
__device__ __forceinline__ void panel_steps_block(double (&a)[R][WM], double* red, double* wt, double* rowj,
                                                  double* sc, int tid, int lane, int wave, int rows, int w) {
  if constexpr (J < WM) {
    constexpr int j = J;
    if (j < w) {
    const int bsel = j & 1;
    double* wtb = wt + bsel * 64;
    double* rjb = rowj + bsel * 16;
    double part[WM];
#pragma unroll
    for (int c = 0; c < WM; c++) part[c] = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int i = tid + 256 * k;
      if (i < rows && i >= j) {
        const double x = a[k][j];
        part[j] += x * x;
        if (i > j) {
#pragma unroll
          for (int c = j + 1; c < WM; c++) part[c] += x * a[k][c];
        }
      }
      if (i == j) {
#pragma unroll
        for (int c = 0; c < WM; c++) rjb[c] = a[k][c];
      }
    }
    double* rw = red + (size_t)(wave * 64) * 17;
#pragma unroll
    for (int c = 0; c < WM; c++) rw[lane * 17 + c] = part[c];
    __builtin_amdgcn_wave_barrier();
    {
      constexpr int GL = 64 / WM;            // lanes per value after the transpose
      const int v = lane / GL, g = lane % GL;
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < WM; q++) s4[q & 3] += rw[(g * WM + q) * 17 + v];
      double sv = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      sv += dpp_mov<0xB1>(sv);             // quad_perm [1,0,3,2]
      sv += dpp_mov<0x4E>(sv);             // quad_perm [2,3,0,1]
      if (GL == 8) sv += dpp_mov<0x141>(sv);   // row_half_mirror: the other quad of 8
      if (g == 0) wtb[wave * 16 + v] = sv;
    }
    __syncthreads();
    double tot;
    {
      const int v = lane % WM;
      tot = (wtb[v] + wtb[16 + v]) + (wtb[32 + v] + wtb[48 + v]);
    }
    const double normu = readlane_d(tot, j);
    const double xi1 = rjb[j];
    double tau = 0.0, nu = 0.0, xi = 1.0;
    if (normu != 0.0) {
      nu = copysign(sqrt(normu), xi1);
      xi = xi1 + nu;
      tau = xi / nu;
    }
    const double rxi = 1.0 / xi;
    double wc[WM];
#pragma unroll
    for (int c = 0; c < WM; c++) wc[c] = c > j ? tau * (rjb[c] + readlane_d(tot, c) * rxi) : 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int i = tid + 256 * k;
      if (normu != 0.0 && i < rows) {
        if (i > j) {
          const double vv = a[k][j] * rxi;
          a[k][j] = vv;
#pragma unroll
          for (int c = j + 1; c < WM; c++) a[k][c] -= vv * wc[c];
        } else if (i == j) {
          a[k][j] = -nu;
#pragma unroll
          for (int c = j + 1; c < WM; c++) a[k][c] -= wc[c];
        }
      }
    }
    if (tid == 0) sc[j] = tau;
    }
    panel_steps_block<J + 1, R, WM>(a, red, wt, rowj, sc, tid, lane, wave, rows, w);
  }
}

// Householder panel of one item: pivot choice for the panel, column moves, factorisation
// with the panel rows in registers (R rows x WM columns per thread), the taus.
// Per column step one block reduction gives ||x||^2 and x^T a_c for the later columns:
// with v = (x + nu e_j) / (x_j + nu), v^T a_c = a_jc + (x_{>j}^T a_c{>j}) / (x_j + nu).
template <int R, int WM, bool TAB>
__global__ __launch_bounds__(kQrThreads) void udt_panel_kernel(UdtBatch batch, int p0a, int wa, int lasta) {
  constexpr bool tab = TAB;
  const UdtItem it = tab ? batch.items[blockIdx.x] : batch.it[blockIdx.x];
  const int n = tab ? it.n : batch.n;
  const int p0 = tab ? it.p0 : p0a, w = tab ? it.w : wa, last = tab ? it.last : lasta;
  const int rows = n - p0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  extern __shared__ __align__(16) unsigned char smem[];
  double* nrm = reinterpret_cast<double*>(smem);
  double* P = nrm + n;
  double* sc = P + kPanelDoubles;   // sc[0..15] taus of the panel
  double* red = P;               // [4][64][17] reduction transposes
  double* wt = sc + 32;          // [2][4][16] wave totals, by step parity
  double* rowj = wt + 2 * 64;    // [2][16] row j, by step parity
  double* candv = rowj + 32;     // [64] top-16 norms of each wave
  int* candi = reinterpret_cast<int*>(candv + 64);   // [64] their columns
  int* perm = candi + 64;
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

  // the w largest residual norms, ties to the lower index (reference: strict ">"); wave 0
  // lane t < w ends with the t-th choice in selk
  int selk = p0;
  if (rows <= kQrThreads) {
    // bitonic sort of (norm, column) best-first inside each wave; the waves' top 16 are
    // merged in wave 0, which keeps the result in registers.  With the odd waves' lists
    // stored reversed, the 64 candidates are in the state the bitonic network reaches after
    // its size-16 stage (runs alternately best-first and reversed), so only the size-32 and
    // size-64 merges remain.
    const int k = p0 + tid;
    double kv = k < n ? nrm[k] : -HUGE_VAL;
    int ki = k < n ? k : n + tid;
    bitonic64(kv, ki, lane);
    if (lane < 16) {
      const int slot = wave * 16 + ((wave & 1) ? 15 - lane : lane);
      candv[slot] = kv;
      candi[slot] = ki;
    }
    __syncthreads();
    if (wave == 0) {
      kv = candv[lane];
      ki = candi[lane];
      bitonic_stage<32, 16>(kv, ki, lane);
      bitonic_stage<64, 32>(kv, ki, lane);
      selk = ki;
    }
  } else {
    for (int k = p0 + tid; k < n; k += kQrThreads) {
      const double v = nrm[k];
      int cnt = 0;
      for (int k0 = p0; k0 < n; k0 += 16) {
        double u[16];
#pragma unroll
        for (int q = 0; q < 16; q++) u[q] = nrm[k0 + q < n ? k0 + q : n - 1];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int k2 = k0 + q;
          cnt += k2 < n && ((u[q] > v) || (u[q] == v && k2 < k));
        }
      }
      if (cnt < w) selr[cnt] = k;
    }
    __syncthreads();
    if (wave == 0) selk = lane < w ? selr[lane] : p0;
  }

  // Sequential swaps "slot p0+t <-> slot holding the t-th choice" simulated by wave 0 on
  // the at most 2w slots involved: lane q < w is slot p0+q, lane w+q the slot of choice q
  // when that lies outside the panel.  cont = original column now held by the lane's slot.
  if (wave == 0) {
    const int selv = lane < w ? selk : p0;
    const bool ok = __ballot(lane < w && !(selv >= p0 && selv < n)) == 0ull;
    const int selc = ok ? selv : p0 + lane;                    // lane t: t-th choice
    // (all lanes take part in the shuffle; only then is the result selected)
    const int cw = __shfl(selc, (lane + 64 - w) & 63, 64);
    int slot = -1;
    if (lane < w) slot = p0 + lane;
    else if (lane < 2 * w) slot = cw >= p0 + w ? cw : -1;
    int cont = slot;
    int broken = 0;
    auto swap_step = [&](int t) {
      const int c = __builtin_amdgcn_readlane(selc, t);
      const unsigned long long hold = __ballot(slot >= 0 && cont == c);
      // every choice is held by exactly one tracked slot; an empty ballot means a broken
      // selection: the swap is skipped and the event counted as a device error
      broken += hold ? 0 : 1;
      const int src = hold ? __ffsll((long long)hold) - 1 : t;  // lane whose slot holds c
      const int a_cont = __builtin_amdgcn_readlane(cont, t);      // content of slot p0+t
      const int b_cont = __builtin_amdgcn_readlane(cont, src);
      if (lane == t) cont = b_cont;
      if (lane == src && src != t) cont = a_cont;
    };
    if (w == kQrMaxW) {       // full panels: straight-line steps
#pragma unroll
      for (int t = 0; t < kQrMaxW; t++) swap_step(t);
    } else {
      for (int t = 0; t < w; t++) swap_step(t);
    }
    if (broken && lane == 0 && it.err) atomicAdd(it.err, (unsigned long long)broken);
    const bool moved = slot >= 0 && cont != slot;
    const unsigned long long mm = __ballot(moved);
    const int idx = __popcll(mm & ((1ull << lane) - 1ull));
    if (moved) {
      mv_dst[idx] = slot;
      mv_src[idx] = cont;
    }
    if (lane == 0) misc[0] = __popcll(mm);
    // panel column q (lane q < w) comes from original column cont
    if (lane < w) selr[lane] = cont;
  }
  __syncthreads();

  // physical column moves (whole columns) and the panel registers, both read from the
  // source columns before anything is written
  const int nm = misc[0];
  // R == 1 (rows <= 256): wave 0 alone factorises, lane l holding rows l + 64 k (k < 4);
  // otherwise all 256 threads, thread t holding rows t + 256 k (k < R)
  constexpr int RT = R == 1 ? 4 : R;         // rows per participating thread
  constexpr int TP = R == 1 ? 64 : 256;      // participating threads
  // rows at or beyond `rows` are zero (the factorisation relies on it); threads beyond TP
  // hold no rows
  double a[RT][WM];
  if (tid < TP) {
#pragma unroll
    for (int k = 0; k < RT; k++) {
      const int i = tid + TP * k;
      const int ic = i < rows ? i : rows - 1;
#pragma unroll
      for (int c = 0; c < WM; c++) {
        const int src = c < w ? selr[c] : selr[0];
        a[k][c] = A[(size_t)(p0 + ic) + (size_t)src * n];
      }
    }
#pragma unroll
    for (int k = 0; k < RT; k++) {
      const bool ok = tid + TP * k < rows;
#pragma unroll
      for (int c = 0; c < WM; c++) a[k][c] = ok ? a[k][c] : 0.0;
    }
  }
  if constexpr (R == 1) {
  // wave 0 factorises the panel while waves 1-3 move the columns.  Every source read of
  // the panel registers is issued before the barrier and every column store after it
  // (the vector memory operations of one workgroup reach memory in issue order).  The
  // moves leave out the rows >= p0 of the panel's columns, which the panel store writes.
  // Single-wave factorisation: per column one wave reduction (partial sums and row j
  // transposed through LDS, quads combined with DPP); no barrier.
  if (wave == 1) {
    int pn = 0;
    if (lane < nm) pn = perm[mv_src[lane]];
    __builtin_amdgcn_wave_barrier();
    if (lane < nm) perm[mv_dst[lane]] = pn;
  }
  __syncthreads();
  if (wave == 0) {
    double taus[WM];
    panel_steps_wave<0, RT, WM>(a, taus, red, lane, w);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < WM; j++) sc[j] = taus[j];
    }
  } else {
    const int t = tid - 64;
    for (int r0 = 0; r0 < n; r0 += kQrThreads - 64) {
      const int rr = r0 + t;
      const int rc = rr < n ? rr : n - 1;
      double v[2 * kQrMaxW];
#pragma unroll
      for (int q = 0; q < 2 * kQrMaxW; q++) {
        const int src = q < nm ? mv_src[q] : p0;
        v[q] = A[(size_t)rc + (size_t)src * n];
      }
#pragma unroll
      for (int q = 0; q < 2 * kQrMaxW; q++) {
        const int dst = mv_dst[q < nm ? q : 0];
        const int lim = dst < p0 + w ? p0 : n;
        if (q < nm && rr < lim) A[(size_t)rr + (size_t)dst * n] = v[q];
      }
    }
  }
  __syncthreads();
  for (int k = p0 + tid; k < n; k += kQrThreads) it.pivot[k] = perm[k];
  } else {
  int pnew = 0;
  if (tid < nm) pnew = perm[mv_src[tid]];
  for (int r0 = 0; r0 < n; r0 += kQrThreads) {
    const int rr = r0 + tid;
    const int rc = rr < n ? rr : n - 1;
    double v[2 * kQrMaxW];
#pragma unroll
    for (int q = 0; q < 2 * kQrMaxW; q++) {
      const int src = q < nm ? mv_src[q] : p0;
      v[q] = A[(size_t)rc + (size_t)src * n];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2 * kQrMaxW; q++)
      if (q < nm && rr < n) A[(size_t)rr + (size_t)mv_dst[q] * n] = v[q];
  }
  if (tid < nm) perm[mv_dst[tid]] = pnew;
  __syncthreads();
  for (int k = p0 + tid; k < n; k += kQrThreads) it.pivot[k] = perm[k];
  // factorisation, one block reduction per column: partial sums go through LDS within
  // each wave (lane (v, g) adds value v of one group of lanes), the groups are combined
  // with DPP inside quads / half rows, the four wave totals through LDS and one barrier.
  panel_steps_block<0, R, WM>(a, red, wt, rowj, sc, tid, lane, wave, rows, w);
  __syncthreads();
  }
  // panel -> A, and the panel's taus for the reflector launch, which forms S = V^T V and
  // the compact-WY factor T from them (sc holds the taus since the factorisation's barrier)
#pragma unroll
  for (int k = 0; k < RT; k++) {
    const int i = tid + TP * k;
    if (tid < TP && i < rows) {
#pragma unroll
      for (int c = 0; c < WM; c++)
        if (c < w) {
          A[(size_t)(p0 + i) + (size_t)(p0 + c) * n] = a[k][c];
          if (last && c == i) {   // D = |diag R| (0 -> 1) of this panel
            const double x = fabs(a[k][c]);
            it.D[p0 + i] = x == 0.0 ? 1.0 : x;
          }
        }
    }
  }
  if (tid >= kQrThreads - kQrMaxW) {
    const int j = tid - (kQrThreads - kQrMaxW);
    it.tfac[(size_t)p0 * 16 + j] = j < w ? sc[j] : 0.0;
  }
  if (last) {
    // D = |diag R| (0 -> 1) of the earlier panels (final since their launches)
    for (int i = tid; i < p0; i += kQrThreads) {
      const double x = fabs(A[(size_t)i + (size_t)i * n]);
      it.D[i] = x == 0.0 ? 1.0 : x;
    }
  }
}

// One workgroup per 16-column strip: X[p0:n, strip] -= V op(T) V^T X[p0:n, strip].
// mode 0 = trailing update of A with T^T, plus the residual norms of the updated columns
// over rows >= p0 + w; mode 1 = Q accumulation into U with T.  The four waves split the
// 16-row blocks of the strip.  X is held in the f64 MFMA B/D layout (lane l: rows
// l/16 + 4r of a block, column l%16), which is at once the B operand of W = V^T X and
// the accumulator of X - V Y, so each block is loaded once; V is read in both
// orientations.  All loads of a group of blocks are issued before the first product.
constexpr int kReflRB = 4;   // 16-row blocks per wave and group

__device__ __forceinline__ double vpanel(const double* Vp, int n, int rows, int w, int i, int a) {
  if (a >= w || i >= rows || i < a) return 0.0;
  if (i == a) return 1.0;
  return Vp[(size_t)i + (size_t)a * n];
}

__device__ __forceinline__ void reflect_strip(const UdtItem& it, int n, int p0, int w, int cb, int mode,
                                              double (&red)[4][4][64], double (&sred)[4][4][64],
                                              double (&stm)[4][2][256], double (&nred)[4][16]) {
  const int rows = n - p0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const bool trans = mode != 1;
  const double* Vp = it.A + (size_t)p0 * n + p0;
  const int col = cb + li;
  const bool colok = col < n;
  const int colc = colok ? col : n - 1;
  // X[i][c] (i >= p0): column cb + c of A (mode 0) or of U (mode 1), or row cb + c of U
  // (mode 2: X = U^T, so X -= V T^T V^T X is U <- U - U V T V^T on those rows)
  const size_t xs = mode == 2 ? (size_t)n : 1;
  double* Xc = mode == 2 ? it.U + colc + (size_t)p0 * n : (mode == 0 ? it.A : it.U) + (size_t)colc * n + p0;
  // the first panel's Q strips start from the identity (U is not read; every column of the
  // strip is written)
  const bool ident = mode == 2 && p0 == 0;
  // the panel's taus (zero beyond w)
  double tz[16];
#pragma unroll
  for (int j = 0; j < 16; j++) tz[j] = it.tfac[(size_t)p0 * 16 + j];
  const int nblk = (rows + 15) / 16;
  const int ngroups = (nblk + 4 * kReflRB - 1) / (4 * kReflRB);
  double xb[kReflRB][4], vb[kReflRB][4], va[kReflRB][4];
  dx4 Wacc = dx4{0, 0, 0, 0};
  dx4 Sacc = dx4{0, 0, 0, 0};   // S = V^T V: vb is both operands
  for (int g = 0; g < ngroups; g++) {
#pragma unroll
    for (int q = 0; q < kReflRB; q++) {
      const int blk = g * 4 * kReflRB + wave + 4 * q;
#pragma unroll
      for (int kc = 0; kc < 4; kc++) {
        const int i = 16 * blk + 4 * kc + lk;
        const int ic = i < rows ? i : rows - 1;
        const double x = ident ? (ic == colc ? 1.0 : 0.0) : Xc[ic * xs];
        xb[q][kc] = (i < rows && colok) ? x : 0.0;
        vb[q][kc] = vpanel(Vp, n, rows, w, i, li);
      }
    }
    if (ngroups == 1) {      // the second product's operand, loaded with the first's
#pragma unroll
      for (int q = 0; q < kReflRB; q++) {
        const int blk = wave + 4 * q;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) va[q][kc] = vpanel(Vp, n, rows, w, 16 * blk + li, 4 * kc + lk);
      }
    }
#pragma unroll
    for (int q = 0; q < kReflRB; q++)
#pragma unroll
      for (int kc = 0; kc < 4; kc++) {
        Wacc = mfma16(vb[q][kc], xb[q][kc], Wacc);
        Sacc = mfma16(vb[q][kc], vb[q][kc], Sacc);
      }
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    red[wave][r][lane] = Wacc[r];
    sred[wave][r][lane] = Sacc[r];
  }
  __syncthreads();
  dx4 Wt;
#pragma unroll
  for (int r = 0; r < 4; r++) Wt[r] = (red[0][r][lane] + red[1][r][lane]) + (red[2][r][lane] + red[3][r][lane]);
  // compact-WY factor T (upper triangular, 16 x 16, zero beyond w): H_1..H_w = I - V T V^T.
  // Every wave forms it on its own (no workgroup barrier): S row-major into the wave's
  // space, then lane i builds row i of T from its own earlier entries,
  // T[i][j] = -tau_j sum_{k=i}^{j-1} T[i][k] S[k][j] (T[i][k] = 0 for k < i, so the sum may
  // start at 0), with S[k][j] read from LDS, the same address in every lane.
  double ta[4];
  {
    double* Sm = stm[wave][0];
    double* Tm = stm[wave][1];
#pragma unroll
    for (int r = 0; r < 4; r++)
      Sm[(4 * r + lk) * 16 + li] = (sred[0][r][lane] + sred[1][r][lane]) + (sred[2][r][lane] + sred[3][r][lane]);
    __builtin_amdgcn_wave_barrier();
    double row[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < j; k++) acc += row[k] * Sm[k * 16 + j];
      double t = 0.0;
      if (lane < w && j < w) {
        if (j == lane) t = tz[j];
        else if (j > lane) t = -tz[j] * acc;
      }
      row[j] = t;
    }
    if (lane < 16) {
#pragma unroll
      for (int j = 0; j < 16; j++) Tm[lane * 16 + j] = row[j];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
      const int kk = 4 * kc + lk;
      ta[kc] = trans ? Tm[kk * 16 + li] : Tm[li * 16 + kk];
    }
  }
  dx4 Y = dx4{0, 0, 0, 0};
#pragma unroll
  for (int kc = 0; kc < 4; kc++) Y = mfma16(ta[kc], Wt[kc], Y);

  double ss = 0.0;
  for (int g = 0; g < ngroups; g++) {
    if (ngroups > 1) {
#pragma unroll
      for (int q = 0; q < kReflRB; q++) {
        const int blk = g * 4 * kReflRB + wave + 4 * q;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
          const int i = 16 * blk + 4 * kc + lk;
          const int ic = i < rows ? i : rows - 1;
          const double x = ident ? (ic == colc ? 1.0 : 0.0) : Xc[ic * xs];
          xb[q][kc] = (i < rows && colok) ? x : 0.0;
        }
      }
#pragma unroll
      for (int q = 0; q < kReflRB; q++) {
        const int blk = g * 4 * kReflRB + wave + 4 * q;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) va[q][kc] = vpanel(Vp, n, rows, w, 16 * blk + li, 4 * kc + lk);
      }
    }
#pragma unroll
    for (int q = 0; q < kReflRB; q++) {
      const int blk = g * 4 * kReflRB + wave + 4 * q;
      if (16 * blk < rows) {
        dx4 acc;
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] = xb[q][r];
#pragma unroll
        for (int kc = 0; kc < 4; kc++) acc = mfma16(-va[q][kc], Y[kc], acc);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 16 * blk + lk + 4 * r;
          if (i < rows && colok) {
            Xc[i * xs] = acc[r];
            if (i >= w) ss += acc[r] * acc[r];
          }
        }
      }
    }
  }
  if (mode == 0) {
    ss += lane_xor_d<16>(ss);
    ss += lane_xor_d<32>(ss);
    if (lk == 0) nred[wave][li] = ss;
    __syncthreads();
    if (tid < 16 && cb + tid < n)
      it.tau[cb + tid] = (nred[0][tid] + nred[1][tid]) + (nred[2][tid] + nred[3][tid]);
  }
}

// Panel p0's block reflector applied in one launch: workgroups [0, nA) update the trailing
// 16-column strips of A from column c0 (with residual norms), the rest accumulate Q on
// 16-row strips of U (U <- U H, forward: U holds H_1 ... H_k after panel k).
template <bool TAB>
__global__ __launch_bounds__(256) void udt_reflect_kernel(UdtBatch batch, int p0, int w, int c0, int nA) {
  __shared__ double red[4][4][64];
  __shared__ double sred[4][4][64];
  __shared__ double stm[4][2][256];
  __shared__ double nred[4][16];
  const int bx = blockIdx.x;
  if constexpr (TAB) {
    // item table: this item's own panel; workgroups beyond its strips leave at once
    const UdtItem& it = batch.items[blockIdx.y];
    const int n = it.n, c0i = it.p0 + it.w;
    const int nAi = c0i < n ? (n - c0i + 15) / 16 : 0, nQi = (n + 15) / 16;
    if (bx >= nAi + nQi) return;
    if (bx < nAi) reflect_strip(it, n, it.p0, it.w, c0i + 16 * bx, 0, red, sred, stm, nred);
    else reflect_strip(it, n, it.p0, it.w, 16 * (bx - nAi), 2, red, sred, stm, nred);
    return;
  }
  if (bx < nA) reflect_strip(batch.it[blockIdx.y], batch.n, p0, w, c0 + 16 * bx, 0, red, sred, stm, nred);
  else reflect_strip(batch.it[blockIdx.y], batch.n, p0, w, 16 * (bx - nA), 2, red, sred, stm, nred);
}

// T = D^-1 R (pivoted: column j of R goes to column pivot[j]; dirty: the strict lower part
// is copied), one wave per column.
template <bool TAB>
__global__ __launch_bounds__(256) void udt_finish_kernel(UdtBatch batch) {
  constexpr bool tab = TAB;
  const UdtItem it = tab ? batch.items[blockIdx.y] : batch.it[blockIdx.y];
  const int n = tab ? it.n : batch.n;
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


int udt_plan_fixed(int n, int* starts, int* widths, int cap) {
  int np = 0;
  for (int p0 = 0; p0 < n;) {
    const int w = udt_panel_width(n - p0);
    if (np == cap) return -1;
    starts[np] = p0;
    widths[np] = w;
    np++;
    p0 += w;
  }
  return np;
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
  for (int k = 0; k < b.count; k++) {
    // U is written while A's reflectors are read: it must be its own buffer
    if (b.it[k].U == b.it[k].A || b.it[k].U == b.it[k].T) return hipErrorInvalidValue;
    if (b.apply_pivot && b.it[k].T == b.it[k].A) return hipErrorInvalidValue;  // pivoted T is written out of place
  }
  return hipSuccess;
}

hipError_t launch_udt_norms(const UdtBatch& b, hipStream_t s) {
  if (b.items) hipLaunchKernelGGL(udt_norms_kernel<true>, dim3((b.max_n + 3) / 4, b.count), dim3(256), 0, s, b);
  else hipLaunchKernelGGL(udt_norms_kernel<false>, dim3((b.n + 3) / 4, b.count), dim3(256), 0, s, b);
  return hipGetLastError();
}

template <int R, int WM>
static void launch_panel_t(const UdtBatch& b, int p0, int w, int last, hipStream_t s) {
  if (b.items)
    hipLaunchKernelGGL((udt_panel_kernel<R, WM, true>), dim3(b.count), dim3(kQrThreads), udt_panel_lds_bytes(b.n),
                       s, b, p0, w, last);
  else
    hipLaunchKernelGGL((udt_panel_kernel<R, WM, false>), dim3(b.count), dim3(kQrThreads), udt_panel_lds_bytes(b.n),
                       s, b, p0, w, last);
}

hipError_t launch_udt_panel(const UdtBatch& b, int p0, int w, int last, hipStream_t s) {
  const int rows = b.n - p0;
  const int R = udt_rows_per_thread(rows);
  if (w < 1 || w > udt_panel_width(rows)) return hipErrorInvalidValue;
  switch (R) {
    case 1: launch_panel_t<1, 16>(b, p0, w, last, s); break;
    case 2: launch_panel_t<2, 16>(b, p0, w, last, s); break;
    case 4: launch_panel_t<4, 16>(b, p0, w, last, s); break;
    default: launch_panel_t<8, 8>(b, p0, w, last, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_udt_reflect(const UdtBatch& b, int p0, int w, hipStream_t s) {
  const int c0 = p0 + w;
  const int nA = c0 < b.n ? (b.n - c0 + 15) / 16 : 0;
  const int nQ = (b.n + 15) / 16;
  hipLaunchKernelGGL(udt_reflect_kernel<false>, dim3(nA + nQ, b.count), dim3(256), 0, s, b, p0, w, c0, nA);
  return hipGetLastError();
}

hipError_t launch_udt_finish(const UdtBatch& b, hipStream_t s) {
  if (b.items) hipLaunchKernelGGL(udt_finish_kernel<true>, dim3((b.max_n + 3) / 4, b.count), dim3(256), 0, s, b);
  else hipLaunchKernelGGL(udt_finish_kernel<false>, dim3((b.n + 3) / 4, b.count), dim3(256), 0, s, b);
  return hipGetLastError();
}

int udt_rows_class(int rows) { return udt_rows_per_thread(rows); }
int udt_width_for(int rows) { return udt_panel_width(rows); }

hipError_t launch_udt_panel_tab(const UdtBatch& b, int R, hipStream_t s) {
  if (!b.items || b.count < 1 || b.max_n < 1 || b.max_n > 2048) return hipErrorInvalidValue;
  UdtBatch c = b;
  c.n = b.max_n;          // sizes the dynamic LDS (per-item arrays of n entries)
  switch (R) {
    case 1: launch_panel_t<1, 16>(c, 0, 0, 0, s); break;
    case 2: launch_panel_t<2, 16>(c, 0, 0, 0, s); break;
    case 4: launch_panel_t<4, 16>(c, 0, 0, 0, s); break;
    case 8: launch_panel_t<8, 8>(c, 0, 0, 0, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_udt_reflect_tab(const UdtBatch& b, int max_grid, hipStream_t s) {
  if (!b.items || b.count < 1 || max_grid < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(udt_reflect_kernel<true>, dim3(max_grid, b.count), dim3(256), 0, s, b, 0, 0, 0, 0);
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
    if ((e = launch_udt_reflect(b, st[pi], wd[pi], s)) != hipSuccess) return e;
  }
  return launch_udt_finish(b, s);
}

// rdivp: A <- A[:, pivot] * triu(T)^-1 (only the upper triangle of T is read).
// Rows are independent.  Each wave owns 16 rows; for every 16-column panel J the
// coupling to the already solved columns, X_{<J} T_{<J,J}, is a 16 x 16 x j0 product on
// the matrix cores, then the 16 x 16 triangular block is solved by 16 lanes (one row each)
// in the reference's order: x_j = (o_j - sum_{k<j} x_k T_kj) / T_jj.
template <bool TAB>
__global__ __launch_bounds__(256) void rdivp_kernel(RdivpBatch batch) {
  constexpr bool tab = TAB;
  const RdivpItem it = tab ? batch.items[blockIdx.y] : batch.it[blockIdx.y];
  const int n = tab ? it.n : batch.n;
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

// rdivp for n <= kRdivpStripMaxN: one wave per 16-row strip.  The strip's pivoted input
// is gathered into LDS once (X[k][r], column-major per strip) and every solved column
// stays there, so a panel's coupling X_{<J} T_{<J,J} reads only T from global memory.
static size_t rdivp_strip_lds_bytes(int n) { return ((size_t)n * 17 + 2 * 16 * 17) * 8 + (size_t)n * 4; }

template <bool TAB>
__global__ __launch_bounds__(64) void rdivp_strip_kernel(RdivpBatch batch) {
  constexpr bool tab = TAB;
  const RdivpItem it = tab ? batch.items[blockIdx.y] : batch.it[blockIdx.y];
  const int n = tab ? it.n : batch.n;
  const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * 16;
  if (TAB && r0 >= n) return;   // item-table launches size the grid for the largest item
  extern __shared__ __align__(16) unsigned char smem[];
  double* X = reinterpret_cast<double*>(smem);   // [n][17]
  double* Ysh = X + (size_t)n * 17;              // [16][17]
  double* Tsh = Ysh + 16 * 17;                   // [16][17]
  int* pv = reinterpret_cast<int*>(Tsh + 16 * 17);
  double* __restrict__ A = it.A;
  const double* __restrict__ T = it.T;
  const int rr = r0 + li;
  const int rc = rr < n ? rr : n - 1;

  for (int j0 = 0; j0 < n; j0 += 64 * 16) {
    int t[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int j = j0 + lane + 64 * q;
      t[q] = it.pivot[j < n ? j : n - 1];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int j = j0 + lane + 64 * q;
      if (j < n) pv[j] = t[q];
    }
  }
  __syncthreads();
  // O[r][j] = A[r][pivot[j]]: lane (row li, phase lk) takes columns lk + 4 q
  for (int jb = 0; jb < n; jb += 128) {
    double v[32];
#pragma unroll
    for (int q = 0; q < 32; q++) {
      const int j = jb + lk + 4 * q;
      v[q] = A[(size_t)rc + (size_t)pv[j < n ? j : n - 1] * n];
    }
#pragma unroll
    for (int q = 0; q < 32; q++) {
      const int j = jb + lk + 4 * q;
      if (j < n) X[(size_t)j * 17 + li] = v[q];
    }
  }
  __syncthreads();

  // T of the next panel (its diagonal block and coupling rows k < 256) is loaded while the
  // current panel is solved
  constexpr int kPre = 64;                       // prefetched coupling rows k = 4 q + lk
  double td[4], tb[kPre];
  auto load_t = [&](int j0n) {
    const int cbn = j0n + li;
    const int cbcn = cbn < n ? cbn : n - 1;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int a = lk + 4 * q;
      const int ta = j0n + a < n ? j0n + a : n - 1;
      const double x = T[(size_t)ta + (size_t)cbcn * n];
      td[q] = (j0n + a < n && cbn < n) ? x : (a == li ? 1.0 : 0.0);   // out of range: identity
    }
#pragma unroll
    for (int g = 0; g < kPre / 16; g++) {
      if (64 * g < j0n) {                        // uniform: 16 loads at once
#pragma unroll
        for (int q = 16 * g; q < 16 * g + 16; q++) {
          const int k = 4 * q + lk;
          tb[q] = T[(size_t)(k < j0n ? k : 0) + (size_t)cbcn * n];
        }
      }
    }
  };
  load_t(0);
  for (int j0 = 0; j0 < n; j0 += 16) {
    const int w = n - j0 < 16 ? n - j0 : 16;
    const int cb = j0 + li;
    const int cbc = cb < n ? cb : n - 1;
    dx4 acc = dx4{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < kPre; q++)
      if (4 * q < j0) acc = mfma16(X[(size_t)(4 * q + lk) * 17 + li], tb[q], acc);
    for (int kb = 4 * kPre; kb < j0; kb += 128) {
      double tc[32];
#pragma unroll
      for (int q = 0; q < 32; q++) {
        const int k = kb + 4 * q + lk;
        tc[q] = T[(size_t)(k < j0 ? k : 0) + (size_t)cbc * n];
      }
#pragma unroll
      for (int q = 0; q < 32; q++) {
        const int k = kb + 4 * q + lk;
        if (kb + 4 * q < j0) acc = mfma16(X[(size_t)k * 17 + li], tc[q], acc);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) Tsh[(lk + 4 * q) * 17 + li] = td[q];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = mfma_row<double>(lane, r);
      Ysh[row * 17 + li] = cb < n ? X[(size_t)cb * 17 + row] - acc[r] : 0.0;
    }
    if (j0 + 16 < n) load_t(j0 + 16);
    __syncthreads();
    if (lane < 16) {
      // right-looking: x_c = v_c / T_cc, then v_c' -= x_c T_cc' for every later c'.  The
      // quotient comes from the reciprocal y = 1 / T_cc with one correction,
      // x = q + (v - T_cc q) y, q = v y, which is the correctly rounded v / T_cc (Markstein)
      // for the unit-modulus diagonal of a UDT's T; any other diagonal value divides.
      const double dl = Tsh[lane * 17 + lane];
      const double yl = 1.0 / dl;
      double v[16], x[16];
#pragma unroll
      for (int c = 0; c < 16; c++) v[c] = Ysh[lane * 17 + c];
#pragma unroll
      for (int c = 0; c < 16; c++) {
        const double d = readlane_d(dl, c);
        const double y = readlane_d(yl, c);
        double xc;
        if (fabs(d) >= 0.5 && fabs(d) <= 2.0) {
          const double q0 = v[c] * y;
          xc = fma(fma(-d, q0, v[c]), y, q0);
        } else {
          xc = v[c] / d;
        }
        x[c] = xc;
#pragma unroll
        for (int c2 = c + 1; c2 < 16; c2++) v[c2] -= xc * Tsh[c * 17 + c2];
      }
#pragma unroll
      for (int c = 0; c < 16; c++)
        if (c < w) X[(size_t)(j0 + c) * 17 + lane] = x[c];
    }
    __syncthreads();
  }
  if (rr < n) {
    for (int j = lk; j < n; j += 4) A[(size_t)rr + (size_t)j * n] = X[(size_t)j * 17 + li];
  }
}

// rdivp for n <= kRdivpBlockMaxN: four waves per 16-row strip.  Every 16 x 16 diagonal block
// of T is inverted first (thread (b, c) solves U_b x = e_c upwards), beside the loads of
// the strip's pivoted input; then per 16-column panel p the four waves share the update
// Y = X_p - X_{<p} T_{<p,p} on the matrix cores (wave w takes k-steps w, w + 4, ...; T's
// coupling column of the next panel is loaded while the current one is formed), and wave 0
// forms X_p = Y inv(T_pp), one 16 x 16 x 16 product.  In exact arithmetic this is the
// column-by-column substitution of the reference; the products round differently (the
// bound of the kernel tests: 1e-12 of the largest entry).
constexpr int kRdivpBlockMaxN = 512;

static size_t rdivp_block_lds_bytes(int n) {
  const size_t nb = (size_t)(n + 15) / 16;
  return ((size_t)n * 17 + nb * 16 * 17 + 4 * 4 * 64 + 16 * 17) * 8 + (size_t)n * 4;
}

// KS = 16 (n <= 256) or 32 (n <= 512): loads per thread of the gather and k-steps of one wave
// per panel, at most.  Every batch of loads has this fixed count (clamped addresses), so the
// wait for one batch never waits for the prefetch behind it.
template <bool TAB, int KS>
__global__ __launch_bounds__(256) void rdivp_block_kernel(RdivpBatch batch) {
  constexpr int kRdivpBlockSteps = KS;
  const RdivpItem it = TAB ? batch.items[blockIdx.y] : batch.it[blockIdx.y];
  const int n = TAB ? it.n : batch.n;
  const int r0 = blockIdx.x * 16;
  if (TAB && r0 >= n) return;   // item-table launches size the grid for the largest item
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lk = lane >> 4;
  const int nb = (n + 15) / 16;
  extern __shared__ __align__(16) unsigned char smem[];
  double* X = reinterpret_cast<double*>(smem);          // [n][17]: X[j * 17 + row]
  double* Ti = X + (size_t)n * 17;                      // [nb][16][17]: block b, row a, column c
  double* red = Ti + (size_t)nb * 16 * 17;              // [4 waves][4][64]
  double* Ysh = red + 4 * 4 * 64;                       // [16][17]: Y[row][col]
  int* pv = reinterpret_cast<int*>(Ysh + 16 * 17);
  double* __restrict__ A = it.A;
  const double* __restrict__ T = it.T;

  for (int j = tid; j < n; j += 256) pv[j] = it.pivot[j];
  __syncthreads();
  // the strip's pivoted input (thread: row tid % 16, columns tid / 16 + 16 q), held in
  // registers while the diagonal blocks are inverted
  const int grow = tid & 15, gcol = tid >> 4;
  const int grr = r0 + grow, grc = grr < n ? grr : n - 1;
  double gv[kRdivpBlockSteps];
#pragma unroll
  for (int q = 0; q < kRdivpBlockSteps; q++) {
    const int j = gcol + 16 * q;
    gv[q] = A[(size_t)grc + (size_t)pv[j < n ? j : n - 1] * n];
  }
  // diagonal blocks of T: the upper part, the identity beyond n (consecutive threads read
  // consecutive rows of a column)
  for (int e = tid; e < nb * 256; e += 256) {
    const int b = e >> 8, c = (e >> 4) & 15, a = e & 15;
    const int ra = 16 * b + a, cc = 16 * b + c;
    double v;
    if (ra < n && cc < n) v = a <= c ? T[(size_t)ra + (size_t)cc * n] : 0.0;
    else v = a == c ? 1.0 : 0.0;
    Ti[(size_t)(b * 16 + a) * 17 + c] = v;
  }
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 16) {
    const int b = b0 + (tid >> 4), c = tid & 15;
    double x[16];
    if (b < nb) {
      const double* U = Ti + (size_t)b * 16 * 17;
      // x[k] = 0 for k > c, so every row runs the same straight-line sum
#pragma unroll
      for (int i = 15; i >= 0; i--) {
        double sum = i == c ? 1.0 : 0.0;
#pragma unroll
        for (int k = i + 1; k < 16; k++) sum -= U[i * 17 + k] * x[k];
        x[i] = i <= c ? sum / U[i * 17 + i] : 0.0;
      }
    }
    __syncthreads();
    if (b < nb) {
#pragma unroll
      for (int a = 0; a < 16; a++) Ti[(size_t)(b * 16 + a) * 17 + c] = x[a];
    }
  }
#pragma unroll
  for (int q = 0; q < kRdivpBlockSteps; q++) {
    const int j = gcol + 16 * q;
    if (j < n) X[(size_t)j * 17 + grow] = gv[q];
  }
  __syncthreads();

  // T's coupling column of panel j0 for this wave's k-steps q = wave + 4 s (4 q < j0)
  auto load_tb = [&](double (&tb)[KS], int j0) {
    const int cb = j0 + li, cbc = cb < n ? cb : n - 1;
#pragma unroll
    for (int s = 0; s < kRdivpBlockSteps; s++) {
      const int k = 4 * (wave + 4 * s) + lk;
      tb[s] = T[(size_t)(k < n ? k : n - 1) + (size_t)cbc * n];
    }
  };
  auto panel = [&](int p, const double (&tb)[KS]) {
    const int j0 = 16 * p;
    dx4 acc = dx4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < kRdivpBlockSteps; s++) {
      const int q = wave + 4 * s;
      if (4 * q < j0) acc = mfma16(X[(size_t)(4 * q + lk) * 17 + li], tb[s], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) red[(wave * 4 + r) * 64 + lane] = acc[r];
    __syncthreads();
    if (wave == 0) {
      const int cb = j0 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = mfma_row<double>(lane, r);
        const double u = (red[r * 64 + lane] + red[(4 + r) * 64 + lane]) +
                         (red[(8 + r) * 64 + lane] + red[(12 + r) * 64 + lane]);
        Ysh[row * 17 + li] = cb < n ? X[(size_t)cb * 17 + row] - u : 0.0;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      dx4 o = dx4{0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        const int k = 4 * ks + lk;
        o = mfma16(Ysh[li * 17 + k], Ti[(size_t)(p * 16 + k) * 17 + li], o);
      }
      if (cb < n) {
#pragma unroll
        for (int r = 0; r < 4; r++) X[(size_t)cb * 17 + mfma_row<double>(lane, r)] = o[r];
      }
    }
    __syncthreads();
  };
  double tbA[kRdivpBlockSteps], tbB[kRdivpBlockSteps];
  load_tb(tbA, 0);
  for (int p = 0; p < nb; p += 2) {
    if (p + 1 < nb) load_tb(tbB, 16 * (p + 1));
    panel(p, tbA);
    if (p + 1 < nb) {
      if (p + 2 < nb) load_tb(tbA, 16 * (p + 2));
      panel(p + 1, tbB);
    }
  }
  if (grr < n) {
    for (int j = gcol; j < n; j += 16) A[(size_t)grr + (size_t)j * n] = X[(size_t)j * 17 + grow];
  }
}

hipError_t launch_rdivp(const RdivpBatch& b, hipStream_t s) {
  if (b.count <= 0) return hipSuccess;
  if (b.items) {
    // item table: one kernel class per launch (the host keeps n <= kRdivpStripMaxN items
    // and larger ones apart); the grid and LDS are sized for the largest item
    if (b.max_n < 1 || b.max_n > 2048) return hipErrorInvalidValue;
    if (b.max_n <= 256)
      hipLaunchKernelGGL((rdivp_block_kernel<true, 16>), dim3((b.max_n + 15) / 16, b.count), dim3(256),
                         rdivp_block_lds_bytes(b.max_n), s, b);
    else if (b.max_n <= kRdivpBlockMaxN)
      hipLaunchKernelGGL((rdivp_block_kernel<true, 32>), dim3((b.max_n + 15) / 16, b.count), dim3(256),
                         rdivp_block_lds_bytes(b.max_n), s, b);
    else if (b.max_n <= kRdivpStripMaxN)
      hipLaunchKernelGGL(rdivp_strip_kernel<true>, dim3((b.max_n + 15) / 16, b.count), dim3(64),
                         rdivp_strip_lds_bytes(b.max_n), s, b);
    else
      hipLaunchKernelGGL(rdivp_kernel<true>, dim3((b.max_n + 63) / 64, b.count), dim3(256), 0, s, b);
    return hipGetLastError();
  }
  if (b.count > kMaxUdtItems) return hipErrorInvalidValue;
  if (b.n <= 256)
    hipLaunchKernelGGL((rdivp_block_kernel<false, 16>), dim3((b.n + 15) / 16, b.count), dim3(256),
                       rdivp_block_lds_bytes(b.n), s, b);
  else if (b.n <= kRdivpBlockMaxN)
    hipLaunchKernelGGL((rdivp_block_kernel<false, 32>), dim3((b.n + 15) / 16, b.count), dim3(256),
                       rdivp_block_lds_bytes(b.n), s, b);
  else if (b.n <= kRdivpStripMaxN)
    hipLaunchKernelGGL(rdivp_strip_kernel<false>, dim3((b.n + 15) / 16, b.count), dim3(64),
                       rdivp_strip_lds_bytes(b.n), s, b);
  else
    hipLaunchKernelGGL(rdivp_kernel<false>, dim3((b.n + 63) / 64, b.count), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t qr_kernels_init() {
  const void* fns[] = {reinterpret_cast<const void*>(udt_panel_kernel<1, 16, false>),
                       reinterpret_cast<const void*>(udt_panel_kernel<2, 16, false>),
                       reinterpret_cast<const void*>(udt_panel_kernel<4, 16, false>),
                       reinterpret_cast<const void*>(udt_panel_kernel<8, 8, false>),
                       reinterpret_cast<const void*>(udt_panel_kernel<1, 16, true>),
                       reinterpret_cast<const void*>(udt_panel_kernel<2, 16, true>),
                       reinterpret_cast<const void*>(udt_panel_kernel<4, 16, true>),
                       reinterpret_cast<const void*>(udt_panel_kernel<8, 8, true>),
                       reinterpret_cast<const void*>(rdivp_strip_kernel<false>),
                       reinterpret_cast<const void*>(rdivp_strip_kernel<true>),
                       reinterpret_cast<const void*>(rdivp_block_kernel<false, 16>),
                       reinterpret_cast<const void*>(rdivp_block_kernel<true, 16>),
                       reinterpret_cast<const void*>(rdivp_block_kernel<false, 32>),
                       reinterpret_cast<const void*>(rdivp_block_kernel<true, 32>)};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dqmc
