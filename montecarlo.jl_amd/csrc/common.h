// Shared declarations of the gfx950 DQMC library (kernels + host orchestration).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>
#include <stdint.h>

namespace dqmc {

// A diagonal operand d[k] given either as a vector, or generated from a Hirsch
// field column: d[k] = conf[k] > 0 ? pos : neg  (exp(+-alpha) tables, so the
// interaction matrix exp(power * alpha * x) never needs its own buffer).  A vector can be
// read through the reference's elementwise maps of D vectors (linalg/real.jl:143-184):
enum DiagMode {
  kDiagPlain = 0,    // v
  kDiagInv = 1,      // 1 / v                 (vinv!)
  kDiagMaxInv = 2,   // 1 / max(1, v)         (vmaxinv!)
  kDiagMin = 3,      // min(1, v)             (vmin!)
};
struct DiagSpec {
  const double* vec;     // nullptr: not a vector
  const int8_t* conf;    // nullptr: not a field column
  double pos, neg;       // values for conf > 0 / conf < 0
  int mode;              // DiagMode applied to vec[k]
};
__host__ __device__ inline double diag_map(double v, int mode) {
  return mode == kDiagPlain ? v : mode == kDiagInv ? 1.0 / v : mode == kDiagMaxInv ? 1.0 / fmax(1.0, v) : fmin(1.0, v);
}

// Row held in accumulator element r of lane `lane` for the 16x16x4 MFMA tiles
// (column = lane % 16 for both):
//   v_mfma_f64_16x16x4f64: row = lane / 16 + 4 r
//   v_mfma_f32_16x16x4f32: row = 4 (lane / 16) + r
template <typename T> __device__ __forceinline__ int mfma_row(int lane, int r);
template <> __device__ __forceinline__ int mfma_row<double>(int lane, int r) { return (lane >> 4) + 4 * r; }
template <> __device__ __forceinline__ int mfma_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

__host__ __device__ inline DiagSpec diag_none() { return DiagSpec{nullptr, nullptr, 1.0, 1.0, 0}; }
__host__ __device__ inline DiagSpec diag_vec(const double* v, int mode = kDiagPlain) {
  return DiagSpec{v, nullptr, 1.0, 1.0, mode};
}
__host__ __device__ inline DiagSpec diag_field(const int8_t* c, double pos, double neg) {
  return DiagSpec{nullptr, c, pos, neg, 0};
}
__host__ __device__ inline bool diag_active(const DiagSpec& d) { return d.vec || d.conf; }
__device__ inline double diag_at(const DiagSpec& d, int k) {
  if (d.vec) return diag_map(d.vec[k], d.mode);
  if (d.conf) return d.conf[k] > 0 ? d.pos : d.neg;
  return 1.0;
}

// One matrix product of a batched launch:
//   C = alpha * diag(dl) * op(A) * diag(dmid) * op(B) * diag(dr) + beta * C + diag(dadd)
template <typename T>
struct GemmItem {
  const T* A;
  const T* B;
  T* C;
  int lda, ldb, ldc;
  int M, N, K;
  DiagSpec dl, dmid, dr, dadd;
};

constexpr int kMaxGemmItems = 4;

template <typename T>
struct GemmBatch {
  GemmItem<T> it[kMaxGemmItems];
  int count;
  int transA, transB;
  double alpha, beta;
  // many-chain launches: `count` items in device memory (it[] unused); max_m / max_n size
  // the grid (every item carries its own M, N, K)
  const GemmItem<T>* items = nullptr;
  int max_m = 0, max_n = 0;
};

// Launchers (gemm.hip)
hipError_t launch_gemm_f64(const GemmBatch<double>& b, hipStream_t s);
hipError_t launch_gemm_f32(const GemmBatch<float>& b, hipStream_t s);

// Pivoted UDT of a batch of n x n matrices (qr.hip).
struct UdtItem {
  double* A;        // input, destroyed (Householder vectors / R)
  double* U;        // orthogonal factor
  double* D;        // |diag R| (0 -> 1)
  double* T;        // D^-1 R, pivoted (apply_pivot) or "dirty upper" (may alias A then)
  int* pivot;       // column permutation, 0-based
  double* tau;      // n scratch
  double* tfac;     // (n + 16) * 16 scratch (each panel's 16 taus, at 16 * its start column)
  double* swap;     // 32 * n scratch (column moves)
  unsigned long long* err;   // device error counter (may be null): +1 per broken pivot selection
  // many-chain launches (item table in device memory): the item's own size and, for the
  // panel and reflector launches, its panel (start, width, last panel)
  int n, p0, w, last;
};
constexpr int kMaxUdtItems = 4;
struct UdtBatch {
  UdtItem it[kMaxUdtItems];
  int count;
  int n;
  int apply_pivot;
  // many-chain launches: `count` items in device memory, each with its own n (<= max_n)
  // and panel; it[] and n unused
  const UdtItem* items = nullptr;
  int max_n = 0;
};
hipError_t launch_udt(const UdtBatch& b, hipStream_t s);
// The steps of launch_udt, for callers that time them separately: panel starts and widths,
// column norms, panel p (single workgroup per item), strip reflections (mode 0: trailing
// update of A from column c0 with residual norms; mode 1: Q accumulation into U from
// column c0), D/T output.
void udt_plan(int n, std::vector<int>& starts, std::vector<int>& widths);
// The same plan into caller arrays of `cap` entries (no allocation); returns the panel
// count, or -1 if it exceeds cap.
int udt_plan_fixed(int n, int* starts, int* widths, int cap);
hipError_t udt_check(const UdtBatch& b);
hipError_t launch_udt_norms(const UdtBatch& b, hipStream_t s);
hipError_t launch_udt_panel(const UdtBatch& b, int p0, int w, int last, hipStream_t s);
hipError_t launch_udt_reflect(const UdtBatch& b, int p0, int w, hipStream_t s);
hipError_t launch_udt_finish(const UdtBatch& b, hipStream_t s);
// Item-table forms (b.items set): every item's panel (p0, w, last) is in its table entry;
// the panel kernel's rows-per-thread class R (1, 2, 4, 8: rows <= 256 R) is the same for all
// items of one launch.  max_w bounds the items' panel widths.
int udt_rows_class(int rows);
int udt_width_for(int rows);
hipError_t launch_udt_panel_tab(const UdtBatch& b, int R, hipStream_t s);
hipError_t launch_udt_reflect_tab(const UdtBatch& b, int max_grid, hipStream_t s);
// Q accumulation into U (U preset to I) for every panel, in one launch


// rdivp: A <- A[:, pivot] * triu(T)^-1 with O as n*n scratch (qr.hip).
struct RdivpItem {
  double* A;
  const double* T;
  double* O;
  const int* pivot;
  int n;            // item-table launches: the item's size
};
struct RdivpBatch {
  RdivpItem it[kMaxUdtItems];
  int count;
  int n;
  const RdivpItem* items = nullptr;   // many-chain launches: device table, own n <= max_n
  int max_n = 0;
};
constexpr int kRdivpStripMaxN = 1024;  // one wave per 16-row strip up to this size
hipError_t launch_rdivp(const RdivpBatch& b, hipStream_t s);

// Elementwise helpers (util.hip)
hipError_t launch_set_identity(double* A, int n, int nmat, hipStream_t s);
hipError_t launch_set_identity_f32(float* A, int n, int nmat, hipStream_t s);
hipError_t launch_fill(double* x, double v, int64_t count, hipStream_t s);
hipError_t launch_copy(double* dst, const double* src, int64_t count, hipStream_t s);
hipError_t launch_copy3(double* d0, const double* s0, int64_t n0, double* d1, const double* s1, int64_t n1,
                        double* d2, const double* s2, int64_t n2, hipStream_t s);
// field <-> BitArray chunks (64 entries per uint64, least significant bit first)
hipError_t launch_pack_field(const int8_t* conf, int64_t count, uint64_t* chunks, hipStream_t s);
hipError_t launch_unpack_field(const uint64_t* chunks, int64_t count, int8_t* conf, hipStream_t s);
hipError_t launch_f64_to_f32(float* dst, const double* src, int64_t count, hipStream_t s);
hipError_t launch_f32_to_f64(double* dst, const float* src, int64_t count, hipStream_t s);
hipError_t launch_symmetrize(double* A, int n, hipStream_t s);
// out[slot] = max |a - b| over count elements (a may be f32 when a32 != nullptr)
// One time slice of a density susceptibility (measure.hip): acc[dir + N (b1 + B b2)] += w sum_src K
// over the N x B x B outputs; d2st is the Bravais shift table (N x N, row dir).
enum DensityObservable { kDensityCharge = 0, kDensitySpinX = 1, kDensitySpinY = 2, kDensitySpinZ = 3 };
hipError_t launch_density_slice(const double* G00, const double* G0l, const double* Gl0, const double* Gll,
                                const int* d2st, int N, int B, int n, int flv, double total_flavors, int observable,
                                int first_slice, int reduced, double w, double* acc, hipStream_t s);

// Equal-time observables from the physical G (measure.hip); out holds 5 + (flv + 1) n values.
hipError_t launch_measure(const double* G, const double* T, int n, int flv, double U, double* out,
                          hipStream_t s);
hipError_t launch_maxabsdiff(const double* a, const float* a32, const double* b, int64_t count,
                             double* out, hipStream_t s);
hipError_t launch_scale_add_identity(double* A, double scale, double add_diag, int n, hipStream_t s);
hipError_t launch_axpy_matrix(double* Y, const double* X, double a, int64_t count, hipStream_t s);
hipError_t launch_norm1(const double* A, int n, double* out, hipStream_t s);
// out = in * diag(d) for nmat consecutive n x n matrices (d: nmat consecutive vectors)
hipError_t launch_scale_cols(double* out, const double* in, const double* d, int n, int nmat,
                             hipStream_t s);

// out_b = diag(map(dl_b)) in_b diag(map(dr_b)) [+ add_b] and A_b = diag(map(v_b)) for nmat
// consecutive n x n blocks, vectors b n apart (DiagMode maps; a null vector means no scaling,
// a null add no addend)
hipError_t launch_diag_scale(double* out, const double* in, const double* dl, int lmode, const double* dr,
                             int rmode, const double* add, int n, int nmat, hipStream_t s);
hipError_t launch_diag_matrix(double* A, const double* v, int mode, int n, int nmat, hipStream_t s);

// Many-chain elementwise launches: one item per (chain, buffer), in device memory.
enum ElemOp { kElemCopy = 0, kElemIdentity = 1, kElemFill = 2 };
struct ElemItem {
  int op;                 // ElemOp
  int n;                  // identity: matrix size (count = n * n * nmat)
  double* dst;
  const double* src;      // copy
  double v;               // fill value
  int64_t count;          // elements written
};
hipError_t launch_elem_multi(const ElemItem* items, int count, int64_t max_count, hipStream_t s);
// out[k] = max |a_k - b_k| over count_k elements (out[k] preset to +0.0)
struct MaxdiffItem {
  const double* a;
  const double* b;
  int64_t count;
  double* out;
};
hipError_t launch_maxabsdiff_multi(const MaxdiffItem* items, int count, int64_t max_count, hipStream_t s);

// One-time kernel attribute setup (dynamic LDS above 64 KB).
hipError_t kernels_init_once();   // all of the below, once per process (dqmc.hip)
hipError_t qr_kernels_init();
hipError_t gemm_kernels_init();
hipError_t local_kernels_init();

// Local update (local.hip)
struct ChainCounters {
  long long u_pos;        // uniforms consumed
  long long accepted;
  long long proposed;
  long long negative;
  long long ties;
  long long exhausted;
  long long log_len;
  long long flushes;
  long long device_errors;  // kernels that gave up a bounded wait or met a broken invariant
};

// Window kernel of the local update (local.hip): sites [w0, w0 + W) of one slice.  The
// same launch applies the previous window's factors to all of G (copy G -> Gd), and the
// window's own block is formed from G and those factors, so consecutive windows need one
// launch each and the application is off the decisions' critical path.
template <typename GT>
struct WindowArgs {
  const GT* G;            // flv blocks, n x n, ld = n: G before the previous window's flips
  GT* Gd;                 // out: G after the previous window's flips (when Yp != nullptr)
  int n, flv;
  int8_t* conf;           // conf column of the current slice
  int w0, W;              // W in {16, 32, 64}
  int wp;                 // start of the previous window (w0 - W)
  int do_window;          // 0: only the application of the previous window
  double d_first[2];      // Delta block 1, index (x > 0)
  double d_second[2];     // Delta block 2 (magnetic)
  double bweight[2];      // exp(-dE_boson) density, index (x > 0)
  const int* Sp;          // previous window: Sp[0] = flips (nullptr: no previous window)
  const GT* Yp;           // previous window: dense factor, flv x (W x W)
  int* S;                 // out: S[0] = m accepted flips of this window
  GT* Y;                  // out: dense factor over window positions, flv x (W x W) row-major,
                          //      zero in the rows and columns of rejected sites
  const double* uniforms; // supplied stream (nullptr: counter stream)
  long long n_uniforms;
  unsigned long long seed;
  ChainCounters* ctr;
  unsigned char* dec_log;
  double* p_log;
  long long log_cap;
  long long wait_limit;   // polls before a factor wave gives up (counted in ctr->device_errors)
};
hipError_t launch_window_f64(const WindowArgs<double>& a, hipStream_t s);
hipError_t launch_window_f32(const WindowArgs<float>& a, hipStream_t s);
// One launch per slice (one chain, W in {16, 32}): workgroup 0 runs the slice's windows one
// after the other, workgroups 1..napply apply each window's update to G as soon as its factor
// is published.  Window k reads G from B[(k - 1) % 2] (B[0] for k = 0) and apply(k) writes
// B[(k + 1) % 2], so the slice ends in B[nw % 2].  sync holds 1 + 2 nw counters (zero when
// the chain is created or rebuilt; every launch advances them, so replays need no reset).
template <typename GT>
struct WindowSlice {
  WindowArgs<GT> a;            // the slice's fields; G, Gd, w0, wp, Sp, Yp, S, Y are per window
  GT* B[2];                    // the two copies of G
  GT* Y;                       // two factor slots of flv x W x W
  int* S;                      // two flip-count slots of W + 1
  unsigned long long* sync;    // [0] launch epoch, [1 + k] window k published, [1 + nw + k] applications done
  int nw, napply;
};
hipError_t launch_window_slice_f64(const WindowSlice<double>& s, hipStream_t st);
// Many-chain window launch: `count` chains' WindowArgs in device memory (the same W for all,
// each with its own n and flavour count).  Workgroup (0, c) runs chain c's window when its
// do_window is set; workgroups (1 + t, c) apply chain c's previous window to tile t when its
// Yp is set.  max_flv and max_n size the grid.
hipError_t launch_window_multi_f64(const WindowArgs<double>* items, int count, int max_flv, int W, int max_n,
                                   hipStream_t s);
int window_size(int n, int flv);

}  // namespace dqmc
