// Host orchestration of one DQMC chain on the GPU and the C ABI (include/dqmc_hip.h).
//
// The chain state mirrors DQMCStack (reference src/flavors/DQMC/stack.jl:1-74): the
// UDT stack slots, the Green's function, the work matrices and the propagate state
// (current_slice, current_range, direction).  Every matrix lives in device memory for
// the life of the handle; host code only sequences kernels on the chain's stream, so a
// sweep is a stream of launches with no host synchronisation until its statistics are
// collected at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dqmc_hip.h"
#include "common.h"
#include "ensemble.h"

using namespace dqmc;

namespace {
thread_local std::string g_create_error;   // per host thread: chains may be created concurrently
std::once_flag g_init_once;
hipError_t g_init_status = hipSuccess;

void init_kernels() {
  hipError_t e = qr_kernels_init();
  if (e == hipSuccess) e = local_kernels_init();
  if (e == hipSuccess) e = gemm_kernels_init();
  g_init_status = e;
}
}  // namespace

hipError_t dqmc::kernels_init_once() {
  std::call_once(g_init_once, init_kernels);
  return g_init_status;
}

// One kernel of a chain's propagate, recorded instead of launched when the chain runs in
// a many-chain step: the step's launches then cover every chain whose recorded sequence
// has the same kinds (each chain's items in its own order, so its results are those of
// its own launches).
enum OpKind { kOpGemm = 0, kOpUdt = 1, kOpRdivp = 2, kOpElem = 3, kOpMaxdiff = 4 };
struct RecOp {
  int kind;
  int key;           // gemm: 2 transA + transB; udt: apply_pivot
  double alpha, beta;
  int first, count;  // items in the chain's per-kind record
};

struct dqmc_chain_s {
  int dev = 0;
  hipStream_t st = nullptr;
  int n = 0, flv = 1, M = 0, sm = 10, K = 0;
  double dtau = 0.1, alpha = 0.0;
  int prec = DQMC_PRECISION_FP64, check = 1;
  std::vector<int> rs, re;  // chunk ranges, 1-based inclusive (generate_chunks)
  double ev_pos = 1, ev_neg = 1;
  double d_first[2] = {0, 0}, d_second[2] = {0, 0}, bweight[2] = {1, 1};
  size_t nn = 0;
  std::vector<void*> allocs;
  // Device memory of a chain is one allocation (2 MiB granular, so the GPU maps it with
  // large pages), carved into 256-byte aligned buffers: sized in a first pass.
  bool sizing = false;
  size_t arena_need = 0, arena_off = 0, arena_cap = 0;
  char* arena = nullptr;

  double *eTmh = nullptr, *eTph = nullptr, *eT2m = nullptr, *eT2p = nullptr, *dT = nullptr;
  double *G = nullptr, *Gx = nullptr, *Gchk = nullptr;
  double *us = nullptr, *ts = nullptr, *ds = nullptr;
  double *Ul = nullptr, *Ur = nullptr, *Tl = nullptr, *Tr = nullptr, *Dl = nullptr, *Dr = nullptr;
  double *tmp1 = nullptr, *tmp2 = nullptr, *cur = nullptr, *O = nullptr;
  int* piv = nullptr;
  double *tau = nullptr, *tfac = nullptr, *swp = nullptr;
  int8_t* conf = nullptr;
  uint64_t* cpack = nullptr;   // the field as BitArray chunks (ceil(n M / 64))
  float *G32 = nullptr, *Gx32 = nullptr, *Gchk32 = nullptr, *eT2m32 = nullptr, *eT2p32 = nullptr;
  double* Gpp = nullptr;   // second copy of G for the window updates (FP64 form)
  float* G32pp = nullptr;  // second copy of G32 (mixed form)
  double* wY = nullptr;    // window factors Y, two slots of flv x W x W (FP64 or FP32 contents)
  int* wS = nullptr;       // window flip counts, two slots
  int win = 32;
  ChainCounters* ctr = nullptr;
  double* unif = nullptr;
  long long nunif = 0;
  unsigned long long seed = 0;
  unsigned char* dlog = nullptr;
  double* plog = nullptr;
  long long logcap = 0;
  double* errbuf = nullptr;
  long long errcap = 4096, nerr = 0;
  double* scal = nullptr;  // small device scalars
  unsigned long long* wsync = nullptr;   // one-launch-per-slice counters (1 + 2 ceil(n / 16))
  bool win_slice = true;   // FP64 sweeps: one window launch per slice (W <= 32)
  long long wait_limit = 1ll << 22;  // polls of the window kernel's factor waves before giving up
  long long dev_err_seen = 0;        // device errors already reported
  double mixed_tol = 1e-6;           // mixed precision: bound on max|G_wrapped - G_scratch|
  long long prec_viol = 0, prec_viol_seen = 0;
  double prec_viol_max = 0.0;

  int cs = 0, cr = 1, dir = 1;
  bool hopping_set = false;
  bool t_set = false;       // T itself is known (dqmc_set_hopping), not only its exponentials
  long long sweeps = 0, perr_count = 0, stabs = 0;
  double perr_max = 0.0, perr_last = 0.0;
  std::vector<double> stablog;
  std::string err;

  // many-chain steps: record mode collects the chain's kernels (RecOp + items; window
  // launches) instead of launching them
  bool rec = false;
  hipStream_t own_st = nullptr;      // the chain's stream while it runs in an ensemble
  std::vector<RecOp> ops;
  std::vector<GemmItem<double>> rgemm;
  std::vector<UdtItem> rudt;
  std::vector<RdivpItem> rrdivp;
  std::vector<ElemItem> relem;
  std::vector<MaxdiffItem> rmax;
  std::vector<WindowArgs<double>> rwin;
  long long steps = 0;               // slice visits (sweep_spatial + propagate) in ensembles
  bool dg_valid = false;             // Dr holds the D of the last calculate_greens of G itself
  long long fver = 0;                // bumped whenever the field or the hopping may change

  // pinned host mirrors of the counters and of the stabilisation slots: the end of a sweep is
  // two stream-ordered copies and one wait
  ChainCounters* hctr = nullptr;
  double* herr = nullptr;
  bool acc_valid = false;            // acc_last is the device's accepted count
  long long acc_last = 0;

  // one sweep (2 M slice visits) as a HIP graph: captured on first use and replayed while
  // every launch argument it baked in (state, buffer roles, streams of randomness, logs) is
  // unchanged; FP64 chains only
  bool use_graph = true;
  hipGraph_t sg = nullptr;
  hipGraphExec_t sgx = nullptr;
  std::vector<uint64_t> sg_key;
  int sg_cs = 0, sg_dir = 0, sg_cr = 0;
  double *sg_G = nullptr, *sg_Gpp = nullptr;
  long long sg_nerr = 0;

  // unequal-time stack (UnequalTimeStack, unequal_time_stack.jl:1-42), one allocation made on
  // first use: forward and backward UDT slots 1..K+1, inverse slots 1..K (flv blocks each);
  // uU/uD/uT the inverse block of the last G(k <- l), uG its effective G
  double *fu = nullptr, *fd = nullptr, *ft = nullptr, *bu = nullptr, *bd = nullptr, *bt = nullptr;
  double *iu = nullptr, *id = nullptr, *it = nullptr;
  double *uU = nullptr, *uD = nullptr, *uT = nullptr, *uG = nullptr, *uDsave = nullptr;
  double *itm = nullptr, *itv = nullptr;  // greens iterator: 14 matrices, 3 vectors (flv blocks)
  int ut_fwd = 1, ut_bwd = 0;        // forward slots 1..ut_fwd, backward slots ut_bwd..K+1 built
  std::vector<char> ut_inv;          // inverse slot built
  long long ut_ver = -1;             // fver the slots belong to
  int ut_k = -1, ut_l = -1;          // (k, l) of uG

  // optional per-kernel timing with HIP events on the chain's stream
  bool prof = false;
  struct Pending { int kind; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> evpool;
  double kms[DQMC_KERNEL_KINDS] = {0};
  long long kcount[DQMC_KERNEL_KINDS] = {0};
};

#define HIPCHK(h, x)                                                         \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      (h)->err = std::string(#x) + ": " + hipGetErrorString(e_);             \
      return DQMC_E_HIP;                                                     \
    }                                                                        \
  } while (0)

#define CHK(x)                 \
  do {                         \
    int r_ = (x);              \
    if (r_ != DQMC_OK) return r_; \
  } while (0)

namespace {

typedef dqmc_chain_s Chain;

int generate_chunks(int length, int max_chunk, std::vector<int>& s, std::vector<int>& e) {
  // stack.jl:154-158 (Julia round: half to even, as nearbyint in the default mode)
  const int nch = (length + max_chunk - 1) / max_chunk;
  const double step = (double)length / nch;
  s.resize(nch);
  e.resize(nch);
  for (int i = 1; i <= nch; i++) {
    s[i - 1] = (int)std::nearbyint((i - 1) * step) + 1;
    e[i - 1] = (int)std::nearbyint(i * step);
  }
  return nch;
}

template <typename T>
int dalloc(Chain* h, T** p, size_t count) {
  const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
  if (h->sizing) {
    h->arena_need += bytes;
    *p = nullptr;
    return DQMC_OK;
  }
  if (h->arena) {
    if (h->arena_off + bytes > h->arena_cap) {
      h->err = "device arena exhausted";
      return DQMC_E_ALLOC;
    }
    *p = reinterpret_cast<T*>(h->arena + h->arena_off);
    h->arena_off += bytes;
    return DQMC_OK;
  }
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    h->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return DQMC_E_ALLOC;
  }
  h->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return DQMC_OK;
}

// Runs `list` (a sequence of dalloc calls) twice: once to size the arena, once to carve it.
template <typename F>
int arena_alloc(Chain* h, F list) {
  h->sizing = true;
  h->arena_need = 0;
  int r = list();
  h->sizing = false;
  if (r != DQMC_OK) return r;
  const size_t cap = (h->arena_need + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, cap);
  if (e != hipSuccess) {
    h->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return DQMC_E_ALLOC;
  }
  h->allocs.push_back(q);
  h->arena = reinterpret_cast<char*>(q);
  h->arena_cap = cap;
  h->arena_off = 0;
  return list();
}

inline double* blk(double* base, const Chain* h, int b) { return base + (size_t)b * h->nn; }
inline float* blk32(float* base, const Chain* h, int b) { return base + (size_t)b * h->nn; }
inline double* uslot(Chain* h, int s, int b) { return h->us + ((size_t)(s - 1) * h->flv + b) * h->nn; }
inline double* tslot(Chain* h, int s, int b) { return h->ts + ((size_t)(s - 1) * h->flv + b) * h->nn; }
inline double* dslot(Chain* h, int s, int b) { return h->ds + ((size_t)(s - 1) * h->flv + b) * h->n; }

// Interaction matrix exp(power * alpha * x) of block b (interaction_matrix_exp!,
// fields.jl:380-386, 429-438): block 2 of the magnetic field uses -power.
DiagSpec ev(const Chain* h, int b, int slice, double power) {
  const double s = (b == 0 ? 1.0 : -1.0) * power;
  const int8_t* c = h->conf + (size_t)(slice - 1) * h->n;
  return s > 0 ? diag_field(c, h->ev_pos, h->ev_neg) : diag_field(c, h->ev_neg, h->ev_pos);
}

template <typename T>
GemmItem<T> item(const T* A, const T* B, T* C, int n) {
  GemmItem<T> g;
  g.A = A; g.B = B; g.C = C;
  g.lda = g.ldb = g.ldc = n;
  g.M = g.N = g.K = n;
  g.dl = g.dmid = g.dr = g.dadd = diag_none();
  return g;
}

template <typename T>
GemmBatch<T> batch(int ta, int tb, double alpha = 1.0, double beta = 0.0) {
  GemmBatch<T> b;
  b.count = 0; b.transA = ta; b.transB = tb; b.alpha = alpha; b.beta = beta;
  return b;
}

hipEvent_t take_event(Chain* h) {
  if (!h->evpool.empty()) {
    hipEvent_t e = h->evpool.back();
    h->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with events when timing is on (kinds: DQMC_KERNEL_*).
struct Timed {
  Chain* h;
  int kind;
  hipEvent_t a = nullptr;
  Timed(Chain* hh, int k) : h(hh), kind(k) {
    if (h->prof) {
      a = take_event(h);
      if (a) (void)hipEventRecord(a, h->st);
    }
  }
  ~Timed() {
    if (a) {
      hipEvent_t b = take_event(h);
      if (b) {
        (void)hipEventRecord(b, h->st);
        h->pending.push_back({kind, a, b});
      } else {
        h->evpool.push_back(a);
      }
    }
  }
};

void collect_times(Chain* h) {
  for (auto& p : h->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      h->kms[p.kind] += ms;
      h->kcount[p.kind] += 1;
    }
    h->evpool.push_back(p.a);
    h->evpool.push_back(p.b);
  }
  h->pending.clear();
}

int record(Chain* h, int kind, int key, int first, int count, double alpha = 1.0, double beta = 0.0) {
  h->ops.push_back(RecOp{kind, key, alpha, beta, first, count});
  return DQMC_OK;
}

int run(Chain* h, const GemmBatch<double>& b) {
  if (h->rec) {
    const int first = (int)h->rgemm.size();
    for (int i = 0; i < b.count; i++) h->rgemm.push_back(b.it[i]);
    return record(h, kOpGemm, 2 * b.transA + b.transB, first, b.count, b.alpha, b.beta);
  }
  Timed t(h, DQMC_KERNEL_GEMM_F64);
  HIPCHK(h, launch_gemm_f64(b, h->st));
  return DQMC_OK;
}

int elem(Chain* h, const ElemItem& e) {
  h->relem.push_back(e);
  return record(h, kOpElem, 0, (int)h->relem.size() - 1, 1);
}
int run(Chain* h, const GemmBatch<float>& b) {
  Timed t(h, DQMC_KERNEL_GEMM_F32);
  HIPCHK(h, launch_gemm_f32(b, h->st));
  return DQMC_OK;
}

int copyd(Chain* h, double* dst, const double* src, size_t count) {
  if (h->rec) {
    if (count == 0 || dst == src) return DQMC_OK;
    return elem(h, ElemItem{kElemCopy, 0, dst, src, 0.0, (int64_t)count});
  }
  HIPCHK(h, launch_copy(dst, src, (int64_t)count, h->st));
  return DQMC_OK;
}

int identity(Chain* h, double* A, int nmat) {
  if (h->rec) return elem(h, ElemItem{kElemIdentity, h->n, A, nullptr, 0.0, (int64_t)h->nn * nmat});
  HIPCHK(h, launch_set_identity(A, h->n, nmat, h->st));
  return DQMC_OK;
}

int ones(Chain* h, double* d, size_t count) {
  if (h->rec) return elem(h, ElemItem{kElemFill, 0, d, nullptr, 1.0, (int64_t)count});
  HIPCHK(h, launch_fill(d, 1.0, (int64_t)count, h->st));
  return DQMC_OK;
}

// ---- slice matrices (stack.jl:319-367) ------------------------------------------------
// out_b = eT2^T diag(eV_b) in_b [diag(dr_b)]          multiply_slice_matrix_left!
int slice_left(Chain* h, int slice, double* const in[], double* const out[], const double* const dr[]) {
  auto g = batch<double>(1, 0);
  for (int b = 0; b < h->flv; b++) {
    auto it = item<double>(h->eT2m, in[b], out[b], h->n);
    it.dmid = ev(h, b, slice, 1.0);
    if (dr) it.dr = diag_vec(dr[b]);
    g.it[g.count++] = it;
  }
  return run(h, g);
}
// out_b = diag(eV_b) eT2^T in_b [diag(dr_b)]          multiply_daggered_slice_matrix_left!
int slice_dagger_left(Chain* h, int slice, double* const in[], double* const out[],
                      const double* const dr[]) {
  auto g = batch<double>(1, 0);
  for (int b = 0; b < h->flv; b++) {
    auto it = item<double>(h->eT2m, in[b], out[b], h->n);
    it.dl = ev(h, b, slice, 1.0);
    if (dr) it.dr = diag_vec(dr[b]);
    g.it[g.count++] = it;
  }
  return run(h, g);
}

// wrap_greens!(gf, curr_slice, direction), stack.jl:594-603, FP64.  `l` is the slice whose
// matrix is applied: l = curr_slice (direction +1) or curr_slice - 1 (direction -1).
//  +1: G <- eT2 eV(l) G eV(l)^-1 eT2^-1          -1: G <- eV(l)^-1 eT2^-1 G eT2 eV(l)
int wrap64(Chain* h, double* Gb, int l, int dir) {
  auto g1 = dir > 0 ? batch<double>(0, 1) : batch<double>(0, 1);
  auto g2 = batch<double>(1, 0);
  for (int b = 0; b < h->flv; b++) {
    if (dir > 0) {
      auto a = item<double>(blk(Gb, h, b), h->eT2p, blk(h->Gx, h, b), h->n);
      a.dl = ev(h, b, l, 1.0);
      a.dmid = ev(h, b, l, -1.0);
      g1.it[g1.count++] = a;
      g2.it[g2.count++] = item<double>(h->eT2m, blk(h->Gx, h, b), blk(Gb, h, b), h->n);
    } else {
      auto a = item<double>(blk(Gb, h, b), h->eT2m, blk(h->Gx, h, b), h->n);
      a.dr = ev(h, b, l, 1.0);
      g1.it[g1.count++] = a;
      auto c = item<double>(h->eT2p, blk(h->Gx, h, b), blk(Gb, h, b), h->n);
      c.dl = ev(h, b, l, -1.0);
      g2.it[g2.count++] = c;
    }
  }
  CHK(run(h, g1));
  return run(h, g2);
}

// Same in FP32 on the mixed-precision copy of G.
int wrap32(Chain* h, float* Gb, int l, int dir) {
  auto g1 = batch<float>(0, 1);
  auto g2 = batch<float>(1, 0);
  for (int b = 0; b < h->flv; b++) {
    if (dir > 0) {
      auto a = item<float>(blk32(Gb, h, b), h->eT2p32, blk32(h->Gx32, h, b), h->n);
      a.dl = ev(h, b, l, 1.0);
      a.dmid = ev(h, b, l, -1.0);
      g1.it[g1.count++] = a;
      g2.it[g2.count++] = item<float>(h->eT2m32, blk32(h->Gx32, h, b), blk32(Gb, h, b), h->n);
    } else {
      auto a = item<float>(blk32(Gb, h, b), h->eT2m32, blk32(h->Gx32, h, b), h->n);
      a.dr = ev(h, b, l, 1.0);
      g1.it[g1.count++] = a;
      auto c = item<float>(h->eT2p32, blk32(h->Gx32, h, b), blk32(Gb, h, b), h->n);
      c.dl = ev(h, b, l, -1.0);
      g2.it[g2.count++] = c;
    }
  }
  CHK(run(h, g1));
  return run(h, g2);
}

int udt(Chain* h, double* const A[], double* const U[], double* const D[], double* const T[],
        int apply_pivot) {
  UdtBatch u;
  u.count = h->flv;
  u.n = h->n;
  u.apply_pivot = apply_pivot;
  for (int b = 0; b < h->flv; b++) {
    u.it[b].err = h->ctr ? reinterpret_cast<unsigned long long*>(&h->ctr->device_errors) : nullptr;
    u.it[b].n = h->n; u.it[b].p0 = u.it[b].w = u.it[b].last = 0;
    u.it[b].A = A[b]; u.it[b].U = U[b]; u.it[b].D = D[b]; u.it[b].T = T[b];
    u.it[b].pivot = h->piv + (size_t)b * h->n;
    u.it[b].tau = h->tau + (size_t)b * h->n;
    u.it[b].tfac = h->tfac + (size_t)b * (h->n + 16) * 16;
    u.it[b].swap = h->swp + (size_t)b * 32 * h->n;
  }
  HIPCHK(h, udt_check(u));
  if (h->rec) {
    const int first = (int)h->rudt.size();
    for (int b = 0; b < u.count; b++) h->rudt.push_back(u.it[b]);
    return record(h, kOpUdt, apply_pivot, first, u.count);
  }
  std::vector<int> st, wd;
  udt_plan(h->n, st, wd);
  const int np = (int)st.size();
  {
    Timed t(h, DQMC_KERNEL_UDT);
    HIPCHK(h, launch_udt_norms(u, h->st));
  }
  for (int pi = 0; pi < np; pi++) {
    {
      Timed t(h, DQMC_KERNEL_UDT_PANEL);
      HIPCHK(h, launch_udt_panel(u, st[pi], wd[pi], pi == np - 1, h->st));
    }
    {
      // trailing update of A and Q accumulation on U's rows, one launch
      Timed t(h, DQMC_KERNEL_UDT_UPDATE);
      HIPCHK(h, launch_udt_reflect(u, st[pi], wd[pi], h->st));
    }
  }
  Timed t(h, DQMC_KERNEL_UDT);
  HIPCHK(h, launch_udt_finish(u, h->st));
  return DQMC_OK;
}

int rdivp(Chain* h, double* const A[], double* const T[]) {
  RdivpBatch r;
  r.count = h->flv;
  r.n = h->n;
  for (int b = 0; b < h->flv; b++) {
    r.it[b].A = A[b]; r.it[b].T = T[b];
    r.it[b].O = blk(h->O, h, b);
    r.it[b].pivot = h->piv + (size_t)b * h->n;
    r.it[b].n = h->n;
  }
  if (h->rec) {
    const int first = (int)h->rrdivp.size();
    for (int b = 0; b < r.count; b++) h->rrdivp.push_back(r.it[b]);
    return record(h, kOpRdivp, 0, first, r.count);
  }
  Timed t(h, DQMC_KERNEL_RDIVP);
  HIPCHK(h, launch_rdivp(r, h->st));
  return DQMC_OK;
}

struct Blocks {
  double* p[2];
  double* const* get() const { return p; }
};
Blocks B(Chain* h, double* base) { return Blocks{{blk(base, h, 0), h->flv > 1 ? blk(base, h, 1) : nullptr}}; }
Blocks Bd(Chain* h, double* base) { return Blocks{{base, h->flv > 1 ? base + h->n : nullptr}}; }
Blocks Us(Chain* h, int s) { return Blocks{{uslot(h, s, 0), h->flv > 1 ? uslot(h, s, 1) : nullptr}}; }
Blocks Ts(Chain* h, int s) { return Blocks{{tslot(h, s, 0), h->flv > 1 ? tslot(h, s, 1) : nullptr}}; }
Blocks Ds(Chain* h, int s) { return Blocks{{dslot(h, s, 0), h->flv > 1 ? dslot(h, s, 1) : nullptr}}; }

// T_out_b = T1_b * T2_b
int mul_blocks(Chain* h, const Blocks& out, const Blocks& t1, const Blocks& t2) {
  auto g = batch<double>(0, 0);
  for (int b = 0; b < h->flv; b++) g.it[g.count++] = item<double>(t1.p[b], t2.p[b], out.p[b], h->n);
  return run(h, g);
}

// A stack of UDT slots (s = 1, 2, ...; flv blocks per slot): the chain's own stack or one of
// the unequal-time stacks.
struct Slots {
  double *u, *d, *t;
};
Slots chain_slots(Chain* h) { return Slots{h->us, h->ds, h->ts}; }
Blocks SU(Chain* h, const Slots& S, int s) {
  double* p = S.u + (size_t)(s - 1) * h->flv * h->nn;
  return Blocks{{p, h->flv > 1 ? p + h->nn : nullptr}};
}
Blocks ST(Chain* h, const Slots& S, int s) {
  double* p = S.t + (size_t)(s - 1) * h->flv * h->nn;
  return Blocks{{p, h->flv > 1 ? p + h->nn : nullptr}};
}
Blocks SD(Chain* h, const Slots& S, int s) {
  double* p = S.d + (size_t)(s - 1) * h->flv * h->n;
  return Blocks{{p, h->flv > 1 ? p + h->n : nullptr}};
}

// add_slice_sequence_left(mc, idx), stack.jl:377-393 (and the forward build of the
// unequal-time stack, unequal_time_stack.jl:126-138, on its slots)
int add_slice_sequence_left(Chain* h, int idx, const Slots& S) {
  const int r0 = h->rs[idx - 1], r1 = h->re[idx - 1];
  Blocks in = SU(h, S, idx);
  Blocks dsl = SD(h, S, idx);
  for (int s = r0; s <= r1; s++) {
    const bool last = s == r1;
    Blocks out = last ? B(h, h->tmp1) : ((s - r0) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
    CHK(slice_left(h, s, in.get(), out.get(), last ? dsl.get() : nullptr));
    in = out;
  }
  CHK(udt(h, B(h, h->tmp1).get(), SU(h, S, idx + 1).get(), SD(h, S, idx + 1).get(), B(h, h->tmp2).get(), 1));
  return mul_blocks(h, ST(h, S, idx + 1), B(h, h->tmp2), ST(h, S, idx));
}
int add_slice_sequence_left(Chain* h, int idx) { return add_slice_sequence_left(h, idx, chain_slots(h)); }

// add_slice_sequence_right(mc, idx), stack.jl:402-416 (and the backward build of the
// unequal-time stack, unequal_time_stack.jl:141-155)
int add_slice_sequence_right(Chain* h, int idx, const Slots& S) {
  const int r0 = h->rs[idx - 1], r1 = h->re[idx - 1];
  Blocks in = SU(h, S, idx + 1);
  Blocks dsl = SD(h, S, idx + 1);
  for (int s = r1; s >= r0; s--) {
    const bool last = s == r0;
    Blocks out = last ? B(h, h->tmp1) : ((r1 - s) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
    CHK(slice_dagger_left(h, s, in.get(), out.get(), last ? dsl.get() : nullptr));
    in = out;
  }
  CHK(udt(h, B(h, h->tmp1).get(), SU(h, S, idx).get(), SD(h, S, idx).get(), B(h, h->tmp2).get(), 1));
  return mul_blocks(h, ST(h, S, idx), B(h, h->tmp2), ST(h, S, idx + 1));
}
int add_slice_sequence_right(Chain* h, int idx) { return add_slice_sequence_right(h, idx, chain_slots(h)); }

// calculate_greens_AVX!(Ul, Dl, Tl, Ur, Dr, Tr, G), stack.jl:442-496
int calculate_greens(Chain* h, double* out) {
  const int n = h->n;
  // the final D stays in Dr; it belongs to the chain's own G only when that is the output
  h->dg_valid = out == h->G;
  Blocks G = B(h, out), Ul = B(h, h->Ul), Ur = B(h, h->Ur), Tl = B(h, h->Tl), Tr = B(h, h->Tr);
  Blocks Dl = Bd(h, h->Dl), Dr = Bd(h, h->Dr);
  {  // G = Dl (Tl Tr') Dr
    auto g = batch<double>(0, 1);
    for (int b = 0; b < h->flv; b++) {
      auto it = item<double>(Tl.p[b], Tr.p[b], G.p[b], n);
      it.dl = diag_vec(Dl.p[b]);
      it.dr = diag_vec(Dr.p[b]);
      g.it[g.count++] = it;
    }
    CHK(run(h, g));
  }
  CHK(udt(h, G.get(), Tr.get(), Dr.get(), G.get(), 0));      // G -> Tr Dr (G P)
  CHK(mul_blocks(h, Tl, Ul, Tr));                            // Tl = Ul Tr
  CHK(rdivp(h, Ur.get(), G.get()));                          // Ur = Ur (G P)^-1
  {  // Tr = Tl' Ur + Dr
    auto g = batch<double>(1, 0);
    for (int b = 0; b < h->flv; b++) {
      auto it = item<double>(Tl.p[b], Ur.p[b], Tr.p[b], n);
      it.dadd = diag_vec(Dr.p[b]);
      g.it[g.count++] = it;
    }
    CHK(run(h, g));
  }
  CHK(udt(h, Tr.get(), Ul.get(), Dr.get(), Tr.get(), 0));    // Tr -> Ul Dr (Tr P)
  CHK(rdivp(h, Ur.get(), Tr.get()));                         // Ur = Ur (Tr P)^-1
  CHK(mul_blocks(h, Tr, Tl, Ul));                            // Tr = Tl Ul
  {  // G = Ur diag(1 / Dr) Tr'
    auto g = batch<double>(0, 1);
    for (int b = 0; b < h->flv; b++) {
      auto it = item<double>(Ur.p[b], Tr.p[b], G.p[b], n);
      it.dmid = diag_vec(Dr.p[b], kDiagInv);
      g.it[g.count++] = it;
    }
    CHK(run(h, g));
  }
  return DQMC_OK;
}

int slot_to(Chain* h, int s, double* U, double* D, double* T) {
  const size_t m = (size_t)h->flv * h->nn;
  if (!h->rec) {   // one launch for the three
    HIPCHK(h, launch_copy3(U, uslot(h, s, 0), (int64_t)m, T, tslot(h, s, 0), (int64_t)m, D, dslot(h, s, 0),
                           (int64_t)h->flv * h->n, h->st));
    return DQMC_OK;
  }
  CHK(copyd(h, U, uslot(h, s, 0), m));
  CHK(copyd(h, T, tslot(h, s, 0), m));
  return copyd(h, D, dslot(h, s, 0), (size_t)h->flv * h->n);
}
int udt_identity(Chain* h, double* U, double* D, double* T) {
  CHK(identity(h, U, h->flv));
  CHK(identity(h, T, h->flv));
  return ones(h, D, (size_t)h->flv * h->n);
}
int slot_identity(Chain* h, int s) { return udt_identity(h, uslot(h, s, 0), dslot(h, s, 0), tslot(h, s, 0)); }

int to32(Chain* h) {
  if (h->prec != DQMC_PRECISION_MIXED) return DQMC_OK;
  HIPCHK(h, launch_f64_to_f32(h->G32, h->G, (int64_t)h->flv * h->nn, h->st));
  return DQMC_OK;
}

// One stabilisation's max|G_wrapped - G_scratch|: propagation-error statistics (the
// reference's 1e-7 check, stack.jl:644-654) and, in mixed precision, the bound check.
void absorb_error(Chain* h, double x) {
  h->stabs++;
  h->perr_last = x;
  h->stablog.push_back(x);
  if (x > 1e-7 || x != x) { h->perr_count++; if (x > h->perr_max || x != x) h->perr_max = x; }
  if (h->prec == DQMC_PRECISION_MIXED && (x > h->mixed_tol || x != x)) {
    h->prec_viol++;
    if (x > h->prec_viol_max || x != x) h->prec_viol_max = x;
  }
}

int record_error(Chain* h, const double* a, const float* a32, const double* b) {
  if (h->nerr >= h->errcap) {
    // harvest pending errors before reusing the device slots
    std::vector<double> v(h->nerr);
    HIPCHK(h, hipMemcpyAsync(v.data(), h->errbuf, sizeof(double) * h->nerr, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    for (double x : v) absorb_error(h, x);
    h->nerr = 0;
  }
  if (h->rec) {
    // the slot starts at +0.0 (a fill), then the maximum is combined into it
    double* out = h->errbuf + h->nerr;
    h->nerr++;
    CHK(elem(h, ElemItem{kElemFill, 0, out, nullptr, 0.0, 1}));
    h->rmax.push_back(MaxdiffItem{a, b, (int64_t)(h->flv * h->nn), out});
    return record(h, kOpMaxdiff, 0, (int)h->rmax.size() - 1, 1);
  }
  HIPCHK(h, launch_maxabsdiff(a, a32, b, (int64_t)h->flv * h->nn, h->errbuf + h->nerr, h->st));
  h->nerr++;
  return DQMC_OK;
}

// Device errors (a bounded wait that gave up, a broken pivot selection) leave the chain's
// G wrong: every call that waits for the device reports new ones as DQMC_E_DEVICE.
int check_device_errors(Chain* h) {
  if (!h->ctr) return DQMC_OK;
  long long e = 0;
  HIPCHK(h, hipMemcpy(&e, &h->ctr->device_errors, sizeof(e), hipMemcpyDeviceToHost));
  if (e > h->dev_err_seen) {
    h->err = "device error: " + std::to_string(e - h->dev_err_seen) +
             " kernel(s) gave up a bounded wait or met a broken pivot selection; G is not valid";
    h->dev_err_seen = e;
    return DQMC_E_DEVICE;
  }
  return DQMC_OK;
}

int harvest(Chain* h) {
  if (h->hctr && h->ctr) {
    HIPCHK(h, hipMemcpyAsync(h->hctr, h->ctr, sizeof(ChainCounters), hipMemcpyDeviceToHost, h->st));
    if (h->nerr > 0)
      HIPCHK(h, hipMemcpyAsync(h->herr, h->errbuf, sizeof(double) * h->nerr, hipMemcpyDeviceToHost, h->st));
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  collect_times(h);
  if (h->hctr && h->ctr) {
    for (long long i = 0; i < h->nerr; i++) absorb_error(h, h->herr[i]);
    h->nerr = 0;
    h->acc_last = h->hctr->accepted;
    h->acc_valid = true;
    const long long e = h->hctr->device_errors;
    if (e > h->dev_err_seen) {
      h->err = "device error: " + std::to_string(e - h->dev_err_seen) +
               " kernel(s) gave up a bounded wait or met a broken pivot selection; G is not valid";
      h->dev_err_seen = e;
      return DQMC_E_DEVICE;
    }
  } else {
    if (h->nerr > 0) {
      std::vector<double> v(h->nerr);
      HIPCHK(h, hipMemcpy(v.data(), h->errbuf, sizeof(double) * h->nerr, hipMemcpyDeviceToHost));
      for (double x : v) absorb_error(h, x);
      h->nerr = 0;
    }
    CHK(check_device_errors(h));
  }
  if (h->prec_viol > h->prec_viol_seen) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "mixed precision: %lld stabilisation(s) with max|G_wrapped - G_scratch| above the "
                  "tolerance %.3g (largest %.3g); G was replaced by the FP64 recompute each time",
                  h->prec_viol - h->prec_viol_seen, h->mixed_tol, h->prec_viol_max);
    h->err = buf;
    h->prec_viol_seen = h->prec_viol;
    return DQMC_E_PRECISION;
  }
  return DQMC_OK;
}

int wrap_current(Chain* h, int l, int dir) {
  if (h->prec == DQMC_PRECISION_MIXED) return wrap32(h, h->G32, l, dir);
  return wrap64(h, h->G, l, dir);
}

// propagate(mc), stack.jl:605-730
int propagate(Chain* h) {
  const size_t tot = (size_t)h->flv * h->nn;
  const bool mixed = h->prec == DQMC_PRECISION_MIXED;
  h->cs += h->dir;
  if (h->dir == 1) {
    if (h->cs == 1) {
      return slot_identity(h, 1);
    } else if (h->cs - 1 == h->re[h->cr - 1]) {
      const int idx = h->cr;
      CHK(slot_to(h, idx + 1, h->Ur, h->Dr, h->Tr));
      CHK(add_slice_sequence_left(h, idx));
      CHK(slot_to(h, idx + 1, h->Ul, h->Dl, h->Tl));
      if (h->check) {
        if (mixed) {
          HIPCHK(h, hipMemcpyAsync(h->Gchk32, h->G32, sizeof(float) * tot, hipMemcpyDeviceToDevice, h->st));
          CHK(wrap32(h, h->Gchk32, h->cs - 1, 1));
        } else {
          CHK(copyd(h, h->Gchk, h->G, tot));
          CHK(wrap64(h, h->Gchk, h->cs - 1, 1));
        }
      }
      CHK(calculate_greens(h, h->G));
      if (h->check) CHK(record_error(h, h->Gchk, mixed ? h->Gchk32 : nullptr, h->G));
      CHK(to32(h));
      if (h->cr == h->K) {
        h->dir = -1;
        return propagate(h);
      }
      h->cr += 1;
      return DQMC_OK;
    }
    return wrap_current(h, h->cs - 1, 1);
  }
  if (h->cs == h->M) {
    CHK(slot_identity(h, h->K + 1));
    return wrap_current(h, h->cs + 1 - 1, -1);
  } else if (h->cs + 1 == h->rs[h->cr - 1]) {
    const int idx = h->cr;
    CHK(slot_to(h, idx, h->Ul, h->Dl, h->Tl));
    CHK(add_slice_sequence_right(h, idx));
    CHK(slot_to(h, idx, h->Ur, h->Dr, h->Tr));
    if (h->check && !mixed) CHK(copyd(h, h->Gchk, h->G, tot));
    if (h->check && mixed)
      HIPCHK(h, hipMemcpyAsync(h->Gchk32, h->G32, sizeof(float) * tot, hipMemcpyDeviceToDevice, h->st));
    CHK(calculate_greens(h, h->G));
    if (h->check) CHK(record_error(h, h->Gchk, mixed ? h->Gchk32 : nullptr, h->G));
    CHK(to32(h));
    if (h->cr == 1) {
      h->dir = 1;
      return propagate(h);
    }
    CHK(wrap_current(h, h->cs + 1 - 1, -1));
    h->cr -= 1;
    return DQMC_OK;
  }
  return wrap_current(h, h->cs + 1 - 1, -1);
}

// The slice launches' counters start again from zero (a wait that gave up leaves them short).
int reset_window_sync(Chain* h) {
  if (h->rec || !h->wsync) return DQMC_OK;
  HIPCHK(h, launch_fill(reinterpret_cast<double*>(h->wsync), 0.0, 1 + 2 * (int64_t)((h->n + 15) / 16), h->st));
  return DQMC_OK;
}

int reverse_build(Chain* h) {
  CHK(reset_window_sync(h));
  CHK(slot_identity(h, h->K + 1));
  for (int i = h->K; i >= 1; i--) CHK(add_slice_sequence_right(h, i));
  h->cs = 0; h->cr = 1; h->dir = 1;
  CHK(udt_identity(h, h->Ul, h->Dl, h->Tl));
  CHK(slot_to(h, 1, h->Ur, h->Dr, h->Tr));
  CHK(calculate_greens(h, h->G));
  return to32(h);
}

int forward_build(Chain* h) {
  CHK(reset_window_sync(h));
  CHK(slot_identity(h, 1));
  for (int i = 1; i <= h->K; i++) CHK(add_slice_sequence_left(h, i));
  h->cs = h->M + 1; h->cr = h->K; h->dir = -1;
  CHK(slot_to(h, h->K + 1, h->Ul, h->Dl, h->Tl));
  CHK(udt_identity(h, h->Ur, h->Dr, h->Tr));
  CHK(calculate_greens(h, h->G));
  return to32(h);
}

// The windows alternate between the two copies of G (each window reads one and writes
// the other); after an odd number of windows the copies swap roles.
template <typename GT>
int sweep_spatial_t(Chain* h, GT** Gp, GT** Gpp, GT* Y) {
  // Window k's launch also applies window k - 1 (source copy -> other copy); the factors
  // of consecutive windows alternate between two slots.
  const int n = h->n, W = h->win;
  const size_t ysz = (size_t)h->flv * W * W;
  GT* src = *Gp;
  GT* dst = *Gpp;
  WindowArgs<GT> a;
  a.n = n; a.flv = h->flv;
  a.conf = h->conf + (size_t)(h->cs - 1) * n;
  a.W = W;
  for (int k = 0; k < 2; k++) { a.d_first[k] = h->d_first[k]; a.d_second[k] = h->d_second[k]; a.bweight[k] = h->bweight[k]; }
  a.uniforms = h->unif; a.n_uniforms = h->nunif; a.seed = h->seed;
  a.ctr = h->ctr; a.dec_log = h->dlog; a.p_log = h->plog; a.log_cap = h->logcap;
  a.wait_limit = h->wait_limit;
  if constexpr (sizeof(GT) == 8) {
    if (!h->rec && h->win_slice && W <= 32 && h->wsync) {
      // one launch for the slice: windows back to back, applications beside them
      WindowSlice<double> sl;
      sl.a = a;
      sl.a.G = src; sl.a.Gd = dst; sl.a.w0 = 0; sl.a.wp = -W; sl.a.do_window = 1;
      sl.a.Sp = nullptr; sl.a.Yp = nullptr; sl.a.S = nullptr; sl.a.Y = nullptr;
      sl.B[0] = src; sl.B[1] = dst;
      sl.Y = Y; sl.S = h->wS;
      sl.sync = h->wsync;
      sl.nw = (n + W - 1) / W;
      const int nt = (n + 63) / 64;
      sl.napply = std::min(h->flv * nt * nt, 128);
      {
        Timed t(h, DQMC_KERNEL_LOCAL);
        HIPCHK(h, launch_window_slice_f64(sl, h->st));
      }
      if (sl.nw & 1) std::swap(src, dst);
      *Gp = src;
      *Gpp = dst;
      return DQMC_OK;
    }
  }
  int slot = 0;
  auto launch = [&](int kind) -> int {
    if constexpr (sizeof(GT) == 8) {
      if (h->rec) {
        h->rwin.push_back(a);
        return DQMC_OK;
      }
    }
    Timed t(h, kind);
    hipError_t e;
    if constexpr (sizeof(GT) == 8) e = launch_window_f64(a, h->st);
    else e = launch_window_f32(a, h->st);
    HIPCHK(h, e);
    return DQMC_OK;
  };
  for (int w0 = 0; w0 < n; w0 += W) {
    const bool prev = w0 > 0;
    a.G = src; a.Gd = dst;
    a.w0 = w0; a.wp = w0 - W; a.do_window = 1;
    a.Sp = prev ? h->wS + (slot ^ 1) * (W + 1) : nullptr;
    a.Yp = prev ? Y + (slot ^ 1) * ysz : nullptr;
    a.S = h->wS + slot * (W + 1);
    a.Y = Y + slot * ysz;
    CHK(launch(DQMC_KERNEL_LOCAL));
    if (prev) std::swap(src, dst);
    slot ^= 1;
  }
  // the last window's update
  a.G = src; a.Gd = dst;
  a.wp = ((n - 1) / W) * W; a.w0 = a.wp; a.do_window = 0;
  a.Sp = h->wS + (slot ^ 1) * (W + 1);
  a.Yp = Y + (slot ^ 1) * ysz;
  a.S = nullptr; a.Y = nullptr;
  CHK(launch(DQMC_KERNEL_LOCAL_APPLY));
  std::swap(src, dst);
  *Gp = src;
  *Gpp = dst;
  return DQMC_OK;
}

int sweep_spatial(Chain* h) {
  h->fver++;
  h->acc_valid = false;
  if (h->cs < 1 || h->cs > h->M) {
    h->err = "sweep_spatial: current slice outside 1..slices (build the stack first)";
    return DQMC_E_STATE;
  }
  if (h->prec == DQMC_PRECISION_MIXED)
    return sweep_spatial_t<float>(h, &h->G32, &h->G32pp, reinterpret_cast<float*>(h->wY));
  return sweep_spatial_t<double>(h, &h->G, &h->Gpp, h->wY);
}

// exp(c T) on the device: Taylor series of degree 20 after scaling by 2^-s so that the
// 1-norm is <= 1/2, then s squarings; finally symmetrised (T is symmetric, and the
// reference's exp of a Hermitian matrix is exactly Hermitian).
int expm_device(Chain* h, double c, double norm1_T, double* out) {
  const int n = h->n;
  const double nrm = std::fabs(c) * norm1_T;
  int s = 0;
  if (nrm > 0.5) s = (int)std::ceil(std::log2(nrm / 0.5));
  const double scale = c / std::ldexp(1.0, s);
  double* Bm = h->tmp1;      // scaled T
  double* term = h->tmp2;
  double* nxt = h->cur;
  CHK(copyd(h, Bm, h->dT, h->nn));
  HIPCHK(h, launch_scale_add_identity(Bm, scale, 0.0, n, h->st));
  CHK(identity(h, out, 1));
  CHK(identity(h, term, 1));
  for (int k = 1; k <= 20; k++) {
    auto g = batch<double>(0, 0, 1.0 / k);
    g.it[g.count++] = item<double>(term, Bm, nxt, n);
    CHK(run(h, g));
    HIPCHK(h, launch_axpy_matrix(out, nxt, 1.0, (int64_t)h->nn, h->st));
    std::swap(term, nxt);
  }
  for (int q = 0; q < s; q++) {
    auto g = batch<double>(0, 0);
    g.it[g.count++] = item<double>(out, out, h->Gx, n);
    CHK(run(h, g));
    CHK(copyd(h, out, h->Gx, h->nn));
  }
  HIPCHK(h, launch_symmetrize(out, n, h->st));
  return DQMC_OK;
}

int finish_hopping(Chain* h) {
  h->fver++;
  if (h->prec == DQMC_PRECISION_MIXED) {
    HIPCHK(h, launch_f64_to_f32(h->eT2m32, h->eT2m, (int64_t)h->nn, h->st));
    HIPCHK(h, launch_f64_to_f32(h->eT2p32, h->eT2p, (int64_t)h->nn, h->st));
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->hopping_set = true;
  return DQMC_OK;
}

int64_t footprint(const dqmc_config_t* c, int* Kout) {
  std::vector<int> s, e;
  const int K = generate_chunks(c->slices, c->safe_mult, s, e);
  if (Kout) *Kout = K;
  const int64_t n = c->n_sites, f = c->flavors, nn = n * n;
  int64_t b = 0;
  b += 5 * nn * 8;                                   // hopping exps + T
  b += 3 * f * nn * 8;                               // G, Gx, Gchk
  b += (int64_t)(K + 1) * f * (2 * nn + n) * 8;      // stack
  b += 8 * f * nn * 8 + 2 * f * n * 8;               // Ul Ur Tl Tr tmp1 tmp2 cur O, Dl Dr
  b += f * n * (4 + 8 + 32 * 8) + f * (n + 16) * 16 * 8; // pivot, tau, swap, tfac
  b += n * (int64_t)c->slices;                       // conf
  b += (n * (int64_t)c->slices + 63) / 64 * 8;       // conf as BitArray chunks
  if (c->precision == DQMC_PRECISION_MIXED) b += (3 * f * nn + 2 * nn) * 4;
  const int64_t W = window_size((int)n, (int)f);
  const int64_t welem = c->precision == DQMC_PRECISION_MIXED ? 4 : 8;
  b += (f * nn + 2 * f * W * W) * welem + 2 * (W + 1) * 4; // second copy of G, window factors Y, S (two slots)
  b += 4096 * 8 + 4096;                              // error slots, counters
  return b;
}

int read_accepted(Chain* h, long long* acc) {
  ChainCounters c;
  HIPCHK(h, hipMemcpy(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost));
  *acc = c.accepted;
  return DQMC_OK;
}

// calculate_greens(mc, slice, output) from scratch (stack.jl:525-583): the chain of slice
// products with a UDT every safe_mult slices (k % safe_mult == 0) on both sides of `slice`,
// then calculate_greens into `out`.  Uses the stack's temporaries, not its slots.
int greens_from_scratch(Chain* h, int slice, double* out) {
  const int sm = h->sm, M = h->M;
  const size_t m = (size_t)h->flv * h->nn;
  double* ca = h->cur;
  double* cb = h->tmp2;
  auto chain_step = [&](double* Ux, double* Dx, double* Tx, bool final) -> int {
    HIPCHK(h, launch_scale_cols(h->tmp1, ca, Dx, h->n, h->flv, h->st));
    CHK(udt(h, B(h, h->tmp1).get(), B(h, final ? Ux : ca).get(), Bd(h, Dx).get(), B(h, h->O).get(), 1));
    CHK(mul_blocks(h, B(h, h->Gx), B(h, h->O), B(h, Tx)));
    return copyd(h, Tx, h->Gx, m);
  };
  CHK(identity(h, ca, h->flv));
  CHK(udt_identity(h, h->Ur, h->Dr, h->Tr));
  if (slice + 1 <= M) {
    for (int k = M; k >= slice + 1; k--) {
      CHK(slice_dagger_left(h, k, B(h, ca).get(), B(h, cb).get(), nullptr));
      std::swap(ca, cb);
      if (k % sm == 0) CHK(chain_step(nullptr, h->Dr, h->Tr, false));
    }
    CHK(chain_step(h->Ur, h->Dr, h->Tr, true));
  }
  CHK(identity(h, ca, h->flv));
  CHK(udt_identity(h, h->Ul, h->Dl, h->Tl));
  if (slice >= 1) {
    for (int k = 1; k <= slice; k++) {
      CHK(slice_left(h, k, B(h, ca).get(), B(h, cb).get(), nullptr));
      std::swap(ca, cb);
      if (k % sm == 0) CHK(chain_step(nullptr, h->Dl, h->Tl, false));
    }
    CHK(chain_step(h->Ul, h->Dl, h->Tl, true));
  }
  return calculate_greens(h, out);
}

// ---- unequal-time Green's functions (unequal_time_stack.jl) --------------------------------
Slots ut_forward(Chain* h) { return Slots{h->fu, h->fd, h->ft}; }
Slots ut_backward(Chain* h) { return Slots{h->bu, h->bd, h->bt}; }
Slots ut_inverse(Chain* h) { return Slots{h->iu, h->id, h->it}; }

// initialize_stack(mc, ::UnequalTimeStack), unequal_time_stack.jl:64-113: one allocation for
// the three stacks and the block / result buffers, made on first use.
int ut_alloc(Chain* h) {
  if (h->fu) return DQMC_OK;
  const size_t K = (size_t)h->K, f = (size_t)h->flv, nn = h->nn, n = (size_t)h->n;
  const size_t mats = f * nn, vecs = f * n;
  auto up = [](size_t x) { return (x + 31) & ~(size_t)31; };  // 256-byte aligned offsets
  const size_t total = up(2 * (K + 1) * mats) * 2 + up((K + 1) * vecs) * 2 + up(K * mats) * 2 + up(K * vecs) +
                       up(mats) * 3 + up(vecs) * 2 + up(14 * mats) + up(3 * vecs);
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, total * sizeof(double));
  if (e != hipSuccess) {
    h->err = std::string("unequal-time stack: hipMalloc: ") + hipGetErrorString(e);
    return DQMC_E_ALLOC;
  }
  h->allocs.push_back(q);
  double* p = static_cast<double*>(q);
  auto take = [&](size_t count) { double* r = p; p += up(count); return r; };
  h->fu = take((K + 1) * mats); h->ft = take((K + 1) * mats);
  h->bu = take((K + 1) * mats); h->bt = take((K + 1) * mats);
  h->fd = take((K + 1) * vecs); h->bd = take((K + 1) * vecs);
  h->iu = take(K * mats); h->it = take(K * mats); h->id = take(K * vecs);
  h->uU = take(mats); h->uT = take(mats); h->uG = take(mats);
  h->uD = take(vecs); h->uDsave = take(vecs);
  h->itm = take(14 * mats); h->itv = take(3 * vecs);
  h->ut_ver = -1;
  return DQMC_OK;
}

// The stacks hold the current field: after any change of field or hopping the built slots
// are dropped (the last_update test of lazy_build_*, unequal_time_stack.jl:190-196).
int ut_validate(Chain* h) {
  CHK(ut_alloc(h));
  if (h->ut_ver == h->fver) return DQMC_OK;
  h->ut_ver = h->fver;
  h->ut_fwd = 1;
  h->ut_bwd = h->K + 1;
  h->ut_inv.assign(h->K, 0);
  h->ut_k = h->ut_l = -1;
  const Slots F = ut_forward(h), Bk = ut_backward(h);
  CHK(udt_identity(h, SU(h, F, 1).p[0], SD(h, F, 1).p[0], ST(h, F, 1).p[0]));
  return udt_identity(h, SU(h, Bk, h->K + 1).p[0], SD(h, Bk, h->K + 1).p[0], ST(h, Bk, h->K + 1).p[0]);
}

// lazy_build_forward!(mc, s, upto), unequal_time_stack.jl:183-220
int ut_build_forward(Chain* h, int upto) {
  for (int idx = h->ut_fwd; idx <= upto - 1; idx++) CHK(add_slice_sequence_left(h, idx, ut_forward(h)));
  h->ut_fwd = std::max(upto, h->ut_fwd);
  return DQMC_OK;
}

// lazy_build_backward!(mc, s, downto), unequal_time_stack.jl:222-250
int ut_build_backward(Chain* h, int downto) {
  for (int idx = h->ut_bwd - 1; idx >= downto; idx--) CHK(add_slice_sequence_right(h, idx, ut_backward(h)));
  h->ut_bwd = std::min(downto, h->ut_bwd);
  return DQMC_OK;
}

// M_b <- eV(slice)^-1 eT2^-1 M_b              multiply_slice_matrix_inv_left!  (stack.jl:349-357)
int slice_inv_left(Chain* h, int slice, double* const in[], double* const out[], const double* const dr[]) {
  auto g = batch<double>(0, 0);
  for (int b = 0; b < h->flv; b++) {
    auto it = item<double>(h->eT2p, in[b], out[b], h->n);
    it.dl = ev(h, b, slice, -1.0);
    if (dr) it.dr = diag_vec(dr[b]);
    g.it[g.count++] = it;
  }
  return run(h, g);
}
// M_b <- M_b eV(slice)^-1 eT2^-1              multiply_slice_matrix_inv_right! (stack.jl:339-347)
int slice_inv_right(Chain* h, int slice, double* const in[], double* const out[]) {
  auto g = batch<double>(0, 0);
  for (int b = 0; b < h->flv; b++) {
    auto it = item<double>(in[b], h->eT2p, out[b], h->n);
    it.dmid = ev(h, b, slice, -1.0);
    g.it[g.count++] = it;
  }
  return run(h, g);
}

// lazy_build_inv!(mc, s, from, to), unequal_time_stack.jl:252-278: inverse slot idx holds
// B_{r0}^-1 .. B_{r1}^-1 of range idx as a pivoted UDT.
int ut_build_inverse(Chain* h, int from, int to) {
  const Slots I = ut_inverse(h);
  for (int idx = from; idx <= to; idx++) {
    if (h->ut_inv[idx - 1]) continue;
    h->ut_inv[idx - 1] = 1;
    const int r0 = h->rs[idx - 1], r1 = h->re[idx - 1];
    CHK(identity(h, h->cur, h->flv));
    Blocks in = B(h, h->cur);
    for (int sl = r1; sl >= r0; sl--) {
      Blocks out = sl == r0 ? B(h, h->tmp1) : ((r1 - sl) % 2 == 0 ? B(h, h->tmp2) : B(h, h->cur));
      CHK(slice_inv_left(h, sl, in.get(), out.get(), nullptr));
      in = out;
    }
    CHK(udt(h, B(h, h->tmp1).get(), SU(h, I, idx).get(), SD(h, I, idx).get(), ST(h, I, idx).get(), 1));
  }
  return DQMC_OK;
}

// _find_range_with_value(mc, val), unequal_time_stack.jl:338-367: the 1-based range holding
// val; 0 below all ranges, K + 1 above.
int find_range(const Chain* h, int val) {
  if (val < 1) return 0;
  if (val > h->re[h->K - 1]) return h->K + 1;
  for (int i = 0; i < h->K; i++)
    if (h->rs[i] <= val && val <= h->re[i]) return i + 1;
  return h->K + 1;
}

// compute_inverse_udt_block!(mc, s, low, high), unequal_time_stack.jl:383-459:
// uU diag(uD) uT = B_{low+1}^-1 .. B_high^-1
int ut_inverse_block(Chain* h, int low, int high) {
  const int lower = find_range(h, low) + 1;
  const int upper = find_range(h, high + 1) - 1;
  CHK(ut_build_inverse(h, lower, upper));
  const Slots I = ut_inverse(h);
  CHK(udt_identity(h, h->uU, h->uD, h->uT));
  const int n = h->n;
  for (int idx = lower; idx <= upper; idx++) {
    {  // tmp1 = D (T inv_u) inv_d
      auto g = batch<double>(0, 0);
      Blocks iu = SU(h, I, idx), idd = SD(h, I, idx);
      for (int b = 0; b < h->flv; b++) {
        auto it = item<double>(blk(h->uT, h, b), iu.p[b], blk(h->tmp1, h, b), n);
        it.dl = diag_vec(h->uD + (size_t)b * n);
        it.dr = diag_vec(idd.p[b]);
        g.it[g.count++] = it;
      }
      CHK(run(h, g));
    }
    CHK(udt(h, B(h, h->tmp1).get(), B(h, h->tmp2).get(), Bd(h, h->uD).get(), B(h, h->cur).get(), 1));
    CHK(mul_blocks(h, B(h, h->uT), B(h, h->cur), ST(h, I, idx)));   // T = T' inv_t
    CHK(mul_blocks(h, B(h, h->tmp1), B(h, h->uU), B(h, h->tmp2)));  // U = U U'
    CHK(copyd(h, h->uU, h->tmp1, (size_t)h->flv * h->nn));
  }
  const int lower_slice = lower <= h->K ? h->rs[lower - 1] : h->re[h->K - 1] + 1;
  const int upper_slice = upper > 0 ? h->re[upper - 1] : 0;
  const int top = std::min(lower_slice - 1, high);
  if (top >= low + 1) {
    // U <- B_{low+1}^-1 .. B_top^-1 U, the last product scaled by D; then U D T <- UDT(U D) T
    Blocks in = B(h, h->uU);
    for (int sl = top; sl >= low + 1; sl--) {
      const bool last = sl == low + 1;
      Blocks out = last ? B(h, h->tmp1) : ((top - sl) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
      CHK(slice_inv_left(h, sl, in.get(), out.get(), last ? Bd(h, h->uD).get() : nullptr));
      in = out;
    }
    CHK(udt(h, B(h, h->tmp1).get(), B(h, h->uU).get(), Bd(h, h->uD).get(), B(h, h->cur).get(), 1));
    CHK(mul_blocks(h, B(h, h->tmp2), B(h, h->cur), B(h, h->uT)));
    CHK(copyd(h, h->uT, h->tmp2, (size_t)h->flv * h->nn));
  }
  const int first = std::max(upper_slice + 1, top + 1);
  if (first <= high) {
    Blocks in = B(h, h->uT);
    for (int sl = first; sl <= high; sl++) {
      const bool last = sl == high;
      Blocks out = last ? B(h, h->uT) : ((sl - first) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
      if (last && in.p[0] == h->uT) out = B(h, h->tmp1);  // one slice: not in place
      CHK(slice_inv_right(h, sl, in.get(), out.get()));
      in = out;
    }
    if (in.p[0] != h->uT) CHK(copyd(h, h->uT, in.p[0], (size_t)h->flv * h->nn));
  }
  return DQMC_OK;
}

// compute_forward_udt_block!(mc, s, slice), unequal_time_stack.jl:473-494:
// Ul diag(Dl) Tl = B_slice .. B_1
int ut_forward_block(Chain* h, int slice) {
  const int idx = std::max(0, find_range(h, slice) - 1);
  CHK(ut_build_forward(h, idx + 1));
  const Slots F = ut_forward(h);
  const int target = idx > 0 ? h->re[idx - 1] + 1 : 1;
  Blocks in = SU(h, F, idx + 1);
  if (target > slice) {
    HIPCHK(h, launch_scale_cols(h->tmp1, in.p[0], SD(h, F, idx + 1).p[0], h->n, h->flv, h->st));
  } else {
    for (int l = target; l <= slice; l++) {
      const bool last = l == slice;
      Blocks out = last ? B(h, h->tmp1) : ((l - target) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
      CHK(slice_left(h, l, in.get(), out.get(), last ? SD(h, F, idx + 1).get() : nullptr));
      in = out;
    }
  }
  CHK(udt(h, B(h, h->tmp1).get(), B(h, h->Ul).get(), Bd(h, h->Dl).get(), B(h, h->tmp2).get(), 1));
  return mul_blocks(h, B(h, h->Tl), B(h, h->tmp2), ST(h, F, idx + 1));
}

// compute_backward_udt_block!(mc, s, slice), unequal_time_stack.jl:508-531:
// (Ur diag(Dr) Tr)^T = B_M .. B_{slice+1}
int ut_backward_block(Chain* h, int slice) {
  const int idx = find_range(h, slice) + 1;
  CHK(ut_build_backward(h, idx));
  const Slots Bk = ut_backward(h);
  const int target = idx <= h->K ? h->rs[idx - 1] - 1 : h->re[h->K - 1];
  Blocks in = SU(h, Bk, idx);
  if (target < slice + 1) {
    HIPCHK(h, launch_scale_cols(h->tmp1, in.p[0], SD(h, Bk, idx).p[0], h->n, h->flv, h->st));
  } else {
    for (int l = target; l >= slice + 1; l--) {
      const bool last = l == slice + 1;
      Blocks out = last ? B(h, h->tmp1) : ((target - l) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
      CHK(slice_dagger_left(h, l, in.get(), out.get(), last ? SD(h, Bk, idx).get() : nullptr));
      in = out;
    }
  }
  CHK(udt(h, B(h, h->tmp1).get(), B(h, h->Ur).get(), Bd(h, h->Dr).get(), B(h, h->tmp2).get(), 1));
  return mul_blocks(h, B(h, h->Tr), B(h, h->tmp2), ST(h, Bk, idx));
}

// C_b = alpha diag(dl_b) op(A_b) diag(dmid_b) op(B_b) diag(dr_b) over the flavour blocks
int gemm_blocks(Chain* h, int ta, int tb, double alpha, const Blocks& C, const Blocks& A, const Blocks& Bm,
                const double* dl, int lmode, const double* dmid, int mmode, const double* dr, int rmode,
                double beta = 0.0) {
  auto g = batch<double>(ta, tb, alpha, beta);
  for (int b = 0; b < h->flv; b++) {
    auto it = item<double>(A.p[b], Bm.p[b], C.p[b], h->n);
    if (dl) it.dl = diag_vec(dl + (size_t)b * h->n, lmode);
    if (dmid) it.dmid = diag_vec(dmid + (size_t)b * h->n, mmode);
    if (dr) it.dr = diag_vec(dr + (size_t)b * h->n, rmode);
    g.it[g.count++] = it;
  }
  return run(h, g);
}

// calculate_greens_full1!(mc, s, k, l), unequal_time_stack.jl:535-617 (k >= l), into uG
int ut_full1(Chain* h, int k, int l) {
  CHK(ut_inverse_block(h, l, k));
  CHK(ut_forward_block(h, l));
  CHK(ut_backward_block(h, k));
  Blocks G = B(h, h->uG), Ul = B(h, h->Ul), Ur = B(h, h->Ur), Tl = B(h, h->Tl), Tr = B(h, h->Tr);
  Blocks U = B(h, h->uU), T = B(h, h->uT), X = B(h, h->tmp1), Y = B(h, h->tmp2);
  // G = Dl (Tl Tr') Dr
  CHK(gemm_blocks(h, 0, 1, 1.0, G, Tl, Tr, h->Dl, kDiagPlain, nullptr, 0, h->Dr, kDiagPlain));
  CHK(udt(h, G.get(), Tr.get(), Bd(h, h->Dr).get(), G.get(), 0));  // G -> Tr Dr (G P)
  CHK(mul_blocks(h, Tl, Ul, Tr));                                  // Tl = Ul Tr
  CHK(rdivp(h, Ur.get(), G.get()));                                // Ur = Ur / G
  // Tr = 1/max(1, D) U' Tl min(1, Dr)
  CHK(gemm_blocks(h, 1, 0, 1.0, Tr, U, Tl, h->uD, kDiagMaxInv, nullptr, 0, h->Dr, kDiagMin));
  // Tr += min(1, D) T Ur 1/max(1, Dr)
  CHK(gemm_blocks(h, 0, 0, 1.0, Tr, T, Ur, h->uD, kDiagMin, nullptr, 0, h->Dr, kDiagMaxInv, 1.0));
  CHK(udt(h, Tr.get(), Tl.get(), Bd(h, h->Dl).get(), Tr.get(), 0));  // sum -> Tl Dl (Tr P)
  // X = diag(1/max(1, Dr)) / (Tr P)
  HIPCHK(h, launch_diag_matrix(h->tmp1, h->Dr, kDiagMaxInv, h->n, h->flv, h->st));
  CHK(rdivp(h, X.get(), Tr.get()));
  // Y = X diag(1/Dl) Tl' diag(1/max(1, D)); G = Ur (Y U')
  CHK(gemm_blocks(h, 0, 1, 1.0, Y, X, Tl, nullptr, 0, h->Dl, kDiagInv, h->uD, kDiagMaxInv));
  CHK(gemm_blocks(h, 0, 1, 1.0, X, Y, U, nullptr, 0, nullptr, 0, nullptr, 0));
  return gemm_blocks(h, 0, 0, 1.0, G, Ur, X, nullptr, 0, nullptr, 0, nullptr, 0);
}

// calculate_greens_full2!(mc, s, k, l), unequal_time_stack.jl:620-697 (k < l), into uG
int ut_full2(Chain* h, int k, int l) {
  CHK(ut_inverse_block(h, k, l));
  CHK(ut_forward_block(h, k));
  CHK(ut_backward_block(h, l));
  Blocks G = B(h, h->uG), Ul = B(h, h->Ul), Ur = B(h, h->Ur), Tl = B(h, h->Tl), Tr = B(h, h->Tr);
  Blocks U = B(h, h->uU), T = B(h, h->uT), X = B(h, h->tmp1), Y = B(h, h->tmp2);
  CHK(gemm_blocks(h, 0, 1, 1.0, G, Tl, Tr, h->Dl, kDiagPlain, nullptr, 0, h->Dr, kDiagPlain));
  CHK(udt(h, G.get(), Tr.get(), Bd(h, h->Dr).get(), G.get(), 0));  // G -> Tr Dr (G P)
  CHK(mul_blocks(h, Tl, Ul, Tr));                                  // Tl = Ul Tr
  // Ul = 1/max(1, D) U' Tl min(1, Dr)
  CHK(gemm_blocks(h, 1, 0, 1.0, Ul, U, Tl, h->uD, kDiagMaxInv, nullptr, 0, h->Dr, kDiagMin));
  // X = (T Ur) / G; Tr = min(1, D) X 1/max(1, Dr) + Ul
  CHK(mul_blocks(h, X, T, Ur));
  CHK(rdivp(h, X.get(), G.get()));
  HIPCHK(h, launch_diag_scale(h->Tr, h->tmp1, h->uD, kDiagMin, h->Dr, kDiagMaxInv, h->Ul, h->n, h->flv, h->st));
  CHK(udt(h, Tr.get(), Ul.get(), Bd(h, h->Dl).get(), Tr.get(), 0));  // sum -> Ul Dl (Tr P)
  // X = diag(min(1, Dr)) / (Tr P); Ur = X diag(1/Dl) Ul' diag(min(1, D))
  HIPCHK(h, launch_diag_matrix(h->tmp1, h->Dr, kDiagMin, h->n, h->flv, h->st));
  CHK(rdivp(h, X.get(), Tr.get()));
  CHK(gemm_blocks(h, 0, 1, 1.0, Ur, X, Ul, nullptr, 0, h->Dl, kDiagInv, h->uD, kDiagMin));
  // G = -Tl (Ur T)
  CHK(mul_blocks(h, Y, Ur, T));
  return gemm_blocks(h, 0, 0, -1.0, G, Tl, Y, nullptr, 0, nullptr, 0, nullptr, 0);
}

// calculate_greens(mc, k, l), unequal_time_stack.jl:326-338: the effective G(k <- l) in uG,
// reused while (k, l) and the field are unchanged.  The chain's temporaries are scratch here;
// the D of the chain's own G (Dr) is kept for global updates.
int ut_greens(Chain* h, int k, int l) {
  CHK(ut_validate(h));
  if (h->ut_k == k && h->ut_l == l) return DQMC_OK;
  const size_t nd = (size_t)h->flv * h->n;
  if (h->dg_valid) CHK(copyd(h, h->uDsave, h->Dr, nd));
  h->ut_k = h->ut_l = -1;
  CHK(k >= l ? ut_full1(h, k, l) : ut_full2(h, k, l));
  if (h->dg_valid) CHK(copyd(h, h->Dr, h->uDsave, nd));
  h->ut_k = k;
  h->ut_l = l;
  return DQMC_OK;
}

// _greens!(mc, target, source), greens.jl:114-125: out = eT/2+ src eT/2- per flavour block
int physical_form(Chain* h, const double* src, double* out) {
  auto g1 = batch<double>(0, 1);
  auto g2 = batch<double>(1, 0);
  for (int b = 0; b < h->flv; b++) {
    g1.it[g1.count++] = item<double>(src + (size_t)b * h->nn, h->eTmh, blk(h->Gx, h, b), h->n);
    g2.it[g2.count++] = item<double>(h->eTph, blk(h->Gx, h, b), out + (size_t)b * h->nn, h->n);
  }
  CHK(run(h, g1));
  return run(h, g2);
}

// CombinedGreensIterator(mc; start, stop, recalculate), measurements/greens_iterators.jl:
// 181-407: for l = start..stop the physical (G(0, l), G(l, 0), G(l, l)), written by `emit`
// (entry i = l - start, matrix j = 0, 1, 2).  Between recalculations (every `recalculate`
// slices, from the unequal-time stack) the three are carried as UDT products advanced by one
// slice per step, re-decomposed every safe_mult slices.
template <typename Emit>
int combined_greens(Chain* h, int start, int stop, int rec, Emit emit) {
  CHK(ut_validate(h));
  const size_t mats = (size_t)h->flv * h->nn, nv = (size_t)h->flv * h->n;
  const int n = h->n;
  double* m[10];
  for (int i = 0; i < 10; i++) m[i] = h->itm + i * mats;
  double *Ul = m[0], *Tl = m[1], *Ur = m[2], *Tr = m[3], *uU = m[4], *uT = m[5];
  double *X = m[6], *Y = m[7], *Z = m[8], *W = m[9];
  double *Dl = h->itv, *Dr = h->itv + nv, *uD = h->itv + 2 * nv;
  auto gemm3 = [&](double* C, const double* A, const double* Bm, const double* dmid) -> int {
    auto g = batch<double>(0, 0);
    for (int b = 0; b < h->flv; b++) {
      auto it = item<double>(A + b * h->nn, Bm + b * h->nn, C + b * h->nn, n);
      if (dmid) it.dmid = diag_vec(dmid + (size_t)b * n);
      g.it[g.count++] = it;
    }
    return run(h, g);
  };
  auto advance = [&](int l) -> int {  // Ul = B_l Ul, Tr = Tr B_l^-1, uU = B_l uU, uT = uT B_l^-1
    CHK(slice_left(h, l, B(h, Ul).get(), B(h, X).get(), nullptr));
    std::swap(Ul, X);
    CHK(slice_inv_right(h, l, B(h, Tr).get(), B(h, X).get()));
    std::swap(Tr, X);
    CHK(slice_left(h, l, B(h, uU).get(), B(h, X).get(), nullptr));
    std::swap(uU, X);
    CHK(slice_inv_right(h, l, B(h, uT).get(), B(h, X).get()));
    std::swap(uT, X);
    return DQMC_OK;
  };
  auto recalc = [&](int l, int i) -> int {
    CHK(ut_greens(h, l, 0));
    CHK(copyd(h, X, h->uG, mats));
    CHK(ut_greens(h, 0, l));
    CHK(copyd(h, Y, h->uG, mats));
    CHK(ut_greens(h, l, l));
    CHK(copyd(h, Z, h->uG, mats));
    CHK(emit(i, 0, Y));
    CHK(emit(i, 1, X));
    CHK(emit(i, 2, Z));
    CHK(udt(h, B(h, Z).get(), B(h, uU).get(), Bd(h, uD).get(), B(h, uT).get(), 1));
    CHK(udt(h, B(h, X).get(), B(h, Ul).get(), Bd(h, Dl).get(), B(h, Tl).get(), 1));
    return udt(h, B(h, Y).get(), B(h, Ur).get(), Bd(h, Dr).get(), B(h, Tr).get(), 1);
  };
  int l;
  if (start == 0 || start == 1) {
    // G(0, 0): the chain's own G at current slice 1, else calculate_greens_full1!(0, 0)
    if (h->cs == 1 && h->prec == DQMC_PRECISION_FP64) {
      CHK(copyd(h, Y, h->G, mats));
    } else {
      CHK(ut_greens(h, 0, 0));
      CHK(copyd(h, Y, h->uG, mats));
    }
    CHK(copyd(h, X, Y, mats));
    for (int b = 0; b < h->flv; b++) HIPCHK(h, launch_scale_add_identity(X + b * h->nn, 1.0, -1.0, n, h->st));
    CHK(copyd(h, Z, Y, mats));
    CHK(udt(h, B(h, Z).get(), B(h, Ul).get(), Bd(h, Dl).get(), B(h, Tl).get(), 1));
    CHK(copyd(h, uU, Ul, mats));
    CHK(copyd(h, uD, Dl, nv));
    CHK(copyd(h, uT, Tl, mats));
    CHK(udt(h, B(h, X).get(), B(h, Ur).get(), Bd(h, Dr).get(), B(h, Tr).get(), 1));  // G(0, 0) - I
    if (start == 0) {
      CHK(emit(0, 0, Y));
      CHK(emit(0, 1, Y));
      CHK(emit(0, 2, Y));
    }
    l = 1;
  } else {
    CHK(recalc(start, 0));
    l = start + 1;
  }
  const int shift = start != 1 ? start : 0;
  for (; l <= stop; l++) {
    const int i = l - start;
    if ((l - shift) % rec == 0) {
      CHK(recalc(l, i));
      continue;
    }
    CHK(advance(l));
    if (((l - shift) % rec) % h->sm == 0) {
      // G(l, 0): Ul Dl Tl, then Ul Dl Tl <- UDT(Ul Dl) Tl
      HIPCHK(h, launch_scale_cols(X, Ul, Dl, n, h->flv, h->st));
      CHK(gemm3(Y, X, Tl, nullptr));
      CHK(emit(i, 1, Y));
      CHK(udt(h, B(h, X).get(), B(h, Ul).get(), Bd(h, Dl).get(), B(h, Z).get(), 1));
      CHK(gemm3(X, Z, Tl, nullptr));
      std::swap(Tl, X);
      // G(0, l): Ur (Dr Tr), then Ur Dr Tr <- Ur UDT(Dr Tr)
      HIPCHK(h, launch_diag_scale(X, Tr, Dr, kDiagPlain, nullptr, 0, nullptr, n, h->flv, h->st));
      CHK(gemm3(Y, Ur, X, nullptr));
      CHK(emit(i, 0, Y));
      CHK(udt(h, B(h, X).get(), B(h, Z).get(), Bd(h, Dr).get(), B(h, Tr).get(), 1));
      CHK(gemm3(X, Ur, Z, nullptr));
      std::swap(Ur, X);
      // G(l, l): (uU uD) uT, then uU uD uT <- C UDT(uD T' uT) with C uD T' = UDT(uU uD)
      HIPCHK(h, launch_scale_cols(X, uU, uD, n, h->flv, h->st));
      CHK(gemm3(Y, X, uT, nullptr));
      CHK(emit(i, 2, Y));
      CHK(udt(h, B(h, X).get(), B(h, Z).get(), Bd(h, uD).get(), B(h, W).get(), 1));
      CHK(gemm3(X, W, uT, nullptr));
      HIPCHK(h, launch_diag_scale(W, X, uD, kDiagPlain, nullptr, 0, nullptr, n, h->flv, h->st));
      CHK(udt(h, B(h, W).get(), B(h, Y).get(), Bd(h, uD).get(), B(h, uT).get(), 1));
      CHK(gemm3(X, Z, Y, nullptr));
      std::swap(uU, X);
    } else {
      CHK(gemm3(X, Ul, Tl, Dl));
      CHK(emit(i, 1, X));
      CHK(gemm3(X, Ur, Tr, Dr));
      CHK(emit(i, 0, X));
      CHK(gemm3(X, uU, uT, uD));
      CHK(emit(i, 2, X));
    }
  }
  return DQMC_OK;
}

// The counter stream of the window kernels, u_k = SplitMix64(seed, k) >> 11 * 2^-53.
double counter_uniform_host(unsigned long long seed, unsigned long long k) {
  unsigned long long z = seed + (k + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * 0x1.0p-53;
}

// ---- many-chain steps -------------------------------------------------------------------
// One step is one slice visit of every chain: sweep_spatial + propagate, the body of
// local_sweep (local_updates.jl:7-14).  Each chain's kernels are recorded, then launched
// together: window launch k covers the k-th window launch of every chain (grouped by
// flavours and window width), and the propagate kernels of the chains whose recorded
// sequences have the same kinds share one launch per position (a chain's items run in its
// own order, so every chain computes exactly what its own launches would).  A step's item
// tables are packed into pinned host memory and reach the device in one copy; a ring of
// slots lets the host pack step t + 1 while the device runs step t.
struct Executor {
  int dev = 0;
  hipStream_t st = nullptr;
  static constexpr int kSlots = 4;
  char* host[kSlots] = {};
  char* devb[kSlots] = {};
  size_t cap[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  bool pending[kSlots] = {};
  int slot = 0;
  std::string err;
  std::vector<char> stage;
  struct Launch {
    int kind;          // 0 window, 1 gemm, 2 udt norms, 3 udt panel, 4 udt reflect, 5 udt finish,
                       // 6 rdivp, 7 elem, 8 maxdiff
    size_t off;
    int count;
    int a, b, c;       // window: flv, W, max_n; gemm: transA, transB, -; udt: max_n, R|apply, grid
    int64_t m1, m2;    // gemm max_m, max_n; elem / maxdiff max count
    double alpha, beta;
  };
  std::vector<Launch> launches;

  ~Executor() {
    if (st) (void)hipStreamSynchronize(st);
    for (int k = 0; k < kSlots; k++) {
      if (host[k]) (void)hipHostFree(host[k]);
      if (devb[k]) (void)hipFree(devb[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
    if (st) (void)hipStreamDestroy(st);
  }

  template <typename T>
  size_t put(const T* p, size_t count) {
    size_t off = (stage.size() + 255) & ~(size_t)255;
    stage.resize(off + sizeof(T) * count);
    std::memcpy(stage.data() + off, p, sizeof(T) * count);
    return off;
  }
};

#define XCHK(x, expr)                                                       \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      (x)->err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
      return DQMC_E_HIP;                                                    \
    }                                                                       \
  } while (0)

int executor_init(Executor* x, int dev) {
  x->dev = dev;
  XCHK(x, hipSetDevice(dev));
  XCHK(x, hipStreamCreateWithFlags(&x->st, hipStreamNonBlocking));
  for (int k = 0; k < Executor::kSlots; k++) XCHK(x, hipEventCreateWithFlags(&x->ev[k], hipEventDisableTiming));
  return DQMC_OK;
}

// A chain joins: its queued work finishes, then its kernels go to the executor's stream.
int executor_attach(Executor* x, Chain* h) {
  if (h->own_st) { x->err = "chain already runs in an ensemble"; return DQMC_E_STATE; }
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->own_st = h->st;
  h->st = x->st;
  return DQMC_OK;
}

// A chain leaves: everything queued for it has run, and it has its own stream again.
int executor_detach(Executor* x, Chain* h) {
  if (!h->own_st) return DQMC_OK;
  XCHK(x, hipStreamSynchronize(x->st));
  h->st = h->own_st;
  h->own_st = nullptr;
  return DQMC_OK;
}

// Item tables of one UDT op for all chains of a group (each item with its own n): norms,
// per panel index the panel launches (one per rows-per-thread class) and the reflector
// launch, then D and T.
void pack_udt(Executor* x, std::vector<UdtItem>& items, int apply_pivot) {
  int max_n = 0;
  for (auto& it : items) max_n = std::max(max_n, it.n);
  x->launches.push_back({2, x->put(items.data(), items.size()), (int)items.size(), max_n, 0, 0, 0, 0, 0, 0});
  std::vector<std::vector<int>> st(items.size()), wd(items.size());
  int np = 0;
  for (size_t i = 0; i < items.size(); i++) {
    udt_plan(items[i].n, st[i], wd[i]);
    np = std::max(np, (int)st[i].size());
  }
  std::vector<UdtItem> tab;
  for (int p = 0; p < np; p++) {
    tab.clear();
    for (int R : {1, 2, 4, 8})
      for (size_t i = 0; i < items.size(); i++) {
        if (p >= (int)st[i].size()) continue;
        UdtItem u = items[i];
        u.p0 = st[i][p];
        u.w = wd[i][p];
        u.last = p == (int)st[i].size() - 1;
        if (udt_rows_class(u.n - u.p0) == R) tab.push_back(u);
      }
    const size_t off = x->put(tab.data(), tab.size());
    int grid = 0, i0 = 0;
    while (i0 < (int)tab.size()) {
      const int R = udt_rows_class(tab[i0].n - tab[i0].p0);
      int i1 = i0, mn = 0;
      while (i1 < (int)tab.size() && udt_rows_class(tab[i1].n - tab[i1].p0) == R) mn = std::max(mn, tab[i1++].n);
      x->launches.push_back({3, off + sizeof(UdtItem) * i0, i1 - i0, mn, R, 0, 0, 0, 0, 0});
      i0 = i1;
    }
    for (auto& u : tab) {
      const int c0 = u.p0 + u.w;
      grid = std::max(grid, (c0 < u.n ? (u.n - c0 + 15) / 16 : 0) + (u.n + 15) / 16);
    }
    x->launches.push_back({4, off, (int)tab.size(), 0, 0, grid, 0, 0, 0, 0});
  }
  x->launches.push_back({5, x->put(items.data(), items.size()), (int)items.size(), max_n, apply_pivot, 0, 0, 0, 0, 0});
}

int executor_launch(Executor* x, const Executor::Launch& L, const char* base) {
  const void* p = base + L.off;
  switch (L.kind) {
    case 0:
      XCHK(x, launch_window_multi_f64(static_cast<const WindowArgs<double>*>(p), L.count, L.a, L.b, L.c, x->st));
      break;
    case 1: {
      GemmBatch<double> g;
      g.count = L.count; g.transA = L.a; g.transB = L.b; g.alpha = L.alpha; g.beta = L.beta;
      g.items = static_cast<const GemmItem<double>*>(p);
      g.max_m = (int)L.m1; g.max_n = (int)L.m2;
      XCHK(x, launch_gemm_f64(g, x->st));
      break;
    }
    case 2: case 3: case 4: case 5: {
      UdtBatch u;
      u.count = L.count; u.n = 0; u.apply_pivot = L.kind == 5 ? L.b : 0;
      u.items = static_cast<const UdtItem*>(p);
      u.max_n = L.a;
      if (L.kind == 2) XCHK(x, launch_udt_norms(u, x->st));
      else if (L.kind == 3) XCHK(x, launch_udt_panel_tab(u, L.b, x->st));
      else if (L.kind == 4) XCHK(x, launch_udt_reflect_tab(u, L.c, x->st));
      else XCHK(x, launch_udt_finish(u, x->st));
      break;
    }
    case 6: {
      RdivpBatch r;
      r.count = L.count; r.n = 0;
      r.items = static_cast<const RdivpItem*>(p);
      r.max_n = L.a;
      XCHK(x, launch_rdivp(r, x->st));
      break;
    }
    case 7: XCHK(x, launch_elem_multi(static_cast<const ElemItem*>(p), L.count, L.m1, x->st)); break;
    case 8: XCHK(x, launch_maxabsdiff_multi(static_cast<const MaxdiffItem*>(p), L.count, L.m1, x->st)); break;
    default: x->err = "unknown launch kind"; return DQMC_E_STATE;
  }
  return DQMC_OK;
}

// One slice visit of every chain in `cs` (all attached, FP64, stack built).
int executor_step(Executor* x, Chain* const* cs, int nc) {
  x->stage.clear();
  x->launches.clear();
  for (int i = 0; i < nc; i++) {
    Chain* h = cs[i];
    h->ops.clear(); h->rgemm.clear(); h->rudt.clear(); h->rrdivp.clear(); h->relem.clear(); h->rmax.clear();
    h->rwin.clear();
    h->rec = true;
    int r = sweep_spatial(h);
    if (r == DQMC_OK) r = propagate(h);
    h->rec = false;
    if (r != DQMC_OK) { x->err = h->err; return r; }
  }
  // window launches: position k of every chain, grouped by window width (one and two
  // flavours share a launch)
  size_t kmax = 0;
  for (int i = 0; i < nc; i++) kmax = std::max(kmax, cs[i]->rwin.size());
  std::vector<WindowArgs<double>> wtab;
  for (size_t k = 0; k < kmax; k++)
    for (int W : {16, 32, 64}) {
      wtab.clear();
      int mn = 0, mf = 0;
      for (int i = 0; i < nc; i++)
        if (cs[i]->rwin.size() > k && cs[i]->win == W) {
          wtab.push_back(cs[i]->rwin[k]);
          mn = std::max(mn, cs[i]->n);
          mf = std::max(mf, cs[i]->flv);
        }
      if (!wtab.empty())
        x->launches.push_back({0, x->put(wtab.data(), wtab.size()), (int)wtab.size(), mf, W, mn, 0, 0, 0, 0});
    }
  // propagate: chains grouped by the kinds of their recorded sequences
  std::map<std::vector<int64_t>, std::vector<int>> groups;
  for (int i = 0; i < nc; i++) {
    std::vector<int64_t> sig;
    sig.reserve(cs[i]->ops.size() * 3);
    for (const RecOp& o : cs[i]->ops) {
      sig.push_back(o.kind * 8 + o.key);
      int64_t ab[2];
      std::memcpy(ab, &o.alpha, 8);
      std::memcpy(ab + 1, &o.beta, 8);
      sig.push_back(ab[0]);
      sig.push_back(ab[1]);
    }
    groups[sig].push_back(i);
  }
  std::vector<GemmItem<double>> gt;
  std::vector<UdtItem> ut;
  std::vector<RdivpItem> rt, rt2;
  std::vector<ElemItem> et;
  std::vector<MaxdiffItem> mt;
  for (auto& kv : groups) {
    const std::vector<int>& mem = kv.second;
    const size_t nops = cs[mem[0]]->ops.size();
    for (size_t q = 0; q < nops; q++) {
      const RecOp& o0 = cs[mem[0]]->ops[q];
      switch (o0.kind) {
        case kOpGemm: {
          gt.clear();
          int64_t mm = 0, mnn = 0;
          for (int i : mem) {
            const RecOp& o = cs[i]->ops[q];
            for (int k = 0; k < o.count; k++) {
              const GemmItem<double>& g = cs[i]->rgemm[o.first + k];
              gt.push_back(g);
              mm = std::max<int64_t>(mm, g.M);
              mnn = std::max<int64_t>(mnn, g.N);
            }
          }
          x->launches.push_back({1, x->put(gt.data(), gt.size()), (int)gt.size(), o0.key >> 1, o0.key & 1, 0,
                                 mm, mnn, o0.alpha, o0.beta});
          break;
        }
        case kOpUdt: {
          ut.clear();
          for (int i : mem) {
            const RecOp& o = cs[i]->ops[q];
            for (int k = 0; k < o.count; k++) ut.push_back(cs[i]->rudt[o.first + k]);
          }
          pack_udt(x, ut, o0.key);
          break;
        }
        case kOpRdivp: {
          rt.clear(); rt2.clear();
          int m1 = 0, m2 = 0;
          for (int i : mem) {
            const RecOp& o = cs[i]->ops[q];
            for (int k = 0; k < o.count; k++) {
              const RdivpItem& r = cs[i]->rrdivp[o.first + k];
              if (r.n <= kRdivpStripMaxN) { rt.push_back(r); m1 = std::max(m1, r.n); }
              else { rt2.push_back(r); m2 = std::max(m2, r.n); }
            }
          }
          if (!rt.empty()) x->launches.push_back({6, x->put(rt.data(), rt.size()), (int)rt.size(), m1, 0, 0, 0, 0, 0, 0});
          if (!rt2.empty()) x->launches.push_back({6, x->put(rt2.data(), rt2.size()), (int)rt2.size(), m2, 0, 0, 0, 0, 0, 0});
          break;
        }
        case kOpElem: {
          et.clear();
          int64_t mc = 0;
          for (int i : mem) {
            const ElemItem& e = cs[i]->relem[cs[i]->ops[q].first];
            et.push_back(e);
            mc = std::max<int64_t>(mc, e.count);
          }
          x->launches.push_back({7, x->put(et.data(), et.size()), (int)et.size(), 0, 0, 0, mc, 0, 0, 0});
          break;
        }
        case kOpMaxdiff: {
          mt.clear();
          int64_t mc = 0;
          for (int i : mem) {
            const MaxdiffItem& m = cs[i]->rmax[cs[i]->ops[q].first];
            mt.push_back(m);
            mc = std::max<int64_t>(mc, m.count);
          }
          x->launches.push_back({8, x->put(mt.data(), mt.size()), (int)mt.size(), 0, 0, 0, mc, 0, 0, 0});
          break;
        }
      }
    }
  }
  // the step's tables: pinned slot -> device slot, one copy; then the launches
  const int k = x->slot;
  x->slot = (x->slot + 1) % Executor::kSlots;
  if (x->pending[k]) XCHK(x, hipEventSynchronize(x->ev[k]));
  const size_t need = std::max<size_t>(x->stage.size(), 256);
  if (need > x->cap[k]) {
    if (x->host[k]) XCHK(x, hipHostFree(x->host[k]));
    if (x->devb[k]) XCHK(x, hipFree(x->devb[k]));
    x->host[k] = nullptr;
    x->devb[k] = nullptr;
    const size_t c = std::max<size_t>(need * 2, (size_t)1 << 20);
    XCHK(x, hipHostMalloc(reinterpret_cast<void**>(&x->host[k]), c, hipHostMallocDefault));
    XCHK(x, hipMalloc(reinterpret_cast<void**>(&x->devb[k]), c));
    x->cap[k] = c;
  }
  std::memcpy(x->host[k], x->stage.data(), x->stage.size());
  XCHK(x, hipMemcpyAsync(x->devb[k], x->host[k], x->stage.size(), hipMemcpyHostToDevice, x->st));
  for (const auto& L : x->launches) CHK(executor_launch(x, L, x->devb[k]));
  XCHK(x, hipEventRecord(x->ev[k], x->st));
  x->pending[k] = true;
  for (int i = 0; i < nc; i++) {
    Chain* h = cs[i];
    h->steps++;
    if (h->steps % (2 * h->M) == 0) h->sweeps++;
  }
  return DQMC_OK;
}

// Runs `nsweeps` sweeps of every chain in lockstep slice visits on one stream: each chain
// makes exactly its own 2 M nsweeps visits (chains with fewer slices finish earlier).
int run_lockstep(Chain* const* hs, int count, int nsweeps, int64_t* accepted, std::string& err) {
  Executor x;
  int r = executor_init(&x, hs[0]->dev);
  std::vector<long long> a0(count, 0), target(count, 0), start(count, 0);
  std::vector<Chain*> att;
  for (int i = 0; i < count && r == DQMC_OK; i++) {
    r = executor_attach(&x, hs[i]);
    if (r == DQMC_OK) {
      att.push_back(hs[i]);
      r = read_accepted(hs[i], &a0[i]);
      start[i] = hs[i]->steps;
      target[i] = start[i] + 2LL * hs[i]->M * nsweeps;
    }
  }
  std::vector<Chain*> act;
  while (r == DQMC_OK) {
    act.clear();
    for (int i = 0; i < count; i++)
      if (hs[i]->steps < target[i]) act.push_back(hs[i]);
    if (act.empty()) break;
    // leave room in every chain's error slots for this step's stabilisation
    for (Chain* h : act)
      if (h->nerr + 2 >= h->errcap && r == DQMC_OK) r = harvest(h);
    if (r == DQMC_OK) r = executor_step(&x, act.data(), (int)act.size());
  }
  if (r != DQMC_OK && err.empty()) err = x.err;
  int r2 = DQMC_OK;
  for (Chain* h : att) {
    int q = executor_detach(&x, h);
    if (q == DQMC_OK) q = harvest(h);
    if (q != DQMC_OK && r2 == DQMC_OK) { r2 = q; if (err.empty()) err = h->err; }
  }
  if (r == DQMC_OK && accepted)
    for (int i = 0; i < count && r == DQMC_OK; i++) {
      long long a1 = 0;
      r = read_accepted(hs[i], &a1);
      accepted[i] = a1 - a0[i];
    }
  return r != DQMC_OK ? r : r2;
}

}  // namespace

// Internal interface of the executor for the serving loop (ensemble.h).
struct dqmc::Ensemble {
  Executor x;
};

int dqmc::ensemble_create(int dev, Ensemble** out) {
  auto* e = new Ensemble();
  const int r = executor_init(&e->x, dev);
  if (r != DQMC_OK) { delete e; return r; }
  *out = e;
  return DQMC_OK;
}
void dqmc::ensemble_destroy(Ensemble* e) { delete e; }
const char* dqmc::ensemble_error(Ensemble* e) { return e->x.err.c_str(); }
int dqmc::ensemble_attach(Ensemble* e, dqmc_chain_t h) {
  if (h->prec != DQMC_PRECISION_FP64 || h->dev != e->x.dev || h->cs < 1 || h->cs > h->M) {
    e->x.err = "only built FP64 chains of the ensemble's device can join it";
    return DQMC_E_STATE;
  }
  if (hipSetDevice(h->dev) != hipSuccess) return DQMC_E_HIP;
  return executor_attach(&e->x, h);
}
int dqmc::ensemble_detach(Ensemble* e, dqmc_chain_t h) { return executor_detach(&e->x, h); }
int dqmc::ensemble_step(Ensemble* e, dqmc_chain_t* hs, int count) {
  for (int i = 0; i < count; i++)
    if (hs[i]->nerr + 2 >= hs[i]->errcap) {
      const int r = harvest(hs[i]);
      if (r != DQMC_OK) { e->x.err = hs[i]->err; return r; }
    }
  return executor_step(&e->x, hs, count);
}
long long dqmc::chain_steps(dqmc_chain_t h) { return h->steps; }
int dqmc::chain_slices(dqmc_chain_t h) { return h->M; }

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

const char* dqmc_last_error(dqmc_chain_t h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int dqmc_footprint_bytes(const dqmc_config_t* cfg, int64_t* bytes) {
  if (!cfg || !bytes || cfg->n_sites < 1 || cfg->slices < 1 || cfg->safe_mult < 1) return DQMC_E_INVALID;
  *bytes = footprint(cfg, nullptr);
  return DQMC_OK;
}

int dqmc_create(const dqmc_config_t* cfg, dqmc_chain_t* out) {
  if (!cfg || !out) { g_create_error = "dqmc_create: null argument"; return DQMC_E_INVALID; }
  *out = nullptr;
  if (cfg->n_sites < 1 || cfg->n_sites > 2048 || (cfg->flavors != 1 && cfg->flavors != 2) ||
      cfg->slices < 1 || cfg->safe_mult < 1 || !(cfg->delta_tau > 0) ||
      (cfg->precision != DQMC_PRECISION_FP64 && cfg->precision != DQMC_PRECISION_MIXED) ||
      !(cfg->mixed_tolerance >= 0)) {
    g_create_error = "dqmc_create: invalid configuration (1 <= n_sites <= 2048, flavors 1|2, "
                     "slices >= 1, safe_mult >= 1, delta_tau > 0, precision 0|1, "
                     "mixed_tolerance >= 0)";
    return DQMC_E_INVALID;
  }
  hipError_t e = hipSetDevice(cfg->device);
  if (e != hipSuccess) { g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e); return DQMC_E_HIP; }
  std::call_once(g_init_once, init_kernels);
  if (g_init_status != hipSuccess) {
    g_create_error = std::string("kernel attribute setup: ") + hipGetErrorString(g_init_status);
    return DQMC_E_HIP;
  }
  Chain* h = new Chain();
  h->dev = cfg->device;
  h->n = cfg->n_sites; h->flv = cfg->flavors; h->M = cfg->slices; h->sm = cfg->safe_mult;
  h->dtau = cfg->delta_tau; h->alpha = cfg->alpha; h->prec = cfg->precision;
  h->check = cfg->check_propagation_error ? 1 : 0;
  if (cfg->mixed_tolerance > 0) h->mixed_tol = cfg->mixed_tolerance;
  h->K = generate_chunks(h->M, h->sm, h->rs, h->re);
  h->nn = (size_t)h->n * h->n;
  const double a = h->alpha;
  h->ev_pos = std::exp(a);
  h->ev_neg = std::exp(-a);
  // Density / first block: dE = -2 a x, Delta = exp(dE) - 1           (fields.jl:389-390)
  h->d_first[0] = std::exp(-2.0 * a * -1.0) - 1.0;
  h->d_first[1] = std::exp(-2.0 * a * 1.0) - 1.0;
  // Magnetic second block: Delta = exp(-dE) - 1                          (fields.jl:441-443)
  h->d_second[0] = std::exp(-(-2.0 * a * -1.0)) - 1.0;
  h->d_second[1] = std::exp(-(-2.0 * a * 1.0)) - 1.0;
  // Density boson weight exp(-dE)                                        (local_updates.jl:31)
  h->bweight[0] = std::exp(-(-2.0 * a * -1.0));
  h->bweight[1] = std::exp(-(-2.0 * a * 1.0));

  h->win = window_size(h->n, h->flv);
  if (cfg->delay_depth > 0) h->win = cfg->delay_depth <= 16 ? 16 : cfg->delay_depth <= 32 ? 32 : 64;

  const size_t f = h->flv, nn = h->nn, n = h->n;
  int r = DQMC_OK;
  auto A = [&](auto** p, size_t count) { if (r == DQMC_OK) r = dalloc(h, p, count); };
  r = arena_alloc(h, [&]() {
  r = DQMC_OK;
  A(&h->eTmh, nn); A(&h->eTph, nn); A(&h->eT2m, nn); A(&h->eT2p, nn); A(&h->dT, nn);
  A(&h->G, f * nn); A(&h->Gx, f * nn); A(&h->Gchk, f * nn);
  A(&h->us, (size_t)(h->K + 1) * f * nn); A(&h->ts, (size_t)(h->K + 1) * f * nn);
  A(&h->ds, (size_t)(h->K + 1) * f * n);
  A(&h->Ul, f * nn); A(&h->Ur, f * nn); A(&h->Tl, f * nn); A(&h->Tr, f * nn);
  A(&h->Dl, f * n); A(&h->Dr, f * n);
  A(&h->tmp1, f * nn); A(&h->tmp2, f * nn); A(&h->cur, f * nn); A(&h->O, f * nn);
  A(&h->piv, f * n); A(&h->tau, f * n); A(&h->tfac, f * (n + 16) * 16); A(&h->swp, f * 32 * n);
  A(&h->conf, n * (size_t)h->M);
  A(&h->cpack, (n * (size_t)h->M + 63) / 64);
  if (h->prec == DQMC_PRECISION_MIXED) {
    A(&h->G32, f * nn); A(&h->Gx32, f * nn); A(&h->Gchk32, f * nn); A(&h->G32pp, f * nn);
    A(&h->eT2m32, nn); A(&h->eT2p32, nn);
  } else {
    A(&h->Gpp, f * nn);
  }
  A(&h->wY, 2 * f * h->win * h->win); A(&h->wS, 2 * ((size_t)h->win + 1));
  A(&h->ctr, 1);
  A(&h->errbuf, (size_t)h->errcap);
  A(&h->scal, 16);
  A(&h->wsync, 1 + 2 * (size_t)((n + 15) / 16));
  return r;
  });
  if (r != DQMC_OK) {
    g_create_error = h->err;
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
    return r;
  }
  e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
    return DQMC_E_HIP;
  }
  ChainCounters z;
  std::memset(&z, 0, sizeof(z));
  e = hipMemcpy(h->ctr, &z, sizeof(z), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(h->wsync, 0, sizeof(unsigned long long) * (1 + 2 * (size_t)((n + 15) / 16)));
  std::vector<int8_t> plus(n * (size_t)h->M, 1);
  if (e == hipSuccess) e = hipMemcpy(h->conf, plus.data(), plus.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    g_create_error = std::string("hipMemcpy: ") + hipGetErrorString(e);
    (void)hipStreamDestroy(h->st);
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
    return DQMC_E_HIP;
  }
  // pinned mirrors for harvest (without them harvest copies synchronously)
  void* hp = nullptr;
  if (hipHostMalloc(&hp, sizeof(ChainCounters) + sizeof(double) * (size_t)h->errcap, hipHostMallocDefault) ==
      hipSuccess) {
    h->hctr = static_cast<ChainCounters*>(hp);
    h->herr = reinterpret_cast<double*>(static_cast<char*>(hp) + sizeof(ChainCounters));
  }
  *out = h;
  return DQMC_OK;
}

int dqmc_destroy(dqmc_chain_t h) {
  if (!h) return DQMC_OK;
  (void)hipSetDevice(h->dev);
  if (h->st) (void)hipStreamSynchronize(h->st);
  collect_times(h);
  for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
  if (h->sgx) (void)hipGraphExecDestroy(h->sgx);
  if (h->sg) (void)hipGraphDestroy(h->sg);
  if (h->hctr) (void)hipHostFree(h->hctr);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->unif) (void)hipFree(h->unif);
  if (h->dlog) (void)hipFree(h->dlog);
  if (h->plog) (void)hipFree(h->plog);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return DQMC_OK;
}

#define ENTER(h)                                   \
  if (!(h)) return DQMC_E_INVALID;                 \
  do {                                             \
    hipError_t e__ = hipSetDevice((h)->dev);       \
    if (e__ != hipSuccess) {                       \
      (h)->err = hipGetErrorString(e__);           \
      return DQMC_E_HIP;                           \
    }                                              \
  } while (0)

static int upload_matrix(Chain* h, double* dst, const double* src, int ld) {
  HIPCHK(h, hipMemcpy2DAsync(dst, sizeof(double) * h->n, src, sizeof(double) * ld, sizeof(double) * h->n,
                             h->n, hipMemcpyHostToDevice, h->st));
  return DQMC_OK;
}
static int download_matrix(Chain* h, double* dst, int ld, const double* src, int cols) {
  HIPCHK(h, hipMemcpy2DAsync(dst, sizeof(double) * ld, src, sizeof(double) * h->n, sizeof(double) * h->n,
                             cols, hipMemcpyDeviceToHost, h->st));
  return DQMC_OK;
}

int dqmc_set_hopping(dqmc_chain_t h, const double* T, int ld) {
  ENTER(h);
  if (!T || ld < h->n) { h->err = "dqmc_set_hopping: null T or ld < n"; return DQMC_E_INVALID; }
  const int n = h->n;
  double norm1 = 0.0;
  for (int c = 0; c < n; c++) {
    double s = 0.0;
    for (int r = 0; r < n; r++) s += std::fabs(T[(size_t)r + (size_t)c * ld]);
    norm1 = std::max(norm1, s);
  }
  CHK(upload_matrix(h, h->dT, T, ld));
  h->t_set = true;
  CHK(expm_device(h, -0.5 * h->dtau, norm1, h->eTmh));
  CHK(expm_device(h, +0.5 * h->dtau, norm1, h->eTph));
  CHK(expm_device(h, -h->dtau, norm1, h->eT2m));
  CHK(expm_device(h, +h->dtau, norm1, h->eT2p));
  return finish_hopping(h);
}

int dqmc_set_hopping_exp(dqmc_chain_t h, const double* e_mh, const double* e_ph, const double* e_m,
                         const double* e_p, int ld) {
  ENTER(h);
  if (!e_mh || !e_ph || !e_m || !e_p || ld < h->n) { h->err = "dqmc_set_hopping_exp: bad arguments"; return DQMC_E_INVALID; }
  CHK(upload_matrix(h, h->eTmh, e_mh, ld));
  CHK(upload_matrix(h, h->eTph, e_ph, ld));
  CHK(upload_matrix(h, h->eT2m, e_m, ld));
  CHK(upload_matrix(h, h->eT2p, e_p, ld));
  return finish_hopping(h);
}

int dqmc_get_hopping_exp(dqmc_chain_t h, double* e_mh, double* e_ph, double* e_m, double* e_p, int ld) {
  ENTER(h);
  if (ld < h->n) { h->err = "ld < n"; return DQMC_E_INVALID; }
  if (e_mh) CHK(download_matrix(h, e_mh, ld, h->eTmh, h->n));
  if (e_ph) CHK(download_matrix(h, e_ph, ld, h->eTph, h->n));
  if (e_m) CHK(download_matrix(h, e_m, ld, h->eT2m, h->n));
  if (e_p) CHK(download_matrix(h, e_p, ld, h->eT2p, h->n));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

int dqmc_set_field(dqmc_chain_t h, const int8_t* conf, int ld) {
  ENTER(h);
  if (!conf || ld < h->n) { h->err = "dqmc_set_field: null conf or ld < n"; return DQMC_E_INVALID; }
  for (int l = 0; l < h->M; l++)
    for (int i = 0; i < h->n; i++) {
      const int8_t x = conf[(size_t)i + (size_t)l * ld];
      if (x != 1 && x != -1) { h->err = "dqmc_set_field: entries must be +-1"; return DQMC_E_INVALID; }
    }
  h->fver++;
  HIPCHK(h, hipMemcpy2DAsync(h->conf, h->n, conf, ld, h->n, h->M, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

int dqmc_get_field(dqmc_chain_t h, int8_t* conf, int ld) {
  ENTER(h);
  if (!conf || ld < h->n) { h->err = "dqmc_get_field: null conf or ld < n"; return DQMC_E_INVALID; }
  HIPCHK(h, hipMemcpy2DAsync(conf, ld, h->conf, h->n, h->n, h->M, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

int dqmc_get_field_packed(dqmc_chain_t h, uint64_t* chunks, int64_t nchunks) {
  ENTER(h);
  const int64_t count = (int64_t)h->n * h->M, need = (count + 63) / 64;
  if (!chunks || nchunks != need) {
    h->err = "dqmc_get_field_packed: null chunks or nchunks != ceil(n_sites * slices / 64)";
    return DQMC_E_INVALID;
  }
  HIPCHK(h, launch_pack_field(h->conf, count, h->cpack, h->st));
  HIPCHK(h, hipMemcpyAsync(chunks, h->cpack, sizeof(uint64_t) * need, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

int dqmc_set_field_packed(dqmc_chain_t h, const uint64_t* chunks, int64_t nchunks) {
  ENTER(h);
  const int64_t count = (int64_t)h->n * h->M, need = (count + 63) / 64;
  if (!chunks || nchunks != need) {
    h->err = "dqmc_set_field_packed: null chunks or nchunks != ceil(n_sites * slices / 64)";
    return DQMC_E_INVALID;
  }
  const int tail = (int)(count % 64);
  if (tail && (chunks[need - 1] >> tail) != 0) {
    h->err = "dqmc_set_field_packed: bits set beyond the field in the last chunk";
    return DQMC_E_INVALID;
  }
  h->fver++;
  HIPCHK(h, hipMemcpyAsync(h->cpack, chunks, sizeof(uint64_t) * need, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, launch_unpack_field(h->cpack, count, h->conf, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

static int set_upos(Chain* h, long long pos) {
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(&h->ctr->u_pos, &pos, sizeof(pos), hipMemcpyHostToDevice));
  return DQMC_OK;
}

int dqmc_set_uniforms(dqmc_chain_t h, const double* u, int64_t count) {
  ENTER(h);
  if (count < 0 || (count > 0 && !u)) { h->err = "dqmc_set_uniforms: bad arguments"; return DQMC_E_INVALID; }
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (h->unif) { HIPCHK(h, hipFree(h->unif)); h->unif = nullptr; }
  HIPCHK(h, hipMalloc(&h->unif, sizeof(double) * std::max<int64_t>(count, 1)));
  if (count > 0) HIPCHK(h, hipMemcpy(h->unif, u, sizeof(double) * count, hipMemcpyHostToDevice));
  h->nunif = count;
  return set_upos(h, 0);
}

int dqmc_set_seed(dqmc_chain_t h, uint64_t seed, uint64_t position) {
  ENTER(h);
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (h->unif) { HIPCHK(h, hipFree(h->unif)); h->unif = nullptr; }
  h->nunif = 0;
  h->seed = seed;
  return set_upos(h, (long long)position);
}

int dqmc_set_decision_log(dqmc_chain_t h, int64_t capacity) {
  ENTER(h);
  if (capacity < 0) return DQMC_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (h->dlog) { HIPCHK(h, hipFree(h->dlog)); h->dlog = nullptr; }
  if (h->plog) { HIPCHK(h, hipFree(h->plog)); h->plog = nullptr; }
  h->logcap = capacity;
  if (capacity > 0) {
    HIPCHK(h, hipMalloc(&h->dlog, capacity));
    HIPCHK(h, hipMalloc(&h->plog, sizeof(double) * capacity));
  }
  long long zero = 0;
  HIPCHK(h, hipMemcpy(&h->ctr->log_len, &zero, sizeof(zero), hipMemcpyHostToDevice));
  return DQMC_OK;
}

int dqmc_get_decision_log(dqmc_chain_t h, uint8_t* accepted, double* p, int64_t capacity, int64_t* length) {
  ENTER(h);
  HIPCHK(h, hipStreamSynchronize(h->st));
  ChainCounters c;
  HIPCHK(h, hipMemcpy(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost));
  int64_t len = std::min<int64_t>(c.log_len, h->logcap);
  if (length) *length = len;
  len = std::min<int64_t>(len, capacity);
  if (len > 0 && accepted) HIPCHK(h, hipMemcpy(accepted, h->dlog, len, hipMemcpyDeviceToHost));
  if (len > 0 && p) HIPCHK(h, hipMemcpy(p, h->plog, sizeof(double) * len, hipMemcpyDeviceToHost));
  return DQMC_OK;
}

#define NEED_HOPPING(h)                                                     \
  if (!(h)->hopping_set) {                                                  \
    (h)->err = "hopping exponentials not set (dqmc_set_hopping first)";     \
    return DQMC_E_STATE;                                                    \
  }

int dqmc_build_stack(dqmc_chain_t h) {
  ENTER(h);
  NEED_HOPPING(h);
  CHK(reverse_build(h));
  CHK(propagate(h));
  return harvest(h);
}

int dqmc_build_stack_forward(dqmc_chain_t h) {
  ENTER(h);
  NEED_HOPPING(h);
  CHK(forward_build(h));
  CHK(propagate(h));
  return harvest(h);
}

int dqmc_reverse_build_stack(dqmc_chain_t h) {
  ENTER(h);
  NEED_HOPPING(h);
  CHK(reverse_build(h));
  return harvest(h);
}

int dqmc_propagate(dqmc_chain_t h) {
  ENTER(h);
  NEED_HOPPING(h);
  CHK(propagate(h));
  return harvest(h);
}


int dqmc_sweep_spatial(dqmc_chain_t h, int64_t* accepted) {
  ENTER(h);
  NEED_HOPPING(h);
  long long a0 = 0, a1 = 0;
  if (accepted) { HIPCHK(h, hipStreamSynchronize(h->st)); CHK(read_accepted(h, &a0)); }
  CHK(sweep_spatial(h));
  CHK(harvest(h));
  if (accepted) { CHK(read_accepted(h, &a1)); *accepted = a1 - a0; }
  return DQMC_OK;
}

static int sweep_eager(Chain* h) {
  for (int k = 0; k < 2 * h->M; k++) {
    CHK(sweep_spatial(h));
    CHK(propagate(h));
  }
  return DQMC_OK;
}

static void drop_sweep_graph(Chain* h) {
  if (h->sgx) (void)hipGraphExecDestroy(h->sgx);
  if (h->sg) (void)hipGraphDestroy(h->sg);
  h->sgx = nullptr;
  h->sg = nullptr;
  h->sg_key.clear();
}

// Everything a sweep's launches take from the host: the propagate state, which copy of G is
// which, the error slots' position, the randomness and log buffers, the wait bound.
static std::vector<uint64_t> sweep_key(const Chain* h) {
  auto u = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  return {(uint64_t)h->cs, (uint64_t)(int64_t)h->dir, (uint64_t)h->cr, u(h->G), u(h->Gpp), (uint64_t)h->nerr,
          (uint64_t)h->seed, u(h->unif), (uint64_t)h->nunif, u(h->dlog), u(h->plog), (uint64_t)h->logcap,
          (uint64_t)h->wait_limit, (uint64_t)h->win};
}

// One sweep: replayed from the captured graph when its key matches, captured (then launched)
// when it does not, and launched kernel by kernel when graphs are off or capture fails.
static int sweep_once(Chain* h) {
  const bool graph = h->use_graph && h->prec == DQMC_PRECISION_FP64 && !h->prof && !h->rec &&
                     h->nerr + 4 * (long long)h->K + 8 < h->errcap;
  if (!graph) return sweep_eager(h);
  std::vector<uint64_t> key = sweep_key(h);
  if (h->sgx && key == h->sg_key) {
    HIPCHK(h, hipGraphLaunch(h->sgx, h->st));
    h->fver++;
    h->acc_valid = false;
    h->cs = h->sg_cs; h->dir = h->sg_dir; h->cr = h->sg_cr;
    h->G = h->sg_G; h->Gpp = h->sg_Gpp;
    h->nerr = h->sg_nerr;
    return DQMC_OK;
  }
  drop_sweep_graph(h);
  const int cs = h->cs, dir = h->dir, cr = h->cr;
  double* const G = h->G;
  double* const Gpp = h->Gpp;
  const long long nerr = h->nerr;
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  int r = DQMC_OK;
  hipError_t e = hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    r = sweep_eager(h);
    e = hipStreamEndCapture(h->st, &g);
    if (r == DQMC_OK && e == hipSuccess && g) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  }
  if (r != DQMC_OK || e != hipSuccess || !g || !x) {
    // capture is not available: this chain launches kernel by kernel from now on
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    h->use_graph = false;
    h->cs = cs; h->dir = dir; h->cr = cr; h->G = G; h->Gpp = Gpp; h->nerr = nerr;
    return sweep_eager(h);
  }
  h->sg = g;
  h->sgx = x;
  h->sg_key = key;
  h->sg_cs = h->cs; h->sg_dir = h->dir; h->sg_cr = h->cr;
  h->sg_G = h->G; h->sg_Gpp = h->Gpp;
  h->sg_nerr = h->nerr;
  HIPCHK(h, hipGraphLaunch(x, h->st));
  return DQMC_OK;
}

int dqmc_set_sweep_graph(dqmc_chain_t h, int enable) {
  if (!h) return DQMC_E_INVALID;
  h->use_graph = enable != 0;
  if (!h->use_graph) drop_sweep_graph(h);
  return DQMC_OK;
}

int dqmc_set_slice_launch(dqmc_chain_t h, int enable) {
  if (!h) return DQMC_E_INVALID;
  if (h->win_slice != (enable != 0)) drop_sweep_graph(h);
  h->win_slice = enable != 0;
  return DQMC_OK;
}

int dqmc_sweep(dqmc_chain_t h, int nsweeps, int64_t* accepted) {
  ENTER(h);
  NEED_HOPPING(h);
  if (nsweeps < 0) { h->err = "nsweeps < 0"; return DQMC_E_INVALID; }
  if (h->cs < 1 || h->cs > h->M) { h->err = "dqmc_sweep: build the stack first"; return DQMC_E_STATE; }
  long long a0 = 0, a1 = 0;
  if (accepted) {
    if (h->acc_valid) {
      a0 = h->acc_last;
    } else {
      HIPCHK(h, hipStreamSynchronize(h->st));
      CHK(read_accepted(h, &a0));
    }
  }
  for (int s = 0; s < nsweeps; s++) {
    CHK(sweep_once(h));
    h->sweeps++;
    // a captured sweep writes fixed error slots: collect them before the next one
    if (h->sgx && s + 1 < nsweeps) CHK(harvest(h));
  }
  CHK(harvest(h));
  if (accepted) {
    if (h->acc_valid) a1 = h->acc_last;
    else CHK(read_accepted(h, &a1));
    *accepted = a1 - a0;
  }
  return DQMC_OK;
}

int dqmc_batch_sweep(dqmc_chain_t* hs, int count, int nsweeps, int64_t* accepted) {
  if (!hs || count < 1 || nsweeps < 0) return DQMC_E_INVALID;
  for (int i = 0; i < count; i++)
    for (int j = 0; j < i; j++)
      if (!hs[i] || hs[i] == hs[j]) return DQMC_E_INVALID;   // a chain appears once
  bool lockstep = true;
  for (int i = 0; i < count; i++) {
    Chain* h = hs[i];
    if (!h->hopping_set || h->cs < 1 || h->cs > h->M) {
      h->err = "dqmc_batch_sweep: set the hopping matrix and build the stack first";
      return DQMC_E_STATE;
    }
    if (h->prec != DQMC_PRECISION_FP64 || h->dev != hs[0]->dev || h->prof) lockstep = false;
  }
  if (lockstep) {
    // FP64 chains on one device: one launch per kernel kind covers all of them
    if (hipSetDevice(hs[0]->dev) != hipSuccess) return DQMC_E_HIP;
    std::string err;
    const int r = run_lockstep(hs, count, nsweeps, accepted, err);
    if (r != DQMC_OK)
      for (int i = 0; i < count; i++) hs[i]->err = "dqmc_batch_sweep: " + err;
    return r;
  }
  // otherwise (mixed precision, several devices, kernel timing on) one host thread and
  // stream per chain
  std::vector<int> rc(count, DQMC_OK);
  std::vector<std::thread> th;
  th.reserve(count);
  for (int i = 0; i < count; i++)
    th.emplace_back([&, i]() { rc[i] = dqmc_sweep(hs[i], nsweeps, accepted ? &accepted[i] : nullptr); });
  for (auto& t : th) t.join();
  for (int i = 0; i < count; i++)
    if (rc[i] != DQMC_OK) return rc[i];
  return DQMC_OK;
}

// Effective (physical = 0) or physical G on the device: greens.jl:114-125,
// G_phys = e^{+dtau T/2} G e^{-dtau T/2}.  The physical form is written to O (Gx is scratch).
int device_greens(Chain* h, int physical, const double** out) {
  const double* src = h->G;
  if (h->prec == DQMC_PRECISION_MIXED) {
    HIPCHK(h, launch_f32_to_f64(h->Gchk, h->G32, (int64_t)h->flv * h->nn, h->st));
    src = h->Gchk;
  }
  if (physical) {
    auto g1 = batch<double>(0, 1);
    auto g2 = batch<double>(1, 0);
    for (int b = 0; b < h->flv; b++) {
      g1.it[g1.count++] = item<double>(src + (size_t)b * h->nn, h->eTmh, blk(h->Gx, h, b), h->n);
      g2.it[g2.count++] = item<double>(h->eTph, blk(h->Gx, h, b), blk(h->O, h, b), h->n);
    }
    CHK(run(h, g1));
    CHK(run(h, g2));
    src = h->O;
  }
  *out = src;
  return DQMC_OK;
}

int dqmc_get_greens(dqmc_chain_t h, double* G, int ld, int physical) {
  ENTER(h);
  if (!G || ld < h->n) { h->err = "dqmc_get_greens: null G or ld < n"; return DQMC_E_INVALID; }
  const double* src = nullptr;
  CHK(device_greens(h, physical, &src));
  CHK(download_matrix(h, G, ld, src, h->flv * h->n));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

int dqmc_measure_equal_time(dqmc_chain_t h, double U, double* out, int nout) {
  ENTER(h);
  NEED_HOPPING(h);
  const int need = 5 + (h->flv + 1) * h->n;
  if (!out || nout < need) { h->err = "dqmc_measure_equal_time: out needs 5 + (flavors + 1) n values"; return DQMC_E_INVALID; }
  if (!h->t_set) { h->err = "dqmc_measure_equal_time: needs T (dqmc_set_hopping), not only its exponentials"; return DQMC_E_STATE; }
  const double* G = nullptr;
  CHK(device_greens(h, 1, &G));
  double* dout = h->Gx;   // free after the physical transform
  HIPCHK(h, launch_measure(G, h->dT, h->n, h->flv, U, dout, h->st));
  HIPCHK(h, hipMemcpyAsync(out, dout, sizeof(double) * need, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

int dqmc_set_greens(dqmc_chain_t h, const double* G, int ld) {
  ENTER(h);
  if (!G || ld < h->n) { h->err = "dqmc_set_greens: null G or ld < n"; return DQMC_E_INVALID; }
  HIPCHK(h, hipMemcpy2DAsync(h->G, sizeof(double) * h->n, G, sizeof(double) * ld, sizeof(double) * h->n,
                             (size_t)h->flv * h->n, hipMemcpyHostToDevice, h->st));
  CHK(to32(h));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

// calculate_greens(mc, slice, output), stack.jl:525-583
int dqmc_calculate_greens(dqmc_chain_t h, int slice, double* Gout, int ld) {
  ENTER(h);
  NEED_HOPPING(h);
  if (!Gout || ld < h->n || slice < 0 || slice > h->M) { h->err = "dqmc_calculate_greens: bad arguments"; return DQMC_E_INVALID; }
  CHK(greens_from_scratch(h, slice, h->Gchk));
  CHK(download_matrix(h, Gout, ld, h->Gchk, h->flv * h->n));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

// greens(mc, k, l) / calculate_greens(mc, k, l), unequal_time_stack.jl:297-338: G(k <- l) from
// the unequal-time stack of the current field (physical: eT/2+ G eT/2-, greens.jl:114-125).
int dqmc_unequal_time_greens(dqmc_chain_t h, int k, int l, double* G, int ld, int physical) {
  ENTER(h);
  NEED_HOPPING(h);
  if (!G || ld < h->n || k < 0 || k > h->M || l < 0 || l > h->M) {
    h->err = "dqmc_unequal_time_greens: null G, ld < n or k, l outside 0..slices";
    return DQMC_E_INVALID;
  }
  CHK(ut_greens(h, k, l));
  const double* src = h->uG;
  if (physical) {
    CHK(physical_form(h, h->uG, h->tmp1));
    src = h->tmp1;
  }
  CHK(download_matrix(h, G, ld, src, h->flv * h->n));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return check_device_errors(h);
}

// CombinedGreensIterator / TimeIntegral iterated over l = start..stop
// (measurements/greens_iterators.jl:83-407): out holds 3 (stop - start + 1) matrices of
// n x (flavors n), G(0, l), G(l, 0), G(l, l) for each l in turn (physical forms).
int dqmc_combined_greens(dqmc_chain_t h, int start, int stop, int recalculate, double* out, int ld,
                         int out_on_device) {
  ENTER(h);
  NEED_HOPPING(h);
  if (!out || start < 0 || stop > h->M || start > stop || recalculate < 1 ||
      (out_on_device ? ld != h->n : ld < h->n)) {
    h->err = "dqmc_combined_greens: need 0 <= start <= stop <= slices, recalculate >= 1, out != NULL, "
             "ld >= n (ld == n for device output)";
    return DQMC_E_INVALID;
  }
  const size_t per = (size_t)ld * h->flv * h->n;
  auto emit = [&](int i, int j, const double* src) -> int {
    double* dst = out + ((size_t)i * 3 + j) * per;
    if (out_on_device) return physical_form(h, src, dst);
    CHK(physical_form(h, src, h->O));
    return download_matrix(h, dst, ld, h->O, h->flv * h->n);
  };
  CHK(combined_greens(h, start, stop, recalculate, emit));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return check_device_errors(h);
}

// charge_density_susceptibility / spin_density_susceptibility(mc, model, dir) measured once
// for the current field (constructors/charge_density.jl:47, spin_density.jl:54;
// apply!(::TimeIntegral), measurements/generic.jl:337-370; EachSitePairByDistance, :434-460;
// finish!, :574-586): the TimeIntegral iteration l = 0..M feeds one accumulation launch per
// slice, and G never leaves the device.
int dqmc_density_susceptibility(dqmc_chain_t h, int observable, int recalculate, const int32_t* dir2srctrg,
                                int n_bravais, int n_basis, int total_flavors, int reduced, double* out) {
  ENTER(h);
  NEED_HOPPING(h);
  if (!dir2srctrg || !out || n_bravais < 1 || n_basis < 1 || (int64_t)n_bravais * n_basis != h->n ||
      recalculate < 1 || total_flavors < 1 || observable < DQMC_CHARGE_DENSITY || observable > DQMC_SPIN_DENSITY_Z) {
    h->err = "dqmc_density_susceptibility: need a known observable, n_bravais * n_basis == sites, "
             "recalculate >= 1, total_flavors >= 1 and non-null tables";
    return DQMC_E_INVALID;
  }
  const int N = n_bravais, Bn = n_basis;
  for (int64_t k = 0; k < (int64_t)N * N; k++)
    if (dir2srctrg[k] < 0 || dir2srctrg[k] >= N) {
      h->err = "dqmc_density_susceptibility: dir2srctrg entries must lie in 0..n_bravais-1";
      return DQMC_E_INVALID;
    }
  CHK(ut_validate(h));
  const size_t mats = (size_t)h->flv * h->nn, nout = (size_t)N * Bn * Bn;
  double* G00 = h->itm + 10 * mats;
  double* ring[3] = {h->itm + 11 * mats, h->itm + 12 * mats, h->itm + 13 * mats};
  void* dmem = nullptr;
  HIPCHK(h, hipMalloc(&dmem, sizeof(double) * nout + sizeof(int32_t) * (size_t)N * N));
  double* acc = static_cast<double*>(dmem);
  int* d2st = reinterpret_cast<int*>(acc + nout);
  int r = DQMC_OK;
  hipError_t e = hipMemcpyAsync(d2st, dir2srctrg, sizeof(int32_t) * (size_t)N * N, hipMemcpyHostToDevice, h->st);
  if (e == hipSuccess) e = launch_fill(acc, 0.0, (int64_t)nout, h->st);
  if (e != hipSuccess) { h->err = hipGetErrorString(e); r = DQMC_E_HIP; }
  if (r == DQMC_OK) {
    const double* src = nullptr;
    r = device_greens(h, 1, &src);
    if (r == DQMC_OK) r = copyd(h, G00, src, mats);
  }
  const int M = h->M;
  auto emit = [&](int i, int j, const double* g) -> int {
    CHK(physical_form(h, g, ring[j]));
    if (j != 2) return DQMC_OK;
    const double w = (i == 0 || i == M ? 0.5 : 1.0) * h->dtau;
    HIPCHK(h, launch_density_slice(G00, ring[0], ring[1], ring[2], d2st, N, Bn, h->n, h->flv,
                                   (double)total_flavors, observable, i == 0, reduced, w, acc, h->st));
    return DQMC_OK;
  };
  if (r == DQMC_OK) r = combined_greens(h, 0, M, recalculate, emit);
  std::vector<double> host(nout);
  if (r == DQMC_OK) {
    e = hipMemcpyAsync(host.data(), acc, sizeof(double) * nout, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e != hipSuccess) { h->err = hipGetErrorString(e); r = DQMC_E_HIP; }
  } else {
    (void)hipStreamSynchronize(h->st);
  }
  (void)hipFree(dmem);
  if (r != DQMC_OK) return r;
  for (size_t k = 0; k < nout; k++) out[k] = host[k] / (double)h->n;
  return check_device_errors(h);
}

// global_update(mc, model, field) with `proposal` as the new conf (global_updates.jl:191-207):
//   propose_global_from_conf (:136-169): the inverse D of the last calculate_greens of G
//   (here 1 / Dr) and inv_det at slice current_slice - 1 = 0 (:64-121), the chain from scratch
//   with the proposal, whose calculate_inv_greens_udt (:25-47) leaves the new D in Dr (the
//   steps of calculate_greens up to its last UDT); detratio = prod_i (1 / D_old[i]) D_new[i]
//   in the reference's order, squared for one unique flavour; dE from energy_boson
//   (fields.jl:395, 451); p = |exp(-dE) detratio|; accepted iff p > 1 || u < p, u from the
//   chain's uniform stream (consumed only when p <= 1);
//   accept_global! (:171-188): reverse_build_stack + propagate; otherwise the old conf (and
//   D) are restored and G is untouched.
int dqmc_global_update(dqmc_chain_t h, const int8_t* proposal, int ld, int* accepted, double* p_out) {
  ENTER(h);
  NEED_HOPPING(h);
  if (!proposal || ld < h->n) { h->err = "dqmc_global_update: null proposal or ld < n"; return DQMC_E_INVALID; }
  if (h->cs != 1 || h->dir != 1 || !h->dg_valid) {
    h->err = "dqmc_global_update: needs current_slice == 1, direction == 1 right after a sweep or a "
             "stack build (propose_global_from_conf, global_updates.jl:136-140)";
    return DQMC_E_STATE;
  }
  const int n = h->n, M = h->M;
  const size_t nd = (size_t)h->flv * n, nc = (size_t)n * M;
  std::vector<int8_t> cnew(nc), cold(nc);
  for (int l = 0; l < M; l++)
    for (int i = 0; i < n; i++) {
      const int8_t x = proposal[(size_t)i + (size_t)l * ld];
      if (x != 1 && x != -1) { h->err = "dqmc_global_update: entries must be +-1"; return DQMC_E_INVALID; }
      cnew[(size_t)i + (size_t)l * n] = x;
    }
  std::vector<double> dold(nd), dnew(nd);
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(dold.data(), h->Dr, sizeof(double) * nd, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(cold.data(), h->conf, nc, hipMemcpyDeviceToHost));
  h->fver++;
  HIPCHK(h, hipMemcpy(h->conf, cnew.data(), nc, hipMemcpyHostToDevice));
  CHK(greens_from_scratch(h, 0, h->Gchk));
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(dnew.data(), h->Dr, sizeof(double) * nd, hipMemcpyDeviceToHost));
  double detratio = 1.0;
  for (size_t i = 0; i < nd; i++) detratio *= (1.0 / dold[i]) * dnew[i];
  if (h->flv == 1) detratio = detratio * detratio;
  double dE = 0.0;
  if (h->flv == 1) {
    long long sn = 0, so = 0;
    for (size_t i = 0; i < nc; i++) { sn += cnew[i]; so += cold[i]; }
    dE = h->alpha * (double)sn - h->alpha * (double)so;
  }
  const double p = std::fabs(std::exp(-dE) * detratio);
  bool acc = p > 1.0;
  if (!acc) {
    // one uniform of the chain's stream (the Metropolis test of local_updates.jl:53 as well)
    ChainCounters c;
    HIPCHK(h, hipMemcpy(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost));
    double u = 1.0;
    if (h->unif) {
      if (c.u_pos < h->nunif) HIPCHK(h, hipMemcpy(&u, h->unif + c.u_pos, sizeof(double), hipMemcpyDeviceToHost));
      else c.exhausted = 1;
    } else {
      u = counter_uniform_host(h->seed, (unsigned long long)c.u_pos);
    }
    if (std::fabs(u - p) < 1e-12) c.ties += 1;
    c.u_pos += 1;
    HIPCHK(h, hipMemcpy(h->ctr, &c, sizeof(c), hipMemcpyHostToDevice));
    acc = u < p;
  }
  if (acc) {
    CHK(reverse_build(h));
    CHK(propagate(h));
    CHK(harvest(h));
  } else {
    HIPCHK(h, hipMemcpy(h->conf, cold.data(), nc, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->Dr, dold.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
    h->dg_valid = true;
  }
  if (accepted) *accepted = acc ? 1 : 0;
  if (p_out) *p_out = p;
  return DQMC_OK;
}

int dqmc_get_state(dqmc_chain_t h, int* cs, int* dir, int* cr) {
  if (!h) return DQMC_E_INVALID;
  if (cs) *cs = h->cs;
  if (dir) *dir = h->dir;
  if (cr) *cr = h->cr;
  return DQMC_OK;
}

int dqmc_get_stats(dqmc_chain_t h, dqmc_stats_t* s) {
  ENTER(h);
  if (!s) return DQMC_E_INVALID;
  const int hr = harvest(h);
  // device errors and precision violations are reported in the stats themselves
  if (hr != DQMC_OK && hr != DQMC_E_DEVICE && hr != DQMC_E_PRECISION) return hr;
  ChainCounters c;
  HIPCHK(h, hipMemcpy(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost));
  s->accepted = c.accepted;
  s->proposed = c.proposed;
  s->uniforms_used = c.u_pos;
  s->negative_probability = c.negative;
  s->ties = c.ties;
  s->stream_exhausted = c.exhausted;
  s->propagation_error_count = h->perr_count;
  s->stabilizations = h->stabs;
  s->sweeps = h->sweeps;
  s->delayed_flushes = c.flushes;
  s->propagation_error_max = h->perr_max;
  s->propagation_error_last = h->perr_last;
  s->device_errors = c.device_errors;
  s->precision_violations = h->prec_viol;
  return DQMC_OK;
}

int dqmc_set_wait_limit(dqmc_chain_t h, int64_t polls) {
  ENTER(h);
  if (polls < 0) { h->err = "dqmc_set_wait_limit: polls < 0"; return DQMC_E_INVALID; }
  h->wait_limit = polls;
  return DQMC_OK;
}

int dqmc_get_stabilization_log(dqmc_chain_t h, double* out, int64_t capacity, int64_t* length) {
  ENTER(h);
  CHK(harvest(h));
  const int64_t len = (int64_t)h->stablog.size();
  if (length) *length = len;
  const int64_t k = std::min(len, capacity);
  if (out && k > 0) std::memcpy(out, h->stablog.data(), sizeof(double) * k);
  h->stablog.clear();
  return DQMC_OK;
}

int dqmc_set_kernel_timing(dqmc_chain_t h, int enable) {
  ENTER(h);
  CHK(harvest(h));
  h->prof = enable != 0;
  for (int k = 0; k < DQMC_KERNEL_KINDS; k++) { h->kms[k] = 0.0; h->kcount[k] = 0; }
  return DQMC_OK;
}

int dqmc_get_kernel_times(dqmc_chain_t h, double* total_ms, int64_t* launches, int kinds) {
  ENTER(h);
  CHK(harvest(h));
  for (int k = 0; k < kinds && k < DQMC_KERNEL_KINDS; k++) {
    if (total_ms) total_ms[k] = h->kms[k];
    if (launches) launches[k] = h->kcount[k];
  }
  return DQMC_OK;
}

int dqmc_synchronize(dqmc_chain_t h) {
  ENTER(h);
  HIPCHK(h, hipStreamSynchronize(h->st));
  return DQMC_OK;
}

// ---- linear algebra on host arrays (kernel-level parity) ------------------------------
namespace {
struct LaCtx {
  dqmc_chain_s c;
  bool ok = true;
};
int la_make(LaCtx& x, int n, int nbuf) {
  std::call_once(g_init_once, init_kernels);
  if (g_init_status != hipSuccess) return DQMC_E_HIP;
  Chain* h = &x.c;
  h->n = n; h->flv = 1; h->nn = (size_t)n * n;
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) return DQMC_E_HIP;
  int r = DQMC_OK;
  auto A = [&](auto** p, size_t count) { if (r == DQMC_OK) r = dalloc(h, p, count); };
  r = arena_alloc(h, [&]() {
    r = DQMC_OK;
    A(&h->tmp1, h->nn); A(&h->tmp2, h->nn); A(&h->cur, h->nn); A(&h->O, h->nn);
    A(&h->Gx, h->nn); A(&h->Dl, (size_t)n); A(&h->piv, (size_t)n); A(&h->tau, (size_t)n);
    A(&h->tfac, (size_t)(n + 16) * 16); A(&h->swp, (size_t)n * 32);
    A(&h->Ul, h->nn); A(&h->Ur, h->nn); A(&h->Tl, h->nn); A(&h->Tr, h->nn); A(&h->Dr, (size_t)n);
    A(&h->G, h->nn);
    A(&h->ctr, 1);
    return r;
  });
  (void)nbuf;
  if (r == DQMC_OK && hipMemsetAsync(h->ctr, 0, sizeof(ChainCounters), h->st) != hipSuccess) r = DQMC_E_HIP;
  return r;
}
void la_free(LaCtx& x) {
  Chain* h = &x.c;
  if (h->st) (void)hipStreamSynchronize(h->st);
  for (void* p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  if (h->st) (void)hipStreamDestroy(h->st);
  h->st = nullptr;
}
}  // namespace

int dqmc_la_udt(int n, const double* X, int apply_pivot, double* U, double* D, double* T, int32_t* pivot) {
  if (n < 1 || n > 2048 || !X || !U || !D || !T || !pivot) return DQMC_E_INVALID;
  LaCtx x;
  int r = la_make(x, n, 0);
  Chain* h = &x.c;
  if (r == DQMC_OK) {
    const size_t bytes = sizeof(double) * h->nn;
    if (hipMemcpy(h->tmp1, X, bytes, hipMemcpyHostToDevice) != hipSuccess) r = DQMC_E_HIP;
    double* Tout = apply_pivot ? h->tmp2 : h->tmp1;
    if (r == DQMC_OK) r = udt(h, B(h, h->tmp1).get(), B(h, h->cur).get(), Bd(h, h->Dl).get(), B(h, Tout).get(), apply_pivot);
    if (r == DQMC_OK && hipStreamSynchronize(h->st) != hipSuccess) r = DQMC_E_HIP;
    if (r == DQMC_OK) r = check_device_errors(h);
    if (r == DQMC_OK) {
      if (hipMemcpy(U, h->cur, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(T, Tout, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(D, h->Dl, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(pivot, h->piv, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
        r = DQMC_E_HIP;
    }
  }
  if (r != DQMC_OK) g_create_error = h->err.empty() ? "dqmc_la_udt failed" : h->err;
  la_free(x);
  return r;
}

int dqmc_la_rdivp(int n, double* Ah, const double* Th, const int32_t* pivot) {
  if (n < 1 || !Ah || !Th || !pivot) return DQMC_E_INVALID;
  LaCtx x;
  int r = la_make(x, n, 0);
  Chain* h = &x.c;
  const size_t bytes = sizeof(double) * h->nn;
  if (r == DQMC_OK) {
    if (hipMemcpy(h->tmp1, Ah, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->tmp2, Th, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->piv, pivot, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess)
      r = DQMC_E_HIP;
  }
  if (r == DQMC_OK) r = rdivp(h, B(h, h->tmp1).get(), B(h, h->tmp2).get());
  if (r == DQMC_OK && hipStreamSynchronize(h->st) != hipSuccess) r = DQMC_E_HIP;
  if (r == DQMC_OK && hipMemcpy(Ah, h->tmp1, bytes, hipMemcpyDeviceToHost) != hipSuccess) r = DQMC_E_HIP;
  if (r != DQMC_OK) g_create_error = h->err.empty() ? "dqmc_la_rdivp failed" : h->err;
  la_free(x);
  return r;
}

int dqmc_la_calc_greens_udt(int n, const double* Ulh, const double* Dlh, const double* Tlh, const double* Urh,
                            const double* Drh, const double* Trh, double* Gh) {
  if (n < 1 || !Ulh || !Dlh || !Tlh || !Urh || !Drh || !Trh || !Gh) return DQMC_E_INVALID;
  LaCtx x;
  int r = la_make(x, n, 0);
  Chain* h = &x.c;
  const size_t bytes = sizeof(double) * h->nn;
  if (r == DQMC_OK) {
    if (hipMemcpy(h->Ul, Ulh, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->Tl, Tlh, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->Ur, Urh, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->Tr, Trh, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->Dl, Dlh, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->Dr, Drh, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
      r = DQMC_E_HIP;
  }
  if (r == DQMC_OK) r = calculate_greens(h, h->G);
  if (r == DQMC_OK && hipStreamSynchronize(h->st) != hipSuccess) r = DQMC_E_HIP;
  if (r == DQMC_OK && hipMemcpy(Gh, h->G, bytes, hipMemcpyDeviceToHost) != hipSuccess) r = DQMC_E_HIP;
  if (r != DQMC_OK) g_create_error = h->err.empty() ? "dqmc_la_calc_greens_udt failed" : h->err;
  la_free(x);
  return r;
}

int dqmc_la_gemm(int n, int ta, int tb, const double* Ah, const double* Bh, double* Ch, int precision) {
  if (n < 1 || !Ah || !Bh || !Ch) return DQMC_E_INVALID;
  LaCtx x;
  int r = la_make(x, n, 0);
  Chain* h = &x.c;
  const size_t bytes = sizeof(double) * h->nn;
  if (r == DQMC_OK) {
    if (hipMemcpy(h->tmp1, Ah, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->tmp2, Bh, bytes, hipMemcpyHostToDevice) != hipSuccess)
      r = DQMC_E_HIP;
  }
  if (r == DQMC_OK) {
    if (precision == DQMC_PRECISION_MIXED) {
      float* a32 = reinterpret_cast<float*>(h->cur);
      float* b32 = reinterpret_cast<float*>(h->O);
      float* c32 = reinterpret_cast<float*>(h->Gx);
      if (launch_f64_to_f32(a32, h->tmp1, (int64_t)h->nn, h->st) != hipSuccess ||
          launch_f64_to_f32(b32, h->tmp2, (int64_t)h->nn, h->st) != hipSuccess) r = DQMC_E_HIP;
      auto g = batch<float>(ta, tb);
      g.it[g.count++] = item<float>(a32, b32, c32, n);
      if (r == DQMC_OK) r = run(h, g);
      if (r == DQMC_OK && launch_f32_to_f64(h->G, c32, (int64_t)h->nn, h->st) != hipSuccess) r = DQMC_E_HIP;
    } else {
      auto g = batch<double>(ta, tb);
      g.it[g.count++] = item<double>(h->tmp1, h->tmp2, h->G, n);
      r = run(h, g);
    }
  }
  if (r == DQMC_OK && hipStreamSynchronize(h->st) != hipSuccess) r = DQMC_E_HIP;
  if (r == DQMC_OK && hipMemcpy(Ch, h->G, bytes, hipMemcpyDeviceToHost) != hipSuccess) r = DQMC_E_HIP;
  if (r != DQMC_OK) g_create_error = h->err.empty() ? "dqmc_la_gemm failed" : h->err;
  la_free(x);
  return r;
}

}  // extern "C"
