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
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dqmc_hip.h"
#include "common.h"

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

int run(Chain* h, const GemmBatch<double>& b) {
  Timed t(h, DQMC_KERNEL_GEMM_F64);
  HIPCHK(h, launch_gemm_f64(b, h->st));
  return DQMC_OK;
}
int run(Chain* h, const GemmBatch<float>& b) {
  Timed t(h, DQMC_KERNEL_GEMM_F32);
  HIPCHK(h, launch_gemm_f32(b, h->st));
  return DQMC_OK;
}

int copyd(Chain* h, double* dst, const double* src, size_t count) {
  HIPCHK(h, launch_copy(dst, src, (int64_t)count, h->st));
  return DQMC_OK;
}

int identity(Chain* h, double* A, int nmat) {
  HIPCHK(h, launch_set_identity(A, h->n, nmat, h->st));
  return DQMC_OK;
}

int ones(Chain* h, double* d, size_t count) {
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
  u.err = h->ctr ? reinterpret_cast<unsigned long long*>(&h->ctr->device_errors) : nullptr;
  for (int b = 0; b < h->flv; b++) {
    u.it[b].A = A[b]; u.it[b].U = U[b]; u.it[b].D = D[b]; u.it[b].T = T[b];
    u.it[b].pivot = h->piv + (size_t)b * h->n;
    u.it[b].tau = h->tau + (size_t)b * h->n;
    u.it[b].tfac = h->tfac + (size_t)b * (h->n + 16) * 16;
    u.it[b].swap = h->swp + (size_t)b * 32 * h->n;
  }
  HIPCHK(h, udt_check(u));
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

// add_slice_sequence_left(mc, idx), stack.jl:377-393
int add_slice_sequence_left(Chain* h, int idx) {
  const int r0 = h->rs[idx - 1], r1 = h->re[idx - 1];
  Blocks in = Us(h, idx);
  Blocks dsl = Ds(h, idx);
  for (int s = r0; s <= r1; s++) {
    const bool last = s == r1;
    Blocks out = last ? B(h, h->tmp1) : ((s - r0) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
    CHK(slice_left(h, s, in.get(), out.get(), last ? dsl.get() : nullptr));
    in = out;
  }
  CHK(udt(h, B(h, h->tmp1).get(), Us(h, idx + 1).get(), Ds(h, idx + 1).get(), B(h, h->tmp2).get(), 1));
  return mul_blocks(h, Ts(h, idx + 1), B(h, h->tmp2), Ts(h, idx));
}

// add_slice_sequence_right(mc, idx), stack.jl:402-416
int add_slice_sequence_right(Chain* h, int idx) {
  const int r0 = h->rs[idx - 1], r1 = h->re[idx - 1];
  Blocks in = Us(h, idx + 1);
  Blocks dsl = Ds(h, idx + 1);
  for (int s = r1; s >= r0; s--) {
    const bool last = s == r0;
    Blocks out = last ? B(h, h->tmp1) : ((r1 - s) % 2 == 0 ? B(h, h->cur) : B(h, h->tmp2));
    CHK(slice_dagger_left(h, s, in.get(), out.get(), last ? dsl.get() : nullptr));
    in = out;
  }
  CHK(udt(h, B(h, h->tmp1).get(), Us(h, idx).get(), Ds(h, idx).get(), B(h, h->tmp2).get(), 1));
  return mul_blocks(h, Ts(h, idx), B(h, h->tmp2), Ts(h, idx + 1));
}

// calculate_greens_AVX!(Ul, Dl, Tl, Ur, Dr, Tr, G), stack.jl:442-496
int calculate_greens(Chain* h, double* out) {
  const int n = h->n;
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
      it.dmid = diag_vec(Dr.p[b], 1);
      g.it[g.count++] = it;
    }
    CHK(run(h, g));
  }
  return DQMC_OK;
}

int slot_to(Chain* h, int s, double* U, double* D, double* T) {
  const size_t m = (size_t)h->flv * h->nn;
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
  HIPCHK(h, hipStreamSynchronize(h->st));
  collect_times(h);
  if (h->nerr > 0) {
    std::vector<double> v(h->nerr);
    HIPCHK(h, hipMemcpy(v.data(), h->errbuf, sizeof(double) * h->nerr, hipMemcpyDeviceToHost));
    for (double x : v) absorb_error(h, x);
    h->nerr = 0;
  }
  CHK(check_device_errors(h));
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

int reverse_build(Chain* h) {
  CHK(slot_identity(h, h->K + 1));
  for (int i = h->K; i >= 1; i--) CHK(add_slice_sequence_right(h, i));
  h->cs = 0; h->cr = 1; h->dir = 1;
  CHK(udt_identity(h, h->Ul, h->Dl, h->Tl));
  CHK(slot_to(h, 1, h->Ur, h->Dr, h->Tr));
  CHK(calculate_greens(h, h->G));
  return to32(h);
}

int forward_build(Chain* h) {
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
  int slot = 0;
  auto launch = [&](int kind) -> int {
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

}  // namespace

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
  std::vector<int8_t> plus(n * (size_t)h->M, 1);
  if (e == hipSuccess) e = hipMemcpy(h->conf, plus.data(), plus.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    g_create_error = std::string("hipMemcpy: ") + hipGetErrorString(e);
    (void)hipStreamDestroy(h->st);
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
    return DQMC_E_HIP;
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

static int read_accepted(Chain* h, long long* acc) {
  ChainCounters c;
  HIPCHK(h, hipMemcpy(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost));
  *acc = c.accepted;
  return DQMC_OK;
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

int dqmc_sweep(dqmc_chain_t h, int nsweeps, int64_t* accepted) {
  ENTER(h);
  NEED_HOPPING(h);
  if (nsweeps < 0) { h->err = "nsweeps < 0"; return DQMC_E_INVALID; }
  if (h->cs < 1 || h->cs > h->M) { h->err = "dqmc_sweep: build the stack first"; return DQMC_E_STATE; }
  long long a0 = 0, a1 = 0;
  if (accepted) { HIPCHK(h, hipStreamSynchronize(h->st)); CHK(read_accepted(h, &a0)); }
  for (int s = 0; s < nsweeps; s++) {
    for (int k = 0; k < 2 * h->M; k++) {
      CHK(sweep_spatial(h));
      CHK(propagate(h));
    }
    h->sweeps++;
  }
  CHK(harvest(h));
  if (accepted) { CHK(read_accepted(h, &a1)); *accepted = a1 - a0; }
  return DQMC_OK;
}

int dqmc_batch_sweep(dqmc_chain_t* hs, int count, int nsweeps, int64_t* accepted) {
  if (!hs || count < 1 || nsweeps < 0) return DQMC_E_INVALID;
  for (int i = 0; i < count; i++)
    for (int j = 0; j < i; j++)
      if (!hs[i] || hs[i] == hs[j]) return DQMC_E_INVALID;   // a chain is not shared by two threads
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
  CHK(calculate_greens(h, h->Gchk));
  CHK(download_matrix(h, Gout, ld, h->Gchk, h->flv * h->n));
  HIPCHK(h, hipStreamSynchronize(h->st));
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
