// UDT of one matrix that is already in device memory, with every scratch buffer in a
// static device array: the entry point allocates nothing.  It is for callers that cannot
// call hipMalloc (code under HIP graph capture, serving loops, harnesses that forbid it).
//
// The scratch array (about 8.4 MB: n_max^2 + 49 n_max + 256 doubles for n_max = 1024) is
// held by every GPU the library runs on until the process ends.  Its device address is
// looked up once per device with hipGetSymbolAddress; host code uses only that address.
// Calls that overlap on one GPU share the array and race: serialise them (one stream, or
// events).  Nothing in the array needs to start at zero: every kernel writes what it reads.
#include <mutex>

#include "../../include/dqmc_hip.h"
#include "common.h"

namespace {

constexpr int kMaxN = DQMC_NOALLOC_MAX_N;
constexpr size_t kScratchDoubles = (size_t)kMaxN * kMaxN + 49 * (size_t)kMaxN + 256;
constexpr int kMaxDevices = 64;
constexpr int kMaxPanels = 1024;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// The device's copy of `scratch`, looked up with that device current (once per device).
hipError_t scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int dqmc_la_udt_noalloc(int n, const double* X, int apply_pivot, double* U, double* D,
                                   double* T, int32_t* pivot, uint64_t* device_errors, void* stream) {
  using namespace dqmc;
  if (n < 1 || n > kMaxN || !X || !U || !D || !T || !pivot) return DQMC_E_INVALID;
  if (U == X || U == T || T == X) return DQMC_E_INVALID;
  if (kernels_init_once() != hipSuccess) return DQMC_E_HIP;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DQMC_E_HIP;
  double* base = nullptr;
  if (scratch_ptr(dev, &base) != hipSuccess) return DQMC_E_HIP;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t nn = (size_t)n * n;
  // carve the scratch: A (the input's working copy, n^2), tau (n), tfac ((n + 16) 16),
  // swap (32 n); 256-byte aligned pieces
  auto up = [](size_t d) { return (d + 31) & ~(size_t)31; };
  double* A = base;
  double* tau = A + up(nn);
  double* tfac = tau + up(n);
  double* swp = tfac + up((size_t)(n + 16) * 16);
  if (swp + 32 * (size_t)n > base + kScratchDoubles) return DQMC_E_INVALID;

  UdtBatch u;
  u.count = 1;
  u.n = n;
  u.apply_pivot = apply_pivot ? 1 : 0;
  u.it[0].err = reinterpret_cast<unsigned long long*>(device_errors);
  u.it[0].n = n;
  u.it[0].p0 = u.it[0].w = u.it[0].last = 0;
  u.it[0].A = A;
  u.it[0].U = U;
  u.it[0].D = D;
  u.it[0].T = apply_pivot ? T : A;     // the "dirty upper" T is formed in place, then copied
  u.it[0].pivot = pivot;
  u.it[0].tau = tau;
  u.it[0].tfac = tfac;
  u.it[0].swap = swp;
  if (udt_check(u) != hipSuccess) return DQMC_E_INVALID;
  int starts[kMaxPanels], widths[kMaxPanels];
  const int np = udt_plan_fixed(n, starts, widths, kMaxPanels);
  if (np < 1) return DQMC_E_INVALID;
  if (launch_copy(A, X, (int64_t)nn, st) != hipSuccess) return DQMC_E_HIP;
  if (launch_udt_norms(u, st) != hipSuccess) return DQMC_E_HIP;
  for (int pi = 0; pi < np; pi++) {
    if (launch_udt_panel(u, starts[pi], widths[pi], pi == np - 1, st) != hipSuccess) return DQMC_E_HIP;
    if (launch_udt_reflect(u, starts[pi], widths[pi], st) != hipSuccess) return DQMC_E_HIP;
  }
  if (launch_udt_finish(u, st) != hipSuccess) return DQMC_E_HIP;
  if (!apply_pivot && launch_copy(T, A, (int64_t)nn, st) != hipSuccess) return DQMC_E_HIP;
  return DQMC_OK;
}
