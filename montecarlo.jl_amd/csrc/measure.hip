// Equal-time observables of the Hubbard model from the physical Green's function, on the
// device (G never leaves HBM).
//
// Reference kernels: occupation (measurements/constructors/occupation.jl:44-55), kinetic and
// total energy (measurements/constructors/energy.jl:119-166), interaction energy
// (models/HubbardModel.jl:160-182).  Double occupancy (1 - G_up_ii)(1 - G_dn_ii) holds at
// fixed auxiliary field because the two spin sectors are independent there.
//
// One workgroup: the sums are reduced in a fixed order (wave butterflies, then the waves in
// index order), so repeated calls give identical bits.
#include "common.h"

namespace dqmc {

__device__ __forceinline__ double wave_sum_fixed(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// out: [0] density per site (both spins), [1] kinetic, [2] interaction, [3] total energy,
//      [4] mean double occupancy, [5, 5 + flv n) occupation 1 - G_ii (flavour-major),
//      [5 + flv n, 5 + flv n + n) double occupancy per site
__global__ __launch_bounds__(1024) void measure_kernel(const double* __restrict__ G,
                                                       const double* __restrict__ T, int n, int flv, double U,
                                                       double* __restrict__ out) {
  __shared__ double part[4][16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const size_t nn = (size_t)n * n;
  const double* Gd = flv > 1 ? G + nn : G;
  double s_dens = 0.0, s_docc = 0.0, s_int = 0.0, s_kin = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    const double gu = G[(size_t)i * (n + 1)];
    const double gd = Gd[(size_t)i * (n + 1)];
    out[5 + i] = 1.0 - gu;
    if (flv > 1) out[5 + n + i] = 1.0 - gd;
    const double docc = (1.0 - gu) * (1.0 - gd);
    out[5 + (size_t)flv * n + i] = docc;
    s_dens += (1.0 - gu) + (1.0 - gd);
    s_docc += docc;
    s_int += (gu - 0.5) * (gd - 0.5);
  }
  for (size_t idx = tid; idx < nn; idx += blockDim.x) {
    const size_t i = idx % n, j = idx / n;
    const double t = T[idx];
    const double e = i == j ? 1.0 : 0.0;
    s_kin += t * (e - G[idx]);
    if (flv > 1) s_kin += t * (e - G[nn + idx]);
  }
  s_dens = wave_sum_fixed(s_dens);
  s_docc = wave_sum_fixed(s_docc);
  s_int = wave_sum_fixed(s_int);
  s_kin = wave_sum_fixed(s_kin);
  if (lane == 0) {
    part[0][wave] = s_dens;
    part[1][wave] = s_docc;
    part[2][wave] = s_int;
    part[3][wave] = s_kin;
  }
  __syncthreads();
  if (tid == 0) {
    double d = 0.0, o = 0.0, it = 0.0, k = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      d += part[0][w];
      o += part[1][w];
      it += part[2][w];
      k += part[3][w];
    }
    const double kin = flv == 1 ? 2.0 * k : k;   // density field: both spins share the block
    const double inter = -U * it;
    out[0] = d / n;
    out[1] = kin;
    out[2] = inter;
    out[3] = kin + inter;
    out[4] = o / n;
  }
}

hipError_t launch_measure(const double* G, const double* T, int n, int flv, double U, double* out,
                          hipStream_t s) {
  hipLaunchKernelGGL(measure_kernel, dim3(1), dim3(1024), 0, s, G, T, n, flv, U, out);
  return hipGetLastError();
}

// One time slice of a density susceptibility (apply! of measurements/generic.jl:337-370 with
// EachSitePairByDistance, :434-460): acc[dir + N (b1 + B b2)] += w sum_src K(i, j),
// i = src + b1 N, j = dir2srctrg[dir][src] + b2 N, for physical G00 (the chain's current G),
// G0l, Gl0, Gll (flv blocks).  K by observable (kDensityCharge: full_cdc_kernel /
// reduced_cdc_kernel, constructors/charge_density.jl:62-170, over the flavour pairs;
// kDensitySpinX / Y / Z: full_sdc_*_kernel / reduced_sdc_*_kernel,
// constructors/spin_density.jl:68-330, called once).  One flavour block is the reference's
// DiagonallyRepeatingMatrix form (tf = total flavours), two its BlockDiagonal form.  One wave
// per output entry; each lane sums its src terms in order and a fixed butterfly reduces them,
// so the result is reproducible bit for bit.
__global__ __launch_bounds__(64) void density_slice_kernel(
    const double* __restrict__ G00, const double* __restrict__ G0l, const double* __restrict__ Gl0,
    const double* __restrict__ Gll, const int* __restrict__ d2st, int N, int B, int n, int flv, double tf,
    int observable, int first_slice, int reduced, double w, double* __restrict__ acc) {
  const int e = blockIdx.x;               // dir + N (b1 + B b2)
  const int dir = e % N, b1 = (e / N) % B, b2 = e / (N * B);
  const size_t nn = (size_t)n * n;
  double sum = 0.0;
  for (int s = threadIdx.x; s < N; s += 64) {
    const int i = s + b1 * N;
    const int j = d2st[(size_t)dir * N + s] + b2 * N;
    const double id = (i == j && first_slice) ? 1.0 : 0.0;
    const size_t ij = (size_t)i + (size_t)j * n, ji = (size_t)j + (size_t)i * n;
    const size_t ii = (size_t)i * (n + 1), jj = (size_t)j * (n + 1);
    double v = 0.0;
    if (flv == 1) {
      if (observable == kDensityCharge) {
        v = tf * (id - G0l[ji]) * Gl0[ij];
        if (!reduced) v += tf * tf * (1.0 - Gll[ii]) * (1.0 - G00[jj]);
      } else {
        v = 2.0 * (id - G0l[ji]) * Gl0[ij];
      }
    } else if (observable == kDensityCharge) {
      for (int f1 = 0; f1 < flv; f1++) {
        if (!reduced)
          for (int f2 = 0; f2 < flv; f2++) v += (1.0 - Gll[f1 * nn + ii]) * (1.0 - G00[f2 * nn + jj]);
        v += (id - G0l[f1 * nn + ji]) * Gl0[f1 * nn + ij];
      }
    } else if (observable == kDensitySpinZ) {
      if (!reduced) {
        const double a1 = 1.0 - Gll[ii], a2 = 1.0 - Gll[nn + ii];
        const double c1 = 1.0 - G00[jj], c2 = 1.0 - G00[nn + jj];
        v = a1 * c1 - a1 * c2 - a2 * c1 + a2 * c2;
      }
      v += (id - G0l[ji]) * Gl0[ij] + (id - G0l[nn + ji]) * Gl0[nn + ij];
    } else {  // spin x and y: the same for block-diagonal G
      v = (id - G0l[ji]) * Gl0[nn + ij] + (id - G0l[nn + ji]) * Gl0[ij];
    }
    sum += v;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if (threadIdx.x == 0) acc[e] += w * sum;
}

hipError_t launch_density_slice(const double* G00, const double* G0l, const double* Gl0, const double* Gll,
                                const int* d2st, int N, int B, int n, int flv, double total_flavors, int observable,
                                int first_slice, int reduced, double w, double* acc, hipStream_t s) {
  hipLaunchKernelGGL(density_slice_kernel, dim3((unsigned)(N * B * B)), dim3(64), 0, s, G00, G0l, Gl0, Gll, d2st,
                     N, B, n, flv, total_flavors, observable, first_slice, reduced, w, acc);
  return hipGetLastError();
}

}  // namespace dqmc
