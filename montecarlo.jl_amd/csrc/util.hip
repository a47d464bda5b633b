// Elementwise and reduction helpers: identity / copy / precision conversion, the
// symmetrisation of the hopping exponentials and the max|G_wrapped - G_scratch|
// reduction of the propagation-error check (reference stack.jl:634-654, 693-709).
#include "common.h"

namespace dqmc {

static inline dim3 grid_for(int64_t count, int bs = 256) {
  int64_t g = (count + bs - 1) / bs;
  if (g > 65535 * 8) g = 65535 * 8;
  return dim3((unsigned)g);
}

template <typename T>
__global__ void set_identity_kernel(T* A, int n, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = i % ((int64_t)n * n);
    const int r = (int)(e % n), c = (int)(e / n);
    A[i] = (r == c) ? T(1) : T(0);
  }
}
hipError_t launch_set_identity(double* A, int n, int nmat, hipStream_t s) {
  int64_t total = (int64_t)n * n * nmat;
  hipLaunchKernelGGL(set_identity_kernel<double>, grid_for(total), dim3(256), 0, s, A, n, total);
  return hipGetLastError();
}
hipError_t launch_set_identity_f32(float* A, int n, int nmat, hipStream_t s) {
  int64_t total = (int64_t)n * n * nmat;
  hipLaunchKernelGGL(set_identity_kernel<float>, grid_for(total), dim3(256), 0, s, A, n, total);
  return hipGetLastError();
}

__global__ void fill_kernel(double* x, double v, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
hipError_t launch_fill(double* x, double v, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, grid_for(count), dim3(256), 0, s, x, v, count);
  return hipGetLastError();
}

// The Hirsch field in the reference's compressed form, BitArray(conf .== 1)
// (fields.jl:331-334): bit k % 64 of chunk k / 64 is set where the column-major entry k is
// +1; the high bits of the last chunk beyond the field are zero.  Pack: one wave per chunk,
// lane l reads entry 64 c + l and the wave's ballot is the chunk.
__global__ __launch_bounds__(256) void pack_field_kernel(const int8_t* __restrict__ conf, int64_t count,
                                                         uint64_t* __restrict__ chunks, int64_t nchunks) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * 4;
  for (int64_t c = blockIdx.x * (int64_t)4 + (threadIdx.x >> 6); c < nchunks; c += waves) {
    const int64_t k = c * 64 + lane;
    const int8_t x = conf[k < count ? k : count - 1];
    const unsigned long long bits = __ballot(k < count && x == 1);   // whole wave active
    if (lane == 0) chunks[c] = (uint64_t)bits;
  }
}
__global__ void unpack_field_kernel(const uint64_t* __restrict__ chunks, int64_t count, int8_t* __restrict__ conf) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x)
    conf[k] = ((chunks[k >> 6] >> (k & 63)) & 1ull) ? int8_t(1) : int8_t(-1);
}
hipError_t launch_pack_field(const int8_t* conf, int64_t count, uint64_t* chunks, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int64_t nchunks = (count + 63) / 64;
  int64_t g = (nchunks + 3) / 4;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(pack_field_kernel, dim3((unsigned)g), dim3(256), 0, s, conf, count, chunks, nchunks);
  return hipGetLastError();
}
hipError_t launch_unpack_field(const uint64_t* chunks, int64_t count, int8_t* conf, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(unpack_field_kernel, grid_for(count), dim3(256), 0, s, chunks, count, conf);
  return hipGetLastError();
}

__global__ void copy_kernel(double* __restrict__ dst, const double* __restrict__ src, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
hipError_t launch_copy(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count <= 0 || dst == src) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel, grid_for(count), dim3(256), 0, s, dst, src, count);
  return hipGetLastError();
}
// Up to three copies in one launch (a stack slot's U, T and D): one grid-stride pass over
// the segments laid end to end.
struct CopySegs {
  double* dst[3];
  const double* src[3];
  int64_t count[3];
};
__global__ void copy3_kernel(CopySegs c) {
  const int64_t total = c.count[0] + c.count[1] + c.count[2];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = i < c.count[0] ? 0 : (i < c.count[0] + c.count[1] ? 1 : 2);
    const int64_t j = k == 0 ? i : (k == 1 ? i - c.count[0] : i - c.count[0] - c.count[1]);
    c.dst[k][j] = c.src[k][j];
  }
}
hipError_t launch_copy3(double* d0, const double* s0, int64_t n0, double* d1, const double* s1, int64_t n1,
                        double* d2, const double* s2, int64_t n2, hipStream_t s) {
  if (n0 < 0 || n1 < 0 || n2 < 0) return hipErrorInvalidValue;
  CopySegs c{{d0, d1, d2}, {s0, s1, s2}, {n0, n1, n2}};
  const int64_t total = n0 + n1 + n2;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(copy3_kernel, grid_for(total), dim3(256), 0, s, c);
  return hipGetLastError();
}


__global__ void f64_to_f32_kernel(float* __restrict__ dst, const double* __restrict__ src, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}
__global__ void f32_to_f64_kernel(double* __restrict__ dst, const float* __restrict__ src, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (double)src[i];
}
hipError_t launch_f64_to_f32(float* dst, const double* src, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(f64_to_f32_kernel, grid_for(count), dim3(256), 0, s, dst, src, count);
  return hipGetLastError();
}
hipError_t launch_f32_to_f64(double* dst, const float* src, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(f32_to_f64_kernel, grid_for(count), dim3(256), 0, s, dst, src, count);
  return hipGetLastError();
}

// A <- (A + A') / 2
__global__ void symmetrize_kernel(double* A, int n) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    if (r < c) {
      const double v = 0.5 * (A[r + (size_t)c * n] + A[c + (size_t)r * n]);
      A[r + (size_t)c * n] = v;
      A[c + (size_t)r * n] = v;
    }
  }
}
hipError_t launch_symmetrize(double* A, int n, hipStream_t s) {
  hipLaunchKernelGGL(symmetrize_kernel, grid_for((int64_t)n * n), dim3(256), 0, s, A, n);
  return hipGetLastError();
}

// A <- scale * A + add_diag * I
__global__ void scale_add_identity_kernel(double* A, double scale, double add, int n) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    A[e] = scale * A[e] + (r == c ? add : 0.0);
  }
}
hipError_t launch_scale_add_identity(double* A, double scale, double add_diag, int n, hipStream_t s) {
  hipLaunchKernelGGL(scale_add_identity_kernel, grid_for((int64_t)n * n), dim3(256), 0, s, A, scale,
                     add_diag, n);
  return hipGetLastError();
}

__global__ void axpy_kernel(double* __restrict__ Y, const double* __restrict__ X, double a, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    Y[i] += a * X[i];
}
hipError_t launch_axpy_matrix(double* Y, const double* X, double a, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(axpy_kernel, grid_for(count), dim3(256), 0, s, Y, X, a, count);
  return hipGetLastError();
}

__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// max(x, y) with a NaN in either argument winning
__device__ inline double nan_max(double x, double y) { return (y > x || y != y) ? y : x; }

// max |a - b| over count elements (a from a32 when given): grid-stride partial maxima, one
// per workgroup, combined with a 64-bit atomic max on the bit patterns, which order
// non-negative doubles as their values (a NaN, positive after fabs, above +inf).  The
// maximum is exact, so the result does not depend on the order.  *out must hold +0.0.
__global__ __launch_bounds__(256) void maxabsdiff_kernel(const double* __restrict__ a,
                                                         const float* __restrict__ a32,
                                                         const double* __restrict__ b, int64_t count,
                                                         double* out) {
  __shared__ double part[4];
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) {
    const double av = a32 ? (double)a32[i] : a[i];
    m = nan_max(m, fabs(av - b[i]));
  }
  for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = nan_max(nan_max(part[0], part[1]), nan_max(part[2], part[3]));
    atomicMax(reinterpret_cast<unsigned long long*>(out), (unsigned long long)__double_as_longlong(r));
  }
}
hipError_t launch_maxabsdiff(const double* a, const float* a32, const double* b, int64_t count,
                             double* out, hipStream_t s) {
  hipError_t e = launch_fill(out, 0.0, 1, s);
  if (e != hipSuccess) return e;
  int64_t g = (count + 255) / 256;
  if (g > 128) g = 128;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(maxabsdiff_kernel, dim3((unsigned)g), dim3(256), 0, s, a, a32, b, count, out);
  return hipGetLastError();
}

// out_b = diag(map(dl_b)) in_b diag(map(dr_b)) [+ add_b] for nmat consecutive n x n blocks,
// with dl_b = dl + b n and dr_b = dr + b n (either may be null: no scaling on that side)
__global__ void diag_scale_kernel(double* __restrict__ out, const double* __restrict__ in,
                                  const double* __restrict__ dl, int lmode, const double* __restrict__ dr,
                                  int rmode, const double* __restrict__ add, int n, int64_t total) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / nn, i = (e - b * nn) % n, j = (e - b * nn) / n;
    double v = in[e];
    if (dl) v = diag_map(dl[b * n + i], lmode) * v;
    if (dr) v = v * diag_map(dr[b * n + j], rmode);
    if (add) v = v + add[e];
    out[e] = v;
  }
}
hipError_t launch_diag_scale(double* out, const double* in, const double* dl, int lmode, const double* dr,
                             int rmode, const double* add, int n, int nmat, hipStream_t s) {
  const int64_t total = (int64_t)n * n * nmat;
  hipLaunchKernelGGL(diag_scale_kernel, grid_for(total), dim3(256), 0, s, out, in, dl, lmode, dr, rmode,
                     add, n, total);
  return hipGetLastError();
}

// A_b = diag(map(v_b)) for nmat consecutive n x n blocks (v_b = v + b n)
__global__ void diag_matrix_kernel(double* __restrict__ A, const double* __restrict__ v, int mode, int n,
                                   int64_t total) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / nn, i = (e - b * nn) % n, j = (e - b * nn) / n;
    A[e] = i == j ? diag_map(v[b * n + i], mode) : 0.0;
  }
}
hipError_t launch_diag_matrix(double* A, const double* v, int mode, int n, int nmat, hipStream_t s) {
  const int64_t total = (int64_t)n * n * nmat;
  hipLaunchKernelGGL(diag_matrix_kernel, grid_for(total), dim3(256), 0, s, A, v, mode, n, total);
  return hipGetLastError();
}

// Many-chain forms: item k of the table is handled by the workgroups with blockIdx.y = k.
__global__ __launch_bounds__(256) void elem_multi_kernel(const ElemItem* __restrict__ items) {
  const ElemItem& it = items[blockIdx.y];
  const int64_t count = it.count;
  double* __restrict__ dst = it.dst;
  const int64_t stride = (int64_t)gridDim.x * 256;
  if (it.op == kElemCopy) {
    const double* __restrict__ src = it.src;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < count; i += stride) dst[i] = src[i];
  } else if (it.op == kElemIdentity) {
    const int64_t nn = (int64_t)it.n * it.n;
    const int n = it.n;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < count; i += stride) {
      const int64_t e = i % nn;
      dst[i] = (e % n) == (e / n) ? 1.0 : 0.0;
    }
  } else {
    const double v = it.v;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < count; i += stride) dst[i] = v;
  }
}
hipError_t launch_elem_multi(const ElemItem* items, int count, int64_t max_count, hipStream_t s) {
  if (count <= 0 || max_count <= 0) return hipSuccess;
  if (count > 65535) return hipErrorInvalidValue;
  int64_t g = (max_count + 255) / 256;
  if (g > 256) g = 256;
  hipLaunchKernelGGL(elem_multi_kernel, dim3((unsigned)g, count), dim3(256), 0, s, items);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void maxabsdiff_multi_kernel(const MaxdiffItem* __restrict__ items) {
  __shared__ double part[4];
  const MaxdiffItem& it = items[blockIdx.y];
  const double* __restrict__ a = it.a;
  const double* __restrict__ b = it.b;
  const int64_t count = it.count;
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256)
    m = nan_max(m, fabs(a[i] - b[i]));
  for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = nan_max(nan_max(part[0], part[1]), nan_max(part[2], part[3]));
    atomicMax(reinterpret_cast<unsigned long long*>(it.out), (unsigned long long)__double_as_longlong(r));
  }
}
hipError_t launch_maxabsdiff_multi(const MaxdiffItem* items, int count, int64_t max_count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count > 65535) return hipErrorInvalidValue;
  int64_t g = (max_count + 255) / 256;
  if (g > 128) g = 128;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(maxabsdiff_multi_kernel, dim3((unsigned)g, count), dim3(256), 0, s, items);
  return hipGetLastError();
}

// max column sum |A| (1-norm), used to pick the scaling of the matrix exponential
__global__ __launch_bounds__(1024) void norm1_kernel(const double* __restrict__ A, int n, double* out) {
  __shared__ double part[16];
  double m = 0.0;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < n; r++) s += fabs(A[r + (size_t)c * n]);
    m = fmax(m, s);
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = fmax(r, part[w]);
    *out = r;
  }
}
hipError_t launch_norm1(const double* A, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(norm1_kernel, dim3(1), dim3(1024), 0, s, A, n, out);
  return hipGetLastError();
}

__global__ void scale_cols_kernel(double* __restrict__ out, const double* __restrict__ in,
                                  const double* __restrict__ d, int n, int64_t total) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e / nn;
    const int c = (int)((e - m * nn) / n);
    out[e] = in[e] * d[m * n + c];
  }
}
hipError_t launch_scale_cols(double* out, const double* in, const double* d, int n, int nmat,
                             hipStream_t s) {
  const int64_t total = (int64_t)n * n * nmat;
  hipLaunchKernelGGL(scale_cols_kernel, grid_for(total), dim3(256), 0, s, out, in, d, n, total);
  return hipGetLastError();
}

}  // namespace dqmc
