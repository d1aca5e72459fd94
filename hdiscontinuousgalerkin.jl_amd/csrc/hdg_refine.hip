// Kernels of hdg_refine_probe: FP64 iterative refinement of the assembled
// trace system whose corrections solve with the trace matrix rounded to
// float, the operator an FP32 condensation and assembly would hand the
// solver at best.
//
//   round_float   vr = (double)(float) vals, entry by entry
//   refine_resid  res = rhs - K u on the free rows (0 on Dirichlet rows) with
//                 the FP64 matrix, and ||res||^2 summed in a fixed order
//   add           u += d
#include <hip/hip_runtime.h>

#include "hdg_internal.h"

namespace hdg {

namespace {

constexpr int kRefBlocks = 1024;

__global__ __launch_bounds__(256) void round_float_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                          int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (double)(float)src[i];
}

__device__ inline double block_sum256(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

template <int NT>
__global__ __launch_bounds__(256) void refine_resid_kernel(DevPattern P, TraceSystem T, const double* __restrict__ u,
                                                           double* __restrict__ res, double* __restrict__ part) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < T.N; r += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    if (!T.dmask[r]) {
      const int64_t f = r / NT;
      s = T.rhs[r];
      if constexpr (kSymBlocks<NT>) {
        s -= sym_row<NT>(P.nbr, T.blk, T.nf, f, (int)(r - f * NT), [&](int64_t g, int c) { return u[g * NT + c]; });
      } else {
#pragma unroll
        for (int sl = 0; sl < kSlots; ++sl) {
          const int64_t g = (int64_t)P.bcol[f * kSlots + sl] * NT;
#pragma unroll
          for (int c = 0; c < NT; ++c) s = fma(-T.vals[(int64_t)(sl * NT + c) * T.N + r], u[g + c], s);
        }
      }
    }
    res[r] = s;
    acc = fma(s, s, acc);
  }
  const double b = block_sum256(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// the partials summed in order by one thread
__global__ void refine_sum_kernel(const double* part, int n, double* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = s;
  }
}

__global__ __launch_bounds__(256) void add_kernel(double* __restrict__ u, const double* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] += d[i];
}

int grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (int)g;
}

}  // namespace

int refine_partials() { return kRefBlocks; }

int launch_round_float(const double* src, double* dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(round_float_kernel, dim3(grid_of(n)), dim3(256), 0, s, src, dst, n);
  return (int)hipGetLastError();
}

int launch_refine_residual(const DevPattern& P, const TraceSystem& T, const double* u, double* res, double* part,
                           double* out2, hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
  switch (T.nt) {
#define RES(NT)                                                                                                  \
  case NT:                                                                                                       \
    hipLaunchKernelGGL(refine_resid_kernel<NT>, dim3(kRefBlocks), dim3(256), 0, s, P, T, u, res, part); \
    break;
    RES(2)
    RES(3)
    RES(4)
    RES(5)
#undef RES
    default: return (int)hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(refine_sum_kernel, dim3(1), dim3(64), 0, s, part, kRefBlocks, out2);
  return (int)hipGetLastError();
}

int launch_add(double* u, const double* d, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(add_kernel, dim3(grid_of(n)), dim3(256), 0, s, u, d, n);
  return (int)hipGetLastError();
}

}  // namespace hdg
