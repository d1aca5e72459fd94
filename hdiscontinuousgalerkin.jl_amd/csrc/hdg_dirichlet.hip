// Trace Dirichlet data and apply! on the device.
//
//   dirichlet_project  Dirichlet(u_hat, mesh, "boundary", g) (src/boundary.jl:
//                      11-42): the projection of g onto the trace basis of
//                      every face with one cell, optionally with the
//                      reference's own weights and accumulator
//   apply              apply!(K, f, dbc) (src/boundary.jl:121-175) on a copy
//                      of the trace matrix and rhs: rhs lifted by the
//                      prescribed values, Dirichlet rows and columns zeroed,
//                      meandiag(K) on their diagonal, f[d] = v m
#include <hip/hip_runtime.h>

#include "hdg_internal.h"

namespace hdg {

namespace {

// One thread per face.  The trace parameter of a face runs from its lower
// to its higher global node id (face_orientation, src/mesh.jl:51-54): point
// q of the 1-D rule lies at parameter x_q along the first cell's local edge
// when that edge runs from the lower node id, at 1 - x_q otherwise
// (spatial_coordinate, src/TraceFunctionSpaces.jl:47-56).
__global__ void dirichlet_project_kernel(DevMesh m, int nt, DirichletRule R, int g_kind, double c0, double c1,
                                         double c2, int flags, unsigned char* dmask, double* uD) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf) return;
  const int cl = m.fcell[2 * f], c2nd = m.fcell[2 * f + 1];
  if (c2nd >= 0) {
    for (int i = 0; i < nt; ++i) {
      dmask[f * nt + i] = 0;
      uD[f * nt + i] = 0.0;
    }
    return;
  }
  int lf = 0;
  for (int l = 0; l < 3; ++l)
    if (m.cf[3 * (int64_t)cl + l] == f) lf = l;
  // local edge lf runs from local vertex (lf+1)%3 to (lf+2)%3
  const int na = m.cn[3 * (int64_t)cl + (lf + 1) % 3], nb = m.cn[3 * (int64_t)cl + (lf + 2) % 3];
  const bool ori = nb > na;
  const double xa = m.xy[2 * (int64_t)na], ya = m.xy[2 * (int64_t)na + 1];
  const double xb = m.xy[2 * (int64_t)nb], yb = m.xy[2 * (int64_t)nb + 1];
  const bool ref = (flags & 1) != 0;
  double acc = 0.0;
  for (int i = 0; i < nt; ++i) {
    if (!ref) acc = 0.0;  // the reference never resets it (boundary.jl:27-37)
    for (int q = 0; q < R.nfq; ++q) {
      const double t = R.x[ori ? q : R.nfq - 1 - q];
      const double x = (1.0 - t) * xa + t * xb, y = (1.0 - t) * ya + t * yb;
      double g = 0.0;
      if (g_kind == 0) g = c0 + c1 * x + c2 * y;
      else g = sin(3.14159265358979311600 * x) * sin(3.14159265358979311600 * y);
      // the reference weights face point q with the cell rule's weight q
      // (fs.fs.qr_weights, boundary.jl:33)
      acc += (ref ? R.wc[q] : R.w[q]) * g * R.leg[i * R.nfq + q];
    }
    dmask[f * nt + i] = 1;
    uD[f * nt + i] = acc;
  }
}

// Partial sums of |K(r, r)| over rows, one per block (fixed order).
__global__ __launch_bounds__(256) void absdiag_partial_kernel(DevPattern P, const double* vals, int nt, int64_t N,
                                                              double* part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = r / nt;
    const int a = (int)(r - f * nt);
    s += fabs(vals[((int64_t)P.self[f] * nt + a) * N + r]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// meandiag: the partials summed in order by one thread, divided by N.
__global__ void meandiag_kernel(const double* part, int n, int64_t N, double* m) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *m = s / (double)N;
  }
}

// Row r of the applied system: the lift f_r - sum_d v_d K(r, d) over the
// prescribed columns d in ascending order (the order of apply!'s loop over
// the prescribed dofs), zeros in prescribed rows and columns, meandiag on
// the prescribed diagonal and f_d = v_d m.
__global__ void apply_kernel(DevPattern P, TraceSystem T, const double* md, double* vals_out, double* rhs_out) {
  const int64_t N = T.N;
  const int nt = T.nt;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  const int64_t f = r / nt;
  const double mdiag = *md;
  const bool dr = T.dmask[r] != 0;
  double rhs = T.rhs[r];
  for (int s = 0; s < kSlots; ++s) {
    const int64_t g = P.bcol[f * kSlots + s];
    const bool real = s < P.nslot[f];
    for (int b = 0; b < nt; ++b) {
      const int64_t col = g * nt + b;
      const int64_t idx = ((int64_t)s * nt + b) * N + r;
      const double kv = T.vals[idx];
      const bool dc = T.dmask[col] != 0;
      if (real && dc) {
        const double v = T.uD[col];
        if (v != 0.0) rhs -= v * kv;
      }
      double out = (dr || dc) ? 0.0 : kv;
      if (dr && real && col == r) out = mdiag;
      vals_out[idx] = out;
    }
  }
  rhs_out[r] = dr ? T.uD[r] * mdiag : rhs;
}

}  // namespace

int launch_dirichlet_project(const DevMesh& m, int nt, const DirichletRule& R, int g_kind, const double* par,
                             int flags, unsigned char* dmask, double* uD, hipStream_t s) {
  const int64_t g = (m.nf + 255) / 256;
  hipLaunchKernelGGL(dirichlet_project_kernel, dim3((int)(g > 0 ? g : 1)), dim3(256), 0, s, m, nt, R, g_kind,
                     par[0], par[1], par[2], flags, dmask, uD);
  return (int)hipGetLastError();
}

int launch_apply(const DevPattern& P, const TraceSystem& T, double* part, double* md, double* vals_out,
                 double* rhs_out, hipStream_t s) {
  constexpr int kBlocks = 1024;
  hipLaunchKernelGGL(absdiag_partial_kernel, dim3(kBlocks), dim3(256), 0, s, P, T.vals, T.nt, T.N, part);
  hipLaunchKernelGGL(meandiag_kernel, dim3(1), dim3(64), 0, s, part, kBlocks, T.N, md);
  const int64_t g = (T.N + 255) / 256;
  hipLaunchKernelGGL(apply_kernel, dim3((int)(g > 0 ? g : 1)), dim3(256), 0, s, P, T, md, vals_out, rhs_out);
  return (int)hipGetLastError();
}

}  // namespace hdg
