// rectangle_mesh(TriangleCell, (nx, ny), LL, UR) generated on the device
// (src/generate_mesh.jl:1-18,110-143): the same node coordinates,
// cells, face numbering and boundary classification as the host generator
// (hdg_mesh.cpp), bit for bit, plus the structures the solver derives from
// them (face end points, vertex -> face lists, the four vertices around each
// face) in closed form.  Coordinates are rounded operation by operation,
// as the host file is compiled without FMA contraction: contraction is off
// for this file (device code otherwise contracts a * b + c into one fma).
#include <hip/hip_runtime.h>

#include "hdg_internal.h"

#pragma clang fp contract(off)

namespace hdg {

namespace {

__device__ inline uint64_t splitmix64_d(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// 1-based ids of the faces first created by grid cell (i, j), in the order
// rectangle_mesh meets them: diagonal, left, bottom of the lower triangle;
// top, right of the upper one (the closed form of hdg_mesh.cpp's sweep).
struct SweepD {
  int64_t nx, ny;
  __device__ int64_t base(int64_t j) const { return j == 1 ? 0 : (4 * nx + 1) + (j - 2) * (3 * nx + 1); }
  __device__ int64_t D(int64_t i, int64_t j) const {
    if (i == 1) return base(j) + 1;
    return base(j) + (j == 1 ? 5 + 4 * (i - 2) : 4 + 3 * (i - 2)) + 1;
  }
  __device__ int64_t T(int64_t i, int64_t j) const {
    if (j == 1) return i == 1 ? base(j) + 4 : base(j) + 5 + 4 * (i - 2) + 3;
    return i == 1 ? base(j) + 3 : base(j) + 4 + 3 * (i - 2) + 2;
  }
  __device__ int64_t R(int64_t i, int64_t j) const {
    if (j == 1) return i == 1 ? base(j) + 5 : base(j) + 5 + 4 * (i - 2) + 4;
    return i == 1 ? base(j) + 4 : base(j) + 4 + 3 * (i - 2) + 3;
  }
  __device__ int64_t L(int64_t i, int64_t j) const { return i == 1 ? base(j) + 2 : R(i - 1, j); }
  __device__ int64_t B(int64_t i, int64_t j) const {
    if (j > 1) return T(i, j - 1);
    return i == 1 ? base(j) + 3 : base(j) + 5 + 4 * (i - 2) + 2;
  }
};

// a (1 - r) + r b, rounded as the host evaluates it (no contraction here)
__device__ inline double lerp_rn(double a, double b, double r) { return a * (1.0 - r) + r * b; }

// node coordinates, then the jitter of interior nodes
__global__ void rm_nodes_kernel(RectMeshArgs a) {
  const int64_t nnx = a.nx + 1, nny = a.ny + 1;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nnx * nny) return;
  const int64_t i = v / nnx, j = v % nnx;  // row, column (0-based), as the host loops
  const double rb = (double)i / (double)(nny - 1);
  const double xa = lerp_rn(a.x0, a.x0, rb), xb = lerp_rn(a.x1, a.x1, rb);
  const double ya = lerp_rn(a.y0, a.y1, rb), yb = lerp_rn(a.y0, a.y1, rb);
  const double ratio = (double)j / (double)(nnx - 1);
  double x = lerp_rn(xa, xb, ratio), y = lerp_rn(ya, yb, ratio);
  if (a.jitter > 0.0 && i >= 1 && i < a.ny && j >= 1 && j < a.nx) {
    const double hx = (a.x1 - a.x0) / (double)a.nx, hy = (a.y1 - a.y0) / (double)a.ny;
    const uint64_t n = (uint64_t)v;
    const double ux = (double)(splitmix64_d(a.seed + 2 * n) >> 11) * (1.0 / 9007199254740992.0);
    const double uy = (double)(splitmix64_d(a.seed + 2 * n + 1) >> 11) * (1.0 / 9007199254740992.0);
    x = x + a.jitter * hx * (2.0 * ux - 1.0);
    y = y + a.jitter * hy * (2.0 * uy - 1.0);
  }
  a.xy[2 * (a.node0 + v)] = x;
  a.xy[2 * (a.node0 + v) + 1] = y;
}

// cells (nodes, faces) of grid cell (i, j) and the faces it creates
__global__ void rm_cells_kernel(RectMeshArgs a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nx * a.ny) return;
  const int64_t i = g % a.nx + 1, j = g / a.nx + 1;
  const int64_t nnx = a.nx + 1;
  const SweepD S{a.nx, a.ny};
  auto node = [&](int64_t ii, int64_t jj) { return a.node0 + ii + (jj - 1) * nnx - 1; };  // 0-based, shifted
  const int64_t lo = 2 * ((j - 1) * a.nx + (i - 1)), up = lo + 1;  // 0-based, local
  const int64_t cl = a.cell0 + lo, cu = a.cell0 + up;
  const int64_t fD = S.D(i, j), fT = S.T(i, j), fR = S.R(i, j), fL = S.L(i, j), fB = S.B(i, j);
  int* cn = a.cn;
  int* cf = a.cf;
  cn[3 * cl] = (int)node(i, j);
  cn[3 * cl + 1] = (int)node(i + 1, j);
  cn[3 * cl + 2] = (int)node(i, j + 1);
  cf[3 * cl] = (int)(a.face0 + fD - 1);
  cf[3 * cl + 1] = (int)(a.face0 + fL - 1);
  cf[3 * cl + 2] = (int)(a.face0 + fB - 1);
  cn[3 * cu] = (int)node(i + 1, j);
  cn[3 * cu + 1] = (int)node(i + 1, j + 1);
  cn[3 * cu + 2] = (int)node(i, j + 1);
  cf[3 * cu] = (int)(a.face0 + fT - 1);
  cf[3 * cu + 1] = (int)(a.face0 + fD - 1);
  cf[3 * cu + 2] = (int)(a.face0 + fR - 1);
  if (a.cell_inst) {
    a.cell_inst[cl] = a.inst;
    a.cell_inst[cu] = a.inst;
  }
  // face f (1-based, local): end points v1, v2, cells c1, c2 (0-based, global; -1 none)
  auto setf = [&](int64_t f, int64_t v1, int64_t v2, int64_t c1, int64_t c2) {
    const int64_t F = a.face0 + f - 1;
    a.fcell[2 * F] = (int)c1;
    a.fcell[2 * F + 1] = (int)c2;
    a.fv[2 * F] = (int)min(v1, v2);
    a.fv[2 * F + 1] = (int)max(v1, v2);
  };
  setf(fD, node(i + 1, j), node(i, j + 1), cl, cu);
  if (i == 1) setf(fL, node(i, j + 1), node(i, j), cl, -1);
  if (j == 1) setf(fB, node(i, j), node(i + 1, j), cl, -1);
  setf(fT, node(i + 1, j + 1), node(i, j + 1), cu, j < a.ny ? a.cell0 + 2 * (j * a.nx + (i - 1)) : -1);
  setf(fR, node(i + 1, j), node(i + 1, j + 1), cu, i < a.nx ? a.cell0 + 2 * ((j - 1) * a.nx + i) : -1);
}

// A face with one cell must lie on a side of the box: both end points on
// the bottom, right, top or left side, by exact coordinate equality
// (_get_rectangular_boundary_sets, generate_mesh.jl:60-89); else bit 4 of
// the error flag (HDG_ERR_BOUNDARY on the host).  Also the four vertices
// around every face (lower, higher end point, opposite vertex of cell 1, of
// cell 2 or cell 1 again).
__global__ void rm_faces_kernel(RectMeshArgs a) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.nf) return;
  const int64_t F = a.face0 + f;
  const int lo = a.fv[2 * F], hi = a.fv[2 * F + 1];
  const int c1 = a.fcell[2 * F], c2 = a.fcell[2 * F + 1];
  if (c2 < 0) {
    const double *p = a.xy + 2 * (int64_t)lo, *q = a.xy + 2 * (int64_t)hi;
    const bool ok = (p[1] == a.y0 && q[1] == a.y0) || (p[0] == a.x1 && q[0] == a.x1) ||
                    (p[1] == a.y1 && q[1] == a.y1) || (p[0] == a.x0 && q[0] == a.x0);
    if (!ok) atomicOr(a.err, 4);
  }
  if (!a.fq4) return;
  auto opp = [&](int c) {
    int o = -1;
    for (int w = 0; w < 3; ++w) {
      const int nw = a.cn[3 * (int64_t)c + w];
      if (nw != lo && nw != hi) o = nw;
    }
    return o;
  };
  const int o1 = opp(c1);
  a.fq4[4 * F] = lo;
  a.fq4[4 * F + 1] = hi;
  a.fq4[4 * F + 2] = o1;
  a.fq4[4 * F + 3] = c2 < 0 ? o1 : opp(c2);
}

// Vertex -> face lists: node (i, j) (0-based column, row) meets the faces
// to its right, left, top and bottom neighbours and the two diagonal ones
// that exist; listed in ascending face order, as the host builds them.
__global__ void rm_vertex_faces_kernel(RectMeshArgs a) {
  const int64_t nnx = a.nx + 1, nny = a.ny + 1;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nnx * nny) return;
  const int64_t i0 = v % nnx, j0 = v / nnx;
  const int64_t i = i0 + 1, j = j0 + 1;  // 1-based grid position
  const SweepD S{a.nx, a.ny};
  int64_t f[6];
  int d = 0;
  // edge along x from node (i, j) to (i + 1, j): bottom edge of cell (i, j)
  // or, on the top row, top edge of cell (i, ny)
  auto hedge = [&](int64_t ii) { return j <= a.ny ? S.B(ii, j) : S.T(ii, a.ny); };
  // edge along y from (i, j) to (i, j + 1): left edge of cell (i, j) or, on
  // the last column, right edge of cell (nx, j)
  auto vedge = [&](int64_t jj) { return i <= a.nx ? S.L(i, jj) : S.R(a.nx, jj); };
  if (i0 < a.nx) f[d++] = hedge(i);
  if (i0 > 0) f[d++] = hedge(i - 1);
  if (j0 < a.ny) f[d++] = vedge(j);
  if (j0 > 0) f[d++] = vedge(j - 1);
  if (i0 > 0 && j0 < a.ny) f[d++] = S.D(i - 1, j);
  if (i0 < a.nx && j0 > 0) f[d++] = S.D(i, j - 1);
  // ascending (insertion sort of at most 6)
  for (int p = 1; p < d; ++p)
    for (int q = p; q > 0 && f[q - 1] > f[q]; --q) {
      const int64_t t = f[q];
      f[q] = f[q - 1];
      f[q - 1] = t;
    }
  // entries before node v: whole rows 0..j0-1, then columns 0..i0-1 of row j0
  const int64_t nx = a.nx, ny = a.ny;
  const int64_t rows = 2 * nx * j0 + (2 * nx + 1) * (min(j0, ny) + max(j0 - 1, (int64_t)0));
  const int64_t top = j0 < ny, bot = j0 > 0;
  const int64_t cols = min(i0, nx) + max(i0 - 1, (int64_t)0) + i0 * top + i0 * bot +
                       max(i0 - 1, (int64_t)0) * top + min(i0, nx) * bot;
  const int64_t p0 = 2 * a.face0 + rows + cols;
  a.vf_ptr[a.node0 + v] = (int)p0;
  for (int k = 0; k < d; ++k) a.vf_idx[p0 + k] = (int)(a.face0 + f[k] - 1);
  if (a.last && v == nnx * nny - 1) a.vf_ptr[a.node0 + v + 1] = (int)(p0 + d);
}

// Every dof of every face with one cell is prescribed, the others are free.
// u = value on the boundary: the trace basis is the orthonormal Legendre
// basis on [0, 1] with P0 = 1, so the projection of a constant is `value`
// on mode 0 and 0 on the higher modes.
__global__ void boundary_dirichlet_kernel(const int* fcell, int64_t face0, int64_t nf, int nt, unsigned char* dmask,
                                          double* uD, double value) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nf * nt) return;
  const int64_t F = face0 + r / nt, R = face0 * nt + r;
  const bool b = fcell[2 * F + 1] < 0;
  dmask[R] = b ? 1 : 0;
  if (uD) uD[R] = (b && r % nt == 0) ? value : 0.0;
}

int blocks_for(int64_t n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

int launch_rect_mesh(const RectMeshArgs& a, hipStream_t s) {
  const int64_t nn = (a.nx + 1) * (a.ny + 1);
  hipLaunchKernelGGL(rm_nodes_kernel, dim3(blocks_for(nn)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(rm_cells_kernel, dim3(blocks_for(a.nx * a.ny)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(rm_faces_kernel, dim3(blocks_for(a.nf)), dim3(256), 0, s, a);
  if (a.vf_ptr) hipLaunchKernelGGL(rm_vertex_faces_kernel, dim3(blocks_for(nn)), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

int launch_boundary_dirichlet(const int* fcell, int64_t face0, int64_t nf, int nt, unsigned char* dmask, double* uD,
                              double value, hipStream_t s) {
  hipLaunchKernelGGL(boundary_dirichlet_kernel, dim3(blocks_for(nf * nt)), dim3(256), 0, s, fcell, face0, nf, nt,
                     dmask, uD, value);
  return (int)hipGetLastError();
}

}  // namespace hdg
