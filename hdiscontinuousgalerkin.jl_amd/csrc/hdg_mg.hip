// Multigrid preconditioner for the HDG trace system on structured meshes.
//
// Level 0 is the trace space (N dofs) with damped block-Jacobi smoothing.
// Level 1 is the continuous P1 space on the mesh vertices: the prolongation
// maps vertex values to the Legendre modes 0 and 1 of the trace of the
// piecewise-linear function on every face (mode 0 = mean, mode 1 =
// (v_hi - v_lo) / (2 sqrt 3), the face parameter running from the lower to
// the higher node id), and the coarse operator is the Galerkin product
// P^T S P, assembled cell by cell from the condensed blocks.  Levels 2.. are
// the vertex grids of a rectangle_mesh coarsened by two (a side with n
// cells gets (n + 1) / 2: coarse node I sits on fine node min(2I, n), so an
// odd side keeps its last fine column as a one-cell-wide coarse column),
// with linear interpolation on the nested triangulation and Galerkin (RAP)
// operators; every level is a 7-point stencil.  One symmetric V-cycle (damped Jacobi,
// fixed sweep counts, Jacobi sweeps on the coarsest grid) is a fixed SPD
// operator, so it preconditions CG.
//
// Stencil slots: 0 (0,0), 1 (+1,0), 2 (-1,0), 3 (0,+1), 4 (0,-1), 5 (+1,-1),
// 6 (-1,+1), 7 (+1,+1), 8 (-1,-1); node (i, j) of a level with nx cells per
// row is i + j (nx+1).  The vertex operator of the mesh and the Galerkin
// operators of even coarsening use slots 0-6; a one-cell-wide coarse column
// or row is not nested in the coarse triangulation, and RAP across it also
// couples the (+1,+1) neighbours (9 points close under RAP).
#include <hip/hip_runtime.h>

#include "hdg_internal.h"

namespace hdg {

namespace {

constexpr double kS3 = 0.28867513459481288225;  // 1 / (2 sqrt 3)
__constant__ int kOffX[kStencil] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
__constant__ int kOffY[kStencil] = {0, 0, 0, 1, -1, -1, 1, 1, -1};

__device__ inline int slot_of(int dx, int dy) {
  if (dx == 0 && dy == 0) return 0;
  if (dy == 0) return dx == 1 ? 1 : (dx == -1 ? 2 : -1);
  if (dx == 0) return dy == 1 ? 3 : (dy == -1 ? 4 : -1);
  if (dx == 1 && dy == -1) return 5;
  if (dx == -1 && dy == 1) return 6;
  if (dx == 1 && dy == 1) return 7;
  if (dx == -1 && dy == -1) return 8;
  return -1;
}

// Coarse parents of fine index x on a side of n cells: one (x even, or the
// last node of an odd side) or two neighbours at weight 1/2 (odd x < n).
struct Par1 {
  int a, b;  // a == b: single parent
};
__device__ inline Par1 parents1(int x, int n) {
  if (!(x & 1)) return Par1{x >> 1, x >> 1};
  if (x == n) return Par1{(n + 1) >> 1, (n + 1) >> 1};
  return Par1{(x - 1) >> 1, (x + 1) >> 1};
}
// fine index of coarse index I on a side of n fine cells
__device__ inline int fine_of(int I, int n) { return min(2 * I, n); }

// Linear interpolation of the coarse vector ec (row length cnx + 1) at fine
// node (x, y) of a level with fnx x fny cells: along x, along y, or along the
// cell diagonal ((x+1)/2, (y-1)/2) -- ((x-1)/2, (y+1)/2) when both are odd.
__device__ inline double prolong_at(int fnx, int fny, int cnx, const double* ec, int x, int y) {
  const Par1 px = parents1(x, fnx), py = parents1(y, fny);
  const int64_t cs = cnx + 1;
  if (px.a == px.b && py.a == py.b) return ec[py.a * cs + px.a];
  if (py.a == py.b) return 0.5 * (ec[py.a * cs + px.a] + ec[py.a * cs + px.b]);
  if (px.a == px.b) return 0.5 * (ec[py.a * cs + px.a] + ec[py.b * cs + px.a]);
  return 0.5 * (ec[py.a * cs + px.b] + ec[py.b * cs + px.a]);
}

// weight of coarse node (I, J) in prolong_at at fine node (x, y)
__device__ inline double pweight(int fnx, int fny, int x, int y, int I, int J) {
  const Par1 px = parents1(x, fnx), py = parents1(y, fny);
  const bool sx = px.a == px.b, sy = py.a == py.b;
  if (sx && sy) return (px.a == I && py.a == J) ? 1.0 : 0.0;
  if (sy) return (py.a == J && (px.a == I || px.b == I)) ? 0.5 : 0.0;
  if (sx) return (px.a == I && (py.a == J || py.b == J)) ? 0.5 : 0.0;
  return ((px.b == I && py.a == J) || (px.a == I && py.b == J)) ? 0.5 : 0.0;
}

__device__ inline double wave_sum_mg(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int NW>
__device__ inline double block_sum_mg(double v, double* sh) {
  v = wave_sum_mg(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NW; ++i) s += sh[i];
  __syncthreads();
  return s;
}

template <int NW>
__device__ inline double sum_partials_mg(const double* p, int n, double* sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  s = wave_sum_mg(s);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += sh[i];
  __syncthreads();
  return t;
}

// (S x)_row for S = -K on free rows, from the matrix stored as V: the
// symmetric blocks at four modes per face, else the face-ELL values
// (accumulated in FP64; values read once per product by nontemporal loads,
// as in the CG product).
template <int NT, class V>
__device__ inline double s_row(const DevPattern& P, const TraceOp<V>& T, const V* x, int64_t r) {
  const int64_t f = r / NT;
  double acc = 0.0;
  if constexpr (kSymBlocks<NT>) {
    acc = sym_row<NT>(P.nbr, T.blk, T.nf, f, (int)(r - f * NT), [&](int64_t g, int b) { return (double)x[g * NT + b]; });
  } else {
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int64_t g = (int64_t)__ldg(P.bcol + f * kSlots + s) * NT;
#pragma unroll
      for (int b = 0; b < NT; ++b)
        acc = fma((double)__builtin_nontemporal_load(T.vals + (int64_t)(s * NT + b) * T.N + r), (double)x[g + b],
                  acc);
    }
  }
  return T.dmask[r] ? 0.0 : -acc;
}

// ------------------------------------------------------------ set-up
// vertex mask: a vertex is fixed when a face through it has a Dirichlet
// mode-0 dof.
__global__ void mg_vmask_kernel(MgData M, TraceSystem T) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= M.nv) return;
  unsigned char m = 0;
  for (int k = M.vf_ptr[v]; k < M.vf_ptr[v + 1]; ++k) m |= T.dmask[(int64_t)M.vf_idx[k] * T.nt];
  M.lev[1].mask[v] = m;
}

// A1_c = P_c^T (-Ate_c) P_c for every cell, masked.
template <int NT>
__global__ void mg_cell_coarse_kernel(DevMesh mesh, MgData M, TraceSystem T, const double* __restrict__ Ate) {
  constexpr int NTR = 3 * NT;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= mesh.nc) return;
  int vg[3];
  unsigned char vm[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    vg[v] = mesh.cn[3 * c + v];
    vm[v] = M.lev[1].mask[vg[v]];
  }
  // P_c rows of modes 0 and 1 of each local face: weights on local vertices
  // local face lf joins local vertices (lf+1)%3, (lf+2)%3
  double Pw[3][2][3];
#pragma unroll
  for (int lf = 0; lf < 3; ++lf) {
    const int a = (lf + 1) % 3, b = (lf + 2) % 3;
    const int lo = vg[a] < vg[b] ? a : b, hi = vg[a] < vg[b] ? b : a;
    const int64_t f = mesh.cf[3 * c + lf];
    const bool d0 = T.dmask[f * NT] != 0, d1 = T.dmask[f * NT + 1] != 0;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      const double keep = vm[w] ? 0.0 : 1.0;
      Pw[lf][0][w] = d0 ? 0.0 : keep * ((w == a || w == b) ? 0.5 : 0.0);
      Pw[lf][1][w] = d1 ? 0.0 : keep * (w == hi ? kS3 : (w == lo ? -kS3 : 0.0));
    }
  }
  const double* at = Ate + c * NTR * NTR;
  // Q = (-Ate) P_c  (NTR x 3), then A = P_c^T Q
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int lf = 0; lf < 3; ++lf)
#pragma unroll
    for (int md = 0; md < 2; ++md) {
      const int row = lf * NT + md;
      double q[3] = {0, 0, 0};
#pragma unroll
      for (int lg = 0; lg < 3; ++lg)
#pragma unroll
        for (int me = 0; me < 2; ++me) {
          const double s = -at[row * NTR + lg * NT + me];
#pragma unroll
          for (int w = 0; w < 3; ++w) q[w] = fma(s, Pw[lg][me][w], q[w]);
        }
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int w = 0; w < 3; ++w) A[u][w] = fma(Pw[lf][md][u], q[w], A[u][w]);
    }
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int w = 0; w < 3; ++w) M.cellA[c * 9 + u * 3 + w] = A[u][w];
}

// Level-1 stencils gathered from the (up to) six cells around each vertex.
// One hierarchy (map null: Ms[0] over the grid) or one per block via map.
__global__ void mg_assemble_l1_kernel(const MgData* __restrict__ Ms, const BlockMap* __restrict__ map) {
  const BlockMap bm = map ? map[blockIdx.x] : BlockMap{0, (int)blockIdx.x};
  const MgData& M = Ms[bm.inst];
  const int64_t v = (int64_t)bm.lb * blockDim.x + threadIdx.x;
  const MgLevel& L = M.lev[1];
  if (v >= L.n) return;
  const int nx = M.nx, ny = M.ny;
  const int i = (int)(v % (nx + 1)), j = (int)(v / (nx + 1));
  double st[kStencil] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (!L.mask[v]) {
    // (cell grid position, upper?, local index of v)
    const int ca[6] = {i, i - 1, i, i - 1, i - 1, i};
    const int cb[6] = {j, j, j - 1, j, j - 1, j - 1};
    const int up[6] = {0, 0, 0, 1, 1, 1};
    const int lv[6] = {0, 1, 2, 0, 1, 2};
    for (int t = 0; t < 6; ++t) {
      const int a = ca[t], b = cb[t];
      if (a < 0 || a >= nx || b < 0 || b >= ny) continue;
      const int64_t c = 2 * ((int64_t)b * nx + a) + up[t];
      // local vertex grid positions of the cell
      int px[3], py[3];
      if (!up[t]) { px[0] = a; py[0] = b; px[1] = a + 1; py[1] = b; px[2] = a; py[2] = b + 1; }
      else { px[0] = a + 1; py[0] = b; px[1] = a + 1; py[1] = b + 1; px[2] = a; py[2] = b + 1; }
      for (int w = 0; w < 3; ++w) {
        const int s = slot_of(px[w] - i, py[w] - j);
        if (s >= 0) st[s] += M.cellA[c * 9 + lv[t] * 3 + w];
      }
    }
  }
  for (int s = 0; s < kStencil; ++s) L.a[s * L.n + v] = L.mask[v] ? (s == 0 ? 1.0 : 0.0) : st[s];
  L.dinv[v] = L.mask[v] ? 0.0 : 1.0 / st[0];
}

// Galerkin coarse stencil R A P on the next level: R = P^T with the weights
// of prolong_at.  Eight lanes per coarse node: lane t < 7 takes the fine
// node at support offset t, and a fixed shuffle tree sums the lanes' stencils.
template <class VS>
__global__ __launch_bounds__(256) void mg_rap_kernel(VS sel) {
  int lb;
  const VRec& R = vrec_of(sel, lb);
  const MgLevel& F = R.F;
  const MgLevel& Cc = R.C;
  const int64_t gid = (int64_t)lb * blockDim.x + threadIdx.x;
  const int64_t C = gid >> 3;
  const int t = (int)(gid & 7);
  const int cnx = Cc.nx, fnx = F.nx, fny = F.ny;
  const bool live = C < Cc.n;
  const int I = live ? (int)(C % (cnx + 1)) : 0, J = live ? (int)(C / (cnx + 1)) : 0;
  const unsigned char cm = live ? Cc.mask[C] : 1;  // written by mg_cmask_kernel
  double st[kStencil] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (!cm && t < 7) {
    const int fx = fine_of(I, fnx) + (t == 1 || t == 5 ? 1 : (t == 2 || t == 6 ? -1 : 0));
    const int fy = fine_of(J, fny) + (t == 3 || t == 6 ? 1 : (t == 4 || t == 5 ? -1 : 0));
    const int64_t fi = (int64_t)fy * (fnx + 1) + fx;
    const bool in = fx >= 0 && fx <= fnx && fy >= 0 && fy <= fny && !F.mask[fi];
    const double w1 = in ? pweight(fnx, fny, fx, fy, I, J) : 0.0;
    if (w1 != 0.0) {
      for (int s = 0; s < F.npts; ++s) {
        const double av = F.a[s * F.n + fi];
        if (av == 0.0) continue;
        const int gx = fx + kOffX[s], gy = fy + kOffY[s];
        if (gx < 0 || gx > fnx || gy < 0 || gy > fny) continue;
        if (F.mask[(int64_t)gy * (fnx + 1) + gx]) continue;
        // coarse parents of fine node g, as prolong_at weighs them
        const Par1 px = parents1(gx, fnx), py = parents1(gy, fny);
        int qx[2], qy[2], np = 2;
        double w2 = 0.5;
        if (px.a == px.b && py.a == py.b) { qx[0] = px.a; qy[0] = py.a; np = 1; w2 = 1.0; }
        else if (py.a == py.b) { qx[0] = px.a; qx[1] = px.b; qy[0] = qy[1] = py.a; }
        else if (px.a == px.b) { qx[0] = qx[1] = px.a; qy[0] = py.a; qy[1] = py.b; }
        else { qx[0] = px.b; qy[0] = py.a; qx[1] = px.a; qy[1] = py.b; }
        for (int q = 0; q < np; ++q) {
          if (Cc.mask[(int64_t)qy[q] * (cnx + 1) + qx[q]]) continue;  // zero column of P
          const int sl = slot_of(qx[q] - I, qy[q] - J);
#pragma unroll
          for (int u = 0; u < kStencil; ++u)
            if (u == sl) st[u] = fma(w1 * av, w2, st[u]);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kStencil; ++u) {
    st[u] += __shfl_xor(st[u], 1, 8);
    st[u] += __shfl_xor(st[u], 2, 8);
    st[u] += __shfl_xor(st[u], 4, 8);
  }
  if (t != 0 || !live) return;
#pragma unroll
  for (int u = 0; u < kStencil; ++u) Cc.a[u * Cc.n + C] = cm ? (u == 0 ? 1.0 : 0.0) : st[u];
  Cc.dinv[C] = cm ? 0.0 : 1.0 / st[0];
}

// Inverse of the coarsest operator (at most kCoarseDirect nodes), by
// Gauss-Jordan with partial pivoting on [A | I] in LDS: one thread per entry,
// the pivot row chosen by thread 0.  Fixed nodes have identity rows.
// Block b inverts the coarsest level of Ms[b] (skipped above kCoarseDirect).
__global__ __launch_bounds__(1024) void mg_coarse_inverse_kernel(const MgData* __restrict__ Ms) {
  constexpr int W = 2 * kCoarseDirect;
  __shared__ double m[kCoarseDirect * W];
  __shared__ int piv;
  const MgData& M = Ms[blockIdx.x];
  const MgLevel& L = M.lev[M.nlev - 1];
  double* const cinv = M.cinv;
  const int n = (int)L.n;
  if (n > kCoarseDirect) return;
  for (int e = threadIdx.x; e < n * W; e += blockDim.x) m[e] = 0.0;
  __syncthreads();
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int i = v % (L.nx + 1), j = v / (L.nx + 1);
    m[v * W + n + v] = 1.0;
    for (int s = 0; s < L.npts; ++s) {
      const int gx = i + kOffX[s], gy = j + kOffY[s];
      if (gx < 0 || gx > L.nx || gy < 0 || gy > L.ny) continue;
      m[v * W + gy * (L.nx + 1) + gx] += L.a[s * L.n + v];
    }
  }
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    if (threadIdx.x == 0) {
      int best = c;
      for (int r = c + 1; r < n; ++r)
        if (fabs(m[r * W + c]) > fabs(m[best * W + c])) best = r;
      piv = best;
    }
    __syncthreads();
    const int pr = piv;
    if (pr != c)
      for (int e = threadIdx.x; e < 2 * n; e += blockDim.x) {
        const double a = m[c * W + e];
        m[c * W + e] = m[pr * W + e];
        m[pr * W + e] = a;
      }
    __syncthreads();
    const double pv = m[c * W + c];
    // each thread owns one (row, column) entry of the update
    double upd = 0.0;
    const int row = threadIdx.x / (2 * n), col = threadIdx.x % (2 * n);
    const bool own = threadIdx.x < n * 2 * n;
    if (own) {
      const double pc = m[c * W + col] / pv;
      upd = row == c ? pc : m[row * W + col] - m[row * W + c] * pc;
    }
    __syncthreads();
    if (own) m[row * W + col] = upd;
    __syncthreads();
  }
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) cinv[e] = m[(e / n) * W + n + e % n];
}

// LDS image of the tail levels: per level its stencils (kStencil n) then
// dinv (n); blockIdx.y selects the hierarchy Ms[y].
__global__ void mg_tailpack_kernel(const MgData* __restrict__ Ms) {
  const MgData& M = Ms[blockIdx.y];
  int o = 0;
  for (int l = M.tail; l < M.nlev; ++l) {
    const int n = (int)M.lev[l].n;
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < (kStencil + 1) * n; u += gridDim.x * blockDim.x)
      M.tailpack[o + u] = u < kStencil * n ? M.lev[l].a[u] : M.lev[l].dinv[u - kStencil * n];
    o += (kStencil + 1) * n;
  }
}

// Coarse masks (a coarse node is fixed when its fine node is), written
// before mg_rap_kernel reads the masks of neighbouring nodes.
template <class VS>
__global__ void mg_cmask_kernel(VS sel) {
  int lb;
  const VRec& R = vrec_of(sel, lb);
  const MgLevel& F = R.F;
  const MgLevel& Cc = R.C;
  const int64_t C = (int64_t)lb * blockDim.x + threadIdx.x;
  if (C >= Cc.n) return;
  const int I = (int)(C % (Cc.nx + 1)), J = (int)(C / (Cc.nx + 1));
  Cc.mask[C] = F.mask[(int64_t)fine_of(J, F.ny) * (F.nx + 1) + fine_of(I, F.nx)];
}

// S P0 restricted to each row: sp[j N + r] = sum over the free gathered
// columns (g, b < 2) of S[r, (g, b)] P0[(g, b), fq4[4 f + j]], where
// P0[(g, 0), lo] = P0[(g, 0), hi] = 1/2 and P0[(g, 1), hi] = -P0[(g, 1), lo]
// = 1/(2 sqrt 3).  Every face of the two cells of face f has its end points
// among fq4[4f .. 4f+3].
template <int NT>
__global__ __launch_bounds__(256) void mg_sp_kernel(DevPattern P, TraceSystem T, MgData M, float* sp32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T.N) return;
  const int64_t f = r / NT;
  int q[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) q[j] = M.fq4[4 * f + j];
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  // entry S[r, (g, b)] = -kv of a mode b < 2 column
  auto add = [&](int64_t g, int b, double kv) {
    if (b >= 2 || T.dmask[g * NT + b]) return;
    const int lo = M.fv[2 * g], hi = M.fv[2 * g + 1];
    const double sv = -kv;
    const double cl = b == 0 ? 0.5 : -kS3, ch = b == 0 ? 0.5 : kS3;
    bool dl = false, dh = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool ml = !dl && q[j] == lo, mh = !dh && q[j] == hi;
      w[j] = ml ? fma(sv, cl, w[j]) : w[j];
      w[j] = mh ? fma(sv, ch, w[j]) : w[j];
      dl = dl || ml;
      dh = dh || mh;
    }
  };
  if (!T.dmask[r]) {
    if constexpr (kSymBlocks<NT>) {
      sym_row_entries<NT>(P.nbr, T.blk, T.nf, f, (int)(r - f * NT), add);
    } else {
      for (int sl = 0; sl < kSlots; ++sl) {
        const int64_t g = P.bcol[f * kSlots + sl];
#pragma unroll
        for (int b = 0; b < 2; ++b) add(g, b, T.vals[(int64_t)(sl * NT + b) * T.N + r]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    M.sp[(int64_t)j * T.N + r] = w[j];
    if (sp32) sp32[(int64_t)j * T.N + r] = (float)w[j];
  }
}

// ------------------------------------------------------- level 0 (trace)
// update(i) with the multigrid preconditioner: x += alpha p; r -= alpha q;
// e0 = omega Dinv r (first smoothing step, preconditioner storage W);
// partial (r, r).
template <int NT, class W, class G>
__global__ __launch_bounds__(64 * NT) void mg_update_kernel(TraceOp<W> T, Krylov<double> v, MgVecs<W> mv, G gsel) {
  constexpr int NTH = 64 * NT;
  __shared__ double rs[NTH];
  __shared__ double sh[NT];
  const Seg S = seg_of(gsel);
  CgState* st = S.st;
  if (st->done) return;
  const int i = st->i_upd;
  const int lb = blockIdx.x - S.b0;
  if (lb == 0 && threadIdx.x == 0) st->i_pre = i;
  const size_t nbt = (size_t)v.nblocks;
  const double pq = sum_partials_mg<NT>(v.part + S.b0, S.nb, sh);
  const double rz = st->rz[i & 1];
  if (!(pq > 0.0)) {
    if (lb == 0 && threadIdx.x == 0) {
      st->done = 3;
      st->it = i - 1;
    }
    return;
  }
  const double alpha = rz / pq, om = st->omega;
  const double* __restrict__ p = (i & 1) ? v.p1 : v.p0;
  double prr = 0.0;
  for (int64_t r0 = S.lo + (int64_t)lb * NTH; r0 < S.hi; r0 += (int64_t)S.nb * NTH) {
    const int64_t r = r0 + threadIdx.x;
    double rn = 0.0;
    if (r < S.hi) {
      v.x[r] = fma(alpha, p[r], v.x[r]);
      rn = fma(-alpha, v.q[r], v.r[r]);
      v.r[r] = rn;
      prr = fma(rn, rn, prr);
    }
    jacobi_apply<NT>(T.Dinv, T.nf, rs, rn, r, r < S.hi, [&](double z) { mv.e0[r] = (W)(om * z); });
  }
  const double b = block_sum_mg<NT>(prr, sh);
  if (threadIdx.x == 0) v.part[nbt * 3 + blockIdx.x] = b;
}

// init (update(0)) for the multigrid path: r0 = b~ = -(rhs - K uD), x = p = 0,
// e0 = omega Dinv r0.  uD is zero off the Dirichlet dofs, so the lift reads
// only the matrix columns of Dirichlet dofs (the mask decides per column).
template <int NT, class W, class G>
__global__ __launch_bounds__(64 * NT) void mg_init_kernel(DevPattern P, TraceSystem T64, TraceOp<W> T,
                                                          Krylov<double> v, MgVecs<W> mv, G gsel) {
  constexpr int NTH = 64 * NT;
  __shared__ double rs[NTH];
  __shared__ double sh[NT];
  const Seg S = seg_of(gsel);
  const int lb = blockIdx.x - S.b0;
  const double om = S.st->omega;
  double prr = 0.0;
  for (int64_t r0 = S.lo + (int64_t)lb * NTH; r0 < S.hi; r0 += (int64_t)S.nb * NTH) {
    const int64_t r = r0 + threadIdx.x;
    double b = 0.0;
    if (r < S.hi) {
      if (!T64.dmask[r]) {
        const int64_t f = r / NT;
        double s = T64.rhs[r];
        if constexpr (kSymBlocks<NT>) {
          s -= sym_row_lift<NT>(P.nbr, T64.blk, T64.nf, T64.dmask, T64.uD, f, (int)(r - f * NT));
        } else {
#pragma unroll
          for (int sl = 0; sl < kSlots; ++sl) {
            const int64_t g = (int64_t)P.bcol[f * kSlots + sl] * NT;
#pragma unroll
            for (int c = 0; c < NT; ++c)
              if (T64.dmask[g + c]) s -= T64.vals[(int64_t)(sl * NT + c) * T64.N + r] * T64.uD[g + c];
          }
        }
        b = -s;
      }
      v.r[r] = b;
      v.x[r] = 0.0;
      v.p0[r] = 0.0;
      v.p1[r] = 0.0;
      prr = fma(b, b, prr);
    }
    jacobi_apply<NT>(T.Dinv, T.nf, rs, b, r, r < S.hi, [&](double z) { mv.e0[r] = (W)(om * z); });
  }
  const double bb = block_sum_mg<NT>(prr, sh);
  if (threadIdx.x == 0) v.part[(size_t)v.nblocks * 3 + blockIdx.x] = bb;
  if (lb == 0 && threadIdx.x == 0) {
    S.st->done = 0;
    S.st->it = 0;
    S.st->pq = 0.0;
    S.st->i_pre = 0;  // the first smoother call completes update(0)
  }
}

// res0 = r - S e0 over all rows (of every instance when packed; v.st null:
// no convergence check).  For even nt the lane of mode 0 of a face also
// writes the face's restriction weights fc (P0^T res0 at its lower and
// higher end: 0.5 res0_0 -+ res0_1 / (2 sqrt 3)), with mode 1 taken from the
// next lane (a face's rows are consecutive lanes of one wavefront).
template <int NT, class W>
__global__ __launch_bounds__(256) void mg_resid0_kernel(DevPattern P, TraceOp<W> T, Krylov<double> v,
                                                        MgVecs<W> mv, double* __restrict__ fc) {
  if (v.st && v.st->done) return;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < T.N; r += (int64_t)gridDim.x * blockDim.x) {
    // every row's product, the Dirichlet rows masked afterwards: the loads
    // do not wait for the mask
    const double sr = s_row<NT, W>(P, T, mv.e0, r);
    const W res = T.dmask[r] ? (W)0 : (W)(v.r[r] - sr);
    mv.res0[r] = res;
    if constexpr (NT % 2 == 0) {
      // the grid is a multiple of 64 threads and N of NT: lane r + 1 of a
      // mode-0 lane r is in the same wavefront and active
      const double r1 = (double)__shfl_down((double)res, 1, 64);
      if (r % NT == 0) {
        const double r0 = (double)res;
        fc[2 * (r / NT)] = 0.5 * r0 + (-kS3) * r1;
        fc[2 * (r / NT) + 1] = 0.5 * r0 + kS3 * r1;
      }
    }
  }
}

// vf2: the vertex -> face entries with the side of the vertex on the face.
__global__ void mg_vf2_kernel(MgData M) {
  const int64_t vv = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (vv >= M.nv) return;
  for (int k = M.vf_ptr[vv]; k < M.vf_ptr[vv + 1]; ++k) {
    const int f = M.vf_idx[k];
    M.vf2[k] = 2 * f + (M.fv[2 * f] == vv ? 0 : 1);
  }
}

// r1 = P0^T res0, vertex by vertex over the incident faces.
template <int NT, class W>
__global__ __launch_bounds__(256) void mg_restrict0_kernel(Krylov<double> v, MgData M, MgVecs<W> mv) {
  if (v.st && v.st->done) return;
  const MgLevel L = M.lev[1];
  const int64_t vv = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (vv >= L.n) return;
  double s = 0.0;
  if (!L.mask[vv]) {
    if constexpr (NT % 2 == 0) {
      // the weights the residual kernel wrote per face end, summed in the
      // vertex's face order; the first 8 indices are loaded before any
      // weight, so that the gathers are in flight together
      const int k0 = M.vf_ptr[vv], k1 = M.vf_ptr[vv + 1];
      int id[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) id[u] = k0 + u < k1 ? M.vf2[k0 + u] : -1;
      double w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) w[u] = id[u] >= 0 ? M.fc[id[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (id[u] >= 0) s += w[u];
      for (int k = k0 + 8; k < k1; ++k) s += M.fc[M.vf2[k]];
    } else {
      // modes 0 and 1 of the incident faces, in the vertex's face order,
      // the first 8 faces' loads issued together
      const int k0 = M.vf_ptr[vv], k1 = M.vf_ptr[vv + 1];
      int64_t fa[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) fa[u] = k0 + u < k1 ? (int64_t)M.vf_idx[k0 + u] : -1;
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        t[u] = 0.0;
        if (fa[u] >= 0) {
          const int64_t f = fa[u];
          const bool lo = M.fv[2 * f] == vv;
          t[u] = 0.5 * (double)mv.res0[f * NT] + (lo ? -kS3 : kS3) * (double)mv.res0[f * NT + 1];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (fa[u] >= 0) s += t[u];
      for (int k = k0 + 8; k < k1; ++k) {
        const int64_t f = M.vf_idx[k];
        const bool lo = M.fv[2 * f] == vv;
        s += 0.5 * (double)mv.res0[f * NT] + (lo ? -kS3 : kS3) * (double)mv.res0[f * NT + 1];
      }
    }
  }
  L.r[vv] = s;
}

// z = e0 + P0 e1 + omega Dinv (res0 - (S P0) e1): the post-smoothing step.
// S (e0 + P0 e1) = S e0 + (S P0) e1 and res0 = r - S e0, so the trace matrix
// is not read again; and e0 = omega Dinv r, so z = P0 e1 + omega Dinv (r +
// res0 - (S P0) e1) and e0 is not read either (r is, for the partial (r, z)
// of the next iteration).
template <int NT, class W, class G>
__global__ __launch_bounds__(64 * NT) void mg_smooth0_kernel(TraceOp<W> T, Krylov<double> v, MgData M,
                                                             MgVecs<W> mv, G gsel) {
  constexpr int NTH = 64 * NT;
  __shared__ double rs[NTH];
  __shared__ double sh[NT];
  const Seg S = seg_of(gsel);
  CgState* st = S.st;
  if (st->done) return;
  const int i = st->i_pre;
  const int lb = blockIdx.x - S.b0;
  if (lb == 0 && threadIdx.x == 0) st->i_spmv = i + 1;
  const double om = st->omega;
  const double* __restrict__ e1 = M.lev[1].e;
  double prz = 0.0;
  for (int64_t r0 = S.lo + (int64_t)lb * NTH; r0 < S.hi; r0 += (int64_t)S.nb * NTH) {
    const int64_t r = r0 + threadIdx.x;
    double rv = 0.0, rr = 0.0, pe = 0.0;
    if (r < S.hi) {
      // every row formed, the Dirichlet rows masked afterwards (the loads do
      // not wait for the mask)
      rv = v.r[r];
      const int64_t f = r / NT;
      double s = rv + (double)mv.res0[r];
      // the face's 4 vertices, its lower and higher end point first
      const int4 q = *reinterpret_cast<const int4*>(M.fq4 + 4 * f);
      const double ev[4] = {e1[q.x], e1[q.y], e1[q.z], e1[q.w]};
#pragma unroll
      for (int j = 0; j < 4; ++j) s = fma(-(double)mv.sp[(int64_t)j * T.N + r], ev[j], s);
      const int a = (int)(r - f * NT);
      const double pc = a == 0 ? 0.5 * (ev[0] + ev[1]) : (a == 1 ? kS3 * (ev[1] - ev[0]) : 0.0);
      if (!T.dmask[r]) {
        rr = s;
        pe = pc;
      }
    }
    jacobi_apply<NT>(T.Dinv, T.nf, rs, rr, r, r < S.hi, [&](double z) {
      const double zv = fma(om, z, pe);
      v.z[r] = zv;
      prz = fma(rv, zv, prz);
    });
  }
  const double b = block_sum_mg<NT>(prz, sh);
  if (threadIdx.x == 0) v.part[(size_t)v.nblocks * (1 + ((i + 1) & 1)) + blockIdx.x] = b;
}

// ------------------------------------------------------- vertex levels
// (A x)_v on a level, with x given as a function of (index, i, j), so that
// fused kernels can form x on the fly.
template <class X>
__device__ inline double stencil_apply(const MgLevel& L, int64_t v, int i, int j, const X& xv) {
  const int64_t row = L.nx + 1;
  double s = L.a[v] * xv(v, i, j);
  if (i < L.nx) s = fma(L.a[L.n + v], xv(v + 1, i + 1, j), s);
  if (i > 0) s = fma(L.a[2 * L.n + v], xv(v - 1, i - 1, j), s);
  if (j < L.ny) s = fma(L.a[3 * L.n + v], xv(v + row, i, j + 1), s);
  if (j > 0) s = fma(L.a[4 * L.n + v], xv(v - row, i, j - 1), s);
  if (i < L.nx && j > 0) s = fma(L.a[5 * L.n + v], xv(v + 1 - row, i + 1, j - 1), s);
  if (i > 0 && j < L.ny) s = fma(L.a[6 * L.n + v], xv(v - 1 + row, i - 1, j + 1), s);
  if (L.npts > 7) {
    if (i < L.nx && j < L.ny) s = fma(L.a[7 * L.n + v], xv(v + 1 + row, i + 1, j + 1), s);
    if (i > 0 && j > 0) s = fma(L.a[8 * L.n + v], xv(v - 1 - row, i - 1, j - 1), s);
  }
  return s;
}

// The vertex-level kernels take a selector (VOne, VGroup) that names the
// instance of the block: its fine level F, coarse level C and CG state.

// two pre-smoothing sweeps from zero in one pass: e1 = om dinv r is formed at
// the neighbours on the fly, e = e1 + om dinv (r - A e1)
template <class VS>
__global__ __launch_bounds__(256) void v_pre_kernel(VS sel, double om) {
  int lb;
  const VRec& R = vrec_of(sel, lb);
  if (R.st->done) return;
  const MgLevel& L = R.F;
  const int64_t v = (int64_t)lb * blockDim.x + threadIdx.x;
  if (v >= L.n) return;
  const int i = (int)(v % (L.nx + 1)), j = (int)(v / (L.nx + 1));
  auto e1 = [&](int64_t u, int, int) { return om * L.dinv[u] * L.r[u]; };
  L.e[v] = e1(v, i, j) + om * L.dinv[v] * (L.r[v] - stencil_apply(L, v, i, j, e1));
}

// second post-smoothing sweep: e = t + om dinv (r - A t)
template <class VS>
__global__ __launch_bounds__(256) void v_smooth_kernel(VS sel, double om) {
  int lb;
  const VRec& R = vrec_of(sel, lb);
  if (R.st->done) return;
  const MgLevel& L = R.F;
  const int64_t v = (int64_t)lb * blockDim.x + threadIdx.x;
  if (v >= L.n) return;
  const int i = (int)(v % (L.nx + 1)), j = (int)(v / (L.nx + 1));
  const double* __restrict__ in = L.t;
  auto xv = [&](int64_t u, int, int) { return in[u]; };
  L.e[v] = in[v] + om * L.dinv[v] * (L.r[v] - stencil_apply(L, v, i, j, xv));
}

// fine residual r - A e at (fx, fy); 0 outside the grid and on fixed nodes.
// With one_sweep the smoothed iterate is the single sweep from zero,
// e = om dinv r, formed on the fly.
__device__ inline double fine_res(const MgLevel& F, int fx, int fy, bool one_sweep, double om) {
  if (fx < 0 || fx > F.nx || fy < 0 || fy > F.ny) return 0.0;
  const int64_t fi = (int64_t)fy * (F.nx + 1) + fx;
  if (F.mask[fi]) return 0.0;
  if (one_sweep) {
    auto e1 = [&](int64_t u, int, int) { return om * F.dinv[u] * F.r[u]; };
    return F.r[fi] - stencil_apply(F, fi, fx, fy, e1);
  }
  auto ev = [&](int64_t u, int, int) { return F.e[u]; };
  return F.r[fi] - stencil_apply(F, fi, fx, fy, ev);
}

// coarse r = R (r - A e), R = P^T: the fine node of C and its six stencil
// neighbours, weighted as prolong_at weighs C there.  Eight lanes per
// coarse node: lane t < 7 forms the fine residual at stencil offset t, and a
// fixed shuffle tree sums them.
template <class VS>
__global__ __launch_bounds__(256) void v_restrict_kernel(VS sel, int one_sweep, double om) {
  int lb;
  const VRec& R = vrec_of(sel, lb);
  if (R.st->done) return;
  const MgLevel& F = R.F;
  const MgLevel& Cc = R.C;
  const int64_t g = (int64_t)lb * blockDim.x + threadIdx.x;
  const int64_t C = g >> 3;
  const int t = (int)(g & 7);
  double val = 0.0;
  if (C < Cc.n && t < 7 && !Cc.mask[C]) {
    const int I = (int)(C % (Cc.nx + 1)), J = (int)(C / (Cc.nx + 1));
    const int x = fine_of(I, F.nx) + (t == 1 || t == 5 ? 1 : (t == 2 || t == 6 ? -1 : 0));
    const int y = fine_of(J, F.ny) + (t == 3 || t == 6 ? 1 : (t == 4 || t == 5 ? -1 : 0));
    const double w = (x < 0 || x > F.nx || y < 0 || y > F.ny) ? 0.0 : pweight(F.nx, F.ny, x, y, I, J);
    if (w != 0.0) val = w * fine_res(F, x, y, one_sweep != 0, om);
  }
  val += __shfl_xor(val, 1, 8);
  val += __shfl_xor(val, 2, 8);
  val += __shfl_xor(val, 4, 8);
  if (t == 0 && C < Cc.n) Cc.r[C] = val;
}

// prolongation and the first post-smoothing sweep in one pass:
// e' = e + P ec on free nodes (formed at the neighbours on the fly),
// out = e' + om dinv (r - A e').  With one_sweep the pre-smoothed iterate is
// e = om dinv r (formed on the fly) and out is the level's final e.
template <class VS>
__global__ __launch_bounds__(256) void v_post_kernel(VS sel, int one_sweep, double om) {
  int lb;
  const VRec& R = vrec_of(sel, lb);
  if (R.st->done) return;
  const MgLevel& F = R.F;
  const MgLevel& Cc = R.C;
  const int64_t v = (int64_t)lb * blockDim.x + threadIdx.x;
  if (v >= F.n) return;
  const int i = (int)(v % (F.nx + 1)), j = (int)(v / (F.nx + 1));
  if (one_sweep) {
    auto ep = [&](int64_t u, int x, int y) {
      const double e = om * F.dinv[u] * F.r[u];
      return F.mask[u] ? e : e + prolong_at(F.nx, F.ny, Cc.nx, Cc.e, x, y);
    };
    F.e[v] = ep(v, i, j) + om * F.dinv[v] * (F.r[v] - stencil_apply(F, v, i, j, ep));
  } else {
    auto ep = [&](int64_t u, int x, int y) {
      return F.mask[u] ? F.e[u] : F.e[u] + prolong_at(F.nx, F.ny, Cc.nx, Cc.e, x, y);
    };
    F.t[v] = ep(v, i, j) + om * F.dinv[v] * (F.r[v] - stencil_apply(F, v, i, j, ep));
  }
}

// Coarsest level too large for v_tail_kernel: `sweeps` damped Jacobi sweeps
// from zero in one block, vectors in global memory.
__global__ __launch_bounds__(1024) void v_coarsest_kernel(MgLevel L, double om, int sweeps, const CgState* st) {
  if (st->done) return;
  for (int64_t v = threadIdx.x; v < L.n; v += blockDim.x) L.e[v] = om * L.dinv[v] * L.r[v];
  __syncthreads();
  for (int it = 1; it < sweeps; ++it) {
    for (int64_t v = threadIdx.x; v < L.n; v += blockDim.x) {
      const int i = (int)(v % (L.nx + 1)), j = (int)(v / (L.nx + 1));
      auto xv = [&](int64_t u, int, int) { return L.e[u]; };
      L.t[v] = L.e[v] + om * L.dinv[v] * (L.r[v] - stencil_apply(L, v, i, j, xv));
    }
    __syncthreads();
    for (int64_t v = threadIdx.x; v < L.n; v += blockDim.x) L.e[v] = L.t[v];
    __syncthreads();
  }
}

// Level of the one-workgroup tail: stencils, diagonal inverse and vectors
// all in LDS.  A fixed node has dinv = 0 (its stencil row is the identity).
struct TailLevel {
  int nx, ny, n, npts;
  const double* a;     // a[s*n + v]
  const double* dinv;
};

__device__ inline double tail_apply(const TailLevel& L, const double* x, int v, int i, int j) {
  const int row = L.nx + 1;
  double s = L.a[v] * x[v];
  if (i < L.nx) s = fma(L.a[L.n + v], x[v + 1], s);
  if (i > 0) s = fma(L.a[2 * L.n + v], x[v - 1], s);
  if (j < L.ny) s = fma(L.a[3 * L.n + v], x[v + row], s);
  if (j > 0) s = fma(L.a[4 * L.n + v], x[v - row], s);
  if (i < L.nx && j > 0) s = fma(L.a[5 * L.n + v], x[v + 1 - row], s);
  if (i > 0 && j < L.ny) s = fma(L.a[6 * L.n + v], x[v - 1 + row], s);
  if (L.npts > 7) {
    if (i < L.nx && j < L.ny) s = fma(L.a[7 * L.n + v], x[v + 1 + row], s);
    if (i > 0 && j > 0) s = fma(L.a[8 * L.n + v], x[v - 1 - row], s);
  }
  return s;
}

// one Jacobi sweep: out = in + om dinv (r - A in)
__device__ inline void tail_sweep(const TailLevel& L, const double* in, double* out, const double* r, double om) {
  for (int v = threadIdx.x; v < L.n; v += blockDim.x) {
    const int i = v % (L.nx + 1), j = v / (L.nx + 1);
    out[v] = in[v] + om * L.dinv[v] * (r[v] - tail_apply(L, in, v, i, j));
  }
}

// Levels lt = tail .. nlev-1 of the V-cycle in one workgroup (their sizes
// sum to at most kTailNodes), everything in LDS: reads lev[lt].r, writes
// lev[lt].e.  Block b serves the hierarchy Ms[b] with CG state sts[b] (one
// block for a single problem, one per instance for a group).
// Same sweeps as the per-level kernels: two pre-smoothing sweeps,
// restriction of the residual, coarse_sweeps sweeps on the coarsest level,
// prolongation and two post-smoothing sweeps.
__global__ __launch_bounds__(kTailThreads) void v_tail_kernel(const MgData* __restrict__ Ms, const CgState* sts) {
  __shared__ double pool[(kStencil + 4) * kTailNodes];
  const MgData& M = Ms[blockIdx.x];
  if (sts[blockIdx.x].done) return;
  const int lt = M.tail;
  const double om = M.omega_v;
  double* const E = pool;                    // scratch / residual
  double* const Tt = pool + kTailNodes;      // iterate
  double* const R = pool + 2 * kTailNodes;   // right-hand sides
  double* const SA = pool + 3 * kTailNodes;  // stencils (kStencil n) and dinv (n) per level
  {
    int tn = 0;
    for (int l = lt; l < M.nlev; ++l) tn += (int)M.lev[l].n;
    const double* __restrict__ src = M.tailpack;
#pragma unroll 8
    for (int u = threadIdx.x; u < (kStencil + 1) * tn; u += blockDim.x) SA[u] = src[u];
  }
  {
    const MgLevel L = M.lev[lt];
    for (int v = threadIdx.x; v < (int)L.n; v += blockDim.x) R[v] = L.r[v];
  }
  __syncthreads();
  // level l's stencil block starts at (kStencil + 1) * (nodes of levels lt..l-1); its
  // vectors at the same node offset o
  auto lvl = [&](int l, int o) {
    TailLevel L;
    L.nx = M.lev[l].nx;
    L.ny = M.lev[l].ny;
    L.n = (int)M.lev[l].n;
    L.npts = M.lev[l].npts;
    L.a = SA + (kStencil + 1) * o;
    L.dinv = SA + (kStencil + 1) * o + kStencil * L.n;
    return L;
  };
  int o = 0;
  for (int l = lt; l < M.nlev - 1; ++l) {
    const TailLevel L = lvl(l, o);
    const TailLevel Cc = lvl(l + 1, o + L.n);
    double *e = E + o, *t = Tt + o, *r = R + o;
    if (M.vsweeps > 1) {
      for (int v = threadIdx.x; v < L.n; v += blockDim.x) e[v] = om * L.dinv[v] * r[v];
      __syncthreads();
      tail_sweep(L, e, t, r, om);
    } else {
      for (int v = threadIdx.x; v < L.n; v += blockDim.x) t[v] = om * L.dinv[v] * r[v];
    }
    __syncthreads();
    for (int v = threadIdx.x; v < L.n; v += blockDim.x) {
      const int i = v % (L.nx + 1), j = v / (L.nx + 1);
      e[v] = L.dinv[v] == 0.0 ? 0.0 : r[v] - tail_apply(L, t, v, i, j);
    }
    __syncthreads();
    double* rc = R + o + L.n;
    for (int C = threadIdx.x; C < Cc.n; C += blockDim.x) {
      double s = 0.0;
      if (Cc.dinv[C] != 0.0) {
        const int I = C % (Cc.nx + 1), J = C / (Cc.nx + 1);
        const int x = fine_of(I, L.nx), y = fine_of(J, L.ny);
        auto res = [&](int fx, int fy) {
          return (fx < 0 || fx > L.nx || fy < 0 || fy > L.ny)
                     ? 0.0
                     : pweight(L.nx, L.ny, fx, fy, I, J) * e[fy * (L.nx + 1) + fx];
        };
        s = res(x, y) + res(x + 1, y) + res(x - 1, y) + res(x, y + 1) + res(x, y - 1) + res(x + 1, y - 1) +
            res(x - 1, y + 1);
      }
      rc[C] = s;
    }
    __syncthreads();
    o += L.n;
  }
  // coarsest level
  double* ec;
  {
    const TailLevel L = lvl(M.nlev - 1, o);
    double *e = E + o, *t = Tt + o;
    const double* r = R + o;
    if (L.n <= kCoarseDirect) {
      // exact solve with the inverse computed at set-up
      for (int v = threadIdx.x; v < L.n; v += blockDim.x) {
        double acc = 0.0;
        for (int u = 0; u < L.n; ++u) acc = fma(M.cinv[v * L.n + u], r[u], acc);
        t[v] = acc;
      }
      __syncthreads();
    } else {
      for (int v = threadIdx.x; v < L.n; v += blockDim.x) t[v] = om * L.dinv[v] * r[v];
      __syncthreads();
      for (int it = 1; it < M.coarse_sweeps; ++it) {
        tail_sweep(L, t, e, r, om);
        __syncthreads();
        double* sw = e;
        e = t;
        t = sw;
      }
    }
    ec = t;
  }
  for (int l = M.nlev - 2; l >= lt; --l) {
    const int oc = o;
    o -= (int)M.lev[l].n;
    const TailLevel L = lvl(l, o);
    const TailLevel Cc = lvl(l + 1, oc);
    double *e = E + o, *t = Tt + o;
    const double* r = R + o;
    for (int v = threadIdx.x; v < L.n; v += blockDim.x) {
      if (L.dinv[v] == 0.0) continue;
      t[v] += prolong_at(L.nx, L.ny, Cc.nx, ec, v % (L.nx + 1), v / (L.nx + 1));
    }
    __syncthreads();
    if (M.vsweeps > 1) {
      tail_sweep(L, t, e, r, om);
      __syncthreads();
      tail_sweep(L, e, t, r, om);
      __syncthreads();
      ec = t;
    } else {
      tail_sweep(L, t, e, r, om);
      __syncthreads();
      ec = e;
    }
  }
  {
    const MgLevel L = M.lev[lt];
    for (int v = threadIdx.x; v < (int)L.n; v += blockDim.x) L.e[v] = ec[v];
  }
}

int grid1(int64_t n, int t = 256) {
  int64_t g = (n + t - 1) / t;
  if (g < 1) g = 1;
  if (g > (1 << 30)) g = 1 << 30;
  return (int)g;
}

#define MG_NT_DISPATCH(nt, CALL) \
  switch (nt) {                  \
    case 2: CALL(2); break;      \
    case 3: CALL(3); break;      \
    case 4: CALL(4); break;      \
    case 5: CALL(5); break;      \
    default: return (int)hipErrorInvalidValue; \
  }

// V-cycle from level l down; the result is left in lev[l].e.  dM is the
// device copy of M that the tail kernel reads.
int vcycle(const MgData& M, const MgData* dM, int l, const CgState* st, hipStream_t s) {
  if (l >= M.tail) {
    hipLaunchKernelGGL(v_tail_kernel, dim3(1), dim3(kTailThreads), 0, s, dM, st);
    return (int)hipGetLastError();
  }
  const MgLevel& L = M.lev[l];
  const double om = M.omega_v;
  if (l == M.nlev - 1) {
    hipLaunchKernelGGL(v_coarsest_kernel, dim3(1), dim3(1024), 0, s, L, om, M.coarse_sweeps, st);
    return (int)hipGetLastError();
  }
  const MgLevel& Cc = M.lev[l + 1];
  const VOne one_rec{VRec{L, Cc, st}};
  const int g = grid1(L.n);
  const int one = M.vsweeps > 1 ? 0 : 1;
  if (!one) hipLaunchKernelGGL(v_pre_kernel<VOne>, dim3(g), dim3(256), 0, s, one_rec, om);
  hipLaunchKernelGGL(v_restrict_kernel<VOne>, dim3(grid1(8 * Cc.n)), dim3(256), 0, s, one_rec, one, om);
  int e = vcycle(M, dM, l + 1, st, s);
  if (e) return e;
  hipLaunchKernelGGL(v_post_kernel<VOne>, dim3(g), dim3(256), 0, s, one_rec, one, om);
  if (!one) hipLaunchKernelGGL(v_smooth_kernel<VOne>, dim3(g), dim3(256), 0, s, one_rec, om);
  return (int)hipGetLastError();
}

// The same V-cycle for every instance of a group: depth by depth down to the
// tails, every tail in one launch (one workgroup each), depth by depth up.
// An instance whose tail starts at depth t takes part in depths 1..t-1 only.
int vcycle_group(const MgGroup& G, hipStream_t s) {
  const int one = G.vsweeps > 1 ? 0 : 1;
  const double om = G.omega_v;
  for (int d = 1; d < G.ndepth; ++d) {
    if (!one) hipLaunchKernelGGL(v_pre_kernel<VGroup>, dim3(G.gnode[d]), dim3(256), 0, s, G.node[d], om);
    hipLaunchKernelGGL(v_restrict_kernel<VGroup>, dim3(G.gres[d]), dim3(256), 0, s, G.res[d], one, om);
  }
  hipLaunchKernelGGL(v_tail_kernel, dim3(G.ninst), dim3(kTailThreads), 0, s, G.dM, (const CgState*)G.st);
  for (int d = G.ndepth - 1; d >= 1; --d) {
    hipLaunchKernelGGL(v_post_kernel<VGroup>, dim3(G.gnode[d]), dim3(256), 0, s, G.node[d], one, om);
    if (!one) hipLaunchKernelGGL(v_smooth_kernel<VGroup>, dim3(G.gnode[d]), dim3(256), 0, s, G.node[d], om);
  }
  return (int)hipGetLastError();
}

}  // namespace

int mg_setup_cells(const DevMesh& mesh, const MgData& M, const TraceSystem& T, const double* Ate, hipStream_t s) {
  hipLaunchKernelGGL(mg_vmask_kernel, dim3(grid1(M.nv)), dim3(256), 0, s, M, T);
  if (T.nt % 2 == 0) hipLaunchKernelGGL(mg_vf2_kernel, dim3(grid1(M.nv)), dim3(256), 0, s, M);
#define CELL(NT) hipLaunchKernelGGL(mg_cell_coarse_kernel<NT>, dim3(grid1(mesh.nc)), dim3(256), 0, s, mesh, M, T, Ate);
  MG_NT_DISPATCH(T.nt, CELL)
#undef CELL
  return (int)hipGetLastError();
}

int mg_setup_sp(const MgData& M, const TraceSystem& T, const DevPattern& P, float* sp32, hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
#define SPK(NT) hipLaunchKernelGGL(mg_sp_kernel<NT>, dim3(grid1(T.N)), dim3(256), 0, s, P, T, M, sp32);
  MG_NT_DISPATCH(T.nt, SPK)
#undef SPK
  return (int)hipGetLastError();
}

int mg_setup_mesh(const DevMesh& mesh, const MgData& M, const TraceSystem& T, const DevPattern& P,
                  const double* Ate, float* sp32, hipStream_t s) {
  const int e = mg_setup_cells(mesh, M, T, Ate, s);
  if (e) return e;
  return mg_setup_sp(M, T, P, sp32, s);
}

int mg_setup_levels(const MgData& M, const MgData* dM, hipStream_t s) {
  hipLaunchKernelGGL(mg_assemble_l1_kernel, dim3(grid1(M.lev[1].n)), dim3(256), 0, s, dM, (const BlockMap*)nullptr);
  for (int l = 1; l + 1 < M.nlev; ++l) {
    const VOne rec{VRec{M.lev[l], M.lev[l + 1], nullptr}};
    hipLaunchKernelGGL(mg_cmask_kernel<VOne>, dim3(grid1(M.lev[l + 1].n)), dim3(256), 0, s, rec);
    hipLaunchKernelGGL(mg_rap_kernel<VOne>, dim3(grid1(8 * M.lev[l + 1].n)), dim3(256), 0, s, rec);
  }
  hipLaunchKernelGGL(mg_coarse_inverse_kernel, dim3(1), dim3(1024), 0, s, dM);
  if (M.tail < M.nlev) hipLaunchKernelGGL(mg_tailpack_kernel, dim3(16, 1), dim3(256), 0, s, dM);
  return (int)hipGetLastError();
}

int mg_setup_levels_group(const MgSetupGroup& S, hipStream_t s) {
  hipLaunchKernelGGL(mg_assemble_l1_kernel, dim3(S.gl1), dim3(256), 0, s, S.dM, S.l1map);
  for (int d = 1; d < S.ndepth; ++d) {
    hipLaunchKernelGGL(mg_cmask_kernel<VGroup>, dim3(S.gcmask[d]), dim3(256), 0, s, S.cmask[d]);
    hipLaunchKernelGGL(mg_rap_kernel<VGroup>, dim3(S.grap[d]), dim3(256), 0, s, S.rap[d]);
  }
  hipLaunchKernelGGL(mg_coarse_inverse_kernel, dim3(S.ninst), dim3(1024), 0, s, S.dM);
  hipLaunchKernelGGL(mg_tailpack_kernel, dim3(16, S.ninst), dim3(256), 0, s, S.dM);
  return (int)hipGetLastError();
}

int mg_setup_numeric(const DevMesh& mesh, MgData& M, const MgData* dM, TraceSystem& T, const DevPattern& P,
                     const double* Ate, float* sp32, hipStream_t s) {
  const int e = mg_setup_mesh(mesh, M, T, P, Ate, sp32, s);
  if (e) return e;
  return mg_setup_levels(M, dM, s);
}

template <class W, class G>
int mg_init_t(const DevPattern& P, const TraceSystem& T64, const TraceOp<W>& T, const Krylov<double>& v,
              const MgVecs<W>& mv, G g, int grid, hipStream_t s) {
  if (!trace_op_ready(P, T64)) return (int)hipErrorInvalidValue;
#define INIT(NT) \
  hipLaunchKernelGGL((mg_init_kernel<NT, W, G>), dim3(grid), dim3(64 * NT), 0, s, P, T64, T, v, mv, g);
  MG_NT_DISPATCH(T.nt, INIT)
#undef INIT
  return (int)hipGetLastError();
}

template <class W, class G>
int mg_update_t(const TraceOp<W>& T, const Krylov<double>& v, const MgVecs<W>& mv, G g, int grid, hipStream_t s) {
#define UPD(NT) hipLaunchKernelGGL((mg_update_kernel<NT, W, G>), dim3(grid), dim3(64 * NT), 0, s, T, v, mv, g);
  MG_NT_DISPATCH(T.nt, UPD)
#undef UPD
  return (int)hipGetLastError();
}

template <class W>
int launch_mg_init(const DevPattern& P, const TraceSystem& T64, const TraceOp<W>& T, const Krylov<double>& v,
                   const MgVecs<W>& mv, hipStream_t s) {
  return mg_init_t(P, T64, T, v, mv, Whole{T.N, v.st}, v.nblocks, s);
}

template <class W>
int launch_mg_init_group(const DevPattern& P, const TraceSystem& T64, const TraceOp<W>& T, const Krylov<double>& v,
                         const MgVecs<W>& mv, Grouped g, int grid, hipStream_t s) {
  return mg_init_t(P, T64, T, v, mv, g, grid, s);
}

template <class W>
int launch_mg_update(const TraceOp<W>& T, const Krylov<double>& v, const MgVecs<W>& mv, hipStream_t s) {
  return mg_update_t(T, v, mv, Whole{T.N, v.st}, v.nblocks, s);
}

template <class W>
int launch_mg_update_group(const TraceOp<W>& T, const Krylov<double>& v, const MgVecs<W>& mv, Grouped g, int grid,
                           hipStream_t s) {
  return mg_update_t(T, v, mv, g, grid, s);
}

// z = V-cycle(r), starting from e0 = omega Dinv r already in mv.e0; writes
// the (r, z) partials of the next iteration.
template <class W>
int launch_mg_precond(const DevPattern& P, const TraceOp<W>& T, const Krylov<double>& v, const MgData& M,
                      const MgData* dM, const MgVecs<W>& mv, hipStream_t s) {
  const int gr = grid1(T.N, 256);
  const int grows = gr > 16384 ? 16384 : gr;
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
#define RES(NT) hipLaunchKernelGGL((mg_resid0_kernel<NT, W>), dim3(grows), dim3(256), 0, s, P, T, v, mv, M.fc);
  MG_NT_DISPATCH(T.nt, RES)
#undef RES
#define RST(NT) \
  hipLaunchKernelGGL((mg_restrict0_kernel<NT, W>), dim3(grid1(M.lev[1].n)), dim3(256), 0, s, v, M, mv);
  MG_NT_DISPATCH(T.nt, RST)
#undef RST
  int e = vcycle(M, dM, 1, v.st, s);
  if (e) return e;
#define SMO(NT) \
  hipLaunchKernelGGL((mg_smooth0_kernel<NT, W, Whole>), dim3(v.nblocks), dim3(64 * NT), 0, s, T, v, M, mv, \
                     Whole{T.N, v.st});
  MG_NT_DISPATCH(T.nt, SMO)
#undef SMO
  return (int)hipGetLastError();
}

// Grouped: the residual and the restriction to level 1 run over the packed
// mesh (every instance, no convergence check), the vertex levels through
// the group's tables, the smoother per instance.
template <class W>
int launch_mg_precond_group(const DevPattern& P, const TraceOp<W>& T, const Krylov<double>& v, const MgData& Mu,
                            const MgGroup& G, const MgVecs<W>& mv, Grouped g, int grid, hipStream_t s) {
  Krylov<double> vu = v;
  vu.st = nullptr;
  const int gr = grid1(T.N, 256);
  const int grows = gr > 8192 ? 8192 : gr;
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
#define RES(NT) hipLaunchKernelGGL((mg_resid0_kernel<NT, W>), dim3(grows), dim3(256), 0, s, P, T, vu, mv, Mu.fc);
  MG_NT_DISPATCH(T.nt, RES)
#undef RES
#define RST(NT) \
  hipLaunchKernelGGL((mg_restrict0_kernel<NT, W>), dim3(grid1(Mu.lev[1].n)), dim3(256), 0, s, vu, Mu, mv);
  MG_NT_DISPATCH(T.nt, RST)
#undef RST
  int e = vcycle_group(G, s);
  if (e) return e;
#define SMO(NT) \
  hipLaunchKernelGGL((mg_smooth0_kernel<NT, W, Grouped>), dim3(grid), dim3(64 * NT), 0, s, T, v, Mu, mv, g);
  MG_NT_DISPATCH(T.nt, SMO)
#undef SMO
  return (int)hipGetLastError();
}

#define MG_INSTANTIATE(W)                                                                                      \
  template int launch_mg_init<W>(const DevPattern&, const TraceSystem&, const TraceOp<W>&, const Krylov<double>&, \
                                 const MgVecs<W>&, hipStream_t);                                                \
  template int launch_mg_init_group<W>(const DevPattern&, const TraceSystem&, const TraceOp<W>&,                 \
                                       const Krylov<double>&, const MgVecs<W>&, Grouped, int, hipStream_t);      \
  template int launch_mg_update<W>(const TraceOp<W>&, const Krylov<double>&, const MgVecs<W>&, hipStream_t);   \
  template int launch_mg_update_group<W>(const TraceOp<W>&, const Krylov<double>&, const MgVecs<W>&, Grouped,  \
                                         int, hipStream_t);                                                     \
  template int launch_mg_precond<W>(const DevPattern&, const TraceOp<W>&, const Krylov<double>&, const MgData&,  \
                                    const MgData*, const MgVecs<W>&, hipStream_t);                              \
  template int launch_mg_precond_group<W>(const DevPattern&, const TraceOp<W>&, const Krylov<double>&,           \
                                          const MgData&, const MgGroup&, const MgVecs<W>&, Grouped, int,        \
                                          hipStream_t);
MG_INSTANTIATE(double)
MG_INSTANTIATE(float)
#undef MG_INSTANTIATE

}  // namespace hdg
