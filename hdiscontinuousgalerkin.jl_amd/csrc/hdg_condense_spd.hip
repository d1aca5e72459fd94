// Static condensation (examples/poisson2D_HDG.jl:155-174) through the Schur
// complement on the u block, for kappa = 1 and a reference mass that the
// quadrature leaves SPD (k = 1 at degree 2, k = 3 at degree 6, k <= 2 at
// degree 4).
//
// With Me = [A -B; B' C], EFe = [-E; F], Bte = [0; be] and A = detJ
// diag(Mref, Mref) on an affine cell:
//   S   = C + B' A^-1 B                      (nu x nu, SPD)
//   R   = F + B' A^-1 E,  U = S^-1 R,  Q = A^-1 (B U - E)     K_e = [Q; U]
//   Ub  = S^-1 be,  Qb = A^-1 B Ub                            b_e = [Qb; Ub]
//   Ate = E'Q + F'U - H = R' S^-1 R - E' A^-1 E - H
//   bte = -(E'Qb + F'Ub) = -R' S^-1 be
// Every product through A^-1 is a combination of constant reference
// matrices (SpdLayout, hdg_tables.h) with the metric detJ J^-1 J^-T, the
// scaled normals J^-1 n |e| and the edge lengths, so a cell costs a 10 x 10
// Cholesky factorisation (k = 3) and small matrix-vector products instead of
// the elimination of its 30 x 30 local matrix.
//
// Mapping: 16 lanes per element (4 per wavefront); lane t < ntr carries
// trace column t of R, U and K_e, lane ntr the rhs column (be, b_e); every
// lane forms S and factors it, row by row, with the factor kept in LDS.  The
// columns L^-1 R meet in LDS for the products of Ate.
#include <hip/hip_runtime.h>

#include "hdg_device.h"
#include "hdg_internal.h"

namespace hdg {

namespace {

template <int K>
struct SpdDims {
  static constexpr int NU = (K + 1) * (K + 2) / 2, NQ = 2 * NU, M = 3 * NU, NT = K + 1, NTR = 3 * NT;
  static constexpr int NL = NU * (NU + 1) / 2;  // packed lower triangle
  static constexpr int YS = NU * (NTR + 1);     // the L^-1 [R be] columns of one element
  // table offsets (SpdLayout; the quadrature tables Nw, Mg follow at Nw)
  static constexpr int R = 0, C = 3 * NU * NU, Gam = 6 * NU * NU, Z = Gam + 6 * NU * NT, V = Z + 12 * NU * NT,
                       P = V + 6 * NU * NT, Om = P + 2 * NU * NU, H = Om + 36 * NT * NT, Nw = H + NT * NT;
};

constexpr int kSeg = 16;  // lanes per element
constexpr int kEpw = 64 / kSeg;

__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

// Entry ((fl,tl),(gs,us)) of Y'Y - E'A^-1 E - H for the lane's column y:
// ys is column (gs,us) of Y in LDS, cg* the factors |e_f| |e_g| (n_f.n_g) /
// detJ of the three faces g.
template <int K>
__device__ __forceinline__ double ate_entry(const double* T, const double* ys, const double (&y)[SpdDims<K>::NU],
                                            int gs, int us, int fo, int fl, int tl, double el, double cg0,
                                            double cg1, double cg2, const int (&ori)[3]) {
  using D = SpdDims<K>;
  constexpr int NU = D::NU, NT = D::NT;
  double dot = 0.0;
#pragma unroll
  for (int i = 0; i < NU; ++i) dot += ys[i] * y[i];
  const double cgs = gs == 0 ? cg0 : (gs == 1 ? cg1 : cg2);
  const int og = gs == 0 ? ori[0] : (gs == 1 ? ori[1] : ori[2]);
  const double om = T[D::Om + ((fo * 6 + gs * 2 + og) * NT + tl) * NT + us];
  const double h = gs == fl ? el * T[D::H + tl * NT + us] : 0.0;
  return dot - cgs * om - h;
}

// WPE: waves per SIMD the register budget is cut for (2: 256 VGPRs, no
// spills at k <= 2 and 60 B per lane at k = 3; 3: 168 VGPRs, 12-40 B of
// spills per lane); the autotuner picks one, 2 by default.  GROUPED:
// several instances packed into one mesh, the source of cell c is
// srcs[cell_inst[c]].
template <int K, int WPE, bool GROUPED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
void condense_spd_kernel(CondenseArgs a) {
  using D = SpdDims<K>;
  constexpr int NU = D::NU, M = D::M, NT = D::NT, NTR = D::NTR, NL = D::NL;
  static_assert(NTR + 1 <= kSeg, "segment too small");
  extern __shared__ double sh[];
  const SpdLayoutDev L = a.spd;
  for (int i = threadIdx.x; i < L.total; i += blockDim.x) sh[i] = __ldg(a.spd_tables + i);
  __syncthreads();
  const double* T = sh;

  const int lane = threadIdx.x & 63;
  const int seg = lane / kSeg, ls = lane - seg * kSeg, sbase = seg * kSeg;
  // per element: the columns L^-1 [R be] (YS doubles), then L (NL)
  double* Y = sh + L.total + ((threadIdx.x >> 6) * kEpw + seg) * (D::YS + NL);
  double* Lb = Y + D::YS;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t count = a.mesh.nc;
  const int nqp = L.nqp;
  const double tau = a.tau;
  // this lane's trace column: face fl, mode tl
  const int fl = ls < NTR ? ls / NT : 0, tl = ls < NTR ? ls - fl * NT : 0;

  for (int64_t base = wid * kEpw; base < count; base += nw * kEpw) {
    const bool valid = base + seg < count;
    const int64_t cell = valid ? base + seg : count - 1;
    Geo g;
    cell_geometry(a.mesh, cell, g);
    const SrcParams& S = GROUPED ? a.srcs[a.cell_inst[cell]] : a.src;

    // be_i = sum_q w_q N_i(q) detJ f(x_q): lane q evaluates the source at
    // point q, every lane sums over the points
    double rc[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) rc[i] = 0.0;
    for (int q0 = 0; q0 < nqp; q0 += kSeg) {
      const int q = q0 + ls;
      double fv = 0.0;
      if (q < nqp) {
        double xq = 0.0, yq = 0.0;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          xq += T[L.Mg + v * nqp + q] * g.x[2 * v];
          yq += T[L.Mg + v * nqp + q] * g.x[2 * v + 1];
        }
        fv = g.detJ * (S.kind == 3 ? __ldg(S.f_qp + cell + a.mesh.nc * (int64_t)q) : Src::f(S, xq, yq));
      }
      const int qn = min(kSeg, nqp - q0);
      for (int qq = 0; qq < qn; ++qq) {
        const double fq = __shfl(fv, sbase + qq, 64);
#pragma unroll
        for (int i = 0; i < NU; ++i) rc[i] += T[L.Nw + i * nqp + q0 + qq] * fq;
      }
    }

    // the lane's face: length, unit normal, orientation (blended, so that
    // the lane-dependent choice does not index the geometry)
    const double w0 = (double)(fl == 0), w1 = (double)(fl == 1), w2 = (double)(fl == 2);
    const double el = w0 * g.len[0] + w1 * g.len[1] + w2 * g.len[2];
    const double nx = w0 * g.nrm[0] + w1 * g.nrm[2] + w2 * g.nrm[4];
    const double ny = w0 * g.nrm[1] + w1 * g.nrm[3] + w2 * g.nrm[5];
    const int of = (fl == 0) * g.ori[0] + (fl == 1) * g.ori[1] + (fl == 2) * g.ori[2];
    const int fo = fl * 2 + of;
    // trace column: R = tau |e_f| Gam_fo + s0 Z_0fo + s1 Z_1fo with
    // s_d = |e_f| (J^-1 n_f)_d
    if (ls < NTR) {
      const double s0 = el * (g.Ji[0] * nx + g.Ji[2] * ny), s1 = el * (g.Ji[1] * nx + g.Ji[3] * ny);
      const double tf = tau * el;
#pragma unroll
      for (int i = 0; i < NU; ++i) {
        rc[i] = tf * T[D::Gam + (fo * NU + i) * NT + tl] + s0 * T[D::Z + (fo * NU + i) * NT + tl] +
                s1 * T[D::Z + ((6 + fo) * NU + i) * NT + tl];
        __builtin_amdgcn_sched_barrier(0);
      }
    } else if (ls > NTR) {
#pragma unroll
      for (int i = 0; i < NU; ++i) rc[i] = 0.0;
    }

    // S = tau sum_f |e_f| Cref_f + detJ (G00 R00 + G01 R01 + G11 R11),
    // G = J^-1 J^-T: lane i forms row i in the element's LDS buffer (in the
    // space of the Y columns, which are written after the factorisation)
    {
      const double c0 = tau * g.len[0], c1 = tau * g.len[1], c2 = tau * g.len[2];
      const double r0 = g.detJ * (g.Ji[0] * g.Ji[0] + g.Ji[2] * g.Ji[2]);
      const double r1 = g.detJ * (g.Ji[0] * g.Ji[1] + g.Ji[2] * g.Ji[3]);
      const double r2 = g.detJ * (g.Ji[1] * g.Ji[1] + g.Ji[3] * g.Ji[3]);
      if (ls < NU) {
#pragma unroll
        for (int j = 0; j < NU; ++j) {
          const int x = ls * NU + j;
          Y[x] = c0 * T[D::C + x] + c1 * T[D::C + NU * NU + x] + c2 * T[D::C + 2 * NU * NU + x] +
                 r0 * T[D::R + x] + r1 * T[D::R + NU * NU + x] + r2 * T[D::R + 2 * NU * NU + x];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // S = L L', row by row (left-looking): every lane of the element
    // computes the same rows, lane 0 keeps them in LDS (row i packed,
    // 1 / L_ii in place of L_ii); the forward solve rc <- L^-1 rc follows
    // each row, in place
    double dmax = 0.0;
#pragma unroll
    for (int i = 0; i < NU; ++i) dmax = fmax(dmax, Y[i * NU + i]);
    bool bad = !(g.detJ > 0.0);
#pragma unroll
    for (int i = 0; i < NU; ++i) {
      double li[NU];
      double d = Y[i * NU + i];
#pragma unroll
      for (int j = 0; j < i; ++j) {
        double v = Y[i * NU + j];
#pragma unroll
        for (int t = 0; t < j; ++t) v -= li[t] * Lb[tri(j, t)];
        li[j] = v * Lb[tri(j, j)];
        d -= li[j] * li[j];
        __builtin_amdgcn_sched_barrier(0);
      }
      bad |= !(d > 1e-10 * dmax);
      const double il = 1.0 / (d > 0.0 ? sqrt(d) : 1.0);
      double v = rc[i];
#pragma unroll
      for (int t = 0; t < i; ++t) v -= li[t] * rc[t];
      rc[i] = v * il;
      if (ls == 0) {
#pragma unroll
        for (int t = 0; t < i; ++t) Lb[tri(i, t)] = li[t];
        Lb[tri(i, i)] = il;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // Ate = Y'Y - E'A^-1 E - H and bte = -Y' y_b, Y = L^-1 R: the columns
    // meet in the element's LDS buffer (a wavefront's LDS instructions run
    // in order, so the reads below see the writes)
    if (ls <= NTR) {
#pragma unroll
      for (int i = 0; i < NU; ++i) Y[ls * NU + i] = rc[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {
      // E'A^-1 E entry ((f,t),(g,s)) = |e_f| |e_g| (n_f . n_g) / detJ Om_fo,gp(t,s)
      const double sc = el / g.detJ;
      const double cg0 = sc * g.len[0] * (nx * g.nrm[0] + ny * g.nrm[1]);
      const double cg1 = sc * g.len[1] * (nx * g.nrm[2] + ny * g.nrm[3]);
      const double cg2 = sc * g.len[2] * (nx * g.nrm[4] + ny * g.nrm[5]);
      // Constant traces: for c = 1 on mode 0 of every face the local
      // problem is solved by u = const, q = 0, so in exact arithmetic
      // Ate c = F'u - H c = (tau - 1) H c (H carries no tau,
      // poisson2D_HDG.jl:149).  The rounding of the shared tables breaks
      // this the same way in every cell, an error the trace solve amplifies
      // by cond(K) ~ h^-2 (1024^2 cells, P3: an L2 error of 6e-10 instead of
      // 5e-13).  Ate is projected onto it, with d = Ate c - (tau - 1) H c:
      //   Ate - (d c' + c d') / 3 + (c'd) c c' / 9,
      // a symmetric change of the order of the rounding (~1e-15 relative).
      double r_own = -(tau - 1.0) * el * T[D::H + tl * NT];
#pragma unroll
      for (int gs = 0; gs < 3; ++gs)
        r_own += ate_entry<K>(T, Y + gs * NT * NU, rc, gs, 0, fo, fl, tl, el, cg0, cg1, cg2, g.ori);
      const double r0 = __shfl(r_own, sbase + 0 * NT, 64), r1 = __shfl(r_own, sbase + 1 * NT, 64);
      const double r2 = __shfl(r_own, sbase + 2 * NT, 64);
      const double rsum = r0 + r1 + r2;
      const double ct = tl == 0 ? 1.0 : 0.0;
      if (valid && ls < NTR) {
        double* at = a.Ate + cell * (int64_t)(NTR * NTR) + ls * NTR;
#pragma unroll 1
        for (int s = 0; s < NTR; ++s) {
          const int gs = s / NT, us = s - gs * NT;
          const double cs = us == 0 ? 1.0 : 0.0;
          const double rs = gs == 0 ? r0 : (gs == 1 ? r1 : r2);
          at[s] = ate_entry<K>(T, Y + s * NU, rc, gs, us, fo, fl, tl, el, cg0, cg1, cg2, g.ori) -
                  (r_own * cs + ct * rs) * (1.0 / 3.0) + rsum * ct * cs * (1.0 / 9.0);
        }
        double bt = 0.0;
#pragma unroll
        for (int i = 0; i < NU; ++i) bt += Y[NTR * NU + i] * rc[i];
        a.bte[cell * NTR + ls] = -bt;
      }
    }
    // rc <- L^-T rc (U, Ub), in place
#pragma unroll
    for (int i = NU - 1; i >= 0; --i) {
      double v = rc[i];
#pragma unroll
      for (int t = i + 1; t < NU; ++t) v -= Lb[tri(t, i)] * rc[t];
      rc[i] = v * Lb[tri(i, i)];
      __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_wave_barrier();

    // the u rows of K_e (U, trace columns) and of b_e (Ub, rhs column),
    // compact at the start of the cell's K_e and b_e blocks (U column-major
    // nu x ntr); the q rows Q = A^-1 (B U - E) follow from them and the
    // cell's geometry, and are formed where they are needed
    // (recover_spd_kernel for the recovered q, expand_ke_spd_kernel for
    // hdg_get_local_ops, which also restores the layout of K_e): applied to
    // one trace vector at recovery they cost an eighth of forming them here
    // for every column, and their stores were a fifth of this kernel's time
    if (valid && ls <= NTR) {
      double* out = ls < NTR ? a.Ke + cell * (int64_t)(M * NTR) + ls * NU : a.be + cell * (int64_t)M;
#pragma unroll
      for (int i = 0; i < NU; ++i) out[i] = rc[i];
    }
    if (valid && bad && ls == 0) atomicOr(a.err + (GROUPED ? a.cell_inst[cell] : 0), g.detJ > 0.0 ? 2 : 1);
  }
}

// K_e and b_e in their exported layout (column t of K_e: q rows then u rows)
// from the compact u rows the Schur kernel stores (U column-major at the
// start of the cell's block, Ub at the start of b_e's), with the q rows
//   Q_t = A^-1 (B U_t - E_t) = sum_d (J^-1)_dc P_d U_t - |e_f| n_fc / detJ V_fo(:, s)
// for trace column t = (face f, mode s), without the E term for the rhs.
// One thread per cell, in place: column t (written at t m) reads compact
// column t (at t nu) before writing, and the columns run from the last, so
// every compact column a write covers has been read already.  Tables
// staged in LDS.
template <int K>
__global__ __launch_bounds__(256) void expand_ke_spd_kernel(DevMesh m, const double* __restrict__ tab,
                                                            double* __restrict__ Ke, double* __restrict__ be) {
  using D = SpdDims<K>;
  constexpr int NU = D::NU, NQ = D::NQ, M = D::M, NT = D::NT, NTR = D::NTR;
  __shared__ double T[D::Om - D::V];  // V then P
  for (int i = threadIdx.x; i < D::Om - D::V; i += blockDim.x) T[i] = __ldg(tab + D::V + i);
  __syncthreads();
  const double* V = T;
  const double* Pm = T + (D::P - D::V);
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= m.nc) return;
  Geo g;
  cell_geometry(m, cell, g);
  double* ke = Ke + cell * (int64_t)(M * NTR);
#pragma unroll 1
  for (int col = NTR; col >= 0; --col) {
    // col == NTR: the rhs column (b_e)
    const double* in = col < NTR ? ke + col * NU : be + cell * (int64_t)M;
    double* out = col < NTR ? ke + col * M : be + cell * (int64_t)M;
    double u[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) u[i] = in[i];
    const int fl = col < NTR ? col / NT : 0, tl = col < NTR ? col - fl * NT : 0;
    const double el = g.len[fl];
    const double nx = g.nrm[2 * fl], ny = g.nrm[2 * fl + 1];
    const int fo = fl * 2 + g.ori[fl];
    const double vs = col < NTR ? el / g.detJ : 0.0;
#pragma unroll 1
    for (int i = 0; i < NU; ++i) {
      double t0 = 0.0, t1 = 0.0;
#pragma unroll
      for (int j = 0; j < NU; ++j) {
        t0 += Pm[i * NU + j] * u[j];
        t1 += Pm[(NU + i) * NU + j] * u[j];
      }
      const double v = vs * V[(fo * NU + i) * NT + tl];
      out[i] = g.Ji[0] * t0 + g.Ji[1] * t1 - nx * v;
      out[NU + i] = g.Ji[2] * t0 + g.Ji[3] * t1 - ny * v;
    }
#pragma unroll
    for (int i = 0; i < NU; ++i) out[NQ + i] = u[i];
  }
}

// Recovery (examples/poisson2D_HDG.jl:197-212) from the u rows of K_e:
// u = U uhat_c + Ub, q = sum_d (J^-1)_dc P_d u - sum_f |e_f| n_fc / detJ
// V_fo uhat_f, i.e. K_e uhat_c + b_e with the q rows applied to the one
// trace vector.  16 lanes per cell (4 per wavefront): lane i < nu forms
// u_i (reading row i of U: the lanes of a cell read runs of nu doubles),
// then q_i from u, which the lanes of the cell exchange by shuffles.
// Tables staged in LDS.
template <int K>
__global__ __launch_bounds__(256) void recover_spd_kernel(DevMesh m, const double* __restrict__ tab,
                                                          const double* __restrict__ Ke,
                                                          const double* __restrict__ be,
                                                          const double* __restrict__ uhat, double* __restrict__ qu) {
  using D = SpdDims<K>;
  constexpr int NU = D::NU, NQ = D::NQ, M = D::M, NT = D::NT, NTR = D::NTR;
  static_assert(NU <= kSeg, "segment too small");
  __shared__ double T[D::Om - D::V];  // V then P
  for (int i = threadIdx.x; i < D::Om - D::V; i += blockDim.x) T[i] = __ldg(tab + D::V + i);
  __syncthreads();
  const double* V = T;
  const double* Pm = T + (D::P - D::V);
  const int lane = threadIdx.x & 63;
  const int sbase = (lane / kSeg) * kSeg, ls = lane - sbase;
  // every lane of a segment has the same cell: segments leave together
  const int64_t cell = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kSeg;
  if (cell >= m.nc) return;
  Geo g;
  cell_geometry(m, cell, g);
  double uh[NTR];
#pragma unroll
  for (int lf = 0; lf < 3; ++lf) {
    const int64_t f = m.cf[3 * cell + lf];
#pragma unroll
    for (int b = 0; b < NT; ++b) uh[lf * NT + b] = uhat[f * NT + b];
  }
  const int i = ls < NU ? ls : NU - 1;  // lanes past nu repeat the last row and store nothing
  // the compact u rows: U column-major (nu x ntr) at the start of the cell's
  // K_e block, Ub at the start of its b_e block
  const double* ke = Ke + cell * (int64_t)(M * NTR) + i;
  double u = be[cell * M + i];
#pragma unroll
  for (int t = 0; t < NTR; ++t) u = fma(ke[t * NU], uh[t], u);
  double wx = 0.0, wy = 0.0;
#pragma unroll
  for (int lf = 0; lf < 3; ++lf) {
    const double sc = g.len[lf] / g.detJ;
    const double sx = sc * g.nrm[2 * lf], sy = sc * g.nrm[2 * lf + 1];
    const int fo = lf * 2 + g.ori[lf];
    double v = 0.0;
#pragma unroll
    for (int b = 0; b < NT; ++b) v = fma(V[(fo * NU + i) * NT + b], uh[lf * NT + b], v);
    wx = fma(sx, v, wx);
    wy = fma(sy, v, wy);
  }
  double t0 = 0.0, t1 = 0.0;
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    const double uj = __shfl(u, sbase + j, 64);
    t0 = fma(Pm[i * NU + j], uj, t0);
    t1 = fma(Pm[(NU + i) * NU + j], uj, t1);
  }
  if (ls < NU) {
    double* o = qu + cell * (int64_t)M;
    o[i] = g.Ji[0] * t0 + g.Ji[1] * t1 - wx;
    o[NU + i] = g.Ji[2] * t0 + g.Ji[3] * t1 - wy;
    o[NQ + i] = u;
  }
}

template <int K>
int spd_launch(const CondenseArgs& a, int wpe, int grid_cap, hipStream_t s) {
  using D = SpdDims<K>;
  // the kernel addresses the constant tables at compile-time offsets
  const SpdLayoutDev& L = a.spd;
  if (L.nu != D::NU || L.nt != D::NT || L.R != D::R || L.C != D::C || L.Gam != D::Gam || L.Z != D::Z ||
      L.V != D::V || L.P != D::P || L.Om != D::Om || L.H != D::H || L.Nw != D::Nw ||
      L.Mg != D::Nw + D::NU * L.nqp || L.total != L.Mg + 3 * L.nqp)
    return (int)hipErrorInvalidValue;
  const int64_t n = a.mesh.nc;
  const int64_t per_block = 4 * kEpw;
  int64_t grid = (n + per_block - 1) / per_block;
  const int64_t cap = grid_cap > 0 ? grid_cap : 16384;
  if (grid < 1) grid = 1;
  if (grid > cap) grid = cap;
  const size_t lds = sizeof(double) * (a.spd.total + 4 * kEpw * (SpdDims<K>::YS + SpdDims<K>::NL));
  if (wpe != 3) wpe = 2;
  if (a.srcs) {
    if (wpe == 2)
      hipLaunchKernelGGL((condense_spd_kernel<K, 2, true>), dim3((int)grid), dim3(256), lds, s, a);
    else
      hipLaunchKernelGGL((condense_spd_kernel<K, 3, true>), dim3((int)grid), dim3(256), lds, s, a);
  } else {
    if (wpe == 2)
      hipLaunchKernelGGL((condense_spd_kernel<K, 2, false>), dim3((int)grid), dim3(256), lds, s, a);
    else
      hipLaunchKernelGGL((condense_spd_kernel<K, 3, false>), dim3((int)grid), dim3(256), lds, s, a);
  }
  return (int)hipGetLastError();
}

}  // namespace

int launch_condense_spd(int k, const CondenseArgs& a, const CondenseSched& sc, hipStream_t s) {
  if (!a.spd_tables) return (int)hipErrorInvalidValue;
  switch (k) {
    case 1: return spd_launch<1>(a, sc.spd_wpe, sc.spd_grid, s);
    case 2: return spd_launch<2>(a, sc.spd_wpe, sc.spd_grid, s);
    case 3: return spd_launch<3>(a, sc.spd_wpe, sc.spd_grid, s);
    default: return (int)hipErrorInvalidValue;
  }
}

int launch_recover_spd(int k, const DevMesh& m, const double* tab, const double* Ke, const double* be,
                       const double* uhat, double* qu, hipStream_t s) {
  const int64_t g = (m.nc * kSeg + 255) / 256;
  const dim3 grid((unsigned)(g > 0 ? g : 1));
  switch (k) {
    case 1: hipLaunchKernelGGL(recover_spd_kernel<1>, grid, dim3(256), 0, s, m, tab, Ke, be, uhat, qu); break;
    case 2: hipLaunchKernelGGL(recover_spd_kernel<2>, grid, dim3(256), 0, s, m, tab, Ke, be, uhat, qu); break;
    case 3: hipLaunchKernelGGL(recover_spd_kernel<3>, grid, dim3(256), 0, s, m, tab, Ke, be, uhat, qu); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

int launch_expand_ke_spd(int k, const DevMesh& m, const double* tab, double* Ke, double* be, hipStream_t s) {
  const int64_t g = (m.nc + 255) / 256;
  const dim3 grid((unsigned)(g > 0 ? g : 1));
  switch (k) {
    case 1: hipLaunchKernelGGL(expand_ke_spd_kernel<1>, grid, dim3(256), 0, s, m, tab, Ke, be); break;
    case 2: hipLaunchKernelGGL(expand_ke_spd_kernel<2>, grid, dim3(256), 0, s, m, tab, Ke, be); break;
    case 3: hipLaunchKernelGGL(expand_ke_spd_kernel<3>, grid, dim3(256), 0, s, m, tab, Ke, be); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

int tune_condense_spd(int k, const CondenseArgs& a, hipStream_t s, CondenseSched* best, double* best_us,
                      double* default_us) {
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return (int)e;
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return (int)e;
  }
  const int grids[] = {2048, 4096, 8192, 16384};
  constexpr int kReps = 3;
  double bt = 1e300, dt = 0.0;
  CondenseSched bs = *best;
  int rc = 0;
  for (int wpe = 3; wpe >= 2 && rc == 0; --wpe)
    for (int g : grids) {
      CondenseSched sc = *best;
      sc.spd_wpe = wpe;
      sc.spd_grid = g;
      rc = launch_condense_spd(k, a, sc, s);  // warm-up
      if (rc) break;
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < kReps && rc == 0; ++r) rc = launch_condense_spd(k, a, sc, s);
      (void)hipEventRecord(e1, s);
      if (rc || (rc = (int)hipEventSynchronize(e1))) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double us = 1e3 * ms / kReps;
      if (us < bt) {
        bt = us;
        bs = sc;
      }
      if (wpe == 2 && g == 16384) dt = us;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *best = bs;
  if (best_us) *best_us = bt;
  if (default_us) *default_us = dt;
  return 0;
}

}  // namespace hdg
