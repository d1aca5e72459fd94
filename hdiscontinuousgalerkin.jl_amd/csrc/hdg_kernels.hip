// Device kernels of the HDG Poisson pipeline for gfx950 (CDNA4).
//
//   pattern    face-level sparsity of the trace matrix (once per mesh)
//   condense   element build + static condensation: one segment of lanes
//              per element, one lane per row of the bordered local system
//              [[Me, EFe, Bte], [Ge, H, 0]]; Gauss-Jordan with partial
//              pivoting in registers leaves K_e = Me^-1 EFe, b_e = Me^-1 Bte
//              in the pivot lanes and -Ate, bte in the border lanes
//              (examples/poisson2D_HDG.jl:77-174)
//   fill       gather of element blocks into the trace matrix (no atomics:
//              each entry has at most two contributions, summed in order)
//   jacobi     inverse of each face's diagonal block (block-Jacobi)
//   cg_*       preconditioned CG on S = -K_II, two kernels per iteration,
//              deterministic reductions without atomics or fences
//   recover    [q;u]_e = K_e uhat_e + b_e (poisson2D_HDG.jl:197-212)
//   errornorm  squared L2 error (src/DiscreteFunctions.jl:97-120)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hdg_device.h"
#include "hdg_internal.h"

namespace hdg {

namespace {

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum in a fixed order; result valid in thread 0.
template <int NW>
__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NW; ++i) s += sh[i];
  __syncthreads();
  return s;
}

template <int K>
struct Dims {
  static constexpr int NU = (K + 1) * (K + 2) / 2;
  static constexpr int NQ = 2 * NU, M = 3 * NU, NT = K + 1, NTR = 3 * NT;
  static constexpr int ROWS = M + NTR, NC = M + NTR + 1;
  static constexpr int SEG = ROWS <= 16 ? 16 : (ROWS <= 32 ? 32 : 64);
  static constexpr int EPW = 64 / SEG;  // elements per wavefront
};

// Max of an int over each segment of SEG lanes, in every lane: DPP stages
// inside each row of 16 lanes (pairs by xor 1 and 2 within quads, then
// mirrored halves and mirrored rows: after each stage every lane holds the
// max of a group twice as large), then one readlane per row combines the
// rows of a 32- or 64-lane segment.  Every lane of the wavefront must be
// active.
template <int SEG>
__device__ inline int seg_max_int(int k, int seg) {
  k = max(k, __builtin_amdgcn_update_dpp(k, k, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  k = max(k, __builtin_amdgcn_update_dpp(k, k, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  k = max(k, __builtin_amdgcn_update_dpp(k, k, 0x141, 0xF, 0xF, false));  // row_half_mirror
  k = max(k, __builtin_amdgcn_update_dpp(k, k, 0x140, 0xF, 0xF, false));  // row_mirror
  if constexpr (SEG == 16) {
    return k;
  } else {
    const int r0 = __builtin_amdgcn_readlane(k, 0), r1 = __builtin_amdgcn_readlane(k, 16);
    const int r2 = __builtin_amdgcn_readlane(k, 32), r3 = __builtin_amdgcn_readlane(k, 48);
    if constexpr (SEG == 32) return seg ? max(r2, r3) : max(r0, r1);
    else return max(max(r0, r1), max(r2, r3));
  }
}

// Broadcast of lane `src` (absolute lane in the wavefront) within a segment.
template <int SEG>
__device__ inline double seg_bcast(double v, int src) {
  if constexpr (SEG == 64) {
    const int s = __builtin_amdgcn_readfirstlane(src);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), s);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), s);
    return __hiloint2double(hi, lo);
  } else {
    return __shfl(v, src, 64);
  }
}

// Broadcast for a 32-lane segment by readlanes: the value of lane s0 to
// the lower half-wave and of lane s1 to the upper one (s0, s1 wave-uniform),
// four readlanes and a select per double, VALU work instead of the two LDS
// permutes of seg_bcast.  Every lane of the wavefront must be active.
__device__ inline double halves_bcast(double v, int s0, int s1, int seg) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int l0 = __builtin_amdgcn_readlane(lo, s0), l1 = __builtin_amdgcn_readlane(lo, s1);
  const int h0 = __builtin_amdgcn_readlane(hi, s0), h1 = __builtin_amdgcn_readlane(hi, s1);
  return __hiloint2double(seg ? h1 : h0, seg ? l1 : l0);
}

// ------------------------------------------------------------- condense
// v0 or v1 (i = 0, 1) and v0, v1 or v2 (i = 0, 1, 2).  BLEND forms them as
// sums of products with 0/1 weights, so that the choice stays a choice
// between values (exact for finite values up to the sign of a zero); the
// plain ternaries may become a lane-dependent address into the cell's
// geometry, which then lives in scratch.  At k = 4 the blend costs the
// registers that keep two waves per SIMD (condensation of 512^2 cells:
// 18.0 ms blended vs 11.7 ms with ternaries and scratch), so k = 4 keeps
// the ternaries.
template <bool BLEND>
__device__ __forceinline__ double pick2(int i, double v0, double v1) {
  if constexpr (BLEND) return (double)(i == 0) * v0 + (double)(i == 1) * v1;
  else return i ? v1 : v0;
}
template <bool BLEND>
__device__ __forceinline__ double pick3(int i, double v0, double v1, double v2) {
  if constexpr (BLEND) return (double)(i == 0) * v0 + (double)(i == 1) * v1 + (double)(i == 2) * v2;
  else return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

// One row of the bordered local system, for lane row r of the segment.
//   r <  NQ : [ A(r,:) | -B(r,:) | -E(r,:) | 0      ]   q test functions
//   r <  M  : [ B(:,i)'|  C(i,:) |  F(i,:) | be(i)  ]   u test functions
//   r < ROWS: [ E(:,a)'|  F(:,a)'|  H(a,:) | 0      ]   trace test functions
// with A = (kappa^-1 v_j, v_i), B = (u_j, div v_i), C = tau<u_j, w_i>,
// E = <uhat_j, v_i.n>, F = tau<uhat_j, w_i>, H = <uhat_j, mu_i>
// (examples/poisson2D_HDG.jl:88-153; the vector basis is block-by-component,
// VectorFunctionSpaces.jl:28-57).
template <int K>
__device__ __forceinline__ void build_row(int r, const double* T, const TabLayoutDev& L, const Geo& g,
                                 double tau, const double* Aacc, double beacc,
                                 double (&row)[Dims<K>::NC]) {
  using D = Dims<K>;
  constexpr int NU = D::NU, NQ = D::NQ, M = D::M, NT = D::NT, NTR = D::NTR;
  // lane-dependent choices among the fields of g go through pick2 / pick3
  constexpr bool BL = K <= 3;
  if (r < NQ) {
    const int c = r / NU, i = r - c * NU;
#pragma unroll
    for (int s = 0; s < NQ; ++s) row[s] = (s / NU == c) ? Aacc[s % NU] : 0.0;
    const double j0 = pick2<BL>(c, g.Ji[0], g.Ji[2]), j1 = pick2<BL>(c, g.Ji[1], g.Ji[3]);
#pragma unroll
    for (int j = 0; j < NU; ++j)
      row[NQ + j] = -g.detJ * (j0 * T[L.Dref + i * NU + j] + j1 * T[L.Dref + NU * NU + i * NU + j]);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const double sc = -g.len[f] * pick2<BL>(c, g.nrm[2 * f], g.nrm[2 * f + 1]);
      const double* Gf = T + L.G + ((f * 2 + g.ori[f]) * NU + i) * NT;
#pragma unroll
      for (int t = 0; t < NT; ++t) row[M + f * NT + t] = sc * Gf[t];
    }
    row[M + NTR] = 0.0;
  } else if (r < M) {
    const int i = r - NQ;
#pragma unroll
    for (int s = 0; s < NQ; ++s) {
      const int cs = s / NU, is = s % NU;
      row[s] = g.detJ * (g.Ji[0 + 2 * cs] * T[L.Dref + is * NU + i] +
                         g.Ji[1 + 2 * cs] * T[L.Dref + NU * NU + is * NU + i]);
    }
#pragma unroll
    for (int j = 0; j < NU; ++j)
      row[NQ + j] = tau * (g.len[0] * T[L.Cref + (0 * NU + i) * NU + j] +
                           g.len[1] * T[L.Cref + (1 * NU + i) * NU + j] +
                           g.len[2] * T[L.Cref + (2 * NU + i) * NU + j]);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const double sc = tau * g.len[f];
      const double* Gf = T + L.G + ((f * 2 + g.ori[f]) * NU + i) * NT;
#pragma unroll
      for (int t = 0; t < NT; ++t) row[M + f * NT + t] = sc * Gf[t];
    }
    row[M + NTR] = beacc;
  } else if (r < D::ROWS) {
    const int a = r - M, fa = a / NT, ta = a - fa * NT;
    const int of = BL ? (fa == 0) * g.ori[0] + (fa == 1) * g.ori[1] + (fa == 2) * g.ori[2]
                      : (fa == 0 ? g.ori[0] : (fa == 1 ? g.ori[1] : g.ori[2]));
    const double* Gf = T + L.G + (fa * 2 + of) * NU * NT;
    const double el = pick3<BL>(fa, g.len[0], g.len[1], g.len[2]);
    const double nx = el * pick3<BL>(fa, g.nrm[0], g.nrm[2], g.nrm[4]);
    const double ny = el * pick3<BL>(fa, g.nrm[1], g.nrm[3], g.nrm[5]);
#pragma unroll
    for (int s = 0; s < NQ; ++s) row[s] = (s < NU ? nx : ny) * Gf[(s % NU) * NT + ta];
#pragma unroll
    for (int j = 0; j < NU; ++j) row[NQ + j] = tau * el * Gf[j * NT + ta];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        row[M + f * NT + t] = (f == fa) ? el * T[L.Href + ta * NT + t] : 0.0;
    row[M + NTR] = 0.0;
  } else {
#pragma unroll
    for (int c = 0; c < D::NC; ++c) row[c] = 0.0;
  }
}

template <bool GROUPED>
__device__ __forceinline__ const SrcParams& src_of(const CondenseArgs& a, int64_t cell) {
  if constexpr (GROUPED) return a.srcs[a.cell_inst[cell]];
  else return a.src;
}

// SEG: lanes per element (a power of two >= the rows of the bordered
// system); EPW = 64 / SEG elements per wavefront.  The result does not
// depend on SEG (the pivot key below is the same for every SEG).
// Occupancy: left alone the compiler keeps two waves per SIMD (175-230
// VGPRs); the kernel issues VALU work back to back and a third wave hides
// more of its latency than the few spilled registers cost at k <= 3
// (condensation of 512^2 cells: k=1 0.53 -> 0.46 ms, k=2 2.04 -> 1.88 ms,
// k=3 6.07 -> 5.31 ms).  At k = 4 the row holds 56 doubles and the spills
// of a third wave cost more than it hides (10.2 -> 11.6 ms).
// NAT: the tables' mass is SPD (CondenseArgs::natural), so A is SPD on every
// cell and Gauss-Jordan runs in natural order: row j is the pivot of column
// j, no pivot search, and a q pivot row (all zero in the other vector
// component's q columns, which stay zero: A is block diagonal and B^T rows
// are never pivots before the q columns are done) skips those columns.
template <int K, bool BLOCKS, bool GROUPED, int SEG, bool NAT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(!BLOCKS && K <= 3 ? 3 : 1)))
void condense_kernel(CondenseArgs a) {
  using D = Dims<K>;
  static_assert(SEG >= D::ROWS && 64 % SEG == 0, "segment too small");
  constexpr int NU = D::NU, NQ = D::NQ, M = D::M, NTR = D::NTR, EPW = 64 / SEG;
  // columns from CS on reach the other lanes through LDS (see below), and
  // the stride of a segment's pivot-row buffer (16-byte aligned, and an odd
  // number of 16-byte units so the segments of a wavefront start in
  // different banks).  CS = M puts the whole border (trace block and rhs)
  // there.  Natural order on 64-lane segments has no pivot search, and LDS
  // issue becomes the limit, so 8 border columns go back to readlanes
  // (k = 3 on 512^2 cells: 4.52 -> 3.88 ms; the pivoting kernels and
  // 16/32-lane segments are fastest with the whole border in LDS).
  constexpr int CS = (NAT && SEG == 64 && M + 8 < D::NC) ? M + 8 : M;
  constexpr int NB = D::NC - CS, PBS = ((NB + 1) / 2 | 1) * 2;
  extern __shared__ double T[];
  __shared__ __align__(16) double prow[4 * EPW * PBS];
  const TabLayoutDev L = a.lay;
  for (int i = threadIdx.x; i < L.total; i += blockDim.x) T[i] = __ldg(a.tables + i);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int seg = lane / SEG, ls = lane - seg * SEG, sbase = seg * SEG;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t first = BLOCKS ? a.first : 0;
  const int64_t count = BLOCKS ? a.count : a.mesh.nc;
  const int nqp = L.nqp;

  for (int64_t base = wid * EPW; base < count; base += nw * EPW) {
    const bool valid = base + seg < count;
    const int64_t cell = first + (valid ? base + seg : count - 1);
    Geo g;
    cell_geometry(a.mesh, cell, g);
    const SrcParams& S = src_of<GROUPED>(a, cell);
    // quadrature-point data shared by the segment: lane q evaluates
    // detJ * f(x_q) and detJ / kappa(x_q); rows collect them by shuffles
    const int ui = (ls < NQ) ? ls % NU : (ls < M ? ls - NQ : 0);  // scalar basis index of this row
    double Aacc[NU];
#pragma unroll
    for (int j = 0; j < NU; ++j) Aacc[j] = g.detJ * T[L.Mref + ui * NU + j];
    double beacc = 0.0;
    for (int q0 = 0; q0 < nqp; q0 += SEG) {
      const int q = q0 + ls;
      double fv = 0.0, kv = 0.0;
      if (q < nqp) {
        double xq = 0.0, yq = 0.0;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          xq += T[L.Mg + v * nqp + q] * g.x[2 * v];
          yq += T[L.Mg + v * nqp + q] * g.x[2 * v + 1];
        }
        fv = g.detJ * (S.kind == 3 ? __ldg(S.f_qp + cell + a.mesh.nc * (int64_t)q) : Src::f(S, xq, yq));
        kv = g.detJ / Src::kappa(S, xq, yq);
      }
      const int qn = min(SEG, nqp - q0);
      for (int qq = 0; qq < qn; ++qq) {
        const double fq = __shfl(fv, sbase + qq, 64);
        beacc += T[L.Nw + ui * nqp + q0 + qq] * fq;
        if (S.kappa_on) {
          const double kq = __shfl(kv, sbase + qq, 64);
          const double wn = T[L.Nw + ui * nqp + q0 + qq] * kq;
          if (qq == 0 && q0 == 0) {
#pragma unroll
            for (int j = 0; j < NU; ++j) Aacc[j] = 0.0;
          }
#pragma unroll
          for (int j = 0; j < NU; ++j) Aacc[j] += wn * T[L.N + j * nqp + q0 + qq];
        }
      }
    }
    double row[D::NC];
    build_row<K>(ls, T, L, g, a.tau, Aacc, beacc, row);

    if constexpr (BLOCKS) {
      if (valid) {
        const int64_t e = cell - first;
        if (ls < NQ) {
          for (int s = 0; s < NQ; ++s) if (a.A) a.A[e * NQ * NQ + ls + NQ * s] = row[s];
          for (int j = 0; j < NU; ++j) if (a.B) a.B[e * NQ * NU + ls + NQ * j] = -row[NQ + j];
          for (int t = 0; t < NTR; ++t) if (a.E) a.E[e * NQ * NTR + ls + NQ * t] = -row[M + t];
        } else if (ls < M) {
          const int i = ls - NQ;
          for (int j = 0; j < NU; ++j) if (a.C) a.C[e * NU * NU + i + NU * j] = row[NQ + j];
          for (int t = 0; t < NTR; ++t) if (a.F) a.F[e * NU * NTR + i + NU * t] = row[M + t];
          if (a.bev) a.bev[e * NU + i] = row[M + NTR];
        } else if (ls < D::ROWS) {
          const int r = ls - M;
          for (int t = 0; t < NTR; ++t) if (a.H) a.H[e * NTR * NTR + r + NTR * t] = row[M + t];
        }
      }
      continue;
    } else {
      // scale of Me for the singularity test
      double sc = 0.0;
      if (ls < M) {
#pragma unroll
        for (int c = 0; c < M; ++c) sc = fmax(sc, fabs(row[c]));
      }
#pragma unroll
      for (int off = SEG / 2; off > 0; off >>= 1) sc = fmax(sc, __shfl_xor(sc, off, SEG));
      const double thr = 1e-10 * sc;
      bool bad = !(g.detJ > 0.0);
      bool used = !(ls < M);
      int mycol = 0;
      double mypiv = 1.0;
      // Gauss-Jordan on the first M columns with partial pivoting among
      // the rows not yet used as pivots (implicit row exchanges), or with
      // NAT row j as the pivot of column j.  Pivot rows
      // stay unscaled (each row is divided by its own pivot at the end), so
      // an elimination step is one fma per column.  The pivot search reduces
      // a 32-bit key: the upper word of |a_ij| with its low 6 bits replaced
      // by 63 - (row index) (ties and near-ties go to the lower row, the
      // same choice for every SEG).
      // The pivot row reaches the other lanes in two ways: columns below CS
      // by cross-lane moves (readlane or bpermute, VALU / LDS-permute work
      // per column), the border columns from CS on (present in every step)
      // through the segment's row in LDS: the pivot lane writes them, every
      // lane reads them back as broadcast 16-byte loads.  LDS instructions
      // of one wavefront execute in order, so the reads see the write and
      // the next step's write follows these reads.
      double* pb = prow + ((threadIdx.x >> 6) * EPW + seg) * PBS;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        int bl = j;
        if constexpr (!NAT) {
          int key = used ? -1 : (int)(((unsigned)__double2hiint(fabs(row[j])) & ~63u) | (unsigned)(63 - ls));
          key = seg_max_int<SEG>(key, seg);
          bl = 63 - (key & 63);
        }
        const int src = sbase + bl;
        const double pv = seg_bcast<SEG>(row[j], src);
        bad |= !(fabs(pv) > thr);
        const bool me = (ls == bl);
        const double f = me ? 0.0 : row[j] / pv;
        // LDS part of the pivot row: columns c0.. of the buffer (pairs)
        const int c0 = (j + 1 > CS ? j + 1 - CS : 0) & ~1;
        if (me) {
#pragma unroll
          for (int c = 0; c < NB; c += 2) {
            if (c < c0) continue;
            double2 v;
            v.x = row[CS + c];
            v.y = c + 1 < NB ? row[CS + c + 1] : 0.0;
            *reinterpret_cast<double2*>(pb + c) = v;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // pivoting on 32-lane segments: the first kRl32 (q,u) columns by
        // readlane pairs, the others by permutes, which balances VALU and
        // LDS issue (k = 2 on 512^2 cells: 1.91 -> 1.75 ms; 9 columns 1.79,
        // 16 columns 1.83, all 18 1.95)
        constexpr int kRl32 = (!NAT && SEG == 32) ? 12 : 0;
        int s0 = 0, s1 = 0;
        if constexpr (kRl32 > 0) {
          s0 = __builtin_amdgcn_readlane(src, 0);
          s1 = __builtin_amdgcn_readlane(src, 32);
        }
#pragma unroll
        for (int c = j + 1; c < CS; ++c) {
          if (NAT && c < NQ && c / NU != j / NU) continue;
          const double b = c < kRl32 ? halves_bcast(row[c], s0, s1, seg) : seg_bcast<SEG>(row[c], src);
          row[c] = fma(-f, b, row[c]);
        }
#pragma unroll
        for (int c = 0; c < NB; c += 2) {
          if (c < c0) continue;
          const double2 v = *reinterpret_cast<const double2*>(pb + c);
          if (CS + c > j) row[CS + c] = fma(-f, v.x, row[CS + c]);
          if (c + 1 < NB) row[CS + c + 1] = fma(-f, v.y, row[CS + c + 1]);
        }
        __builtin_amdgcn_wave_barrier();
        if (me) {
          used = true;
          mycol = j;
          mypiv = pv;
        }
      }
      if (ls < M) {
        const double ip = 1.0 / mypiv;
#pragma unroll
        for (int c = M; c < D::NC; ++c) row[c] *= ip;
      }
      if (valid) {
        if (bad && ls == 0) atomicOr(a.err + (GROUPED ? a.cell_inst[cell] : 0), g.detJ > 0.0 ? 2 : 1);
        if (ls < M) {
          double* ke = a.Ke + cell * (int64_t)(M * NTR);
#pragma unroll
          for (int b = 0; b < NTR; ++b) ke[b * M + mycol] = row[M + b];
          a.be[cell * M + mycol] = row[M + NTR];
        } else if (ls < D::ROWS) {
          const int r = ls - M;
          double* at = a.Ate + cell * (int64_t)(NTR * NTR) + r * NTR;
#pragma unroll
          for (int b = 0; b < NTR; ++b) at[b] = -row[M + b];
          a.bte[cell * NTR + r] = row[M + NTR];
        }
      }
    }
  }
}

// -------------------------------------------------------------- pattern
__global__ void pattern_kernel(DevMesh m, DevPattern P) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf) return;
  const int c1 = m.fcell[2 * f], c2 = m.fcell[2 * f + 1];
  int la1 = 0, la2 = 0;
  for (int l = 0; l < 3; ++l) {
    if (m.cf[3 * c1 + l] == f) la1 = l;
    if (c2 >= 0 && m.cf[3 * c2 + l] == f) la2 = l;
  }
  int col[5], s1[5], s2[5], n = 0;
  col[n] = (int)f;
  s1[n] = c1 * 9 + la1 * 3 + la1;
  s2[n] = c2 >= 0 ? c2 * 9 + la2 * 3 + la2 : -1;
  ++n;
  for (int l = 0; l < 3; ++l)
    if (l != la1) { col[n] = m.cf[3 * c1 + l]; s1[n] = c1 * 9 + la1 * 3 + l; s2[n] = -1; ++n; }
  if (c2 >= 0)
    for (int l = 0; l < 3; ++l)
      if (l != la2) { col[n] = m.cf[3 * c2 + l]; s1[n] = c2 * 9 + la2 * 3 + l; s2[n] = -1; ++n; }
  for (int i = 1; i < n; ++i) {  // insertion sort by column
    int cc = col[i], a1 = s1[i], a2 = s2[i], j = i - 1;
    while (j >= 0 && col[j] > cc) { col[j + 1] = col[j]; s1[j + 1] = s1[j]; s2[j + 1] = s2[j]; --j; }
    col[j + 1] = cc; s1[j + 1] = a1; s2[j + 1] = a2;
  }
  int self = 0;
  for (int i = 0; i < n; ++i)
    if (col[i] == f) self = i;
  for (int i = n; i < kSlots; ++i) { col[i] = (int)f; s1[i] = -1; s2[i] = -1; }
  for (int i = 0; i < kSlots; ++i) {
    P.bcol[f * kSlots + i] = col[i];
    P.src[(f * kSlots + i) * 2] = s1[i];
    P.src[(f * kSlots + i) * 2 + 1] = s2[i];
  }
  P.nslot[f] = (unsigned char)n;
  P.self[f] = (unsigned char)self;
  if (P.nbr) {
    // the block face f owns in each cell couples it to local face la + 1;
    // it meets local face la + 2 through that face's block
    int o[4] = {(int)f, (int)f, (int)(2 * f + 1), (int)(2 * f + 1)};
    const int cs[2] = {c1, c2}, ls[2] = {la1, la2};
    for (int k = 0; k < 2; ++k) {
      if (cs[k] < 0) continue;
      o[k] = m.cf[3 * cs[k] + (ls[k] + 1) % 3];
      const int g = m.cf[3 * cs[k] + (ls[k] + 2) % 3];
      o[2 + k] = 2 * g + (m.fcell[2 * g] == cs[k] ? 0 : 1);
    }
    *reinterpret_cast<int4*>(P.nbr + 4 * f) = make_int4(o[0], o[1], o[2], o[3]);
  }
}

__global__ void tslot_kernel(int64_t nf, DevPattern P) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int n = P.nslot[f];
  for (int s = 0; s < kSlots; ++s) {
    unsigned char t = P.self[f];
    if (s < n) {
      const int g = P.bcol[f * kSlots + s];
      const int ng = P.nslot[g];
      for (int u = 0; u < ng; ++u)
        if (P.bcol[(int64_t)g * kSlots + u] == f) t = (unsigned char)u;
    }
    P.tslot[f * kSlots + s] = t;
  }
}

// ----------------------------------------------------------------- fill
// Row (f, a): slot s, entry b of the trace matrix K (assembler.jl:31-49,
// duplicates summed in cell order) and the assembled rhs.
template <int NT>
__global__ __launch_bounds__(256) void fill_kernel(DevPattern P, const double* __restrict__ Ate,
                                                   const double* __restrict__ bte, TraceSystem T) {
  constexpr int NTR = 3 * NT;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T.N) return;
  const int64_t f = r / NT;
  const int a = (int)(r - f * NT);
  const int* sp = P.src + f * kSlots * 2;
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    const int c1 = sp[2 * s], c2 = sp[2 * s + 1];
    double v[NT];
#pragma unroll
    for (int b = 0; b < NT; ++b) v[b] = 0.0;
    if (c1 >= 0) {
      const int64_t cell = c1 / 9;
      const int la = (c1 / 3) % 3, lb = c1 % 3;
      const double* blk = Ate + cell * NTR * NTR + (la * NT + a) * NTR + lb * NT;
#pragma unroll
      for (int b = 0; b < NT; ++b) v[b] = blk[b];
    }
    if (c2 >= 0) {
      const int64_t cell = c2 / 9;
      const int la = (c2 / 3) % 3, lb = c2 % 3;
      const double* blk = Ate + cell * NTR * NTR + (la * NT + a) * NTR + lb * NT;
#pragma unroll
      for (int b = 0; b < NT; ++b) v[b] = v[b] + blk[b];
    }
#pragma unroll
    for (int b = 0; b < NT; ++b) T.vals[(int64_t)(s * NT + b) * T.N + r] = v[b];
  }
  const int sf = P.self[f];
  const int c1 = sp[2 * sf], c2 = sp[2 * sf + 1];
  double rh = bte[(int64_t)(c1 / 9) * NTR + ((c1 / 3) % 3) * NT + a];
  if (c2 >= 0) rh = rh + bte[(int64_t)(c2 / 9) * NTR + ((c2 / 3) % 3) * NT + a];
  T.rhs[r] = rh;
}

// Entry a of the row v held by lane B of the caller's quad.
template <int B>
__device__ __forceinline__ double quad_entry(const double (&v)[4], int a) {
  const double x0 = dpp_quad<quad_ctrl<4, B>()>(v[0]), x1 = dpp_quad<quad_ctrl<4, B>()>(v[1]);
  const double x2 = dpp_quad<quad_ctrl<4, B>()>(v[2]), x3 = dpp_quad<quad_ctrl<4, B>()>(v[3]);
  return a == 0 ? x0 : (a == 1 ? x1 : (a == 2 ? x2 : x3));
}

// The symmetric blocks (kSymBlocks<NT>) and the rhs, cell by cell: thread
// per (cell c, local face l, j, row a), so that a cell's threads read its
// element block Ate together.  j = 1: the block face f = cf[3c + l] owns in
// c (coupling to local face l + 1), into own block k of f, k the side of c
// in fcell.  j = 0, by the cell on side 0 only: the upper triangle of K_ff,
// the two cells' diagonal blocks summed in cell order (as the face-ELL sums
// them), the rhs entry, and a zero own block 1 where f has one cell.  The
// triangle holds (K_ab + K_ba) / 2: the rows a = 0..3 of one (c, l) are the
// lanes of one quad (24 items per cell), so K_ba comes from lane b by DPP,
// and the quad returns or goes on as one.
template <int NT>
__global__ __launch_bounds__(256) void fill_sym_kernel(DevMesh m, const double* __restrict__ Ate,
                                                       const double* __restrict__ bte, TraceSystem T) {
  constexpr int NTR = 3 * NT, B = NT * NT;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.nc * 3 * 2 * NT) return;
  const int64_t c = t / (6 * NT);
  const int rem = (int)(t - c * 6 * NT);
  const int l = rem / (2 * NT), j = (rem / NT) & 1, a = rem % NT;
  const int64_t f = m.cf[3 * c + l];
  const int k = m.fcell[2 * f] == c ? 0 : 1;
  const double* row = Ate + c * NTR * NTR + (l * NT + a) * NTR;
  double v[NT];
  if (j == 1) {
#pragma unroll
    for (int b = 0; b < NT; ++b) v[b] = row[((l + 1) % 3) * NT + b];
    double* out = T.blk + (int64_t)(NT * (NT + 1) / 2) * T.nf + (2 * f + k) * B + a * NT;
#pragma unroll
    for (int b = 0; b < NT; ++b) out[b] = v[b];
    return;
  }
  if (k != 0) return;
#pragma unroll
  for (int b = 0; b < NT; ++b) v[b] = row[l * NT + b];
  double rh = bte[c * NTR + l * NT + a];
  const int64_t c2 = m.fcell[2 * f + 1];
  if (c2 >= 0) {
    int l2 = 0;
    for (int q = 1; q < 3; ++q)
      if (m.cf[3 * c2 + q] == f) l2 = q;
    const double* r2 = Ate + c2 * NTR * NTR + (l2 * NT + a) * NTR + l2 * NT;
#pragma unroll
    for (int b = 0; b < NT; ++b) v[b] = v[b] + r2[b];
    rh = rh + bte[c2 * NTR + l2 * NT + a];
  } else {
    double* own1 = T.blk + (int64_t)(NT * (NT + 1) / 2) * T.nf + (2 * f + 1) * B + a * NT;
#pragma unroll
    for (int b = 0; b < NT; ++b) own1[b] = 0.0;
  }
  // K_ff's upper triangle, the mean of K_ab and K_ba: row a from column a on
  static_assert(NT == 4, "the quad exchange holds the four rows of a face");
  const double w[NT] = {quad_entry<0>(v, a), quad_entry<1>(v, a), quad_entry<2>(v, a), quad_entry<3>(v, a)};
#pragma unroll
  for (int b = 0; b < NT; ++b)
    if (b >= a) T.blk[(int64_t)dinv_slot(NT, a, b) * T.nf + f] = 0.5 * (v[b] + w[b]);
  T.rhs[f * NT + a] = rh;
}

// The face-ELL values of the symmetric blocks (kSymBlocks<NT>), for the
// exports: row (f, a), slot s, entry b = K[(f, a), (bcol, b)]; padding slots
// zero.
template <int NT>
__global__ __launch_bounds__(256) void expand_ell_kernel(DevPattern P, TraceSystem T, double* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T.N) return;
  const int64_t f = r / NT;
  const int a = (int)(r - f * NT);
  double row[5][NT];
  int64_t col[5];
  int e = 0;
  sym_row_entries<NT>(P.nbr, T.blk, T.nf, f, a, [&](int64_t g, int b, double v) {
    row[e / NT][b] = v;
    col[e / NT] = g;
    ++e;
  });
  const int n = P.nslot[f];
  for (int s = 0; s < kSlots; ++s) {
    const int64_t g = P.bcol[f * kSlots + s];
    int m = -1;  // the block of column face g: K_ff first (a boundary face's missing blocks name f)
    for (int q = 4; q >= 0; --q)
      if (col[q] == g) m = q;
#pragma unroll
    for (int b = 0; b < NT; ++b) vals[(int64_t)(s * NT + b) * T.N + r] = (s < n && m >= 0) ? row[m][b] : 0.0;
  }
}

// Block-Jacobi: inverse of -K_ff restricted to the non-Dirichlet dofs.
template <int NT>
__global__ void jacobi_kernel(DevPattern P, TraceSystem T) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= T.nf) return;
  const int s0 = P.self[f];
  double a[NT][2 * NT];
  bool dm[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) dm[i] = T.dmask[f * NT + i] != 0;
  // column j of the face's diagonal block: the NT consecutive rows of
  // slot column (s0, j), read in 16-byte pieces where NT is even (the block
  // starts at a multiple of 2 doubles); with symmetric blocks the column of
  // the row-major K_ff
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    double cv[NT];
    if constexpr (kSymBlocks<NT>) {
#pragma unroll
      for (int i = 0; i < NT; ++i) cv[i] = T.blk[(int64_t)dinv_slot(NT, i, j) * T.nf + f];
    } else if constexpr (NT % 2 == 0) {
      const double* col = T.vals + (int64_t)(s0 * NT + j) * T.N + f * NT;
#pragma unroll
      for (int i = 0; i < NT; i += 2) {
        const double2 w = *reinterpret_cast<const double2*>(col + i);
        cv[i] = w.x;
        cv[i + 1] = w.y;
      }
    } else {
      const double* col = T.vals + (int64_t)(s0 * NT + j) * T.N + f * NT;
#pragma unroll
      for (int i = 0; i < NT; ++i) cv[i] = col[i];
    }
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      double v = -cv[i];
      if (dm[i] || dm[j]) v = (i == j) ? 1.0 : 0.0;
      a[i][j] = v;
      a[i][NT + j] = (i == j) ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int c = 0; c < NT; ++c) {
    int p = c;
    double mx = fabs(a[c][c]);
#pragma unroll
    for (int i = c + 1; i < NT; ++i)
      if (fabs(a[i][c]) > mx) {
        mx = fabs(a[i][c]);
        p = i;
      }
    // row exchange through a fully unrolled select keeps `a` in registers
#pragma unroll
    for (int i = c + 1; i < NT; ++i)
      if (i == p)
#pragma unroll
        for (int j = 0; j < 2 * NT; ++j) { const double t = a[c][j]; a[c][j] = a[i][j]; a[i][j] = t; }
    const double inv = 1.0 / a[c][c];
#pragma unroll
    for (int j = 0; j < 2 * NT; ++j) a[c][j] *= inv;
#pragma unroll
    for (int i = 0; i < NT; ++i)
      if (i != c) {
        const double l = a[i][c];
#pragma unroll
        for (int j = 0; j < 2 * NT; ++j) a[i][j] -= l * a[c][j];
      }
  }
  if constexpr (kQuadRows<NT>) {
    // the upper triangle of the mean of the inverse and its transpose
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = i; j < NT; ++j)
        T.Dinv[(int64_t)dinv_slot(NT, i, j) * T.nf + f] =
            (dm[i] || dm[j]) ? 0.0 : 0.5 * (a[i][NT + j] + a[j][NT + i]);
  } else {
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      double* out = T.Dinv + (int64_t)j * T.N + f * NT;
#pragma unroll
      for (int i = 0; i < NT; ++i) out[i] = (dm[i] || dm[j]) ? 0.0 : a[i][NT + j];
    }
  }
}

// ------------------------------------------------------------------- CG
// Preconditioned CG on S = -K restricted to the non-Dirichlet dofs, two
// kernels per iteration and no atomics or device-wide fences: each kernel
// writes one partial sum per block, and the kernel that consumes a scalar
// re-reduces those partials itself, in a fixed order, so every block gets
// the same bits.  Iteration i (i >= 1), host-numbered:
//   spmv(i):   rz_i = (r, z) and rr_i = (r, r) from update(i-1)'s partials;
//              stop if rr_i <= rtol^2 bb or i-1 >= maxit; beta = rz_i/rz_{i-1}
//              (0 for i = 1); p_i = z + beta p_{i-1}; q = S z + beta q (= S p_i);
//              (p_i, q)
//   update(i): alpha = rz_i/(p_i, q); x += alpha p_i; r -= alpha q;
//              z = Dinv r; partials of (r, z) and (r, r)
// The init kernel plays update(0): r = b~, z = Dinv r, x = p = 0.
// Partials: part[0:nb) (p,q); part[nb(1+par):...) (r,z) by parity; part[3nb:4nb) (r,r).
constexpr int kCgThreads = 256;

__device__ inline double pnew_of(double z, double p, double beta) { return fma(beta, p, z); }

// Sum of n partials in a fixed order, returned to every thread of the block.
template <int NW>
__device__ inline double sum_partials(const double* p, int n, double* sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  s = wave_sum(s);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += sh[i];
  __syncthreads();
  return t;
}

// CG iteration i, first half: convergence test on the (r, r) partials of
// update(i-1), beta, p = z + beta p_old, q = S p, partial (p, q).  G selects
// the instance of the block (one, or one of a group).  S p = S z + beta S p_old
// and S p_old is the previous q, so the product gathers z alone and q is
// carried by that recurrence (its rounding, like that of a direct product, is
// of order eps ||S|| ||p||, and the beta factors damp what earlier steps left:
// the true residual of the result is checked after every solve).  Gathering
// z and p_old for p at every neighbour cost 1024^2 P3: 474 -> 410 us,
// 512^2 P2: 88 -> 76 us.  The matrix values are read once per launch, by
// nontemporal loads, so that they do not displace the gathered z from the
// caches (1024^2 P3 face-ELL: 562 -> 555 us; 512^2 P2: 91 -> 86 us).
template <int NT, class G>
__global__ __launch_bounds__(kCgThreads) void cg_spmv_kernel(DevPattern P, TraceOp<double> T, Krylov<double> v,
                                                             G gsel, int i_fixed, double rtol2, int maxit,
                                                             int force) {
  constexpr int NW = kCgThreads / 64;
  __shared__ double sh[NW];
  const Seg S = seg_of(gsel);
  CgState* st = S.st;
  if (!force && st->done) return;
  const int i = i_fixed > 0 ? i_fixed : st->i_spmv;
  const size_t nbt = (size_t)v.nblocks;  // partial-sum slot stride of the whole launch
  const double rz = sum_partials<NW>(v.part + nbt * (1 + (i & 1)) + S.b0, S.nb, sh);
  const double rr = sum_partials<NW>(v.part + nbt * 3 + S.b0, S.nb, sh);
  const double bb = (i == 1) ? rr : st->bb;
  const int lb = blockIdx.x - S.b0;
  const bool lead = lb == 0 && threadIdx.x == 0;
  if (!force) {
    const bool conv = rr <= rtol2 * bb;
    if (conv || i - 1 >= maxit) {
      if (lead) {
        st->done = conv ? 1 : 2;
        st->it = i - 1;
        st->rr = rr;
        st->bb = bb;
      }
      return;
    }
    if (lead) {
      st->rz[i & 1] = rz;
      st->it = i - 1;
      st->rr = rr;
      if (i == 1) st->bb = rr;
      st->i_upd = i;
      if (i <= kLanczos) st->lz[i] = rz;
    }
  }
  const double beta = (i == 1) ? 0.0 : rz / st->rz[(i - 1) & 1];
  const double* __restrict__ pold = (i & 1) ? v.p0 : v.p1;
  double* __restrict__ pnw = (i & 1) ? v.p1 : v.p0;
  const double* __restrict__ z = v.z;
  double* __restrict__ qo = v.q;
  double part = 0.0;
  for (int64_t r = S.lo + (int64_t)lb * blockDim.x + threadIdx.x; r < S.hi; r += (int64_t)S.nb * blockDim.x) {
    const int64_t f = r / NT;
    double acc = 0.0;
    if constexpr (kSymBlocks<NT>) {
      acc = sym_row<NT>(P.nbr, T.blk, T.nf, f, (int)(r - f * NT), [&](int64_t g, int b) { return z[g * NT + b]; });
    } else {
      const int* bc = P.bcol + f * kSlots;
#pragma unroll
      for (int s = 0; s < kSlots; ++s) {
        const int64_t g = (int64_t)__ldg(bc + s) * NT;
#pragma unroll
        for (int b = 0; b < NT; ++b)
          acc = fma(__builtin_nontemporal_load(T.vals + (int64_t)(s * NT + b) * T.N + r), z[g + b], acc);
      }
    }
    double q = T.dmask[r] ? 0.0 : -acc;
    if (i > 1) q = fma(beta, qo[r], q);  // S p = S z + beta S p_old
    const double pn = pnew_of(z[r], pold[r], beta);
    pnw[r] = pn;
    qo[r] = q;
    part = fma(pn, q, part);
  }
  const double bs = block_sum<NW>(part, sh);
  if (threadIdx.x == 0) v.part[blockIdx.x] = bs;
}

// X false: the iterate x is not carried (the Lanczos run behind the
// smoothing weight needs only the CG coefficients).
template <int NT, class G, bool X = true>
__global__ __launch_bounds__(64 * NT) void cg_update_kernel(TraceSystem T, CgVectors v, G gsel, int i_fixed,
                                                            int force) {
  constexpr int NTH = 64 * NT;
  __shared__ double rs[NTH];
  __shared__ double sh[NT];
  const Seg S = seg_of(gsel);
  CgState* st = S.st;
  if (!force && st->done) return;
  const int i = i_fixed > 0 ? i_fixed : st->i_upd;
  const int lb = blockIdx.x - S.b0;
  if (!force && lb == 0 && threadIdx.x == 0) st->i_spmv = i + 1;
  const size_t nbt = (size_t)v.nblocks;
  const double pq = sum_partials<NT>(v.part + S.b0, S.nb, sh);
  const double rz = st->rz[i & 1];
  if (!force && !(pq > 0.0)) {  // loss of positive curvature or NaN
    if (lb == 0 && threadIdx.x == 0) {
      st->done = 3;
      st->it = i - 1;
    }
    return;
  }
  if (!force && i <= kLanczos && lb == 0 && threadIdx.x == 0) st->lpq[i] = pq;
  const double alpha = rz / pq;
  const double* __restrict__ p = (i & 1) ? v.p1 : v.p0;
  double prz = 0.0, prr = 0.0;
  for (int64_t r0 = S.lo + (int64_t)lb * NTH; r0 < S.hi; r0 += (int64_t)S.nb * NTH) {
    const int64_t r = r0 + threadIdx.x;
    double rn = 0.0;
    if (r < S.hi) {
      rn = fma(-alpha, v.q[r], v.r[r]);
      if (!force) {
        if constexpr (X) v.x[r] = fma(alpha, p[r], v.x[r]);
        v.r[r] = rn;
      }
    }
    jacobi_apply<NT>(T.Dinv, T.nf, rs, rn, r, r < S.hi, [&](double zr) {
      v.z[r] = zr;
      prz = fma(rn, zr, prz);
      prr = fma(rn, rn, prr);
    });
  }
  const double a = block_sum<NT>(prz, sh);
  const double b = block_sum<NT>(prr, sh);
  if (threadIdx.x == 0 && !force) {
    v.part[nbt * (1 + ((i + 1) & 1)) + blockIdx.x] = a;
    v.part[nbt * 3 + blockIdx.x] = b;
  }
}

// update(0): r0 = b~ = -(rhs - K uD) on free dofs, x = 0, z = Dinv r0,
// p = 0; partials of (r0, z0) and (r0, r0).
__device__ inline double hash01_k(uint64_t i) {
  uint64_t z = i * 0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// RAND: the right-hand side is a fixed pseudo-random vector in [-1/2, 1/2)
// on the free dofs (indexed from the instance's first row), for the Lanczos
// estimate.
template <int NT, class G, bool RAND = false>
__global__ __launch_bounds__(64 * NT) void cg_init_kernel(DevPattern P, TraceSystem T, CgVectors v, G gsel) {
  constexpr int NTH = 64 * NT;
  __shared__ double rs[NTH];
  __shared__ double sh[NT];
  const Seg S = seg_of(gsel);
  const int lb = blockIdx.x - S.b0;
  double prz = 0.0, prr = 0.0;
  for (int64_t r0 = S.lo + (int64_t)lb * NTH; r0 < S.hi; r0 += (int64_t)S.nb * NTH) {
    const int64_t r = r0 + threadIdx.x;
    double b = 0.0;
    if (r < S.hi) {
      if (RAND) {
        b = T.dmask[r] ? 0.0 : hash01_k((uint64_t)(r - S.lo)) - 0.5;
      } else if (!T.dmask[r]) {
        const int64_t f = r / NT;
        double s = T.rhs[r];
        if constexpr (kSymBlocks<NT>) {
          s -= sym_row_lift<NT>(P.nbr, T.blk, T.nf, T.dmask, T.uD, f, (int)(r - f * NT));
        } else {
#pragma unroll
          for (int sl = 0; sl < kSlots; ++sl) {
            const int64_t g = (int64_t)P.bcol[f * kSlots + sl] * NT;
#pragma unroll
            for (int c = 0; c < NT; ++c)  // uD is zero off the Dirichlet dofs
              if (T.dmask[g + c]) s -= T.vals[(int64_t)(sl * NT + c) * T.N + r] * T.uD[g + c];
          }
        }
        b = -s;
      }
      v.r[r] = b;
      v.x[r] = 0.0;
      v.p0[r] = 0.0;
      v.p1[r] = 0.0;
    }
    jacobi_apply<NT>(T.Dinv, T.nf, rs, b, r, r < S.hi, [&](double zr) {
      v.z[r] = zr;
      prz = fma(b, zr, prz);
      prr = fma(b, b, prr);
    });
  }
  const double a = block_sum<NT>(prz, sh);
  const double bb = block_sum<NT>(prr, sh);
  if (threadIdx.x == 0) {
    v.part[(size_t)v.nblocks * 2 + blockIdx.x] = a;  // parity of update(0) + 1 = 1
    v.part[(size_t)v.nblocks * 3 + blockIdx.x] = bb;
  }
  if (lb == 0 && threadIdx.x == 0) {
    CgState* st = S.st;
    st->done = 0;
    st->it = 0;
    st->pq = 0.0;
    st->i_spmv = 1;
  }
}

// float copy of a double array
__global__ __launch_bounds__(256) void to_float_kernel(const double* __restrict__ src, float* __restrict__ dst,
                                                       int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// Plain streaming copy (HBM traffic calibration for hardware counters).
__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   int64_t n2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Final sum of n partials by one block (a separate launch, so no fences).
__global__ __launch_bounds__(256) void reduce_kernel(const double* part, int n, double* out) {
  __shared__ double sh[4];
  const double s = sum_partials<4>(part, n, sh);
  if (threadIdx.x == 0) *out = s;
}

__global__ void cg_finish_kernel(TraceSystem T, const double* x, double* uhat) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T.N) return;
  uhat[r] = T.dmask[r] ? T.uD[r] : x[r];
}

// Sum of each instance's partials by one block per instance: out[j] over
// part[seg[j].b0 .. seg[j].b0 + seg[j].nb), in reduce_kernel's order.
__global__ __launch_bounds__(256) void reduce_seg_kernel(const double* part, const Seg* seg, double* out) {
  __shared__ double sh[4];
  const Seg S = seg[blockIdx.x];
  const double s = sum_partials<4>(part + S.b0, S.nb, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ||K uhat - rhs||^2 over the free dofs (the FP64 residual of the solve).
template <int NT, class G>
__global__ __launch_bounds__(kCgThreads) void true_residual_kernel(DevPattern P, TraceSystem T,
                                                                   const double* uhat, double* part, G gsel) {
  __shared__ double sh[kCgThreads / 64];
  const Seg S = seg_of(gsel);
  const int lb = blockIdx.x - S.b0;
  double acc2 = 0.0;
  for (int64_t r = S.lo + (int64_t)lb * blockDim.x + threadIdx.x; r < S.hi; r += (int64_t)S.nb * blockDim.x) {
    if (T.dmask[r]) continue;
    const int64_t f = r / NT;
    double s = -T.rhs[r];
    if constexpr (kSymBlocks<NT>) {
      s += sym_row<NT>(P.nbr, T.blk, T.nf, f, (int)(r - f * NT), [&](int64_t g, int c) { return uhat[g * NT + c]; });
    } else {
#pragma unroll
      for (int sl = 0; sl < kSlots; ++sl) {
        const int64_t g = (int64_t)P.bcol[f * kSlots + sl] * NT;
#pragma unroll
        for (int c = 0; c < NT; ++c) s = fma(T.vals[(int64_t)(sl * NT + c) * T.N + r], uhat[g + c], s);
      }
    }
    acc2 = fma(s, s, acc2);
  }
  const double bs = block_sum<kCgThreads / 64>(acc2, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// -------------------------------------------------------------- recover
template <int K>
__global__ __launch_bounds__(256) void recover_kernel(DevMesh m, const double* __restrict__ Ke,
                                                      const double* __restrict__ be,
                                                      const double* __restrict__ uhat, double* qu) {
  using D = Dims<K>;
  constexpr int M = D::M, NT = D::NT, NTR = D::NTR;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.nc * M) return;
  const int64_t c = t / M;
  const int i = (int)(t - c * M);
  const double* ke = Ke + c * (M * NTR);
  double s = 0.0;
#pragma unroll
  for (int lf = 0; lf < 3; ++lf) {
    const int64_t f = m.cf[3 * c + lf];
#pragma unroll
    for (int b = 0; b < NT; ++b) s = fma(ke[(lf * NT + b) * M + i], uhat[f * NT + b], s);
  }
  qu[t] = s + be[t];
}

// ------------------------------------------------------------ errornorm
// Per block: sum over its cells of int (u_h - u_ex)^2; G selects the
// instance (cells [lo, hi)).
template <int K, class G>
__global__ __launch_bounds__(256) void errornorm_kernel(DevMesh m, const double* __restrict__ Tg,
                                                        TabLayoutDev L, const double* __restrict__ qu,
                                                        int kind, double c0, double c1, double c2,
                                                        double* part, G gsel) {
  using D = Dims<K>;
  constexpr int NU = D::NU, NQ = D::NQ, M = D::M;
  __shared__ double sh[4];
  const Seg S = seg_of(gsel);
  const int lb = blockIdx.x - S.b0;
  double acc = 0.0;
  for (int64_t c = S.lo + (int64_t)lb * blockDim.x + threadIdx.x; c < S.hi; c += (int64_t)S.nb * blockDim.x) {
    Geo g;
    cell_geometry(m, c, g);
    double el = 0.0;
    for (int q = 0; q < L.nqp; ++q) {
      double uh = 0.0;
#pragma unroll
      for (int i = 0; i < NU; ++i) uh = fma(qu[c * M + NQ + i], Tg[L.N + i * L.nqp + q], uh);
      double x = 0.0, y = 0.0;
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        x += Tg[L.Mg + v * L.nqp + q] * g.x[2 * v];
        y += Tg[L.Mg + v * L.nqp + q] * g.x[2 * v + 1];
      }
      const double ue = kind == 1 ? c0 + c1 * x + c2 * y : sin(kPi * x) * sin(kPi * y);
      const double d = uh - ue;
      el += d * d * (g.detJ * Tg[L.W + q]);
    }
    acc += el;
  }
  const double bs = block_sum<4>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// Largest eigenvalue of the Lanczos tridiagonal matrix of one instance's
// run: alpha_j = lz[j] / lpq[j], beta_j = lz[j+1] / lz[j], diagonal
// 1/alpha_j + beta_{j-1}/alpha_{j-1}, off-diagonal sqrt(beta_j)/alpha_j,
// over the m iterations that completed; by bisection on Sturm counts.
// omega = factor / lambda (0.5 when no iteration completed).
__device__ void lanczos_omega(CgState* st, int m, double factor) {
  if (st->done) m = min(m, st->it);
  double d[kLanczos], e2[kLanczos];  // diagonal, squared off-diagonal
  int n = 0;
  for (int j = 1; j <= m; ++j) {
    const double a = st->lz[j] / st->lpq[j];
    double dj = 1.0 / a;
    if (j > 1) dj += (st->lz[j] / st->lz[j - 1]) / (st->lz[j - 1] / st->lpq[j - 1]);
    if (!(dj > 0.0) || !isfinite(dj)) break;
    d[n] = dj;
    e2[n] = j < m ? (st->lz[j + 1] / st->lz[j]) / (a * a) : 0.0;
    if (!isfinite(e2[n])) break;
    ++n;
  }
  if (n == 0) {
    st->omega = 0.5;
    return;
  }
  double hi = 0.0;
  for (int j = 0; j < n; ++j) {
    const double r = (j > 0 ? sqrt(e2[j - 1]) : 0.0) + (j + 1 < n ? sqrt(e2[j]) : 0.0);
    hi = fmax(hi, d[j] + r);
  }
  double lo = 0.0;
  for (int it = 0; it < 200 && hi - lo > 1e-13 * hi; ++it) {
    const double x = 0.5 * (lo + hi);
    int below = 0;  // eigenvalues below x
    double q = 1.0;
    for (int j = 0; j < n; ++j) {
      q = d[j] - x - (j > 0 ? e2[j - 1] / q : 0.0);
      if (q == 0.0) q = -1e-300;
      below += q < 0.0;
    }
    if (below < n) lo = x;  // the largest eigenvalue is above x
    else hi = x;
  }
  st->omega = factor / hi;
}

__global__ void lanczos_omega_kernel(CgState* st, int m, double factor) {
  if (threadIdx.x == 0) lanczos_omega(st, m, factor);
}

__global__ void lanczos_omega_seg_kernel(const Seg* seg, int ninst, int m, double factor) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < ninst) lanczos_omega(seg[j].st, m, factor);
}

int grid_for(int64_t n, int threads, int cap) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

template <int K, int SEG>
int condense_launch_seg(const CondenseArgs& a, bool blocks, int grid_cap, hipStream_t s) {
  const int64_t n = blocks ? a.count : a.mesh.nc;
  const int64_t per_block = 4 * (64 / SEG);
  const int grid = grid_for(n, (int)per_block, grid_cap > 0 ? grid_cap : 8192);
  const size_t lds = sizeof(double) * a.lay.total;
  // natural order only where a rule can make the mass SPD: at k = 4 the
  // rules with positive weights have too few points
  const bool nat = a.natural && K <= 3;
  if (blocks)
    hipLaunchKernelGGL((condense_kernel<K, true, false, SEG, false>), dim3(grid), dim3(256), lds, s, a);
  else if (a.srcs && nat)
    hipLaunchKernelGGL((condense_kernel<K, false, true, SEG, K <= 3>), dim3(grid), dim3(256), lds, s, a);
  else if (a.srcs)
    hipLaunchKernelGGL((condense_kernel<K, false, true, SEG, false>), dim3(grid), dim3(256), lds, s, a);
  else if (nat)
    hipLaunchKernelGGL((condense_kernel<K, false, false, SEG, K <= 3>), dim3(grid), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL((condense_kernel<K, false, false, SEG, false>), dim3(grid), dim3(256), lds, s, a);
  return (int)hipGetLastError();
}

// seg 0 picks the narrowest segment that holds the rows
template <int K>
int condense_launch(const CondenseArgs& a, bool blocks, const CondenseSched& sc, hipStream_t s) {
  constexpr int NAT = Dims<K>::SEG;
  const int seg = sc.seg > 0 ? sc.seg : NAT;
  if (seg < NAT) return (int)hipErrorInvalidValue;
  if constexpr (NAT <= 16) {
    if (seg == 16) return condense_launch_seg<K, 16>(a, blocks, sc.grid, s);
  }
  if constexpr (NAT <= 32) {
    if (seg == 32) return condense_launch_seg<K, 32>(a, blocks, sc.grid, s);
  }
  if (seg == 64) return condense_launch_seg<K, 64>(a, blocks, sc.grid, s);
  return (int)hipErrorInvalidValue;
}

}  // namespace

int cg_blocks(int64_t N, int nt) {
  (void)nt;
  return grid_for(N, kCgThreads, 1024);
}

int launch_pattern(const DevMesh& m, DevPattern& P, hipStream_t s) {
  const int g = grid_for(m.nf, 256, 1 << 30);
  hipLaunchKernelGGL(pattern_kernel, dim3(g), dim3(256), 0, s, m, P);
  hipLaunchKernelGGL(tslot_kernel, dim3(g), dim3(256), 0, s, m.nf, P);
  return (int)hipGetLastError();
}

int launch_condense(int k, const CondenseArgs& a, bool blocks, hipStream_t s, const CondenseSched& sc) {
  switch (k) {
    case 1: return condense_launch<1>(a, blocks, sc, s);
    case 2: return condense_launch<2>(a, blocks, sc, s);
    case 3: return condense_launch<3>(a, blocks, sc, s);
    case 4: return condense_launch<4>(a, blocks, sc, s);
    default: return (int)hipErrorInvalidValue;
  }
}

int tune_condense(int k, const CondenseArgs& a, hipStream_t s, CondenseSched* best, double* best_us,
                  double* default_us) {
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return (int)e;
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return (int)e;
  }
  const int grids[] = {512, 1024, 2048, 4096, 8192, 16384};
  constexpr int kReps = 3;
  double bt = 1e300, dt = 0.0;
  CondenseSched bs;
  int rc = 0;
  for (int seg = condense_native_seg(k); seg <= 64 && rc == 0; seg *= 2)
    for (int g : grids) {
      const CondenseSched sc{seg, g};
      rc = launch_condense(k, a, false, s, sc);  // warm-up
      if (rc) break;
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < kReps && rc == 0; ++r) rc = launch_condense(k, a, false, s, sc);
      (void)hipEventRecord(e1, s);
      if (rc || (rc = (int)hipEventSynchronize(e1))) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double us = 1e3 * ms / kReps;
      if (us < bt) {
        bt = us;
        bs = sc;
      }
      if (seg == condense_native_seg(k) && g == 8192) dt = us;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *best = bs;
  if (best_us) *best_us = bt;
  if (default_us) *default_us = dt;
  return 0;
}

int condense_native_seg(int k) {
  switch (k) {
    case 1: return Dims<1>::SEG;
    case 2: return Dims<2>::SEG;
    case 3: return Dims<3>::SEG;
    case 4: return Dims<4>::SEG;
    default: return 64;
  }
}

#define HDG_NT_DISPATCH(nt, CALL) \
  switch (nt) {                   \
    case 2: CALL(2); break;       \
    case 3: CALL(3); break;       \
    case 4: CALL(4); break;       \
    case 5: CALL(5); break;       \
    default: return (int)hipErrorInvalidValue; \
  }

int launch_fill(const DevMesh& m, const DevPattern& P, const double* Ate, const double* bte, TraceSystem& T,
                hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
  if (sym_blocks(T.nt)) {
    hipLaunchKernelGGL(fill_sym_kernel<4>, dim3(grid_for(m.nc * 6 * 4, 256, 1 << 30)), dim3(256), 0, s, m, Ate, bte,
                       T);
    return (int)hipGetLastError();
  }
  const int gr = grid_for(T.N, 256, 1 << 30);
#define FILL(NT) hipLaunchKernelGGL(fill_kernel<NT>, dim3(gr), dim3(256), 0, s, P, Ate, bte, T);
  HDG_NT_DISPATCH(T.nt, FILL)
#undef FILL
  return (int)hipGetLastError();
}

int launch_expand_ell(const DevPattern& P, const TraceSystem& T, double* vals, hipStream_t s) {
  if (!sym_blocks(T.nt) || !trace_op_ready(P, T) || !vals) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(expand_ell_kernel<4>, dim3(grid_for(T.N, 256, 1 << 30)), dim3(256), 0, s, P, T, vals);
  return (int)hipGetLastError();
}

int launch_jacobi(const DevPattern& P, TraceSystem& T, hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
  const int gf = grid_for(T.nf, 256, 1 << 30);
#define JAC(NT) hipLaunchKernelGGL(jacobi_kernel<NT>, dim3(gf), dim3(256), 0, s, P, T);
  HDG_NT_DISPATCH(T.nt, JAC)
#undef JAC
  return (int)hipGetLastError();
}

template <class G>
int launch_cg_init_t(const DevPattern& P, const TraceSystem& T, CgVectors& v, G g, int grid, hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
#define INIT(NT) hipLaunchKernelGGL((cg_init_kernel<NT, G>), dim3(grid), dim3(64 * NT), 0, s, P, T, v, g);
  HDG_NT_DISPATCH(T.nt, INIT)
#undef INIT
  return (int)hipGetLastError();
}

template <class G>
int launch_cg_spmv_t(const DevPattern& P, const TraceOp<double>& T, const Krylov<double>& v, G g, int grid,
                     int i_fixed, double rtol2, int maxit, int force, hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
#define SPMV(NT)                                                                                          \
  hipLaunchKernelGGL((cg_spmv_kernel<NT, G>), dim3(grid), dim3(kCgThreads), 0, s, P, T, v, g, i_fixed, rtol2, \
                     maxit, force);
  HDG_NT_DISPATCH(T.nt, SPMV)
#undef SPMV
  return (int)hipGetLastError();
}

// The convergence test of the next CG product (the sums, decision and state
// of cg_spmv_kernel's test), taken right after the multigrid update: a solve
// whose residual has converged then skips that iteration's V-cycle, whose
// kernels return at once on done.  One block of kCgThreads per instance, so
// that the (r, r) partials are summed in the product's order.
__device__ inline Seg check_seg(const Whole& w, int nb) { return Seg{0, w.n, 0, nb, w.st}; }
__device__ inline Seg check_seg(const Grouped& g, int) { return g.seg[blockIdx.x]; }
template <class G>
__global__ __launch_bounds__(kCgThreads) void cg_check_kernel(Krylov<double> v, G gsel, int nb, double rtol2,
                                                              int maxit) {
  constexpr int NW = kCgThreads / 64;
  __shared__ double sh[NW];
  const Seg S = check_seg(gsel, nb);
  CgState* st = S.st;
  if (st->done) return;
  const int i = st->i_pre + 1;  // the product that would take the test
  const double rr = sum_partials<NW>(v.part + (size_t)v.nblocks * 3 + S.b0, S.nb, sh);
  const double bb = st->bb;
  const bool conv = rr <= rtol2 * bb;
  if ((conv || i - 1 >= maxit) && threadIdx.x == 0) {
    st->done = conv ? 1 : 2;
    st->it = i - 1;
    st->rr = rr;
    st->bb = bb;
  }
}

int launch_cg_check(const Krylov<double>& v, int64_t N, double rtol2, int maxit, hipStream_t s) {
  hipLaunchKernelGGL((cg_check_kernel<Whole>), dim3(1), dim3(kCgThreads), 0, s, v, Whole{N, v.st}, v.nblocks, rtol2,
                     maxit);
  return (int)hipGetLastError();
}

int launch_cg_check_group(const Krylov<double>& v, Grouped g, int ninst, double rtol2, int maxit, hipStream_t s) {
  hipLaunchKernelGGL((cg_check_kernel<Grouped>), dim3(ninst), dim3(kCgThreads), 0, s, v, g, 0, rtol2, maxit);
  return (int)hipGetLastError();
}

template <class G>
int launch_cg_update_t(const TraceSystem& T, CgVectors& v, G g, int grid, int i_fixed, int force,
                       hipStream_t s) {
#define UPD(NT) \
  hipLaunchKernelGGL((cg_update_kernel<NT, G>), dim3(grid), dim3(64 * NT), 0, s, T, v, g, i_fixed, force);
  HDG_NT_DISPATCH(T.nt, UPD)
#undef UPD
  return (int)hipGetLastError();
}

template <class G>
int launch_true_residual_t(const DevPattern& P, const TraceSystem& T, const double* uhat, double* part, G g,
                           int grid, hipStream_t s) {
  if (!trace_op_ready(P, T)) return (int)hipErrorInvalidValue;
#define TR(NT) \
  hipLaunchKernelGGL((true_residual_kernel<NT, G>), dim3(grid), dim3(kCgThreads), 0, s, P, T, uhat, part, g);
  HDG_NT_DISPATCH(T.nt, TR)
#undef TR
  return (int)hipGetLastError();
}

template <class G>
int launch_errornorm_t(const DevMesh& m, const double* tables, const TabLayoutDev& lay, int k, const double* qu,
                       int uex_kind, const double* par3, double* part, G g, int grid, hipStream_t s) {
#define ERR(KK)                                                                                       \
  hipLaunchKernelGGL((errornorm_kernel<KK, G>), dim3(grid), dim3(256), 0, s, m, tables, lay, qu, uex_kind, \
                     par3[0], par3[1], par3[2], part, g);
  switch (k) {
    case 1: ERR(1); break;
    case 2: ERR(2); break;
    case 3: ERR(3); break;
    case 4: ERR(4); break;
    default: return (int)hipErrorInvalidValue;
  }
#undef ERR
  return (int)hipGetLastError();
}

#define GROUP_INSTANTIATE(G)                                                                                  \
  template int launch_cg_init_t<G>(const DevPattern&, const TraceSystem&, CgVectors&, G, int, hipStream_t);     \
  template int launch_cg_spmv_t<G>(const DevPattern&, const TraceOp<double>&, const Krylov<double>&, G, int, int, \
                                   double, int, int, hipStream_t);                                              \
  template int launch_cg_update_t<G>(const TraceSystem&, CgVectors&, G, int, int, int, hipStream_t);             \
  template int launch_true_residual_t<G>(const DevPattern&, const TraceSystem&, const double*, double*, G, int,  \
                                         hipStream_t);                                                          \
  template int launch_errornorm_t<G>(const DevMesh&, const double*, const TabLayoutDev&, int, const double*,    \
                                     int, const double*, double*, G, int, hipStream_t);
GROUP_INSTANTIATE(Whole)
GROUP_INSTANTIATE(Grouped)
#undef GROUP_INSTANTIATE

template <class G>
int launch_lanczos_omega(const DevPattern& P, const TraceSystem& T, CgVectors& v, G g, int grid, const Seg* seg,
                         int ninst, double factor, int steps, hipStream_t s) {
  steps = std::min(std::max(steps, 1), kLanczos);
#define LINIT(NT) \
  hipLaunchKernelGGL((cg_init_kernel<NT, G, true>), dim3(grid), dim3(64 * NT), 0, s, P, T, v, g);
  HDG_NT_DISPATCH(T.nt, LINIT)
#undef LINIT
  const TraceOp<double> op{T.nt, T.nf, T.N, T.vals, T.Dinv, T.dmask, T.blk};
  const Krylov<double> kv{v.r, v.z, v.p0, v.p1, v.q, v.x, v.part, v.st, v.nblocks};
  for (int i = 0; i < steps; ++i) {
    int e = launch_cg_spmv_t(P, op, kv, g, grid, 0, 0.0, steps, 0, s);
    if (e) return e;
#define LUPD(NT) \
  hipLaunchKernelGGL((cg_update_kernel<NT, G, false>), dim3(grid), dim3(64 * NT), 0, s, T, v, g, 0, 0);
    HDG_NT_DISPATCH(T.nt, LUPD)
#undef LUPD
    e = (int)hipGetLastError();
    if (e) return e;
  }
  if (seg)
    hipLaunchKernelGGL(lanczos_omega_seg_kernel, dim3((ninst + 63) / 64), dim3(64), 0, s, seg, ninst, steps, factor);
  else
    hipLaunchKernelGGL(lanczos_omega_kernel, dim3(1), dim3(64), 0, s, v.st, steps, factor);
  return (int)hipGetLastError();
}
template int launch_lanczos_omega<Whole>(const DevPattern&, const TraceSystem&, CgVectors&, Whole, int, const Seg*,
                                         int, double, int, hipStream_t);
template int launch_lanczos_omega<Grouped>(const DevPattern&, const TraceSystem&, CgVectors&, Grouped, int,
                                           const Seg*, int, double, int, hipStream_t);

int launch_reduce_segs(const double* part, const Seg* seg, int nseg, double* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_seg_kernel, dim3(nseg), dim3(256), 0, s, part, seg, out);
  return (int)hipGetLastError();
}

int launch_cg_init(const DevPattern& P, const TraceSystem& T, CgVectors& v, hipStream_t s) {
  return launch_cg_init_t(P, T, v, Whole{T.N, v.st}, v.nblocks, s);
}

int launch_cg_spmv(const DevPattern& P, const TraceOp<double>& T, const Krylov<double>& v, int i_fixed,
                   double rtol2, int maxit, int force, hipStream_t s) {
  return launch_cg_spmv_t(P, T, v, Whole{T.N, v.st}, v.nblocks, i_fixed, rtol2, maxit, force, s);
}

int launch_cg_update(const TraceSystem& T, CgVectors& v, int i_fixed, int force, hipStream_t s) {
  return launch_cg_update_t(T, v, Whole{T.N, v.st}, v.nblocks, i_fixed, force, s);
}

int launch_to_float(const double* src, float* dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(to_float_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, src, dst, n);
  return (int)hipGetLastError();
}

int launch_copy(const double* src, double* dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, reinterpret_cast<const double2*>(src),
                     reinterpret_cast<double2*>(dst), n / 2);
  return (int)hipGetLastError();
}

int launch_cg_finish(const TraceSystem& T, CgVectors& v, double* uhat, hipStream_t s) {
  hipLaunchKernelGGL(cg_finish_kernel, dim3(grid_for(T.N, 256, 1 << 30)), dim3(256), 0, s, T, v.x, uhat);
  return (int)hipGetLastError();
}

int launch_true_residual(const DevPattern& P, const TraceSystem& T, const double* uhat, double* out2,
                         CgVectors& v, hipStream_t s) {
  const int e = launch_true_residual_t(P, T, uhat, v.part, Whole{T.N, v.st}, v.nblocks, s);
  if (e) return e;
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, v.part, v.nblocks, out2);
  return (int)hipGetLastError();
}

int launch_recover(int k, const DevMesh& m, const double* Ke, const double* be, const double* uhat,
                   double* qu, hipStream_t s) {
#define REC(KK)                                                                             \
  {                                                                                         \
    const int64_t n = m.nc * Dims<KK>::M;                                                   \
    hipLaunchKernelGGL(recover_kernel<KK>, dim3(grid_for(n, 256, 1 << 30)), dim3(256), 0, s, m, Ke, \
                       be, uhat, qu);                                                       \
  }
  switch (k) {
    case 1: REC(1); break;
    case 2: REC(2); break;
    case 3: REC(3); break;
    case 4: REC(4); break;
    default: return (int)hipErrorInvalidValue;
  }
#undef REC
  return (int)hipGetLastError();
}

int launch_errornorm(const DevMesh& m, const double* tables, const TabLayoutDev& lay, int k,
                     const double* qu, int uex_kind, const double* par3, double* part, double* result,
                     hipStream_t s) {
  const int g = errornorm_blocks(m.nc);
  const int e = launch_errornorm_t(m, tables, lay, k, qu, uex_kind, par3, part, Whole{m.nc, nullptr}, g, s);
  if (e) return e;
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, part, g, result);
  return (int)hipGetLastError();
}

int errornorm_blocks(int64_t nc) { return grid_for(nc, 256, 2048); }

}  // namespace hdg
