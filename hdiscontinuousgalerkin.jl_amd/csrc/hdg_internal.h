// Internal interface between the C ABI (hdg_api.hip) and the device
// kernels (hdg_kernels.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <utility>

namespace hdg {

// Serialises every graph call of the process: stream capture through
// hipGraphInstantiate, hipGraphLaunch and hipGraphExecDestroy.  Contexts
// and batch groups are driven from host threads of their own, and the HIP
// runtime keeps process-wide state for capturing streams, graph objects and
// their kernel-argument memory.  (Serving slots, where 16 threads captured
// and replayed at once, ended in a host segmentation fault inside the
// runtime below these calls even with this lock; the slots no longer
// capture at all.)
std::mutex& graph_mutex();

constexpr int kSlots = 5;  // block columns per face row: self + 2 faces of each cell

// Device-resident mesh, 0-based, int32 (meshes up to 2^31 faces).
struct DevMesh {
  int64_t nn = 0, nc = 0, nf = 0;
  const double* xy = nullptr;  // [2*nn]
  const int* cn = nullptr;     // [3*nc] node ids
  const int* cf = nullptr;     // [3*nc] face ids
  const int* fcell = nullptr;  // [2*nf] cell1, cell2 (-1 none)
};

// Face-level pattern of the trace matrix.  Row block f has nslot[f] (3 or
// 5) block columns bcol[f*5+s], sorted ascending; slots beyond nslot repeat
// f with zero values so that every row has 5 slots.  src[(f*5+s)*2+t] is
// the t-th contributing element block, encoded cell*9 + la*3 + lb (local
// face of f, local face of the column face), or -1.  tslot[f*5+s] is the
// slot of f in the row of bcol[f*5+s] (the transpose map).
//
// nbr[4f .. 4f+3] indexes the symmetric block storage (TraceSystem::blk).
// In every cell the face of local index l owns the block coupling it to
// local face (l + 1) % 3, so a face owns one off-diagonal block per cell
// (k = 0 for fcell[2f], 1 for fcell[2f+1]) and meets the other face of each
// of its cells through that face's block, transposed.  nbr[4f+k] is the
// face its block k couples to; nbr[4f+2+k] = 2 g + k' names the block k' of
// face g that holds K_gf.  A boundary face's missing block is zero:
// nbr[4f+1] = f and nbr[4f+3] = 2f + 1.
struct DevPattern {
  int* bcol = nullptr;
  int* src = nullptr;
  unsigned char* nslot = nullptr;
  unsigned char* self = nullptr;
  unsigned char* tslot = nullptr;
  int* nbr = nullptr;
};

// Symmetric block storage for 4 trace modes per face (k = 3): a face's
// block row is two 16-byte pieces per row of the block, and the product
// reads 3 blocks per face from HBM instead of 5 (the transposed reads of a
// neighbour's block come from the caches).  With 2, 3 or 5 modes the face-ELL
// product is faster and is kept.
template <int NT>
constexpr bool kSymBlocks = NT == 4;
inline bool sym_blocks(int nt) { return nt == 4; }

struct SrcParams {
  int kind = 0, kappa_on = 0;
  double ka = 0, kb = 0;
  double coef[16] = {0};
  const double* f_qp = nullptr;  // [ncells x nqp] column-major
};

struct TabLayoutDev {
  int nu, nt, nqp, nfq;
  int Mref, Dref, Cref, G, Href, Nw, N, Mg, W, total;
};

// Offsets of the SPD condensation tables (hdg_tables.h SpdLayout).
struct SpdLayoutDev {
  int nu, nt, nqp;
  int R, C, Gam, Z, V, P, Om, H, Nw, Mg, total;
};

// Element build + condensation.  Per cell outputs: Ate row-major ntr x ntr,
// bte[ntr], Ke column-major m x ntr, be[m].  Element blocks mode writes the
// A..H blocks instead (column-major per cell) for cells [first, first+count).
// A grouped launch (several instances packed into one mesh) takes the source
// of cell c from srcs[cell_inst[c]] instead of src.
struct CondenseArgs {
  DevMesh mesh;
  const double* tables = nullptr;
  TabLayoutDev lay;
  SrcParams src;
  const SrcParams* srcs = nullptr;
  const int* cell_inst = nullptr;
  double tau = 1.0;
  double* Ate = nullptr;
  double* bte = nullptr;
  double* Ke = nullptr;
  double* be = nullptr;
  int* err = nullptr;  // bit 1: det(J) <= 0; bit 2: singular local matrix (grouped: one int per instance)
  int natural = 0;     // the tables' mass is SPD (RefTables::mass_spd): eliminate in natural order
  // SPD condensation (kappa = 1, SPD reference mass): its tables, or null
  const double* spd_tables = nullptr;
  SpdLayoutDev spd{};
  // element-block mode
  int64_t first = 0, count = 0;
  double *A = nullptr, *B = nullptr, *C = nullptr, *E = nullptr, *F = nullptr, *H = nullptr,
         *bev = nullptr;
};

// The 1-D rule of the trace space for the device Dirichlet projection:
// points x and weights w on [0, 1], the cell rule's first weights wc (the
// weights the reference applies to face points, src/boundary.jl:33) and the
// Legendre basis leg[i * nfq + q] at the points.
constexpr int kMaxFacePoints = 16;
struct DirichletRule {
  int nfq = 0;
  double x[kMaxFacePoints], w[kMaxFacePoints], wc[kMaxFacePoints];
  double leg[5 * kMaxFacePoints];
};

// PCG coefficients kept for the Lanczos estimate of lambda_max(Dinv S)
constexpr int kLanczos = 16;

// PCG scalar state, resident on the device; written by block 0 of the
// kernels that own each field, read by later launches.
struct CgState {
  double rz[2];  // (r, z) by iteration parity
  double pq;
  double rr;     // (r, r) at the last convergence test
  double bb;     // (b~, b~)
  double omega;  // block-Jacobi smoothing weight of the multigrid preconditioner
  int it;        // completed iterations when the solve stopped
  int maxit;
  int done;      // 0 running, 1 converged, 2 iteration cap, 3 breakdown
  // iteration number handed from kernel to kernel, so that no launch takes
  // it as an argument and a captured chunk of iterations can be replayed:
  // each field is written by the lead thread of one kernel and read by the
  // next kernel of the chain, never by the kernel that writes it
  int i_spmv;    // for the next CG product (written by the update / smoother)
  int i_upd;     // for the update (written by the product)
  int i_pre;     // for the smoother (written by the multigrid update)
  // iteration i <= kLanczos of a run: (r, z) at its product and (p, S p) at
  // its update; block-Jacobi PCG's coefficients are those of the Lanczos
  // process of Dinv S
  double lz[kLanczos + 1];
  double lpq[kLanczos + 1];
};

// The host loops queue chunks of CG iterations and keep one chunk queued
// behind the one whose read-back they wait for.  From two read-backs of an
// instance's state (prev, then last) the per-iteration log-rate of (r, r)
// forecasts whether the chunk now running reaches (r, r) <= rtol2 (b, b) at
// its last test, n iterations past last.it: the chunk's length, plus one
// where a convergence test follows each multigrid update (the read-back
// holds the product's test of the iteration before).  When it should, the
// loop queues nothing behind it (a chunk there would only run launches that
// return at once).  A wrong forecast costs one read-back round trip; the
// launches that do work and their results are the same either way.
inline bool chunk_converges(const CgState& prev, const CgState& last, int n, double rtol2) {
  if (last.done) return true;
  const int di = last.it - prev.it;
  if (prev.done || di <= 0 || !(last.rr > 0.0) || !(prev.rr > last.rr) || !(last.bb > 0.0)) return false;
  const double rate = std::log(last.rr / prev.rr) / di;
  return rate * n <= std::log(rtol2 * last.bb / last.rr);
}

struct TraceSystem {
  int nt = 0;
  int64_t nf = 0, N = 0;
  double* vals = nullptr;  // ELL, column-major: vals[(s*nt + b)*N + row]
  // symmetric blocks (sym_blocks(nt), else null), sym_blk_count(nt, nf)
  // doubles: first K_ff's upper triangle slot-major, (a, b) of face f at
  // blk[dinv_slot(nt, a, b) nf + f]; then the two blocks face f owns, one in
  // each of its cells k, own[((2f + k) nt + a) nt + b] with own = blk +
  // nt (nt + 1) / 2 nf (row a of f, column b of the neighbour coupled in
  // that cell; DevPattern::nbr)
  double* blk = nullptr;
  double* rhs = nullptr;   // assembled rhs (N)
  double* Dinv = nullptr;  // block-Jacobi inverse (dinv_count; 0 on Dirichlet dofs)
  const unsigned char* dmask = nullptr;  // 1 = Dirichlet dof
  const double* uD = nullptr;            // Dirichlet values (0 elsewhere)
};

// The block-Jacobi inverse of a face is symmetric (K_ff is, and the Dirichlet
// rows and columns are masked alike).  Where the NT rows of a face fill whole
// lane quads (NT = 1, 2, 4: rows f NT .. f NT + NT - 1 in consecutive lanes of
// one wavefront) only its upper triangle is kept, entry (a, b) of face f at
// Dinv[dinv_slot(nt, a, b) * nf + f] (nt (nt + 1) / 2 per face), and the rows
// of a face trade their operands by DPP; at NT = 3 and 5 the whole block is
// kept by columns, Dinv[b N + r], and the operands go through LDS (the
// triangle's reads there touch more cache lines per load than they save).
template <int NT>
constexpr bool kQuadRows = NT == 1 || NT == 2 || NT == 4;
inline bool quad_rows(int nt) { return nt == 1 || nt == 2 || nt == 4; }
__host__ __device__ constexpr int dinv_slot(int nt, int a, int b) {
  return a <= b ? a * nt - a * (a - 1) / 2 + (b - a) : b * nt - b * (b - 1) / 2 + (a - b);
}
inline size_t sym_blk_count(int nt, int64_t nf) { return (size_t)(nt * (nt + 1) / 2 + 2 * nt * nt) * (size_t)nf; }
inline size_t dinv_count(int nt, int64_t nf) {
  return (size_t)(quad_rows(nt) ? nt * (nt + 1) / 2 : nt * nt) * (size_t)nf;
}
// DPP quad_perm control that gives lane j of every quad the value of lane
// (j & ~(NT - 1)) + B: operand b of the caller's face
template <int NT, int B>
__host__ __device__ constexpr int quad_ctrl() {
  int c = 0;
  for (int j = 0; j < 4; ++j) c |= ((j & ~(NT - 1)) + B) << (2 * j);
  return c;
}
template <int CTRL>
__device__ __forceinline__ double dpp_quad(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int NT, class W, int... B>
__device__ __forceinline__ double dinv_lanes(const W* __restrict__ D, int64_t nf, int64_t r, double x,
                                             std::integer_sequence<int, B...>) {
  const int64_t f = r / NT;
  const int a = (int)(r - f * NT);
  double z = 0.0;
  ((z = fma((double)D[(int64_t)dinv_slot(NT, a, B) * nf + f], dpp_quad<quad_ctrl<NT, B>()>(x), z)), ...);
  return z;
}
// fn(sum_b Dinv_f[a][b] x_(f, b)) on the rows with `in` set, x one value per
// row r = f NT + a (one row per thread, a block's rows consecutive from a
// multiple of NT).  Every thread of the block calls it (the LDS form has
// barriers; rs holds blockDim doubles) and the rows of a face share `in`.
template <int NT, class W, class F>
__device__ __forceinline__ void jacobi_apply(const W* __restrict__ D, int64_t nf, double* rs, double x, int64_t r,
                                             bool in, F&& fn) {
  if constexpr (kQuadRows<NT>) {
    if (in) fn(dinv_lanes<NT>(D, nf, r, x, std::make_integer_sequence<int, NT>{}));
  } else {
    rs[threadIdx.x] = x;
    __syncthreads();
    if (in) {
      const double* xs = rs + (threadIdx.x / NT) * NT;
      const int64_t N = nf * NT;
      double z = 0.0;
#pragma unroll
      for (int b = 0; b < NT; ++b) z = fma((double)D[(int64_t)b * N + r], xs[b], z);
      fn(z);
    }
    __syncthreads();
  }
}

// The trace operator's storage is there for the kernels that read it: the
// symmetric blocks and their references at four modes, the face-ELL values
// otherwise (launchers return hipErrorInvalidValue instead of launching).
template <class TS>
inline bool trace_op_ready(const DevPattern& P, const TS& T) {
  return sym_blocks(T.nt) ? (T.blk != nullptr && P.nbr != nullptr) : T.vals != nullptr;
}

struct CgVectors {
  double *x, *r, *z, *p0, *p1, *q;
  double* part;  // 4 * nblocks partial sums
  CgState* st;
  int nblocks;
};

// ------------------------------------------------------ grouped launches
// Several independent problems (instances) packed into one mesh share every
// launch.  Instance j owns rows (or cells) [lo, hi) and the blocks
// [b0, b0 + nb) of the launch; its blocks visit its rows with stride
// nb * blockDim, and its reductions run over its own blocks' partial sums
// only, so an instance gets the bits it would get alone with nb blocks.
// A selector turns the block index into the instance's segment: Whole is
// one instance spanning the grid, Grouped looks the block up in a table.
struct Seg {
  int64_t lo, hi;
  int b0, nb;
  CgState* st;
};
struct Whole {
  int64_t n;
  CgState* st;
};
struct Grouped {
  const Seg* seg;
  const int* binst;  // block -> instance
};
__device__ __forceinline__ Seg seg_of(const Whole& w) { return Seg{0, w.n, 0, (int)gridDim.x, w.st}; }
__device__ __forceinline__ Seg seg_of(const Grouped& g) { return g.seg[g.binst[blockIdx.x]]; }

// Views for the CG kernels.  TraceOp<W> is the trace operator stored as W:
// FP64 for the CG product, FP64 or float for the multigrid preconditioner
// (HDG_MIXED).  Krylov vectors, the iterate and every reduction are FP64.
template <class W>
struct TraceOp {
  int nt;
  int64_t nf, N;
  const W* vals;   // face-ELL, vals[(s*nt + b)*N + row]
  const W* Dinv;   // TraceSystem::Dinv
  const unsigned char* dmask;
  const W* blk;    // symmetric blocks (TraceSystem::blk) where kSymBlocks<nt>
  // every field named at every construction: an operator without its
  // blocks would send the product to a null pointer
  __host__ __device__ TraceOp(int nt_, int64_t nf_, int64_t N_, const W* vals_, const W* Dinv_,
                              const unsigned char* dmask_, const W* blk_)
      : nt(nt_), nf(nf_), N(N_), vals(vals_), Dinv(Dinv_), dmask(dmask_), blk(blk_) {}
};
// The entries of row a of face f of K over the symmetric blocks
// (kSymBlocks<NT>, TraceSystem::blk): fn(g, b, K[(f, a), (g, b)]) block by
// block, K_ff, the face's own blocks, then the two blocks of the neighbours
// that own the coupling.  K_ff comes from its triangle (a wavefront's 16
// faces read 10 consecutive runs of 128 bytes, once per product:
// nontemporal), the own blocks by rows, the neighbours' blocks by columns,
// their lines shared with the owners' own reads through the caches.
template <int NT, class V, class F>
__device__ __forceinline__ void sym_row_entries(const int* __restrict__ nbr, const V* __restrict__ blk, int64_t nf,
                                                int64_t f, int a, const F& fn) {
  constexpr int B = NT * NT;
  const int4 nb = *reinterpret_cast<const int4*>(nbr + 4 * f);
#pragma unroll
  for (int b = 0; b < NT; ++b) fn(f, b, (double)__builtin_nontemporal_load(blk + (int64_t)dinv_slot(NT, a, b) * nf + f));
  const V* own = blk + (int64_t)(NT * (NT + 1) / 2) * nf;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const V* d = own + (2 * f + k) * B + a * NT;
    const int64_t g = k ? nb.y : nb.x;
#pragma unroll
    for (int b = 0; b < NT; ++b) fn(g, b, (double)d[b]);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int code = k ? nb.w : nb.z;
    const int64_t g = code >> 1;
    const V* d = own + (2 * g + (code & 1)) * B + a;
#pragma unroll
    for (int b = 0; b < NT; ++b) fn(g, b, (double)d[b * NT]);
  }
}

// The Dirichlet lift of row a of face f over the symmetric blocks:
// sum over the prescribed columns (g, c) of K[(f, a), (g, c)] uD[(g, c)], in
// sym_row_entries' order; a block's values are read only for prescribed
// columns, so rows away from the boundary read none.
template <int NT>
__device__ __forceinline__ double sym_row_lift(const int* __restrict__ nbr, const double* __restrict__ blk, int64_t nf,
                                               const unsigned char* __restrict__ dmask,
                                               const double* __restrict__ uD, int64_t f, int a) {
  constexpr int B = NT * NT;
  const int4 nb = *reinterpret_cast<const int4*>(nbr + 4 * f);
  const int64_t gs[5] = {f, nb.x, nb.y, nb.z >> 1, nb.w >> 1};
  const int64_t o = (int64_t)(NT * (NT + 1) / 2) * nf;  // the own blocks
  const int64_t base[5] = {0, o + 2 * f * B + a * NT, o + (2 * f + 1) * B + a * NT,
                           o + (2 * (int64_t)(nb.z >> 1) + (nb.z & 1)) * B + a,
                           o + (2 * (int64_t)(nb.w >> 1) + (nb.w & 1)) * B + a};
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int stride = j < 3 ? 1 : NT;
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int64_t col = gs[j] * NT + c;
      const int64_t at = j == 0 ? (int64_t)dinv_slot(NT, a, c) * nf + f : base[j] + c * stride;
      if (dmask[col]) s = fma(blk[at], uD[col], s);
    }
  }
  return s;
}

// Row a of face f of K x over the symmetric blocks, x given as xv(g, b) =
// entry b of face g.
template <int NT, class V, class X>
__device__ __forceinline__ double sym_row(const int* __restrict__ nbr, const V* __restrict__ blk, int64_t nf, int64_t f,
                                          int a, const X& xv) {
  double acc = 0.0;
  sym_row_entries<NT>(nbr, blk, nf, f, a, [&](int64_t g, int b, double v) { acc = fma(v, xv(g, b), acc); });
  return acc;
}

template <class V>
struct Krylov {
  V *r, *z, *p0, *p1, *q;
  double* x;
  double* part;    // partial-sum slots of nblocks doubles: 0 (p,q), 1-2 (r,z) by parity, 3 (r,r)
  CgState* st;
  int nblocks;
};
template <class W>
struct MgVecs {
  W* e0;           // trace-level correction
  W* res0;         // trace-level residual
  const W* sp;     // S P0 rows (4 per row)
};

// Multigrid hierarchy on the vertex grids of a structured mesh (level 0 is
// the trace space, levels 1.. are vertex grids with 7-point stencils).
struct MgLevel {
  int nx = 0, ny = 0;
  int npts = 7;                  // 7, or kStencil once a ragged coarsening is above this level
  int64_t n = 0;                 // (nx+1)(ny+1) nodes
  double* a = nullptr;           // stencils, a[s*n + v], s < kStencil
  double* dinv = nullptr;        // 1 / a[0], 0 on fixed nodes
  unsigned char* mask = nullptr; // fixed (Dirichlet) nodes
  double *r = nullptr, *e = nullptr, *t = nullptr;
};
constexpr int kMaxLevels = 16;
// points of the vertex-level stencils (slots listed in hdg_mg.hip)
constexpr int kStencil = 9;
// node budget of the levels solved inside one workgroup: their stencils,
// diagonal inverses and three vectors live in LDS (13 doubles per node,
// 156 KB of the CU's 160 KB), so that the 33 x 33 grid of a 1024- or
// 512-cell side runs there instead of in 2-4 launches of its own (1024^2
// P3 step -0.45 ms, 512^2 P2 -0.26 ms against 512 nodes, 53 KB); 1024
// threads, one workgroup per CU
constexpr int kTailNodes = 1536;
constexpr int kTailThreads = 1024;
// coarsest grids up to this many nodes are solved exactly (inverse at set-up)
constexpr int kCoarseDirect = 16;
struct MgData {
  int nlev = 0;                  // levels 1..nlev-1 are vertex grids
  int nx = 0, ny = 0;
  int64_t nv = 0;
  MgLevel lev[kMaxLevels];
  int* fv = nullptr;             // [2*nf] lower, higher node id of each face
  int* vf_ptr = nullptr;         // vertex -> incident faces (CSR)
  int* vf_idx = nullptr;
  int* vf2 = nullptr;            // [2*nf] the same entries as 2 f + (1 where the vertex is f's higher end)
  double* fc = nullptr;          // [2*nf] per face: P0^T res0 restricted to the face, at its lower and
                                 // higher end (even nt: written by the level-0 residual)
  double* cellA = nullptr;       // [9*nc] per-cell P^T (-Ate) P
  double* e0 = nullptr;          // trace-level correction
  double* res0 = nullptr;        // trace-level residual
  int* fq4 = nullptr;            // [4*nf] the 4 vertices of the two cells of a face:
                                 // lower, higher end point, opposite vertex of cell 1, of cell 2
                                 // (= that of cell 1 on boundary faces)
  double* sp = nullptr;          // [4*N] S P0 restricted to the rows: sp[j*N + r] couples
                                 // row r with vertex fq4[4*face(r) + j]
  int tail = kMaxLevels;         // levels tail..nlev-1 run in one workgroup (LDS)
  double* cinv = nullptr;        // [kCoarseDirect^2] inverse of the coarsest operator (row-major)
  double* tailpack = nullptr;    // stencils (7 n) and dinv (n) of levels tail.. back to back,
                                 // the LDS image the tail kernel loads
  // vertex-level Jacobi weight: 0.6-0.7 take one CG iteration fewer than
  // 0.8 at 512^2 P2 (37 vs 38), 0.9 four more; 1024^2 P3 stays at 29
  double omega_v = 0.7;
  double omega_factor = 1.6;     // trace-level weight = factor / lambda_max(Dinv S) (Lanczos estimate)
  int coarse_sweeps = 20;
  int vsweeps = 2;               // pre- and post-smoothing sweeps on the vertex levels (1 or 2)
  // block-Jacobi PCG (Lanczos) steps behind omega: 6 estimate lambda_max
  // a little lower than 8, which lands omega * lambda_max between 1.6 and
  // 1.7, still short of the 1.8 where iterations grow
  int lanczos_steps = 6;
};
// per-order choices of the V-cycle, shared by single problems and batches:
// V(1,1) on the vertex levels for k <= 3, V(2,2) at k = 4 (fewest CG
// iterations x cost per iteration measured on 256..1536 meshes: at k = 3
// V(1,1) takes one iteration more and two launches fewer per level, 256^2
// 6.40 -> 5.90 ms, 512^2 16.8 -> 16.0, 1024^2 58.6 -> 57.9, 1536^2 even;
// at k = 4 1024^2 loses 147 -> 150 ms)
inline void mg_order_choices(int k, MgData& M) { M.vsweeps = k <= 3 ? 1 : 2; }

// A fine level, the next coarser one and the CG state of one instance: the
// unit of a vertex-level kernel.  VOne launches one instance over the whole
// grid; VGroup maps each block to (instance, block of that instance).
struct VRec {
  MgLevel F, C;
  const CgState* st;
};
struct BlockMap {
  int inst, lb;
};
struct VOne {
  VRec r;
};
struct VGroup {
  const VRec* rec;
  const BlockMap* map;
};
__device__ __forceinline__ const VRec& vrec_of(const VOne& o, int& lb) {
  lb = blockIdx.x;
  return o.r;
}
__device__ __forceinline__ const VRec& vrec_of(const VGroup& g, int& lb) {
  const BlockMap m = g.map[blockIdx.x];
  lb = m.lb;
  return g.rec[m.inst];
}

// launchers (return hipError_t as int)
int launch_pattern(const DevMesh& m, DevPattern& P, hipStream_t s);
// Schedule of the condensation kernel: lanes per element (16, 32 or 64;
// 0 = the narrowest that holds the rows) and the grid cap (blocks of 256
// threads that stride over the cells; 0 = 8192).  Results do not depend on
// it; the autotuner picks it by measurement.
struct CondenseSched {
  int seg = 0;
  int grid = 0;
  // the Schur-complement kernel: waves per SIMD its register budget is cut
  // for (2 or 3; 0 = 2, the budget without spills) and its grid cap
  // (0 = 16384)
  int spd_wpe = 0;
  int spd_grid = 0;
};
int launch_condense(int k, const CondenseArgs& a, bool blocks_mode, hipStream_t s,
                    const CondenseSched& sc = CondenseSched{});
int condense_native_seg(int k);
// Condensation through the Schur complement on the u block for kappa = 1
// and an SPD reference mass (a.spd_tables set, a.src without kappa): same
// outputs as launch_condense up to rounding.  Its autotuner times the
// (spd_wpe, spd_grid) schedules and keeps the fastest in *best.
int launch_condense_spd(int k, const CondenseArgs& a, const CondenseSched& sc, hipStream_t s);
// The Schur-complement kernel writes only the u rows of K_e and b_e, compact
// at the start of each cell's blocks (U column-major nu x ntr, then Ub).
// launch_recover_spd recovers [q; u] from them (and the SPD tables `tab`);
// launch_expand_ke_spd rewrites K_e and b_e in place in their full layout
// (q rows formed), for the export of K_e.
int launch_recover_spd(int k, const DevMesh& m, const double* tab, const double* Ke, const double* be,
                       const double* uhat, double* qu, hipStream_t s);
int launch_expand_ke_spd(int k, const DevMesh& m, const double* tab, double* Ke, double* be, hipStream_t s);
int tune_condense_spd(int k, const CondenseArgs& a, hipStream_t s, CondenseSched* best, double* best_us,
                      double* default_us);
// Autotuner: times every (seg, grid) schedule of the condensation kernel on
// the cells of a (3 warm launches each, HIP events on stream s) and returns
// the fastest; the default schedule's time in default_us.  Overwrites the
// outputs of a.
int tune_condense(int k, const CondenseArgs& a, hipStream_t s, CondenseSched* best, double* best_us,
                  double* default_us);
int launch_fill(const DevMesh& m, const DevPattern& P, const double* Ate, const double* bte, TraceSystem& T,
                hipStream_t s);
int launch_jacobi(const DevPattern& P, TraceSystem& T, hipStream_t s);
// the face-ELL values (kSlots nt N doubles, TraceSystem::vals layout) of the
// symmetric blocks, for the exports
int launch_expand_ell(const DevPattern& P, const TraceSystem& T, double* vals, hipStream_t s);
int launch_cg_init(const DevPattern& P, const TraceSystem& T, CgVectors& v, hipStream_t s);
// i_fixed = 0 takes the iteration number from the device state; force = 1
// (timing only, with i_fixed > 0) runs the kernel regardless of the
// convergence state and leaves x, r untouched (the CG state is not usable
// afterwards)
int launch_cg_spmv(const DevPattern& P, const TraceOp<double>& T, const Krylov<double>& v, int i_fixed,
                   double rtol2, int maxit, int force, hipStream_t s);
int launch_cg_update(const TraceSystem& T, CgVectors& v, int i_fixed, int force, hipStream_t s);
int launch_cg_finish(const TraceSystem& T, CgVectors& v, double* uhat, hipStream_t s);
// copies n doubles (n even) with a plain streaming kernel
int launch_copy(const double* src, double* dst, int64_t n, hipStream_t s);
int launch_true_residual(const DevPattern& P, const TraceSystem& T, const double* uhat, double* out2,
                         CgVectors& v, hipStream_t s);
int launch_recover(int k, const DevMesh& m, const double* Ke, const double* be, const double* uhat,
                   double* qu, hipStream_t s);
int launch_errornorm(const DevMesh& m, const double* tables, const TabLayoutDev& lay, int k,
                     const double* qu, int uex_kind, const double* par3, double* part,
                     double* result, hipStream_t s);
int cg_blocks(int64_t N, int nt);
int errornorm_blocks(int64_t nc);
// Dirichlet data on every face with one cell (src/boundary.jl:11-42):
// g_kind 0 = par[0] + par[1] x + par[2] y, 1 = sin(pi x) sin(pi y); flags
// bit 1 = the reference's weights and accumulator.  Other faces are freed.
int launch_dirichlet_project(const DevMesh& m, int nt, const DirichletRule& R, int g_kind, const double* par,
                             int flags, unsigned char* dmask, double* uD, hipStream_t s);
// apply! (src/boundary.jl:121-175) into vals_out / rhs_out (shaped like
// T.vals / T.rhs); part holds 1024 partial sums, md receives meandiag(K).
int launch_apply(const DevPattern& P, const TraceSystem& T, double* part, double* md, double* vals_out,
                 double* rhs_out, hipStream_t s);

// Selector-generic forms of the CG and norm launchers (G = Whole or Grouped,
// grid = blocks of the launch; the Whole forms above call these).
template <class G>
int launch_cg_init_t(const DevPattern& P, const TraceSystem& T, CgVectors& v, G g, int grid, hipStream_t s);
// the next product's convergence test right after the multigrid update (one
// instance, or one block per instance of a group)
int launch_cg_check(const Krylov<double>& v, int64_t N, double rtol2, int maxit, hipStream_t s);
int launch_cg_check_group(const Krylov<double>& v, Grouped g, int ninst, double rtol2, int maxit, hipStream_t s);
template <class G>
int launch_cg_spmv_t(const DevPattern& P, const TraceOp<double>& T, const Krylov<double>& v, G g, int grid,
                     int i_fixed, double rtol2, int maxit, int force, hipStream_t s);
template <class G>
int launch_cg_update_t(const TraceSystem& T, CgVectors& v, G g, int grid, int i_fixed, int force, hipStream_t s);
// per-block partials of ||K uhat - rhs||^2
template <class G>
int launch_true_residual_t(const DevPattern& P, const TraceSystem& T, const double* uhat, double* part, G g,
                           int grid, hipStream_t s);
// per-block partials of the squared L2 error
template <class G>
int launch_errornorm_t(const DevMesh& m, const double* tables, const TabLayoutDev& lay, int k, const double* qu,
                       int uex_kind, const double* par3, double* part, G g, int grid, hipStream_t s);
// out[j] = sum of part over the blocks of segment j
int launch_reduce_segs(const double* part, const Seg* seg, int nseg, double* out, hipStream_t s);

// vertex hierarchy and S P0 (FP64, and a float copy of S P0 when sp32 is set)
int mg_setup_numeric(const DevMesh& mesh, MgData& M, const MgData* dM, TraceSystem& T, const DevPattern& P,
                     const double* Ate, float* sp32, hipStream_t s);
// the two halves of mg_setup_numeric: the parts that run over the whole
// (possibly packed) mesh -- vertex masks, per-cell coarse blocks, S P0 --
// and the vertex levels of one structured instance (M describes it, with
// its arrays at its own offsets; dM is its device copy)
int mg_setup_mesh(const DevMesh& mesh, const MgData& M, const TraceSystem& T, const DevPattern& P,
                  const double* Ate, float* sp32, hipStream_t s);
// mg_setup_mesh in two parts: what needs only the condensed blocks (vertex
// masks, per-cell coarse blocks) and what reads the assembled trace matrix
// (S P0), so that the first can run beside the assembly
int mg_setup_cells(const DevMesh& mesh, const MgData& M, const TraceSystem& T, const double* Ate, hipStream_t s);
int mg_setup_sp(const MgData& M, const TraceSystem& T, const DevPattern& P, float* sp32, hipStream_t s);
int mg_setup_levels(const MgData& M, const MgData* dM, hipStream_t s);
// omega = factor / lambda_max(Dinv S), lambda_max the largest Ritz value of
// `steps` block-Jacobi PCG iterations from a random right-hand side (the
// Lanczos tridiagonal matrix of Dinv S built from their coefficients); uses
// the CG vectors as scratch.  Whole: one problem (st); Grouped: one estimate
// per instance (seg[0 .. ninst)).
template <class G>
int launch_lanczos_omega(const DevPattern& P, const TraceSystem& T, CgVectors& v, G g, int grid, const Seg* seg,
                         int ninst, double factor, int steps, hipStream_t s);
// r = b~, x = p = 0, e0 = omega Dinv r (preconditioner storage W)
template <class W>
int launch_mg_init(const DevPattern& P, const TraceSystem& T64, const TraceOp<W>& T, const Krylov<double>& v,
                   const MgVecs<W>& mv, hipStream_t s);
template <class W>
int launch_mg_update(const TraceOp<W>& T, const Krylov<double>& v, const MgVecs<W>& mv, hipStream_t s);
// dM: a device copy of M (read by the one-workgroup tail kernel)
template <class W>
int launch_mg_precond(const DevPattern& P, const TraceOp<W>& T, const Krylov<double>& v, const MgData& M,
                      const MgData* dM, const MgVecs<W>& mv, hipStream_t s);

// Launch tables of the multigrid preconditioner of a group of instances:
// per vertex depth d (1 <= d < ndepth), the instances whose level d lies
// above their one-workgroup tail, with block maps for the fine-node kernels
// (node, gnode blocks) and the restriction (res, gres blocks); then one
// tail workgroup per instance.
struct MgGroup {
  int ninst = 0;
  const MgData* dM = nullptr;  // [ninst] per-instance hierarchies (device copies)
  CgState* st = nullptr;       // [ninst] CG states
  int ndepth = 1;
  VGroup node[kMaxLevels];
  int gnode[kMaxLevels];
  VGroup res[kMaxLevels];
  int gres[kMaxLevels];
  int vsweeps = 2;
  double omega_v = 0.7;
};
// Set-up tables of a group's vertex hierarchies: level-1 assembly blocks
// (instance, block) and, per depth d, the instances with a level d+1, with
// block maps over the coarse nodes (cmask) and 8 lanes per coarse node (RAP).
struct MgSetupGroup {
  int ninst = 0;
  const MgData* dM = nullptr;
  const BlockMap* l1map = nullptr;
  int gl1 = 0;
  int ndepth = 1;
  VGroup cmask[kMaxLevels];
  int gcmask[kMaxLevels];
  VGroup rap[kMaxLevels];
  int grap[kMaxLevels];
};
int mg_setup_levels_group(const MgSetupGroup& S, hipStream_t s);

// grouped forms: rows of each instance from the Grouped selector (grid
// blocks); Mu describes the packed mesh (its level 1 spans every
// instance's vertices)
template <class W>
int launch_mg_init_group(const DevPattern& P, const TraceSystem& T64, const TraceOp<W>& T, const Krylov<double>& v,
                         const MgVecs<W>& mv, Grouped g, int grid, hipStream_t s);
template <class W>
int launch_mg_update_group(const TraceOp<W>& T, const Krylov<double>& v, const MgVecs<W>& mv, Grouped g, int grid,
                           hipStream_t s);
template <class W>
int launch_mg_precond_group(const DevPattern& P, const TraceOp<W>& T, const Krylov<double>& v, const MgData& Mu,
                            const MgGroup& G, const MgVecs<W>& mv, Grouped g, int grid, hipStream_t s);
// float copy of n doubles
int launch_to_float(const double* src, float* dst, int64_t n, hipStream_t s);

// hdg_refine_probe: the trace matrix rounded to float (vr = (double)(float)
// vals), res = rhs - K u on the free rows with ||res||^2 in *out2 (part:
// refine_partials() doubles), u += d.
int refine_partials();
int launch_round_float(const double* src, double* dst, int64_t n, hipStream_t s);
int launch_refine_residual(const DevPattern& P, const TraceSystem& T, const double* u, double* res, double* part,
                           double* out2, hipStream_t s);
int launch_add(double* u, const double* d, int64_t n, hipStream_t s);

// rectangle_mesh generated on the device into (possibly packed) arrays:
// nodes, cells and faces land at node0 / cell0 / face0; ids written are
// global (shifted); cell_inst (optional) receives inst for every cell;
// vf_ptr / vf_idx / fq4 optional; err bit 4: a one-cell face off the box.
struct RectMeshArgs {
  int64_t nx = 0, ny = 0, nf = 0;
  double x0 = 0, y0 = 0, x1 = 1, y1 = 1, jitter = 0;
  uint64_t seed = 0;
  int64_t node0 = 0, cell0 = 0, face0 = 0;
  int inst = 0;
  bool last = true;  // writes the closing entry of vf_ptr
  double* xy = nullptr;
  int *cn = nullptr, *cf = nullptr, *fcell = nullptr, *fv = nullptr;
  int *vf_ptr = nullptr, *vf_idx = nullptr, *fq4 = nullptr, *cell_inst = nullptr;
  int* err = nullptr;
};
int launch_rect_mesh(const RectMeshArgs& a, hipStream_t s);
// dmask / uD (uD may be null) of the faces [face0, face0 + nf): u = value on faces with one cell
int launch_boundary_dirichlet(const int* fcell, int64_t face0, int64_t nf, int nt, unsigned char* dmask, double* uD,
                              double value, hipStream_t s);

}  // namespace hdg
