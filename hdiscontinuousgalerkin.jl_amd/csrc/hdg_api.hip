// C ABI of libhdg (include/hdg.h): context, device memory, orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hdg.h"
#include "hdg_arena.h"
#include "hdg_host.h"
#include "hdg_internal.h"
#include "hdg_tables.h"

using namespace hdg;

std::mutex& hdg::graph_mutex() {
  static std::mutex m;
  return m;
}

struct hdg_ctx {
  int device = 0;
  size_t budget = 0, used = 0;
  std::unordered_map<void*, size_t> allocs;
  hdg_arena* arena = nullptr;  // serving loop's arena, or null
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // second stream for the multigrid set-up beside the Lanczos steps
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_fill = nullptr, ev_join = nullptr;
  // hdg_run queued the multigrid set-up on aux beside the assembly; the
  // solve joins it instead of queueing it
  bool mg_forked = false;
  std::string err;

  // mesh
  bool has_mesh = false;
  DevMesh mesh;
  std::vector<int> h_cf;  // face ids per cell, 0-based
  int64_t nblock_cols = 0;  // stored nt x nt blocks of K
  DevPattern pat;
  bool structured = false;  // cells follow rectangle_mesh's (nx, ny) numbering
  MgData mg;                // vertex hierarchy (structured meshes)
  MgData* d_mg = nullptr;   // its device copy (read by the one-workgroup tail kernel)

  // discretisation
  bool has_setup = false;
  hdg_options opt{};
  int k = 0, nu = 0, nq = 0, m = 0, nt = 0, ntr = 0, nqp = 0, nfq = 0;
  RefTables tab;
  double* d_tab = nullptr;
  TabLayoutDev lay{};
  // SPD condensation tables (kappa = 1 and an SPD reference mass), or null
  double* d_spd = nullptr;
  SpdLayoutDev spd{};
  SrcParams src;
  double* d_fqp = nullptr;

  // Dirichlet
  unsigned char* d_dmask = nullptr;
  double* d_uD = nullptr;
  bool jacobi_dirty = true;

  // per-cell data
  double *d_Ate = nullptr, *d_bte = nullptr, *d_Ke = nullptr, *d_be = nullptr, *d_qu = nullptr;
  // K_e and b_e hold the compact u rows of the Schur form (recovery forms q
  // itself), or (ke_full) their full layout: from the elimination form, or
  // rewritten by hdg_get_local_ops
  bool ke_compact = false, ke_full = true;
  int* d_err = nullptr;

  // trace system and solver
  TraceSystem T;
  CgVectors cg{};
  double* d_uhat = nullptr;
  CgState* h_st = nullptr;  // pinned, 2 slots
  double* d_scal = nullptr; // [0] errornorm, [1] true residual^2
  double* d_part2 = nullptr;

  // mixed precision (HDG_MIXED): the multigrid preconditioner's trace-level
  // data in float: trace matrix, Jacobi blocks, S P0, e0, res0
  float *f_vals = nullptr, *f_blk = nullptr, *f_Dinv = nullptr, *f_sp = nullptr;
  float *f_e0 = nullptr, *f_res0 = nullptr;

  // captured chunks of CG iterations, one per read-back slot (valid until
  // the next hdg_setup / hdg_set_mesh); solves since the set-up
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int solves = 0;
  // condensation schedule (autotuned or default) and its measurements
  CondenseSched csched;
  bool tuned = false;
  double tuned_us = 0.0, default_us = 0.0;

  bool assembled = false, solved = false, recovered = false;
};

namespace {

int fail(hdg_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                                    \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(ctx, HDG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

#define LAUNCHCHK(ctx, call)                                                                 \
  do {                                                                                       \
    int e_ = (call);                                                                         \
    if (e_ != 0)                                                                             \
      return fail(ctx, HDG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString((hipError_t)e_)); \
  } while (0)

template <class T>
int dalloc(hdg_ctx* c, T** p, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (c->budget && c->used + bytes > c->budget)
    return fail(c, HDG_ERR_HBM_BUDGET, "device allocation of %zu bytes exceeds the budget (%zu of %zu used)",
                bytes, c->used, c->budget);
  // the serving loop's arena first; else stream-ordered allocation from the
  // device's pool: unlike hipMalloc / hipFree it does not synchronise the
  // device, so contexts working on other streams keep running
  void* q = c->arena ? c->arena->take(bytes) : nullptr;
  if (!q) {
    hipError_t e = hipMallocAsync(&q, bytes, c->stream);
    if (e != hipSuccess) return fail(c, HDG_ERR_HIP, "hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  c->allocs[q] = bytes;
  c->used += bytes;
  *p = static_cast<T*>(q);
  return HDG_OK;
}

template <class T>
void dfree(hdg_ctx* c, T*& p) {
  if (!p) return;
  auto it = c->allocs.find((void*)p);
  size_t bytes = 0;
  if (it != c->allocs.end()) {
    bytes = it->second;
    c->used -= it->second;
    c->allocs.erase(it);
  }
  if (c->arena && c->arena->owns((const void*)p)) {
    // work queued on the stream may still read the buffer; a pointer this
    // context did not allocate is left alone
    (void)hipStreamSynchronize(c->stream);
    if (bytes) c->arena->give((void*)p, bytes);
  } else {
    (void)hipFreeAsync((void*)p, c->stream);
  }
  p = nullptr;
}

#define TRY(x)             \
  do {                     \
    int r_ = (x);          \
    if (r_ != HDG_OK) return r_; \
  } while (0)

void drop_graphs(hdg_ctx* c) {
  // a replay may still be queued on the stream
  if ((c->gexec[0] || c->gexec[1]) && c->stream) (void)hipStreamSynchronize(c->stream);
  std::lock_guard<std::mutex> lk(graph_mutex());
  for (auto& g : c->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  c->solves = 0;
}

void free_problem(hdg_ctx* c) {
  drop_graphs(c);
  dfree(c, c->d_tab);
  dfree(c, c->d_spd);
  dfree(c, c->d_fqp);
  dfree(c, c->d_dmask);
  dfree(c, c->d_uD);
  dfree(c, c->d_Ate);
  dfree(c, c->d_bte);
  dfree(c, c->d_Ke);
  dfree(c, c->d_be);
  dfree(c, c->d_qu);
  dfree(c, c->T.vals);
  dfree(c, c->T.blk);
  dfree(c, c->T.rhs);
  dfree(c, c->T.Dinv);
  dfree(c, c->cg.x);
  dfree(c, c->cg.r);
  dfree(c, c->cg.z);
  dfree(c, c->cg.p0);
  dfree(c, c->cg.p1);
  dfree(c, c->cg.q);
  dfree(c, c->cg.part);
  dfree(c, c->d_uhat);
  dfree(c, c->d_part2);
  dfree(c, c->mg.cellA);
  dfree(c, c->mg.e0);
  dfree(c, c->mg.res0);
  dfree(c, c->mg.fc);
  dfree(c, c->mg.sp);
  dfree(c, c->d_mg);
  dfree(c, c->f_vals);
  dfree(c, c->f_blk);
  dfree(c, c->f_Dinv);
  dfree(c, c->f_sp);
  dfree(c, c->f_e0);
  dfree(c, c->f_res0);
  c->has_setup = c->assembled = c->solved = c->recovered = false;
  c->csched = CondenseSched{};
  c->tuned = false;
}

void free_mesh(hdg_ctx* c) {
  free_problem(c);
  dfree(c, const_cast<double*&>(c->mesh.xy));
  dfree(c, const_cast<int*&>(c->mesh.cn));
  dfree(c, const_cast<int*&>(c->mesh.cf));
  dfree(c, const_cast<int*&>(c->mesh.fcell));
  dfree(c, c->pat.bcol);
  dfree(c, c->pat.src);
  dfree(c, c->pat.nslot);
  dfree(c, c->pat.self);
  dfree(c, c->pat.tslot);
  dfree(c, c->pat.nbr);
  dfree(c, c->mg.fv);
  dfree(c, c->mg.vf_ptr);
  dfree(c, c->mg.vf_idx);
  dfree(c, c->mg.vf2);
  dfree(c, c->mg.fq4);
  dfree(c, c->mg.tailpack);
  dfree(c, c->mg.cinv);
  for (int l = 0; l < kMaxLevels; ++l) {
    MgLevel& L = c->mg.lev[l];
    dfree(c, L.a);
    dfree(c, L.dinv);
    dfree(c, L.mask);
    dfree(c, L.r);
    dfree(c, L.e);
    dfree(c, L.t);
  }
  c->mg = MgData{};
  c->structured = false;
  c->has_mesh = false;
  c->mesh = DevMesh{};
}

// Cells numbered as rectangle_mesh(TriangleCell, (nx, ny)) numbers them
// (generate_mesh.jl:130-138): lower triangle (n(i,j), n(i+1,j), n(i,j+1)),
// then upper (n(i+1,j), n(i+1,j+1), n(i,j+1)), n(i,j) = i + (j-1)(nx+1).
bool detect_structured(int64_t nn, int64_t nc, const int64_t* cells, int* nx_out, int* ny_out) {
  const int64_t nx = cells[2] - 2;
  if (nx < 1 || nc % (2 * nx) != 0) return false;
  const int64_t ny = nc / (2 * nx);
  if ((nx + 1) * (ny + 1) != nn) return false;
  for (int64_t j = 1; j <= ny; ++j)
    for (int64_t i = 1; i <= nx; ++i) {
      auto n = [&](int64_t a, int64_t b) { return a + (b - 1) * (nx + 1); };
      const int64_t* lo = cells + 6 * (2 * ((j - 1) * nx + (i - 1)));
      const int64_t* up = lo + 6;
      if (lo[0] != n(i, j) || lo[1] != n(i + 1, j) || lo[2] != n(i, j + 1)) return false;
      if (up[0] != n(i + 1, j) || up[1] != n(i + 1, j + 1) || up[2] != n(i, j + 1)) return false;
    }
  *nx_out = (int)nx;
  *ny_out = (int)ny;
  return true;
}

}  // namespace

namespace hdg {

bool face_vertex_maps(int64_t nn, int64_t nf, const int64_t* faces, const int64_t* cells, std::vector<int>& fv,
                      std::vector<int>& vf_ptr, std::vector<int>& vf_idx, std::vector<int>* fq4) {
  fv.assign(2 * nf, 0);
  vf_ptr.assign(nn + 1, 0);
  for (int64_t f = 0; f < nf; ++f) {
    const int a = (int)faces[f] - 1, b = (int)faces[f + nf] - 1;
    fv[2 * f] = std::min(a, b);
    fv[2 * f + 1] = std::max(a, b);
    vf_ptr[a + 1]++;
    vf_ptr[b + 1]++;
  }
  for (int64_t v = 0; v < nn; ++v) vf_ptr[v + 1] += vf_ptr[v];
  vf_idx.assign(vf_ptr[nn], 0);
  std::vector<int> pos(vf_ptr.begin(), vf_ptr.end() - 1);
  for (int64_t f = 0; f < nf; ++f) {
    vf_idx[pos[fv[2 * f]]++] = (int)f;
    vf_idx[pos[fv[2 * f + 1]]++] = (int)f;
  }
  if (!fq4) return true;
  fq4->assign(4 * nf, 0);
  for (int64_t f = 0; f < nf; ++f) {
    const int a = fv[2 * f], b = fv[2 * f + 1];
    int opp[2];
    for (int side = 0; side < 2; ++side) {
      const int64_t cc = faces[f + (2 + side) * nf] - 1;
      opp[side] = -1;
      if (cc < 0) continue;
      for (int w = 0; w < 3; ++w) {
        const int nw = (int)cells[6 * cc + w] - 1;
        if (nw != a && nw != b) opp[side] = nw;
      }
    }
    if (opp[0] < 0) return false;
    (*fq4)[4 * f] = a;
    (*fq4)[4 * f + 1] = b;
    (*fq4)[4 * f + 2] = opp[0];
    (*fq4)[4 * f + 3] = opp[1] < 0 ? opp[0] : opp[1];
  }
  return true;
}

void mg_plan_levels(int nx, int ny, MgData& M) {
  M.nx = nx;
  M.ny = ny;
  int l = 1, lx = nx, ly = ny;
  for (;;) {
    MgLevel& L = M.lev[l];
    L.nx = lx;
    L.ny = ly;
    L.n = (int64_t)(lx + 1) * (ly + 1);
    // RAP across a one-cell-wide coarse column or row adds the (+1,+1) pair
    L.npts = l == 1 ? 7 : ((M.lev[l - 1].npts > 7 || M.lev[l - 1].nx % 2 || M.lev[l - 1].ny % 2) ? kStencil : 7);
    ++l;
    if (std::max(lx, ly) <= 2 || l >= kMaxLevels) break;
    lx = (lx + 1) / 2;  // coarsening with a ragged end on odd sides
    ly = (ly + 1) / 2;
  }
  M.nlev = l;
  // smallest suffix of levels whose nodes fit the one-workgroup budget
  M.tail = M.nlev;
  int64_t acc = 0;
  for (int t = M.nlev - 1; t >= 1; --t) {
    acc += M.lev[t].n;
    if (acc > kTailNodes) break;
    M.tail = t;
  }
}

}  // namespace hdg

namespace {

// face end points and the vertex -> face lists (any mesh), and the vertex
// grid hierarchy when the mesh is structured
int build_vertex_structures(hdg_ctx* c, int64_t nn, int64_t nf, const int64_t* faces, const int64_t* cells,
                            int64_t nc) {
  std::vector<int> fv, cnt, idx, fq4;
  int nx = 0, ny = 0;
  c->structured = detect_structured(nn, nc, cells, &nx, &ny);
  if (!face_vertex_maps(nn, nf, faces, cells, fv, cnt, idx, c->structured ? &fq4 : nullptr))
    return fail(c, HDG_ERR_ARG, "a face has no first cell");
  MgData& M = c->mg;
  TRY(dalloc(c, &M.fv, 2 * nf));
  TRY(dalloc(c, &M.vf_ptr, nn + 1));
  TRY(dalloc(c, &M.vf_idx, idx.size()));
  TRY(dalloc(c, &M.vf2, idx.size()));
  HIPCHK(c, hipMemcpyAsync(M.fv, fv.data(), sizeof(int) * fv.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(M.vf_ptr, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(M.vf_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  M.nv = nn;
  if (!c->structured) return HDG_OK;
  TRY(dalloc(c, &M.fq4, 4 * nf));
  HIPCHK(c, hipMemcpyAsync(M.fq4, fq4.data(), sizeof(int) * fq4.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  mg_plan_levels(nx, ny, M);
  for (int l = 1; l < M.nlev; ++l) {
    MgLevel& L = M.lev[l];
    TRY(dalloc(c, &L.a, kStencil * L.n));
    TRY(dalloc(c, &L.dinv, L.n));
    TRY(dalloc(c, &L.mask, L.n));
    TRY(dalloc(c, &L.r, L.n));
    TRY(dalloc(c, &L.e, L.n));
    TRY(dalloc(c, &L.t, L.n));
  }
  TRY(dalloc(c, &M.cinv, (size_t)kCoarseDirect * kCoarseDirect));
  if (M.tail < M.nlev) {
    int64_t tn = 0;
    for (int t = M.tail; t < M.nlev; ++t) tn += M.lev[t].n;
    TRY(dalloc(c, &M.tailpack, (kStencil + 1) * tn));
  }
  return HDG_OK;
}

int select_device(hdg_ctx* c) {
  HIPCHK(c, hipSetDevice(c->device));
  return HDG_OK;
}

int err_bits(hdg_ctx* c, int flag);

int check_err_flag(hdg_ctx* c) {
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return err_bits(c, flag);
}

// the condensation's error bits as a status
int err_bits(hdg_ctx* c, int flag) {
  if (flag & 1) return fail(c, HDG_ERR_NEG_JACOBIAN, "det(J) is not positive for at least one cell");
  if (flag & 2)
    return fail(c, HDG_ERR_LOCAL_SINGULAR,
                "local (q,u) matrix is singular for at least one cell (k=%d, quad_degree=%d); "
                "raise quad_degree", c->k, c->opt.quad_degree);
  return HDG_OK;
}

CondenseArgs condense_args(const hdg_ctx* c) {
  CondenseArgs a;
  a.mesh = c->mesh;
  a.tables = c->d_tab;
  a.lay = c->lay;
  a.src = c->src;
  a.src.f_qp = c->d_fqp;
  a.tau = c->opt.tau;
  a.Ate = c->d_Ate;
  a.bte = c->d_bte;
  a.Ke = c->d_Ke;
  a.be = c->d_be;
  a.err = c->d_err;
  a.natural = c->tab.mass_spd() ? 1 : 0;
  // the Schur-complement form where A = detJ diag(Mref, Mref): kappa = 1,
  // an SPD reference mass, unless the schedule asks for the elimination
  if (c->d_spd && !c->src.kappa_on && !(c->opt.schedule & HDG_SCHEDULE_LU)) {
    a.spd_tables = c->d_spd;
    a.spd = c->spd;
  }
  return a;
}

// condensation of every cell by the form condense_args selects
int launch_condense_cells(hdg_ctx* c, const CondenseArgs& a) {
  c->ke_compact = a.spd_tables != nullptr;
  c->ke_full = !c->ke_compact;
  if (a.spd_tables) return launch_condense_spd(c->k, a, c->csched, c->stream);
  return launch_condense(c->k, a, false, c->stream, c->csched);
}

// [q; u] = K_e uhat_c + b_e on every cell
int launch_recover_cells(const hdg_ctx* c) {
  if (c->ke_compact && !c->ke_full)
    return launch_recover_spd(c->k, c->mesh, c->d_spd, c->d_Ke, c->d_be, c->d_uhat, c->d_qu, c->stream);
  return launch_recover(c->k, c->mesh, c->d_Ke, c->d_be, c->d_uhat, c->d_qu, c->stream);
}

int autotune_impl(hdg_ctx* c) {
  const CondenseArgs a = condense_args(c);
  if (a.spd_tables)
    LAUNCHCHK(c, tune_condense_spd(c->k, a, c->stream, &c->csched, &c->tuned_us, &c->default_us));
  else
    LAUNCHCHK(c, tune_condense(c->k, a, c->stream, &c->csched, &c->tuned_us, &c->default_us));
  c->tuned = true;
  c->assembled = c->solved = c->recovered = false;
  return HDG_OK;
}

bool use_mg(const hdg_ctx* c);
bool mixed(const hdg_ctx* c);

// The context's second stream and its events, created on first use.
int ensure_aux(hdg_ctx* c) {
  if (c->aux) return HDG_OK;
  HIPCHK(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_fill, hipEventDisableTiming));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  return HDG_OK;
}

// Condensation, trace matrix and Jacobi blocks.  With mg_early (hdg_run on
// a multigrid context with its own streams), the multigrid set-up goes to
// the second stream as soon as its inputs exist: the per-cell coarse blocks
// after the condensation (beside the fill and the Jacobi blocks), S P0 and
// the vertex levels after the fill; the solve joins it.
int assemble_impl(hdg_ctx* c, bool mg_early = false) {
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  if ((c->opt.schedule & HDG_SCHEDULE_AUTOTUNE) && !c->tuned) TRY(autotune_impl(c));
  c->mg_forked = false;
  const bool early = mg_early && use_mg(c) && c->arena == nullptr;
  HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  LAUNCHCHK(c, launch_condense_cells(c, condense_args(c)));
  if (early) {
    TRY(ensure_aux(c));
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    LAUNCHCHK(c, mg_setup_cells(c->mesh, c->mg, c->T, c->d_Ate, c->aux));
  }
  LAUNCHCHK(c, launch_fill(c->mesh, c->pat, c->d_Ate, c->d_bte, c->T, c->stream));
  if (early) {
    HIPCHK(c, hipEventRecord(c->ev_fill, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->aux, c->ev_fill, 0));
    LAUNCHCHK(c, mg_setup_sp(c->mg, c->T, c->pat, mixed(c) ? c->f_sp : nullptr, c->aux));
    LAUNCHCHK(c, mg_setup_levels(c->mg, c->d_mg, c->aux));
    HIPCHK(c, hipEventRecord(c->ev_join, c->aux));
    c->mg_forked = true;
  }
  LAUNCHCHK(c, launch_jacobi(c->pat, c->T, c->stream));
  c->jacobi_dirty = false;
  c->assembled = true;
  c->solved = c->recovered = false;
  return HDG_OK;
}

bool use_mg(const hdg_ctx* c) {
  return c->structured && c->opt.precond != HDG_PC_BLOCK_JACOBI;
}
bool mixed(const hdg_ctx* c) { return use_mg(c) && c->opt.precision == HDG_MIXED; }

// views of the trace system and Krylov vectors for the CG kernels
TraceOp<double> op64(const hdg_ctx* c) {
  return TraceOp<double>{c->T.nt, c->T.nf, c->T.N, c->T.vals, c->T.Dinv, c->T.dmask, c->T.blk};
}
TraceOp<float> op32(const hdg_ctx* c) {
  return TraceOp<float>{c->T.nt, c->T.nf, c->T.N, c->f_vals, c->f_Dinv, c->T.dmask, c->f_blk};
}
Krylov<double> kv64(const hdg_ctx* c) {
  const CgVectors& v = c->cg;
  return Krylov<double>{v.r, v.z, v.p0, v.p1, v.q, v.x, v.part, v.st, v.nblocks};
}
MgVecs<double> mv64(const hdg_ctx* c) { return MgVecs<double>{c->mg.e0, c->mg.res0, c->mg.sp}; }
MgVecs<float> mv32(const hdg_ctx* c) { return MgVecs<float>{c->f_e0, c->f_res0, c->f_sp}; }

// PCG iterations enqueued between convergence checks
int chunk_of(const hdg_ctx* c) { return use_mg(c) ? 4 : 64; }

// one chunk of CG iterations; the kernels take the iteration number from the
// device state, so the same launches serve every chunk (and a graph of them)
int enqueue_chunk(hdg_ctx* c, double rtol2, int maxit) {
  const int n = chunk_of(c);
  for (int it = 0; it < n; ++it) {
    LAUNCHCHK(c, launch_cg_spmv(c->pat, op64(c), kv64(c), 0, rtol2, maxit, 0, c->stream));
    if (mixed(c)) {
      LAUNCHCHK(c, launch_mg_update(op32(c), kv64(c), mv32(c), c->stream));
      LAUNCHCHK(c, launch_cg_check(kv64(c), c->T.N, rtol2, maxit, c->stream));
      LAUNCHCHK(c, launch_mg_precond(c->pat, op32(c), kv64(c), c->mg, c->d_mg, mv32(c), c->stream));
    } else if (use_mg(c)) {
      LAUNCHCHK(c, launch_mg_update(op64(c), kv64(c), mv64(c), c->stream));
      LAUNCHCHK(c, launch_cg_check(kv64(c), c->T.N, rtol2, maxit, c->stream));
      LAUNCHCHK(c, launch_mg_precond(c->pat, op64(c), kv64(c), c->mg, c->d_mg, mv64(c), c->stream));
    } else {
      LAUNCHCHK(c, launch_cg_update(c->T, c->cg, 0, 0, c->stream));
    }
  }
  return HDG_OK;
}

// A chunk of iterations followed by the read-back of the CG state into
// h_st[slot], eagerly or by replaying its captured graph (built once per
// set-up: every launch argument is fixed until the next hdg_setup).  The
// first solve after a set-up runs eagerly: capturing and instantiating the
// graphs costs more than one solve's launches saves, and many contexts
// capturing at once (the serving loop) serialise in the runtime.
int launch_chunk(hdg_ctx* c, int slot, double rtol2, int maxit) {
  const bool eager = (c->opt.schedule & HDG_SCHEDULE_EAGER) != 0 || c->solves == 0;
  if (!eager && !c->gexec[slot]) {
    std::lock_guard<std::mutex> lk(graph_mutex());
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = enqueue_chunk(c, rtol2, maxit);
    hipError_t e = hipMemcpyAsync(&c->h_st[slot], c->cg.st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (r != HDG_OK) return r;
    HIPCHK(c, e);
    HIPCHK(c, e2);
    e = hipGraphInstantiate(&c->gexec[slot], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(c, e);
  }
  if (!eager) {
    std::lock_guard<std::mutex> lk(graph_mutex());
    HIPCHK(c, hipGraphLaunch(c->gexec[slot], c->stream));
    return HDG_OK;
  }
  TRY(enqueue_chunk(c, rtol2, maxit));
  HIPCHK(c, hipMemcpyAsync(&c->h_st[slot], c->cg.st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
  return HDG_OK;
}

int finish_solve(hdg_ctx* c, hdg_solve_info* info, int maxit, int* err_flag);

// err_flag non-null: the recovery and the read-back of the condensation's
// error bits are queued with the solve's own read-back (one synchronisation
// for hdg_run); the recovery counts only if the solve converged.
int solve_impl(hdg_ctx* c, hdg_solve_info* info, int* err_flag = nullptr) {
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  if (c->jacobi_dirty) {
    LAUNCHCHK(c, launch_jacobi(c->pat, c->T, c->stream));
    c->jacobi_dirty = false;
  }
  const double rtol = c->opt.rtol > 0 ? c->opt.rtol : 1e-12;
  const int maxit = c->opt.maxit > 0 ? c->opt.maxit : 200000;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if (use_mg(c)) {
    const bool mx = mixed(c);
    if (mx) {
      const int64_t N = c->T.N;
      if (c->T.blk)
        LAUNCHCHK(c, launch_to_float(c->T.blk, c->f_blk, (int64_t)sym_blk_count(c->nt, c->mesh.nf), c->stream));
      else
        LAUNCHCHK(c, launch_to_float(c->T.vals, c->f_vals, (int64_t)kSlots * c->nt * N, c->stream));
      LAUNCHCHK(c, launch_to_float(c->T.Dinv, c->f_Dinv, (int64_t)dinv_count(c->nt, c->mesh.nf), c->stream));
    }
    // The multigrid set-up (small, latency-bound kernels writing the level
    // arrays and S P0) and the Lanczos estimate of the smoothing weight
    // (bandwidth-bound, writing the CG vectors, which serve as its scratch,
    // and the CG state) only read the trace system in common: the set-up
    // runs on the second stream beside the estimate.
    // A serving slot (a context on the server's arena) stays on one stream:
    // the other slots keep the device busy, and a second stream per slot
    // would outnumber the hardware queues.
    const bool fork = c->arena == nullptr;
    if (!c->mg_forked) {
      if (fork) TRY(ensure_aux(c));
      hipStream_t ss = fork ? c->aux : c->stream;
      if (fork) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
      }
      LAUNCHCHK(c, mg_setup_numeric(c->mesh, c->mg, c->d_mg, c->T, c->pat, c->d_Ate, mx ? c->f_sp : nullptr, ss));
      if (fork) HIPCHK(c, hipEventRecord(c->ev_join, c->aux));
    }
    c->mg_forked = false;
    LAUNCHCHK(c, launch_lanczos_omega(c->pat, c->T, c->cg, Whole{c->T.N, c->cg.st}, c->cg.nblocks, nullptr, 1,
                                      c->mg.omega_factor, c->mg.lanczos_steps, c->stream));
    if (fork) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (mx) {
      LAUNCHCHK(c, launch_mg_init(c->pat, c->T, op32(c), kv64(c), mv32(c), c->stream));
      LAUNCHCHK(c, launch_mg_precond(c->pat, op32(c), kv64(c), c->mg, c->d_mg, mv32(c), c->stream));
    } else {
      LAUNCHCHK(c, launch_mg_init(c->pat, c->T, op64(c), kv64(c), mv64(c), c->stream));
      LAUNCHCHK(c, launch_mg_precond(c->pat, op64(c), kv64(c), c->mg, c->d_mg, mv64(c), c->stream));
    }
  } else {
    LAUNCHCHK(c, launch_cg_init(c->pat, c->T, c->cg, c->stream));
  }
  // one chunk stays queued behind the one whose state is read back, unless
  // that one is forecast to converge (chunk_converges); launches after
  // convergence return at once
  hipEvent_t evs[2];
  HIPCHK(c, hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
  HIPCHK(c, hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
  const double rtol2 = rtol * rtol;
  int slot = 0;
  TRY(launch_chunk(c, slot, rtol2, maxit));
  HIPCHK(c, hipEventRecord(evs[slot], c->stream));
  CgState prev{}, last{};
  int nread = 0;
  bool behind = false;
  for (;;) {
    const int nxt = slot ^ 1;
    behind = nread < 2 || !chunk_converges(prev, last, chunk_of(c) + (use_mg(c) ? 1 : 0), rtol2);
    if (behind) {
      TRY(launch_chunk(c, nxt, rtol2, maxit));
      HIPCHK(c, hipEventRecord(evs[nxt], c->stream));
    }
    HIPCHK(c, hipEventSynchronize(evs[slot]));
    prev = last;
    last = c->h_st[slot];
    ++nread;
    if (last.done) break;
    if (!behind) {  // the forecast missed
      TRY(launch_chunk(c, nxt, rtol2, maxit));
      HIPCHK(c, hipEventRecord(evs[nxt], c->stream));
      behind = true;
    }
    slot = nxt;
  }
  if (behind) HIPCHK(c, hipEventSynchronize(evs[slot ^ 1]));
  (void)hipEventDestroy(evs[0]);
  (void)hipEventDestroy(evs[1]);
  ++c->solves;
  return finish_solve(c, info, maxit, err_flag);
}

// uhat = x (+ Dirichlet values), the FP64 residual, the solve record
int finish_solve(hdg_ctx* c, hdg_solve_info* info, int maxit, int* err_flag) {
  LAUNCHCHK(c, launch_cg_finish(c->T, c->cg, c->d_uhat, c->stream));
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  LAUNCHCHK(c, launch_true_residual(c->pat, c->T, c->d_uhat, c->d_scal + 1, c->cg, c->stream));
  if (err_flag) {
    LAUNCHCHK(c, launch_recover_cells(c));
    HIPCHK(c, hipMemcpyAsync(err_flag, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  }
  CgState st;
  double tr2 = 0.0;
  HIPCHK(c, hipMemcpyAsync(&st, c->cg.st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&tr2, c->d_scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (info) {
    info->iters = st.it;
    info->refinements = 0;
    info->relres = st.bb > 0 ? std::sqrt(st.rr / st.bb) : 0.0;
    info->true_relres = st.bb > 0 ? std::sqrt(tr2 / st.bb) : 0.0;
    info->seconds = ms * 1e-3;
  }
  c->solved = true;
  c->recovered = err_flag != nullptr && st.done == 1;
  if (st.done == 3)
    return fail(c, HDG_ERR_NOT_CONVERGED, "trace solve broke down at iteration %d (non-positive curvature or NaN)",
                st.it);
  if (st.done == 2)
    return fail(c, HDG_ERR_NOT_CONVERGED, "trace solve reached maxit=%d at relres %.3e", maxit,
                st.bb > 0 ? std::sqrt(st.rr / st.bb) : 0.0);
  return HDG_OK;
}

int recover_impl(hdg_ctx* c) {
  if (!c->solved) return fail(c, HDG_ERR_STATE, "hdg_solve has not been called");
  LAUNCHCHK(c, launch_recover_cells(c));
  c->recovered = true;
  return HDG_OK;
}

}  // namespace

extern "C" {

int hdg_create(hdg_ctx** out, int device, size_t hbm_budget) {
  if (!out) return HDG_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HDG_ERR_HIP;
  if (device < 0 || device >= ndev) return HDG_ERR_ARG;
  hdg_ctx* c = new hdg_ctx();
  c->device = device;
  c->budget = hbm_budget;
  int r = select_device(c);
  if (r) { delete c; return r; }
  {
    // freed blocks stay in the device pool for the next allocation instead
    // of going back to the driver at every synchronisation
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipHostMalloc((void**)&c->h_st, 2 * sizeof(CgState), hipHostMallocDefault) != hipSuccess) {
    delete c;
    return HDG_ERR_HIP;
  }
  if (dalloc(c, &c->d_err, 1) || dalloc(c, &c->cg.st, 1) || dalloc(c, &c->d_scal, 4)) {
    delete c;
    return HDG_ERR_HIP;
  }
  if (hipMemsetAsync(c->cg.st, 0, sizeof(CgState), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    hdg_destroy(c);
    return HDG_ERR_HIP;
  }
  *out = c;
  return HDG_OK;
}

void hdg_destroy(hdg_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_mesh(c);
  dfree(c, c->d_err);
  dfree(c, c->cg.st);
  dfree(c, c->d_scal);
  // whatever is still allocated (arena pieces must go back before the arena)
  while (!c->allocs.empty()) {
    void* p = c->allocs.begin()->first;
    dfree(c, p);
  }
  if (c->h_st) (void)hipHostFree(c->h_st);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->aux) {
    (void)hipStreamSynchronize(c->aux);
    (void)hipStreamDestroy(c->aux);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_fill) (void)hipEventDestroy(c->ev_fill);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

const char* hdg_last_error(const hdg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hdg_memory_used(const hdg_ctx* c, size_t* bytes) {
  if (!c || !bytes) return HDG_ERR_ARG;
  *bytes = c->used;
  return HDG_OK;
}

int hdg_set_mesh(hdg_ctx* c, int64_t nn, const double* xy, int64_t nc, const int64_t* cells,
                 int64_t nf, const int64_t* faces) {
  if (!c) return HDG_ERR_ARG;
  if (nn < 3 || nc < 1 || nf < 3 || !xy || !cells || !faces)
    return fail(c, HDG_ERR_ARG, "empty mesh or null array");
  if (nn > INT32_MAX || nf > INT32_MAX || nc > INT32_MAX / 9)
    return fail(c, HDG_ERR_ARG, "mesh too large for 32-bit device indices");
  TRY(select_device(c));
  // validate and convert to 0-based int32
  std::vector<int> cn(3 * nc), cf(3 * nc), fcell(2 * nf);
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t c1 = faces[f + 2 * nf], c2 = faces[f + 3 * nf];
    if (c1 < 1 || c1 > nc || c2 < 0 || c2 > nc || c1 == c2)
      return fail(c, HDG_ERR_ARG, "face %lld has invalid cells (%lld, %lld)", (long long)f + 1,
                  (long long)c1, (long long)c2);
    fcell[2 * f] = (int)(c1 - 1);
    fcell[2 * f + 1] = (int)(c2 - 1);
  }
  int64_t nblk = 0;
  for (int64_t f = 0; f < nf; ++f) nblk += faces[f + 3 * nf] > 0 ? 5 : 3;
  for (int64_t e = 0; e < nc; ++e)
    for (int l = 0; l < 3; ++l) {
      const int64_t nd = cells[6 * e + l], fc = cells[6 * e + 3 + l];
      if (nd < 1 || nd > nn) return fail(c, HDG_ERR_ARG, "cell %lld: node id out of range", (long long)e + 1);
      if (fc < 1 || fc > nf) return fail(c, HDG_ERR_ARG, "cell %lld: face id out of range", (long long)e + 1);
      if (fcell[2 * (fc - 1)] != e && fcell[2 * (fc - 1) + 1] != e)
        return fail(c, HDG_ERR_ARG, "cell %lld: face %lld does not list the cell", (long long)e + 1,
                    (long long)fc);
      cn[3 * e + l] = (int)(nd - 1);
      cf[3 * e + l] = (int)(fc - 1);
    }
  free_mesh(c);
  double* dxy;
  int *dcn, *dcf, *dfc;
  TRY(dalloc(c, &dxy, 2 * nn));
  TRY(dalloc(c, &dcn, 3 * nc));
  TRY(dalloc(c, &dcf, 3 * nc));
  TRY(dalloc(c, &dfc, 2 * nf));
  HIPCHK(c, hipMemcpyAsync(dxy, xy, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dcn, cn.data(), sizeof(int) * 3 * nc, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dcf, cf.data(), sizeof(int) * 3 * nc, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dfc, fcell.data(), sizeof(int) * 2 * nf, hipMemcpyHostToDevice, c->stream));
  c->mesh.nn = nn;
  c->mesh.nc = nc;
  c->mesh.nf = nf;
  c->mesh.xy = dxy;
  c->mesh.cn = dcn;
  c->mesh.cf = dcf;
  c->mesh.fcell = dfc;
  TRY(dalloc(c, &c->pat.bcol, kSlots * nf));
  TRY(dalloc(c, &c->pat.src, 2 * kSlots * nf));
  TRY(dalloc(c, &c->pat.nslot, nf));
  TRY(dalloc(c, &c->pat.self, nf));
  TRY(dalloc(c, &c->pat.tslot, kSlots * nf));
  TRY(dalloc(c, &c->pat.nbr, 4 * nf));
  LAUNCHCHK(c, launch_pattern(c->mesh, c->pat, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->h_cf.swap(cf);
  c->nblock_cols = nblk;
  TRY(build_vertex_structures(c, nn, nf, faces, cells, nc));
  c->has_mesh = true;
  return HDG_OK;
}

int hdg_set_rectangle_mesh(hdg_ctx* c, int64_t nx, int64_t ny, double x0, double y0, double x1, double y1,
                           double jitter, uint64_t seed) {
  if (!c) return HDG_ERR_ARG;
  if (nx < 1 || ny < 1) return fail(c, HDG_ERR_ARG, "rectangle mesh needs nx, ny >= 1");
  // the device generator writes the counter-clockwise numbering; a box given
  // with decreasing sides (whose triangles rectangle_mesh reorders) goes
  // through the host generator and hdg_set_mesh
  if (!(x1 > x0) || !(y1 > y0))
    return fail(c, HDG_ERR_UNSUPPORTED,
                "the device generator needs x1 > x0 and y1 > y0; use hdg_rectangle_mesh + hdg_set_mesh");
  const int64_t nn = (nx + 1) * (ny + 1), nc = 2 * nx * ny, nf = hdg_rectangle_nfaces(nx, ny);
  if (nn > INT32_MAX || 2 * nf > INT32_MAX || nc > INT32_MAX / 9)
    return fail(c, HDG_ERR_ARG, "mesh too large for 32-bit device indices");
  TRY(select_device(c));
  free_mesh(c);
  double* dxy;
  int *dcn, *dcf, *dfc;
  TRY(dalloc(c, &dxy, 2 * nn));
  TRY(dalloc(c, &dcn, 3 * nc));
  TRY(dalloc(c, &dcf, 3 * nc));
  TRY(dalloc(c, &dfc, 2 * nf));
  MgData& M = c->mg;
  TRY(dalloc(c, &M.fv, 2 * nf));
  TRY(dalloc(c, &M.vf_ptr, nn + 1));
  TRY(dalloc(c, &M.vf_idx, 2 * nf));
  TRY(dalloc(c, &M.vf2, 2 * nf));
  TRY(dalloc(c, &M.fq4, 4 * nf));
  HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  RectMeshArgs a;
  a.nx = nx;
  a.ny = ny;
  a.nf = nf;
  a.x0 = x0;
  a.y0 = y0;
  a.x1 = x1;
  a.y1 = y1;
  a.jitter = jitter;
  a.seed = seed;
  a.xy = dxy;
  a.cn = dcn;
  a.cf = dcf;
  a.fcell = dfc;
  a.fv = M.fv;
  a.vf_ptr = M.vf_ptr;
  a.vf_idx = M.vf_idx;
  a.fq4 = M.fq4;
  a.err = c->d_err;
  LAUNCHCHK(c, launch_rect_mesh(a, c->stream));
  c->mesh.nn = nn;
  c->mesh.nc = nc;
  c->mesh.nf = nf;
  c->mesh.xy = dxy;
  c->mesh.cn = dcn;
  c->mesh.cf = dcf;
  c->mesh.fcell = dfc;
  TRY(dalloc(c, &c->pat.bcol, kSlots * nf));
  TRY(dalloc(c, &c->pat.src, 2 * kSlots * nf));
  TRY(dalloc(c, &c->pat.nslot, nf));
  TRY(dalloc(c, &c->pat.self, nf));
  TRY(dalloc(c, &c->pat.tslot, kSlots * nf));
  TRY(dalloc(c, &c->pat.nbr, 4 * nf));
  LAUNCHCHK(c, launch_pattern(c->mesh, c->pat, c->stream));
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flag & 4) {
    free_mesh(c);
    return fail(c, HDG_ERR_BOUNDARY, "a boundary face lies on no side of the box (generate_mesh.jl:80)");
  }
  c->h_cf.clear();  // downloaded when an export needs it
  const int64_t nbdry = 2 * nx + 2 * ny;
  c->nblock_cols = 5 * (nf - nbdry) + 3 * nbdry;
  c->structured = true;
  M.nv = nn;
  mg_plan_levels((int)nx, (int)ny, M);
  for (int l = 1; l < M.nlev; ++l) {
    MgLevel& L = M.lev[l];
    TRY(dalloc(c, &L.a, kStencil * L.n));
    TRY(dalloc(c, &L.dinv, L.n));
    TRY(dalloc(c, &L.mask, L.n));
    TRY(dalloc(c, &L.r, L.n));
    TRY(dalloc(c, &L.e, L.n));
    TRY(dalloc(c, &L.t, L.n));
  }
  TRY(dalloc(c, &M.cinv, (size_t)kCoarseDirect * kCoarseDirect));
  if (M.tail < M.nlev) {
    int64_t tn = 0;
    for (int t = M.tail; t < M.nlev; ++t) tn += M.lev[t].n;
    TRY(dalloc(c, &M.tailpack, (kStencil + 1) * tn));
  }
  c->has_mesh = true;
  return HDG_OK;
}

int hdg_setup(hdg_ctx* c, const hdg_options* opt) {
  if (!c || !opt) return HDG_ERR_ARG;
  if (!c->has_mesh) return fail(c, HDG_ERR_STATE, "hdg_set_mesh has not been called");
  if (opt->k < 1 || opt->k > 4) return fail(c, HDG_ERR_UNSUPPORTED, "k=%d: supported orders are 1..4", opt->k);
  if (opt->precision != HDG_FP64 && opt->precision != HDG_MIXED)
    return fail(c, HDG_ERR_ARG, "unknown precision mode %d", opt->precision);
  if (opt->precond < HDG_PC_AUTO || opt->precond > HDG_PC_MULTIGRID)
    return fail(c, HDG_ERR_ARG, "unknown preconditioner %d", opt->precond);
  if (opt->precond == HDG_PC_MULTIGRID && !c->structured)
    return fail(c, HDG_ERR_UNSUPPORTED, "the multigrid preconditioner needs a structured rectangle_mesh");
  if (opt->precision == HDG_MIXED && (!c->structured || opt->precond == HDG_PC_BLOCK_JACOBI))
    return fail(c, HDG_ERR_UNSUPPORTED,
                "mixed precision runs on the multigrid path: it needs a structured rectangle_mesh");
  TRY(select_device(c));
  hdg_options o = *opt;
  if (o.quad_degree <= 0) o.quad_degree = o.k + 1;
  RefTables tab;
  if (!tab.build(o.k, o.quad_degree))
    return fail(c, HDG_ERR_QUAD, "Quadrature rule of order %d not available", o.quad_degree);
  free_problem(c);
  c->opt = o;
  c->k = o.k;
  mg_order_choices(o.k, c->mg);
  c->nu = tab.lay.nu;
  c->nq = 2 * c->nu;
  c->m = 3 * c->nu;
  c->nt = tab.lay.nt;
  c->ntr = 3 * c->nt;
  c->nqp = tab.lay.nqp;
  c->nfq = tab.lay.nfq;
  c->tab = tab;
  const TableLayout& L = tab.lay;
  c->lay = TabLayoutDev{L.nu, L.nt, L.nqp, L.nfq, L.Mref, L.Dref, L.Cref, L.G, L.Href, L.Nw, L.N, L.Mg, L.W, L.total};
  TRY(dalloc(c, &c->d_tab, L.total));
  HIPCHK(c, hipMemcpyAsync(c->d_tab, tab.data.data(), sizeof(double) * L.total, hipMemcpyHostToDevice, c->stream));
  {
    SpdLayout sl;
    std::vector<double> sd;
    if (o.k <= 3 && build_spd_tables(tab, sl, sd)) {
      c->spd = SpdLayoutDev{sl.nu, sl.nt, sl.nqp, sl.R, sl.C, sl.Gam, sl.Z, sl.V, sl.P, sl.Om, sl.H, sl.Nw, sl.Mg,
                            sl.total};
      TRY(dalloc(c, &c->d_spd, sl.total));
      HIPCHK(c, hipMemcpyAsync(c->d_spd, sd.data(), sizeof(double) * sl.total, hipMemcpyHostToDevice, c->stream));
      // the upload reads sd: finish it before sd goes out of scope
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
  }
  const int64_t nc = c->mesh.nc, nf = c->mesh.nf, N = nf * c->nt;
  TRY(dalloc(c, &c->d_Ate, nc * c->ntr * c->ntr));
  TRY(dalloc(c, &c->d_bte, nc * c->ntr));
  TRY(dalloc(c, &c->d_Ke, nc * c->m * c->ntr));
  TRY(dalloc(c, &c->d_be, nc * c->m));
  TRY(dalloc(c, &c->d_qu, nc * c->m));
  c->T = TraceSystem{};
  c->T.nt = c->nt;
  c->T.nf = nf;
  c->T.N = N;
  // the trace operator: symmetric blocks at four modes per face, else the
  // face-ELL (the exports expand the blocks into a face-ELL on demand)
  if (sym_blocks(c->nt))
    TRY(dalloc(c, &c->T.blk, sym_blk_count(c->nt, c->mesh.nf)));
  else
    TRY(dalloc(c, &c->T.vals, (size_t)kSlots * c->nt * N));
  TRY(dalloc(c, &c->T.rhs, N));
  TRY(dalloc(c, &c->T.Dinv, dinv_count(c->nt, c->mesh.nf)));
  TRY(dalloc(c, &c->d_dmask, N));
  TRY(dalloc(c, &c->d_uD, N));
  HIPCHK(c, hipMemsetAsync(c->d_dmask, 0, N, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_uD, 0, sizeof(double) * N, c->stream));
  c->T.dmask = c->d_dmask;
  c->T.uD = c->d_uD;
  c->cg.nblocks = cg_blocks(N, c->nt);
  TRY(dalloc(c, &c->cg.x, N));
  TRY(dalloc(c, &c->cg.r, N));
  TRY(dalloc(c, &c->cg.z, N));
  TRY(dalloc(c, &c->cg.p0, N));
  TRY(dalloc(c, &c->cg.p1, N));
  TRY(dalloc(c, &c->cg.q, N));
  TRY(dalloc(c, &c->cg.part, 4 * (size_t)c->cg.nblocks + 64));
  TRY(dalloc(c, &c->d_uhat, N));
  TRY(dalloc(c, &c->d_part2, 4096));
  if (c->structured) {
    TRY(dalloc(c, &c->mg.cellA, 9 * nc));
    TRY(dalloc(c, &c->mg.e0, N));
    TRY(dalloc(c, &c->mg.res0, N));
    TRY(dalloc(c, &c->mg.fc, 2 * c->mesh.nf));
    TRY(dalloc(c, &c->mg.sp, 4 * N));
    TRY(dalloc(c, &c->d_mg, 1));
    HIPCHK(c, hipMemcpyAsync(c->d_mg, &c->mg, sizeof(MgData), hipMemcpyHostToDevice, c->stream));
  }
  if (c->structured && o.precision == HDG_MIXED) {
    if (c->T.blk)
      TRY(dalloc(c, &c->f_blk, sym_blk_count(c->nt, c->mesh.nf)));
    else
      TRY(dalloc(c, &c->f_vals, (size_t)kSlots * c->nt * N));
    TRY(dalloc(c, &c->f_Dinv, dinv_count(c->nt, c->mesh.nf)));
    TRY(dalloc(c, &c->f_sp, 4 * N));
    TRY(dalloc(c, &c->f_e0, N));
    TRY(dalloc(c, &c->f_res0, N));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->src = SrcParams{};
  c->jacobi_dirty = true;
  c->has_setup = true;
  return HDG_OK;
}

int hdg_set_source(hdg_ctx* c, const hdg_source* s) {
  if (!c || !s) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  if (s->kind < 0 || s->kind > 3) return fail(c, HDG_ERR_ARG, "unknown source kind %d", s->kind);
  TRY(select_device(c));
  SrcParams p;
  p.kind = s->kind;
  p.kappa_on = s->kappa_on ? 1 : 0;
  p.ka = s->ka;
  p.kb = s->kb;
  for (int i = 0; i < 16; ++i) p.coef[i] = s->coef[i];
  dfree(c, c->d_fqp);
  if (s->kind == 3) {
    if (!s->f_qp) return fail(c, HDG_ERR_ARG, "HDG_SRC_VALUES needs f_qp");
    const size_t n = (size_t)c->mesh.nc * c->nqp;
    TRY(dalloc(c, &c->d_fqp, n));
    HIPCHK(c, hipMemcpyAsync(c->d_fqp, s->f_qp, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->src = p;
  c->assembled = c->solved = c->recovered = false;
  return HDG_OK;
}

int hdg_set_dirichlet(hdg_ctx* c, int64_t n, const int64_t* dofs, const double* vals) {
  if (!c) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  if (n < 0 || (n > 0 && (!dofs || !vals))) return fail(c, HDG_ERR_ARG, "bad Dirichlet arrays");
  TRY(select_device(c));
  const int64_t N = c->T.N;
  std::vector<unsigned char> mask(N, 0);
  std::vector<double> uD(N, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    if (dofs[i] < 1 || dofs[i] > N)
      return fail(c, HDG_ERR_ARG, "Dirichlet dof %lld out of range 1..%lld", (long long)dofs[i], (long long)N);
    mask[dofs[i] - 1] = 1;
    uD[dofs[i] - 1] = vals[i];
  }
  HIPCHK(c, hipMemcpyAsync(c->d_dmask, mask.data(), N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_uD, uD.data(), sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->jacobi_dirty = true;
  c->solved = c->recovered = false;
  return HDG_OK;
}

int hdg_set_dirichlet_boundary(hdg_ctx* c, double value) {
  if (!c) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  TRY(select_device(c));
  LAUNCHCHK(c, launch_boundary_dirichlet(c->mesh.fcell, 0, c->mesh.nf, c->nt, c->d_dmask, c->d_uD, value,
                                         c->stream));
  c->jacobi_dirty = true;
  c->solved = c->recovered = false;
  return HDG_OK;
}

int hdg_set_dirichlet_function(hdg_ctx* c, int g_kind, const double* par, int flags) {
  if (!c) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  if (g_kind != HDG_G_LINEAR && g_kind != HDG_G_SINSIN) return fail(c, HDG_ERR_ARG, "unknown g kind %d", g_kind);
  if (g_kind == HDG_G_LINEAR && !par) return fail(c, HDG_ERR_ARG, "HDG_G_LINEAR needs 3 parameters");
  if (flags & ~HDG_DIRICHLET_REFERENCE) return fail(c, HDG_ERR_ARG, "unknown flags %d", flags);
  TRY(select_device(c));
  const RefTables& t = c->tab;
  const int nfq = (int)t.fx.size();
  if (nfq > kMaxFacePoints) return fail(c, HDG_ERR_UNSUPPORTED, "face rule of %d points", nfq);
  if ((flags & HDG_DIRICHLET_REFERENCE) && (int)t.qw.size() < nfq)
    return fail(c, HDG_ERR_ARG, "the cell rule has %d points, fewer than the face rule's %d",
                (int)t.qw.size(), nfq);
  DirichletRule R;
  R.nfq = nfq;
  for (int q = 0; q < nfq; ++q) {
    R.x[q] = t.fx[q];
    R.w[q] = t.fw[q];
    R.wc[q] = q < (int)t.qw.size() ? t.qw[q] : 0.0;
    double lv[5];
    legendre01_eval(c->nt, t.fx[q], lv);
    for (int i = 0; i < c->nt; ++i) R.leg[i * nfq + q] = lv[i];
  }
  const double p3[3] = {par ? par[0] : 0.0, par ? par[1] : 0.0, par ? par[2] : 0.0};
  LAUNCHCHK(c, launch_dirichlet_project(c->mesh, c->nt, R, g_kind, p3, flags, c->d_dmask, c->d_uD, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->jacobi_dirty = true;
  c->solved = c->recovered = false;
  return HDG_OK;
}

int hdg_query_sizes(const hdg_ctx* c, hdg_sizes* s) {
  if (!c || !s) return HDG_ERR_ARG;
  std::memset(s, 0, sizeof(*s));
  s->n_nodes = c->mesh.nn;
  s->n_cells = c->mesh.nc;
  s->n_faces = c->mesh.nf;
  if (!c->has_setup) return HDG_OK;
  s->n_dofs = c->T.N;
  s->nu = c->nu;
  s->nq = c->nq;
  s->nt = c->nt;
  s->m = c->m;
  s->ntr = c->ntr;
  s->nqp = c->nqp;
  s->nfq = c->nfq;
  // rows of a face with two cells hold 5 block columns, with one cell 3
  s->nnz = c->nblock_cols * (int64_t)c->nt * c->nt;
  s->sym_blocks = c->T.blk ? 1 : 0;
  return HDG_OK;
}

int hdg_assemble_condense(hdg_ctx* c) {
  if (!c) return HDG_ERR_ARG;
  TRY(select_device(c));
  TRY(assemble_impl(c));
  return check_err_flag(c);
}

int hdg_solve(hdg_ctx* c, double* uhat_out, hdg_solve_info* info) {
  if (!c) return HDG_ERR_ARG;
  TRY(select_device(c));
  int r = solve_impl(c, info);
  if (r != HDG_OK && r != HDG_ERR_NOT_CONVERGED) return r;
  if (uhat_out) {
    HIPCHK(c, hipMemcpyAsync(uhat_out, c->d_uhat, sizeof(double) * c->T.N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return r;
}

int hdg_set_trace(hdg_ctx* c, const double* uhat) {
  if (!c || !uhat) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  TRY(select_device(c));
  HIPCHK(c, hipMemcpyAsync(c->d_uhat, uhat, sizeof(double) * c->T.N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->solved = true;
  c->recovered = false;
  return HDG_OK;
}

int hdg_get_trace(hdg_ctx* c, double* uhat) {
  if (!c || !uhat) return HDG_ERR_ARG;
  if (!c->solved) return fail(c, HDG_ERR_STATE, "no trace solution: call hdg_solve or hdg_run");
  TRY(select_device(c));
  HIPCHK(c, hipMemcpyAsync(uhat, c->d_uhat, sizeof(double) * c->T.N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HDG_OK;
}

int hdg_recover(hdg_ctx* c, double* q, double* u, double* uhat_cell) {
  if (!c) return HDG_ERR_ARG;
  TRY(select_device(c));
  if (!c->recovered) TRY(recover_impl(c));
  const int64_t nc = c->mesh.nc;
  const int M = c->m, nq = c->nq, nu = c->nu, nt = c->nt;
  if (q || u) {
    std::vector<double> qu((size_t)nc * M);
    HIPCHK(c, hipMemcpyAsync(qu.data(), c->d_qu, sizeof(double) * qu.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t e = 0; e < nc; ++e) {
      if (q)
        for (int i = 0; i < nq; ++i) q[e + nc * i] = qu[e * M + i];
      if (u)
        for (int i = 0; i < nu; ++i) u[e + nc * i] = qu[e * M + nq + i];
    }
  }
  if (uhat_cell) {
    if ((int64_t)c->h_cf.size() != 3 * nc) {
      c->h_cf.resize(3 * nc);
      HIPCHK(c, hipMemcpyAsync(c->h_cf.data(), c->mesh.cf, sizeof(int) * 3 * nc, hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<double> uh(c->T.N);
    HIPCHK(c, hipMemcpyAsync(uh.data(), c->d_uhat, sizeof(double) * uh.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t e = 0; e < nc; ++e)
      for (int lf = 0; lf < 3; ++lf)
        for (int t = 0; t < nt; ++t) uhat_cell[e + nc * (t + nt * lf)] = uh[(int64_t)c->h_cf[3 * e + lf] * nt + t];
  }
  return HDG_OK;
}

int hdg_errornorm(hdg_ctx* c, int kind, const double* par, double* sq) {
  if (!c || !sq) return HDG_ERR_ARG;
  if (kind != HDG_UEX_SINSIN && kind != HDG_UEX_LINEAR) return fail(c, HDG_ERR_ARG, "unknown exact solution %d", kind);
  if (kind == HDG_UEX_LINEAR && !par) return fail(c, HDG_ERR_ARG, "HDG_UEX_LINEAR needs 3 parameters");
  TRY(select_device(c));
  if (!c->recovered) TRY(recover_impl(c));
  const double p3[3] = {par ? par[0] : 0.0, par ? par[1] : 0.0, par ? par[2] : 0.0};
  LAUNCHCHK(c, launch_errornorm(c->mesh, c->d_tab, c->lay, c->k, c->d_qu, kind, p3, c->d_part2, c->d_scal,
                                c->stream));
  HIPCHK(c, hipMemcpyAsync(sq, c->d_scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HDG_OK;
}

int hdg_run(hdg_ctx* c, hdg_solve_info* info) {
  if (!c) return HDG_ERR_ARG;
  TRY(select_device(c));
  TRY(assemble_impl(c, true));
  int flag = 0;
  const int r = solve_impl(c, info, &flag);
  // a singular local matrix or a negative Jacobian is reported as such, not
  // as the failed solve it causes (the flag was read with the solve's
  // read-back, unless the solve stopped before it)
  if (r != HDG_OK && r != HDG_ERR_NOT_CONVERGED) {
    TRY(check_err_flag(c));
    return r;
  }
  TRY(err_bits(c, flag));
  return r;
}

namespace {

// Julia's CSC of K (the CSR of K^T: column (g, b) holds rows (f, a) for the
// faces f of row g's pattern, which is structurally symmetric, ascending)
// from a face-ELL value array and an rhs in device memory.
int export_csc(hdg_ctx* c, const double* d_vals, const double* d_rhs, int64_t* colptr, int64_t* rowval,
               double* nzval, double* rhs) {
  const int64_t nf = c->mesh.nf, N = c->T.N;
  const int nt = c->nt;
  std::vector<int> bcol(kSlots * nf);
  std::vector<unsigned char> nslot(nf), tslot(kSlots * nf);
  HIPCHK(c, hipMemcpyAsync(bcol.data(), c->pat.bcol, sizeof(int) * bcol.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(nslot.data(), c->pat.nslot, nf, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(tslot.data(), c->pat.tslot, tslot.size(), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> vals;
  if (nzval) {
    vals.resize((size_t)kSlots * nt * N);
    HIPCHK(c, hipMemcpyAsync(vals.data(), d_vals, sizeof(double) * vals.size(), hipMemcpyDeviceToHost, c->stream));
  }
  if (rhs) HIPCHK(c, hipMemcpyAsync(rhs, d_rhs, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t w = 0;
  if (colptr) colptr[0] = 1;
  for (int64_t g = 0; g < nf; ++g)
    for (int b = 0; b < nt; ++b) {
      for (int s = 0; s < nslot[g]; ++s) {
        const int64_t f = bcol[kSlots * g + s];
        const int ts = tslot[kSlots * g + s];
        for (int a = 0; a < nt; ++a) {
          if (rowval) rowval[w] = f * nt + a + 1;
          if (nzval) nzval[w] = vals[(size_t)(ts * nt + b) * N + f * nt + a];
          ++w;
        }
      }
      if (colptr) colptr[g * nt + b + 1] = w + 1;
    }
  return HDG_OK;
}

}  // namespace

int hdg_get_system(hdg_ctx* c, int64_t* colptr, int64_t* rowval, double* nzval, double* rhs) {
  if (!c) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  TRY(select_device(c));
  if (!c->T.blk || !nzval) return export_csc(c, c->T.vals, c->T.rhs, colptr, rowval, nzval, rhs);
  double* vals = nullptr;
  TRY(dalloc(c, &vals, (size_t)kSlots * c->nt * c->T.N));
  int r = HDG_OK;
  const int e = launch_expand_ell(c->pat, c->T, vals, c->stream);
  if (e) r = fail(c, HDG_ERR_HIP, "expanding the trace blocks: %s", hipGetErrorString((hipError_t)e));
  if (r == HDG_OK) r = export_csc(c, vals, c->T.rhs, colptr, rowval, nzval, rhs);
  (void)hipStreamSynchronize(c->stream);
  dfree(c, vals);
  return r;
}

int hdg_get_applied_system(hdg_ctx* c, int64_t* colptr, int64_t* rowval, double* nzval, double* rhs,
                           double* meandiag_out) {
  if (!c) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  TRY(select_device(c));
  const int64_t N = c->T.N;
  double *vals = nullptr, *b = nullptr, *part = nullptr;
  int r = dalloc(c, &vals, (size_t)kSlots * c->nt * N);
  if (r == HDG_OK) r = dalloc(c, &b, N);
  if (r == HDG_OK) r = dalloc(c, &part, 1024 + 1);
  if (r == HDG_OK) {
    // with symmetric blocks apply! runs in place on their face-ELL expansion
    TraceSystem T = c->T;
    int e = 0;
    if (T.blk) {
      e = launch_expand_ell(c->pat, T, vals, c->stream);
      T.vals = vals;
    }
    if (!e) e = launch_apply(c->pat, T, part, part + 1024, vals, b, c->stream);
    if (e) r = fail(c, HDG_ERR_HIP, "apply! kernels: %s", hipGetErrorString((hipError_t)e));
  }
  if (r == HDG_OK && meandiag_out) {
    const hipError_t e = hipMemcpyAsync(meandiag_out, part + 1024, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) r = fail(c, HDG_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
  }
  if (r == HDG_OK) r = export_csc(c, vals, b, colptr, rowval, nzval, rhs);
  (void)hipStreamSynchronize(c->stream);
  dfree(c, part);
  dfree(c, b);
  dfree(c, vals);
  return r;
}

int hdg_get_local_ops(hdg_ctx* c, double* Ke, double* be) {
  if (!c) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  TRY(select_device(c));
  const int64_t nc = c->mesh.nc;
  if ((Ke || be) && !c->ke_full) {
    LAUNCHCHK(c, launch_expand_ke_spd(c->k, c->mesh, c->d_spd, c->d_Ke, c->d_be, c->stream));
    c->ke_full = true;
  }
  if (Ke) HIPCHK(c, hipMemcpyAsync(Ke, c->d_Ke, sizeof(double) * nc * c->m * c->ntr, hipMemcpyDeviceToHost, c->stream));
  if (be) HIPCHK(c, hipMemcpyAsync(be, c->d_be, sizeof(double) * nc * c->m, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HDG_OK;
}

int hdg_get_condensed(hdg_ctx* c, double* Ate, double* bte) {
  if (!c) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  TRY(select_device(c));
  const int64_t nc = c->mesh.nc;
  const int ntr = c->ntr;
  if (Ate) {
    std::vector<double> a((size_t)nc * ntr * ntr);
    HIPCHK(c, hipMemcpyAsync(a.data(), c->d_Ate, sizeof(double) * a.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t e = 0; e < nc; ++e)
      for (int i = 0; i < ntr; ++i)
        for (int j = 0; j < ntr; ++j) Ate[e * ntr * ntr + i + ntr * j] = a[e * ntr * ntr + i * ntr + j];
  }
  if (bte) {
    HIPCHK(c, hipMemcpyAsync(bte, c->d_bte, sizeof(double) * nc * ntr, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return HDG_OK;
}

int hdg_get_element_blocks(hdg_ctx* c, int64_t first, int64_t count, double* A, double* B, double* C,
                           double* E, double* F, double* H, double* be) {
  if (!c) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  if (first < 0 || count < 1 || first + count > c->mesh.nc) return fail(c, HDG_ERR_ARG, "cell range out of bounds");
  TRY(select_device(c));
  const int nq = c->nq, nu = c->nu, ntr = c->ntr;
  const size_t sz[7] = {(size_t)nq * nq, (size_t)nq * nu, (size_t)nu * nu, (size_t)nq * ntr,
                        (size_t)nu * ntr, (size_t)ntr * ntr, (size_t)nu};
  double* host[7] = {A, B, C, E, F, H, be};
  double* dev[7] = {nullptr};
  int rc = HDG_OK;
  for (int i = 0; i < 7 && rc == HDG_OK; ++i)
    if (host[i]) rc = dalloc(c, &dev[i], sz[i] * count);
  if (rc == HDG_OK) {
    HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    CondenseArgs a;
    a.mesh = c->mesh;
    a.tables = c->d_tab;
    a.lay = c->lay;
    a.src = c->src;
    a.src.f_qp = c->d_fqp;
    a.tau = c->opt.tau;
    a.err = c->d_err;
    a.first = first;
    a.count = count;
    a.A = dev[0]; a.B = dev[1]; a.C = dev[2]; a.E = dev[3]; a.F = dev[4]; a.H = dev[5]; a.bev = dev[6];
    int e = launch_condense(c->k, a, true, c->stream);
    if (e) rc = fail(c, HDG_ERR_HIP, "element block kernel failed: %s", hipGetErrorString((hipError_t)e));
    for (int i = 0; i < 7 && rc == HDG_OK; ++i)
      if (host[i] && hipMemcpyAsync(host[i], dev[i], sizeof(double) * sz[i] * count, hipMemcpyDeviceToHost,
                                    c->stream) != hipSuccess)
        rc = fail(c, HDG_ERR_HIP, "copy of element blocks failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == HDG_OK) rc = fail(c, HDG_ERR_HIP, "sync failed");
  }
  for (int i = 0; i < 7; ++i) dfree(c, dev[i]);
  return rc;
}

int hdg_dubiner(int k, double r, double s, double* val, double* grad) {
  if (k < 0 || k > 14 || !val) return HDG_ERR_ARG;
  dubiner_eval(k, r, s, val, grad);
  return HDG_OK;
}

int hdg_quadrature(int dim, int order, int cap, double* pts, double* w) {
  std::vector<double> p, ww;
  if (dim == 2) {
    if (!triangle_rule(order, p, ww)) return -HDG_ERR_QUAD;
  } else if (dim == 1) {
    if (order < 1) return -HDG_ERR_QUAD;
    gauss_legendre01(order, p, ww);
  } else {
    return -HDG_ERR_ARG;
  }
  const int n = (int)ww.size();
  if (n > cap || !pts || !w) return -HDG_ERR_ARG;
  std::copy(p.begin(), p.end(), pts);
  std::copy(ww.begin(), ww.end(), w);
  return n;
}

int hdg_autotune(hdg_ctx* c, hdg_schedule* out) {
  if (!c) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  TRY(select_device(c));
  TRY(autotune_impl(c));
  if (out) {
    const bool schur = condense_args(c).spd_tables != nullptr;
    out->condense_form = schur ? HDG_CONDENSE_SCHUR : HDG_CONDENSE_ELIMINATION;
    out->condense_seg = schur ? 16 : c->csched.seg;
    out->condense_grid = schur ? c->csched.spd_grid : c->csched.grid;
    out->condense_waves = schur ? c->csched.spd_wpe : 0;
    out->condense_us = c->tuned_us;
    out->default_us = c->default_us;
  }
  return HDG_OK;
}

int hdg_set_schedule(hdg_ctx* c, const hdg_schedule* sc) {
  if (!c || !sc) return HDG_ERR_ARG;
  if (!c->has_setup) return fail(c, HDG_ERR_STATE, "hdg_setup has not been called");
  const int seg = sc->condense_seg, nat = condense_native_seg(c->k);
  // the Schur form's 16 lanes (what hdg_autotune reports for it) are no
  // width of the elimination form, which then keeps its narrowest
  const bool schur = condense_args(c).spd_tables != nullptr;
  if (seg != 0 && ((seg != 16 && seg != 32 && seg != 64) || (seg < nat && !schur)))
    return fail(c, HDG_ERR_ARG, "condense_seg=%d: 0 or 16/32/64 lanes, at least %d at k=%d", seg, nat, c->k);
  const int eseg = seg < nat ? 0 : seg;
  if (sc->condense_grid < 0 || sc->condense_grid > (1 << 20))
    return fail(c, HDG_ERR_ARG, "condense_grid=%d out of range", sc->condense_grid);
  const int w = sc->condense_waves;
  if (w != 0 && w != 2 && w != 3) return fail(c, HDG_ERR_ARG, "condense_waves=%d: 0, 2 or 3", w);
  c->csched = CondenseSched{eseg, sc->condense_grid, w, sc->condense_grid};
  c->tuned = true;
  return HDG_OK;
}

int hdg_time_kernel(hdg_ctx* c, int which, int reps, double* avg_us) {
  if (!c || !avg_us || reps < 1) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "nothing assembled to time");
  TRY(select_device(c));
  // make sure the solver state exists (pq, rz) so the kernels divide by
  // finite numbers; timing runs leave x and r untouched
  if (!c->solved) LAUNCHCHK(c, launch_cg_init(c->pat, c->T, c->cg, c->stream));
  const CondenseArgs a = condense_args(c);
  double* scratch = nullptr;
  const int64_t ncopy = (int64_t)kSlots * c->nt * c->T.N;  // the size of the face-ELL values
  // source: the face-ELL values, or (symmetric blocks) a second scratch half
  const double* csrc = c->T.vals;
  if (which == 3) {
    TRY(dalloc(c, &scratch, c->T.vals ? ncopy : 2 * ncopy));
    if (!csrc) {
      csrc = scratch + ncopy;
      HIPCHK(c, hipMemsetAsync(scratch + ncopy, 0, sizeof(double) * ncopy, c->stream));
    }
  }
  auto once = [&]() -> int {
    if (which == 0) return launch_cg_spmv(c->pat, op64(c), kv64(c), 2, 0.0, 0, 1, c->stream);
    if (which == 1) return launch_cg_update(c->T, c->cg, 2, 1, c->stream);
    if (which == 3) return launch_copy(csrc, scratch, ncopy, c->stream);
    return launch_condense_cells(c, a);
  };
  int rc = once();  // warm-up
  if (rc) { dfree(c, scratch); return fail(c, HDG_ERR_HIP, "kernel launch failed"); }
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < reps && rc == 0; ++i) rc = once();
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  dfree(c, scratch);
  if (rc) return fail(c, HDG_ERR_HIP, "kernel launch failed");
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *avg_us = 1e3 * ms / reps;
  c->solved = c->recovered = false;  // the CG state was overwritten
  return HDG_OK;
}

int hdg_refine_probe(hdg_ctx* c, int outer, double inner_rtol, double* relres, int32_t* inner_iters) {
  if (!c || !relres || outer < 0 || outer > 64 || !(inner_rtol >= 0.0)) return HDG_ERR_ARG;
  if (!c->assembled) return fail(c, HDG_ERR_STATE, "hdg_assemble_condense has not been called");
  TRY(select_device(c));
  for (int j = 0; j <= outer; ++j) relres[j] = std::nan("");
  if (inner_iters)
    for (int j = 0; j < outer; ++j) inner_iters[j] = 0;
  const int64_t N = c->T.N, nv = c->T.vals ? (int64_t)kSlots * c->nt * N : 0;
  const int64_t nb = c->T.blk ? (int64_t)sym_blk_count(c->nt, c->mesh.nf) : 0;
  double *vr = nullptr, *vb = nullptr, *u = nullptr, *res = nullptr, *zero = nullptr, *part = nullptr;
  int rc = HDG_OK;
  if ((!nv || (rc = dalloc(c, &vr, nv)) == HDG_OK) && (rc = dalloc(c, &u, N)) == HDG_OK && (rc = dalloc(c, &res, N)) == HDG_OK &&
      (rc = dalloc(c, &zero, N)) == HDG_OK && (!nb || (rc = dalloc(c, &vb, nb)) == HDG_OK))
    rc = dalloc(c, &part, refine_partials() + 1);
  // graphs hold the operator's pointers: none is replayed or kept across the swap
  drop_graphs(c);
  const TraceSystem saved = c->T;
  const hdg_options opt = c->opt;
  c->opt.schedule |= HDG_SCHEDULE_EAGER;
  if (inner_rtol > 0.0) c->opt.rtol = inner_rtol;
  auto step = [&]() -> int {
    HIPCHK(c, hipMemcpyAsync(u, c->T.uD, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(zero, 0, sizeof(double) * N, c->stream));
    if (nv) LAUNCHCHK(c, launch_round_float(c->T.vals, vr, nv, c->stream));
    if (nb) LAUNCHCHK(c, launch_round_float(c->T.blk, vb, nb, c->stream));
    double r0 = 0.0;
    for (int j = 0;; ++j) {
      c->T = saved;
      LAUNCHCHK(c, launch_refine_residual(c->pat, c->T, u, res, part, part + refine_partials(), c->stream));
      double r2 = 0.0;
      HIPCHK(c, hipMemcpyAsync(&r2, part + refine_partials(), sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (j == 0) r0 = std::sqrt(r2);
      relres[j] = r0 > 0.0 ? std::sqrt(r2) / r0 : 0.0;
      if (j == outer) return HDG_OK;
      // K32 d = r_j: rhs r_j, no Dirichlet lift, the rounded values
      if (nv) c->T.vals = vr;
      if (nb) c->T.blk = vb;
      c->T.rhs = res;
      c->T.uD = zero;
      LAUNCHCHK(c, launch_jacobi(c->pat, c->T, c->stream));
      c->jacobi_dirty = false;
      hdg_solve_info info{};
      const int e = solve_impl(c, &info);
      if (inner_iters) inner_iters[j] = info.iters;
      if (e != HDG_OK) return e;
      LAUNCHCHK(c, launch_add(u, c->d_uhat, N, c->stream));
    }
  };
  if (rc == HDG_OK) rc = step();
  // the assembled system as it was, u_outer as the trace
  c->T = saved;
  c->opt = opt;
  drop_graphs(c);
  int rc2 = launch_jacobi(c->pat, c->T, c->stream);
  if (rc == HDG_OK && u) {
    if (rc2 == 0) rc2 = (int)hipMemcpyAsync(c->d_uhat, u, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream);
    if (rc2 == 0) rc2 = (int)hipStreamSynchronize(c->stream);
    if (rc2 != 0) rc = fail(c, HDG_ERR_HIP, "restoring the trace system failed: %s", hipGetErrorString((hipError_t)rc2));
  } else {
    (void)hipStreamSynchronize(c->stream);
  }
  dfree(c, vr);
  dfree(c, vb);
  dfree(c, u);
  dfree(c, res);
  dfree(c, zero);
  dfree(c, part);
  c->jacobi_dirty = false;
  c->solved = rc == HDG_OK;
  c->recovered = false;
  return rc;
}

}  // extern "C"

// ------------------------------------------------------- serving arena
int hdg_arena_create(hdg_arena** out, int device, size_t bytes) {
  if (!out || bytes == 0) return HDG_ERR_ARG;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return HDG_ERR_HIP;
  hdg_arena* a = new hdg_arena();
  a->device = device;
  a->size = hdg_arena::round(bytes);
  if (hipMalloc((void**)&a->base, a->size) != hipSuccess) {
    delete a;
    return HDG_ERR_HIP;
  }
  a->free_[0] = a->size;
  *out = a;
  return HDG_OK;
}

void hdg_arena_destroy(hdg_arena* a) {
  if (!a) return;
  (void)hipSetDevice(a->device);
  (void)hipFree(a->base);
  delete a;
}

int hdg_attach_arena(hdg_ctx* c, hdg_arena* a) {
  if (!c || !a || a->device != c->device) return HDG_ERR_ARG;
  c->arena = a;
  return HDG_OK;
}
