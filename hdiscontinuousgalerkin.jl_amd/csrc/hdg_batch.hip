// Batches of independent Poisson solves in grouped launches (include/hdg.h,
// hdg_batch_*).  Requests with the same order and options form a group whose
// rectangle meshes are packed into one mesh: node, cell, face and trace-row
// ranges of instance j follow those of instance j-1.  Element build,
// condensation, assembly, recovery and the elementwise trace kernels run
// over the packed mesh once per group; the PCG kernels, the block-Jacobi
// smoother and the multigrid vertex levels take their instance from launch
// tables (Seg / VRec), so each instance keeps its own scalars, reductions
// and hierarchy and gets the bits it would get alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/hdg.h"
#include "hdg_arena.h"
#include "hdg_host.h"
#include "hdg_internal.h"
#include "hdg_tables.h"

using namespace hdg;

namespace {

struct Inst {
  int req = 0;                        // request index
  int64_t nx = 0, ny = 0, nn = 0, nc = 0, nf = 0, N = 0;
  int64_t node0 = 0, cell0 = 0, face0 = 0, row0 = 0;
  MgData plan;                        // level sizes and tail (host plan)
  int64_t lev_off[kMaxLevels] = {0};  // offset of its nodes in the group's level-l buffer
  int64_t tail_off = 0;               // offset of its LDS image in the group's tailpack
};

struct LevelBuf {
  int64_t n = 0;  // nodes of every instance at this level
  double *a = nullptr, *dinv = nullptr, *r = nullptr, *e = nullptr, *t = nullptr;
  unsigned char* mask = nullptr;
};

struct Group {
  hdg_options opt{};
  int kappa = 0;  // the instances' sources weight (q, v) with kappa (groups never mix)
  int k = 0, nt = 0, ntr = 0, m = 0;
  bool mg = false, mixed = false, eager = false;
  std::vector<Inst> in;
  int64_t nn = 0, nc = 0, nf = 0, N = 0;
  hipStream_t s = nullptr;
  std::vector<std::pair<void*, size_t>> owned;  // device buffers and their sizes
  size_t bytes = 0;
  // packed mesh and trace system
  DevMesh mesh;
  DevPattern pat;
  TraceSystem T;
  CgVectors cg{};
  RefTables tab;
  TabLayoutDev lay{};
  double* d_spd = nullptr;  // Schur-complement condensation tables (no kappa in the group), or null
  SpdLayoutDev spd{};
  double *d_tab = nullptr, *Ate = nullptr, *bte = nullptr, *Ke = nullptr, *be = nullptr, *qu = nullptr,
         *uhat = nullptr;
  unsigned char* dmask = nullptr;
  double* uD = nullptr;
  int* err = nullptr;  // [ninst] condensation error flags
  SrcParams* srcs = nullptr;
  int* cell_inst = nullptr;
  int *fv = nullptr, *vf_ptr = nullptr, *vf_idx = nullptr, *vf2 = nullptr, *fq4 = nullptr;  // face end points, vertex -> faces
  // launch tables
  CgState* st = nullptr;
  Seg *segs = nullptr, *csegs = nullptr;
  int *binst = nullptr, *cbinst = nullptr;
  int nbt = 0, nbc = 0;
  double* part2 = nullptr;  // error-norm partials
  double* res = nullptr;    // [ninst] ||K uhat - rhs||^2, [ninst] squared L2 error
  // multigrid
  MgData Mu;                // the packed mesh's view (level 1 spans every vertex)
  std::vector<MgData> Mi;   // per-instance hierarchies (views into the level buffers)
  MgData* dMi = nullptr;
  LevelBuf lev[kMaxLevels];
  MgGroup G;
  MgSetupGroup S;
  CondenseSched csched;  // condensation schedule (autotuned at create with HDG_SCHEDULE_AUTOTUNE)
  float *f_vals = nullptr, *f_blk = nullptr, *f_Dinv = nullptr, *f_sp = nullptr, *f_e0 = nullptr, *f_res0 = nullptr;
  // read-back
  CgState* h_st = nullptr;  // pinned, 2 x ninst (chunk slots) + ninst (final)
  double* h_res = nullptr;  // pinned, 2 x ninst, then ninst error flags (ints)
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int runs = 0;  // completed runs (the first runs eagerly, later ones replay graphs)
  double t_done = 0.0;
};

}  // namespace

struct hdg_batch {
  int device = 0;
  size_t budget = 0, used = 0;
  hdg_arena* arena = nullptr;  // serving loop's arena, or null
  std::vector<hdg_request> reqs;
  std::vector<Group> groups;
  std::vector<std::pair<int, int>> where;  // request -> (group, instance)
  bool ran = false;
  std::mutex mu;  // guards err (groups run in host threads of their own)
  std::string err;
};

namespace {

int fail(hdg_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (b) {
    std::lock_guard<std::mutex> lk(b->mu);
    b->err = buf;
  }
  return code;
}

#define BHIP(b, call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return fail(b, HDG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define BLAUNCH(b, call)                                                                     \
  do {                                                                                       \
    int e_ = (call);                                                                         \
    if (e_ != 0)                                                                             \
      return fail(b, HDG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString((hipError_t)e_)); \
  } while (0)

#define BTRY(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != HDG_OK) return r_; \
  } while (0)

template <class T>
int galloc(hdg_batch* b, Group& g, T** p, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (b->budget && b->used + bytes > b->budget)
    return fail(b, HDG_ERR_HBM_BUDGET, "batch needs more device memory than the budget (%zu of %zu bytes used)",
                b->used, b->budget);
  void* q = b->arena ? b->arena->take(bytes) : nullptr;
  if (!q) {
    hipError_t e = hipMallocAsync(&q, bytes, g.s);
    if (e != hipSuccess) return fail(b, HDG_ERR_HIP, "hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  g.owned.push_back({q, bytes});
  g.bytes += bytes;
  b->used += bytes;
  *p = static_cast<T*>(q);
  return HDG_OK;
}

// copy of a host array that may go out of scope on return (set-up only)
template <class T>
int upload(hdg_batch* b, Group& g, T** p, const std::vector<T>& h) {
  BTRY(galloc(b, g, p, h.size()));
  if (!h.empty()) {
    BHIP(b, hipMemcpyAsync(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, g.s));
    BHIP(b, hipStreamSynchronize(g.s));
  }
  return HDG_OK;
}

void free_group(hdg_batch* b, Group& g) {
  if (!g.s) return;
  (void)hipStreamSynchronize(g.s);  // queued work may still read the buffers and replay the graphs
  {
    std::lock_guard<std::mutex> lk(graph_mutex());
    for (auto& x : g.gexec)
      if (x) (void)hipGraphExecDestroy(x);
  }
  for (auto& [p, bytes] : g.owned) {
    if (b->arena && b->arena->owns(p))
      b->arena->give(p, bytes);
    else
      (void)hipFreeAsync(p, g.s);
  }
  (void)hipStreamSynchronize(g.s);
  b->used -= std::min(b->used, g.bytes);
  g.owned.clear();
  for (auto& e : g.ev)
    if (e) (void)hipEventDestroy(e);
  if (g.h_st) (void)hipHostFree(g.h_st);
  if (g.h_res) (void)hipHostFree(g.h_res);
  (void)hipStreamDestroy(g.s);
  g.s = nullptr;
}

bool same_group(const hdg_options& a, const hdg_options& o) {
  return a.k == o.k && a.quad_degree == o.quad_degree && a.tau == o.tau && a.precision == o.precision &&
         a.rtol == o.rtol && a.maxit == o.maxit && a.schedule == o.schedule && a.precond == o.precond;
}

SrcParams src_params(const hdg_source& s) {
  SrcParams p;
  p.kind = s.kind;
  p.kappa_on = s.kappa_on ? 1 : 0;
  p.ka = s.ka;
  p.kb = s.kb;
  for (int i = 0; i < 16; ++i) p.coef[i] = s.coef[i];
  return p;
}

CondenseArgs group_condense_args(const Group& g) {
  CondenseArgs a;
  a.mesh = g.mesh;
  a.tables = g.d_tab;
  a.lay = g.lay;
  a.srcs = g.srcs;
  a.cell_inst = g.cell_inst;
  a.tau = g.opt.tau;
  a.Ate = g.Ate;
  a.bte = g.bte;
  a.Ke = g.Ke;
  a.be = g.be;
  a.err = g.err;
  a.natural = g.tab.mass_spd() ? 1 : 0;
  if (g.d_spd) {
    a.spd_tables = g.d_spd;
    a.spd = g.spd;
  }
  return a;
}

int launch_group_condense(const Group& g, const CondenseArgs& a) {
  if (a.spd_tables) return launch_condense_spd(g.k, a, g.csched, g.s);
  return launch_condense(g.k, a, false, g.s, g.csched);
}

// With HDG_SCHEDULE_AUTOTUNE: the condensation schedule that runs fastest
// on the group's packed mesh (the arrays it writes are rewritten by every
// run).
int tune_group(hdg_batch* b, Group& g) {
  if (!(g.opt.schedule & HDG_SCHEDULE_AUTOTUNE)) return HDG_OK;
  const CondenseArgs a = group_condense_args(g);
  if (a.spd_tables)
    BLAUNCH(b, tune_condense_spd(g.k, a, g.s, &g.csched, nullptr, nullptr));
  else
    BLAUNCH(b, tune_condense(g.k, a, g.s, &g.csched, nullptr, nullptr));
  return HDG_OK;
}

// Lays out one group: offsets, levels, host packing, device arrays, tables.
int build_group(hdg_batch* b, Group& g) {
  const int n = (int)g.in.size();
  RefTables tab;
  if (!tab.build(g.opt.k, g.opt.quad_degree))
    return fail(b, HDG_ERR_QUAD, "Quadrature rule of order %d not available", g.opt.quad_degree);
  g.tab = tab;
  const TableLayout& L = tab.lay;
  g.lay = TabLayoutDev{L.nu, L.nt, L.nqp, L.nfq, L.Mref, L.Dref, L.Cref, L.G, L.Href, L.Nw, L.N, L.Mg, L.W, L.total};
  g.k = g.opt.k;
  g.nt = L.nt;
  g.ntr = 3 * L.nt;
  g.m = 3 * L.nu;
  // offsets of every instance in the packed mesh
  for (Inst& I : g.in) {
    I.nn = (I.nx + 1) * (I.ny + 1);
    I.nc = 2 * I.nx * I.ny;
    I.nf = hdg_rectangle_nfaces(I.nx, I.ny);
    I.N = I.nf * g.nt;
    I.node0 = g.nn;
    I.cell0 = g.nc;
    I.face0 = g.nf;
    I.row0 = g.N;
    g.nn += I.nn;
    g.nc += I.nc;
    g.nf += I.nf;
    g.N += I.N;
  }
  if (g.nn > INT32_MAX || 2 * g.nf > INT32_MAX || g.nc > INT32_MAX / 9)
    return fail(b, HDG_ERR_ARG, "group of order %d too large for 32-bit device indices", g.k);
  BHIP(b, hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
  BHIP(b, hipEventCreateWithFlags(&g.ev[0], hipEventDisableTiming));
  BHIP(b, hipEventCreateWithFlags(&g.ev[1], hipEventDisableTiming));
  BHIP(b, hipHostMalloc((void**)&g.h_st, 3 * (size_t)n * sizeof(CgState), hipHostMallocDefault));
  BHIP(b, hipHostMalloc((void**)&g.h_res, (3 * (size_t)n + 1) * sizeof(double), hipHostMallocDefault));

  // packed mesh, generated on the device instance by instance at its offsets
  double* dxy;
  int *dcn, *dcf, *dfc, *merr;
  BTRY(galloc(b, g, &dxy, 2 * g.nn));
  BTRY(galloc(b, g, &dcn, 3 * g.nc));
  BTRY(galloc(b, g, &dcf, 3 * g.nc));
  BTRY(galloc(b, g, &dfc, 2 * g.nf));
  BTRY(galloc(b, g, &g.cell_inst, g.nc));
  BTRY(galloc(b, g, &g.dmask, g.N));
  BTRY(galloc(b, g, &g.fv, 2 * g.nf));
  BTRY(galloc(b, g, &g.vf_ptr, g.nn + 1));
  BTRY(galloc(b, g, &g.vf_idx, 2 * g.nf));
  BTRY(galloc(b, g, &g.vf2, 2 * g.nf));
  if (g.mg) BTRY(galloc(b, g, &g.fq4, 4 * g.nf));
  BTRY(galloc(b, g, &merr, 1));
  BHIP(b, hipMemsetAsync(merr, 0, sizeof(int), g.s));
  for (int j = 0; j < n; ++j) {
    const Inst& I = g.in[j];
    const hdg_request& q = b->reqs[I.req];
    RectMeshArgs a;
    a.nx = I.nx;
    a.ny = I.ny;
    a.nf = I.nf;
    a.jitter = q.jitter;
    a.seed = q.seed;
    a.node0 = I.node0;
    a.cell0 = I.cell0;
    a.face0 = I.face0;
    a.inst = j;
    a.last = j + 1 == n;
    a.xy = dxy;
    a.cn = dcn;
    a.cf = dcf;
    a.fcell = dfc;
    a.fv = g.fv;
    a.vf_ptr = g.vf_ptr;
    a.vf_idx = g.vf_idx;
    a.fq4 = g.fq4;
    a.cell_inst = g.cell_inst;
    a.err = merr;
    BLAUNCH(b, launch_rect_mesh(a, g.s));
  }
  // u = 0 on every boundary face of every instance
  BLAUNCH(b, launch_boundary_dirichlet(dfc, 0, g.nf, g.nt, g.dmask, nullptr, 0.0, g.s));
  {
    int flag = 0;
    BHIP(b, hipMemcpyAsync(&flag, merr, sizeof(int), hipMemcpyDeviceToHost, g.s));
    BHIP(b, hipStreamSynchronize(g.s));
    if (flag) return fail(b, HDG_ERR_BOUNDARY, "a boundary face lies on no side of the box (order %d group)", g.k);
  }
  g.mesh = DevMesh{g.nn, g.nc, g.nf, dxy, dcn, dcf, dfc};
  BTRY(galloc(b, g, &g.pat.bcol, kSlots * g.nf));
  BTRY(galloc(b, g, &g.pat.src, 2 * kSlots * g.nf));
  BTRY(galloc(b, g, &g.pat.nslot, g.nf));
  BTRY(galloc(b, g, &g.pat.self, g.nf));
  BTRY(galloc(b, g, &g.pat.tslot, kSlots * g.nf));
  BTRY(galloc(b, g, &g.pat.nbr, 4 * g.nf));
  BLAUNCH(b, launch_pattern(g.mesh, g.pat, g.s));
  BTRY(galloc(b, g, &g.d_tab, L.total));
  BHIP(b, hipMemcpyAsync(g.d_tab, tab.data.data(), sizeof(double) * L.total, hipMemcpyHostToDevice, g.s));
  {
    std::vector<SrcParams> sp(n);
    for (int j = 0; j < n; ++j) sp[j] = src_params(b->reqs[g.in[j].req].src);
    BTRY(upload(b, g, &g.srcs, sp));
    // the Schur-complement condensation where the group has no kappa (as a
    // context on its own would condense each of its instances)
    SpdLayout sl;
    std::vector<double> sd;
    if (!g.kappa && g.k <= 3 && !(g.opt.schedule & HDG_SCHEDULE_LU) && build_spd_tables(tab, sl, sd)) {
      g.spd = SpdLayoutDev{sl.nu, sl.nt, sl.nqp, sl.R, sl.C, sl.Gam, sl.Z, sl.V, sl.P, sl.Om, sl.H, sl.Nw, sl.Mg,
                           sl.total};
      BTRY(upload(b, g, &g.d_spd, sd));
    }
  }
  const int64_t nc = g.nc, N = g.N;
  BTRY(galloc(b, g, &g.Ate, nc * g.ntr * g.ntr));
  BTRY(galloc(b, g, &g.bte, nc * g.ntr));
  BTRY(galloc(b, g, &g.Ke, nc * g.m * g.ntr));
  BTRY(galloc(b, g, &g.be, nc * g.m));
  BTRY(galloc(b, g, &g.qu, nc * g.m));
  BTRY(galloc(b, g, &g.uhat, N));
  BTRY(galloc(b, g, &g.uD, N));
  BTRY(galloc(b, g, &g.err, n));
  BHIP(b, hipMemsetAsync(g.uD, 0, sizeof(double) * N, g.s));
  g.T = TraceSystem{};
  g.T.nt = g.nt;
  g.T.nf = g.nf;
  g.T.N = N;
  if (sym_blocks(g.nt))
    BTRY(galloc(b, g, &g.T.blk, sym_blk_count(g.nt, g.nf)));
  else
    BTRY(galloc(b, g, &g.T.vals, (size_t)kSlots * g.nt * N));
  BTRY(galloc(b, g, &g.T.rhs, N));
  BTRY(galloc(b, g, &g.T.Dinv, dinv_count(g.nt, g.nf)));
  g.T.dmask = g.dmask;
  g.T.uD = g.uD;

  // CG segments: instance j gets the blocks it would get alone
  std::vector<Seg> segs(n), csegs(n);
  std::vector<int> binst, cbinst;
  BTRY(galloc(b, g, &g.st, n));
  BHIP(b, hipMemsetAsync(g.st, 0, sizeof(CgState) * n, g.s));
  for (int j = 0; j < n; ++j) {
    const Inst& I = g.in[j];
    const int nb = cg_blocks(I.N, g.nt), nbe = errornorm_blocks(I.nc);
    segs[j] = Seg{I.row0, I.row0 + I.N, (int)binst.size(), nb, g.st + j};
    csegs[j] = Seg{I.cell0, I.cell0 + I.nc, (int)cbinst.size(), nbe, g.st + j};
    binst.insert(binst.end(), nb, j);
    cbinst.insert(cbinst.end(), nbe, j);
  }
  g.nbt = (int)binst.size();
  g.nbc = (int)cbinst.size();
  BTRY(upload(b, g, &g.segs, segs));
  BTRY(upload(b, g, &g.csegs, csegs));
  BTRY(upload(b, g, &g.binst, binst));
  BTRY(upload(b, g, &g.cbinst, cbinst));
  BTRY(galloc(b, g, &g.part2, g.nbc));
  BTRY(galloc(b, g, &g.res, 2 * (size_t)n));
  CgVectors& v = g.cg;
  v.nblocks = g.nbt;
  v.st = g.st;
  BTRY(galloc(b, g, &v.x, N));
  BTRY(galloc(b, g, &v.r, N));
  BTRY(galloc(b, g, &v.z, N));
  BTRY(galloc(b, g, &v.p0, N));
  BTRY(galloc(b, g, &v.p1, N));
  BTRY(galloc(b, g, &v.q, N));
  BTRY(galloc(b, g, &v.part, 4 * (size_t)g.nbt + 64));
  if (!g.mg) {
    BTRY(tune_group(b, g));
    BHIP(b, hipStreamSynchronize(g.s));
    return HDG_OK;
  }

  // multigrid: packed level buffers, per-instance views, launch tables
  int maxlev = 1, ndepth = 1;
  for (Inst& I : g.in) {
    mg_plan_levels((int)I.nx, (int)I.ny, I.plan);
    if (I.plan.tail >= I.plan.nlev) return fail(b, HDG_ERR_UNSUPPORTED, "vertex hierarchy without a tail");
    maxlev = std::max(maxlev, I.plan.nlev);
    ndepth = std::max(ndepth, I.plan.tail);
  }
  for (int l = 1; l < maxlev; ++l) {
    int64_t off = 0;
    for (Inst& I : g.in)
      if (l < I.plan.nlev) {
        I.lev_off[l] = off;
        off += I.plan.lev[l].n;
      }
    LevelBuf& B = g.lev[l];
    B.n = off;
    BTRY(galloc(b, g, &B.a, kStencil * off));
    BTRY(galloc(b, g, &B.dinv, off));
    BTRY(galloc(b, g, &B.mask, off));
    BTRY(galloc(b, g, &B.r, off));
    BTRY(galloc(b, g, &B.e, off));
    BTRY(galloc(b, g, &B.t, off));
  }
  int64_t tailn = 0;
  for (Inst& I : g.in) {
    I.tail_off = tailn;
    for (int t = I.plan.tail; t < I.plan.nlev; ++t) tailn += (kStencil + 1) * I.plan.lev[t].n;
  }
  double *cellA, *cinv, *tailpack;
  BTRY(galloc(b, g, &cellA, 9 * nc));
  BTRY(galloc(b, g, &cinv, (size_t)n * kCoarseDirect * kCoarseDirect));
  BTRY(galloc(b, g, &tailpack, tailn));
  MgData& U = g.Mu;
  U.nlev = 2;
  U.nv = g.nn;
  U.fv = g.fv;
  U.vf_ptr = g.vf_ptr;
  U.vf_idx = g.vf_idx;
  U.vf2 = g.vf2;
  U.fq4 = g.fq4;
  U.cellA = cellA;
  {
    MgLevel& L1 = U.lev[1];
    const LevelBuf& B = g.lev[1];
    L1.n = g.nn;
    L1.a = B.a;
    L1.dinv = B.dinv;
    L1.mask = B.mask;
    L1.r = B.r;
    L1.e = B.e;
    L1.t = B.t;
  }
  BTRY(galloc(b, g, &U.e0, N));
  BTRY(galloc(b, g, &U.res0, N));
  BTRY(galloc(b, g, &U.fc, 2 * g.nf));
  BTRY(galloc(b, g, &U.sp, 4 * N));
  if (g.mixed) {
    if (g.T.blk)
      BTRY(galloc(b, g, &g.f_blk, sym_blk_count(g.nt, g.nf)));
    else
      BTRY(galloc(b, g, &g.f_vals, (size_t)kSlots * g.nt * N));
    BTRY(galloc(b, g, &g.f_Dinv, dinv_count(g.nt, g.nf)));
    BTRY(galloc(b, g, &g.f_sp, 4 * N));
    BTRY(galloc(b, g, &g.f_e0, N));
    BTRY(galloc(b, g, &g.f_res0, N));
  }
  MgData choice;
  mg_order_choices(g.k, choice);  // as hdg_setup chooses for a single problem
  g.Mi.assign(n, MgData{});
  for (int j = 0; j < n; ++j) {
    const Inst& I = g.in[j];
    MgData& M = g.Mi[j];
    M = I.plan;
    M.nv = I.nn;
    M.vsweeps = choice.vsweeps;
    M.cellA = cellA + 9 * I.cell0;
    M.cinv = cinv + (size_t)j * kCoarseDirect * kCoarseDirect;
    M.tailpack = tailpack + I.tail_off;
    for (int l = 1; l < M.nlev; ++l) {
      MgLevel& V = M.lev[l];
      const LevelBuf& B = g.lev[l];
      const int64_t o = I.lev_off[l];
      V.a = B.a + kStencil * o;
      V.dinv = B.dinv + o;
      V.mask = B.mask + o;
      V.r = B.r + o;
      V.e = B.e + o;
      V.t = B.t + o;
    }
    if (I.lev_off[1] != I.node0) return fail(b, HDG_ERR_HIP, "level-1 layout differs from the node numbering");
  }
  BTRY(upload(b, g, &g.dMi, g.Mi));
  MgGroup& G = g.G;
  G.ninst = n;
  G.dM = g.dMi;
  G.st = g.st;
  G.ndepth = ndepth;
  G.vsweeps = choice.vsweeps;
  G.omega_v = g.Mi[0].omega_v;
  for (int d = 1; d < ndepth; ++d) {
    std::vector<VRec> rec;
    std::vector<BlockMap> node, res;
    for (int j = 0; j < n; ++j) {
      const MgData& M = g.Mi[j];
      if (d >= M.tail) continue;
      const int p = (int)rec.size();
      rec.push_back(VRec{M.lev[d], M.lev[d + 1], g.st + j});
      const int64_t bn = (M.lev[d].n + 255) / 256, br = (8 * M.lev[d + 1].n + 255) / 256;
      for (int64_t u = 0; u < bn; ++u) node.push_back(BlockMap{p, (int)u});
      for (int64_t u = 0; u < br; ++u) res.push_back(BlockMap{p, (int)u});
    }
    VRec* drec;
    BlockMap *dnode, *dres;
    BTRY(upload(b, g, &drec, rec));
    BTRY(upload(b, g, &dnode, node));
    BTRY(upload(b, g, &dres, res));
    G.node[d] = VGroup{drec, dnode};
    G.res[d] = VGroup{drec, dres};
    G.gnode[d] = (int)node.size();
    G.gres[d] = (int)res.size();
  }
  BTRY(tune_group(b, g));
  // set-up tables: level-1 assembly, then cmask / RAP for every level pair
  MgSetupGroup& S = g.S;
  S.ninst = n;
  S.dM = g.dMi;
  {
    std::vector<BlockMap> l1;
    for (int j = 0; j < n; ++j)
      for (int64_t u = 0; u < (g.Mi[j].lev[1].n + 255) / 256; ++u) l1.push_back(BlockMap{j, (int)u});
    BlockMap* dl1;
    BTRY(upload(b, g, &dl1, l1));
    S.l1map = dl1;
    S.gl1 = (int)l1.size();
  }
  S.ndepth = maxlev - 1;
  for (int d = 1; d < S.ndepth; ++d) {
    std::vector<VRec> rec;
    std::vector<BlockMap> cm, rp;
    for (int j = 0; j < n; ++j) {
      const MgData& M = g.Mi[j];
      if (d + 1 >= M.nlev) continue;
      const int p = (int)rec.size();
      rec.push_back(VRec{M.lev[d], M.lev[d + 1], nullptr});
      const int64_t bc = (M.lev[d + 1].n + 255) / 256, br = (8 * M.lev[d + 1].n + 255) / 256;
      for (int64_t u = 0; u < bc; ++u) cm.push_back(BlockMap{p, (int)u});
      for (int64_t u = 0; u < br; ++u) rp.push_back(BlockMap{p, (int)u});
    }
    VRec* drec;
    BlockMap *dcm, *drp;
    BTRY(upload(b, g, &drec, rec));
    BTRY(upload(b, g, &dcm, cm));
    BTRY(upload(b, g, &drp, rp));
    S.cmask[d] = VGroup{drec, dcm};
    S.rap[d] = VGroup{drec, drp};
    S.gcmask[d] = (int)cm.size();
    S.grap[d] = (int)rp.size();
  }
  BHIP(b, hipStreamSynchronize(g.s));  // the host arrays go out of scope
  return HDG_OK;
}

Grouped rows(const Group& g) { return Grouped{g.segs, g.binst}; }
TraceOp<double> op64(const Group& g) {
  return TraceOp<double>{g.nt, g.nf, g.N, g.T.vals, g.T.Dinv, g.dmask, g.T.blk};
}
TraceOp<float> op32(const Group& g) {
  return TraceOp<float>{g.nt, g.nf, g.N, g.f_vals, g.f_Dinv, g.dmask, g.f_blk};
}
Krylov<double> kv(const Group& g) {
  const CgVectors& v = g.cg;
  return Krylov<double>{v.r, v.z, v.p0, v.p1, v.q, v.x, v.part, v.st, v.nblocks};
}
MgVecs<double> mv64(const Group& g) { return MgVecs<double>{g.Mu.e0, g.Mu.res0, g.Mu.sp}; }
MgVecs<float> mv32(const Group& g) { return MgVecs<float>{g.f_e0, g.f_res0, g.f_sp}; }

double rtol2_of(const Group& g) {
  const double rtol = g.opt.rtol > 0 ? g.opt.rtol : 1e-12;
  return rtol * rtol;
}
int maxit_of(const Group& g) { return g.opt.maxit > 0 ? g.opt.maxit : 200000; }

// condensation, assembly, preconditioner set-up, update(0)
int enqueue_setup(hdg_batch* b, Group& g) {
  BHIP(b, hipMemsetAsync(g.err, 0, sizeof(int) * g.in.size(), g.s));
  BLAUNCH(b, launch_group_condense(g, group_condense_args(g)));
  BLAUNCH(b, launch_fill(g.mesh, g.pat, g.Ate, g.bte, g.T, g.s));
  BLAUNCH(b, launch_jacobi(g.pat, g.T, g.s));
  const Grouped gr = rows(g);
  const int n = (int)g.in.size();
  if (!g.mg) {
    BLAUNCH(b, launch_cg_init_t(g.pat, g.T, g.cg, gr, g.nbt, g.s));
    return HDG_OK;
  }
  const double factor = g.Mi[0].omega_factor;
  if (g.mixed) {
    if (g.T.blk)
      BLAUNCH(b, launch_to_float(g.T.blk, g.f_blk, (int64_t)sym_blk_count(g.nt, g.nf), g.s));
    else
      BLAUNCH(b, launch_to_float(g.T.vals, g.f_vals, (int64_t)kSlots * g.nt * g.N, g.s));
    BLAUNCH(b, launch_to_float(g.T.Dinv, g.f_Dinv, (int64_t)dinv_count(g.nt, g.nf), g.s));
    BLAUNCH(b, mg_setup_mesh(g.mesh, g.Mu, g.T, g.pat, g.Ate, g.f_sp, g.s));
    BLAUNCH(b, mg_setup_levels_group(g.S, g.s));
    BLAUNCH(b, launch_lanczos_omega(g.pat, g.T, g.cg, gr, g.nbt, g.segs, n, factor, g.Mi[0].lanczos_steps, g.s));
    BLAUNCH(b, launch_mg_init_group(g.pat, g.T, op32(g), kv(g), mv32(g), gr, g.nbt, g.s));
    BLAUNCH(b, launch_mg_precond_group(g.pat, op32(g), kv(g), g.Mu, g.G, mv32(g), gr, g.nbt, g.s));
  } else {
    BLAUNCH(b, mg_setup_mesh(g.mesh, g.Mu, g.T, g.pat, g.Ate, nullptr, g.s));
    BLAUNCH(b, mg_setup_levels_group(g.S, g.s));
    // the CG vectors serve as scratch before the solve initialises them
    BLAUNCH(b, launch_lanczos_omega(g.pat, g.T, g.cg, gr, g.nbt, g.segs, n, factor, g.Mi[0].lanczos_steps, g.s));
    BLAUNCH(b, launch_mg_init_group(g.pat, g.T, op64(g), kv(g), mv64(g), gr, g.nbt, g.s));
    BLAUNCH(b, launch_mg_precond_group(g.pat, op64(g), kv(g), g.Mu, g.G, mv64(g), gr, g.nbt, g.s));
  }
  return HDG_OK;
}

// PCG iterations per chunk
int chunk_iters(const Group& g) { return g.mg ? 4 : 64; }

int enqueue_chunk(hdg_batch* b, Group& g) {
  const int iters = chunk_iters(g);
  const Grouped gr = rows(g);
  for (int it = 0; it < iters; ++it) {
    BLAUNCH(b, launch_cg_spmv_t(g.pat, op64(g), kv(g), gr, g.nbt, 0, rtol2_of(g), maxit_of(g), 0, g.s));
    if (g.mixed) {
      BLAUNCH(b, launch_mg_update_group(op32(g), kv(g), mv32(g), gr, g.nbt, g.s));
      BLAUNCH(b, launch_cg_check_group(kv(g), gr, (int)g.in.size(), rtol2_of(g), maxit_of(g), g.s));
      BLAUNCH(b, launch_mg_precond_group(g.pat, op32(g), kv(g), g.Mu, g.G, mv32(g), gr, g.nbt, g.s));
    } else if (g.mg) {
      BLAUNCH(b, launch_mg_update_group(op64(g), kv(g), mv64(g), gr, g.nbt, g.s));
      BLAUNCH(b, launch_cg_check_group(kv(g), gr, (int)g.in.size(), rtol2_of(g), maxit_of(g), g.s));
      BLAUNCH(b, launch_mg_precond_group(g.pat, op64(g), kv(g), g.Mu, g.G, mv64(g), gr, g.nbt, g.s));
    } else {
      BLAUNCH(b, launch_cg_update_t(g.T, g.cg, gr, g.nbt, 0, 0, g.s));
    }
  }
  return HDG_OK;
}

// One chunk of iterations for every instance of the group, then the states
// into h_st[slot]; captured once as a graph, from the second run on, unless
// HDG_SCHEDULE_EAGER (as for one problem, the first run is eager).
int launch_chunk(hdg_batch* b, Group& g, int slot) {
  const size_t n = g.in.size();
  CgState* dst = g.h_st + slot * n;
  const bool eager = g.eager || g.runs == 0;
  if (!eager && !g.gexec[slot]) {
    std::lock_guard<std::mutex> lk(graph_mutex());
    hipGraph_t gr = nullptr;
    BHIP(b, hipStreamBeginCapture(g.s, hipStreamCaptureModeThreadLocal));
    int r = enqueue_chunk(b, g);
    hipError_t e = hipMemcpyAsync(dst, g.st, sizeof(CgState) * n, hipMemcpyDeviceToHost, g.s);
    hipError_t e2 = hipStreamEndCapture(g.s, &gr);
    if (r != HDG_OK) return r;
    BHIP(b, e);
    BHIP(b, e2);
    e = hipGraphInstantiate(&g.gexec[slot], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    BHIP(b, e);
  }
  if (!eager) {
    std::lock_guard<std::mutex> lk(graph_mutex());
    BHIP(b, hipGraphLaunch(g.gexec[slot], g.s));
  } else {
    BTRY(enqueue_chunk(b, g));
    BHIP(b, hipMemcpyAsync(dst, g.st, sizeof(CgState) * n, hipMemcpyDeviceToHost, g.s));
  }
  BHIP(b, hipEventRecord(g.ev[slot], g.s));
  return HDG_OK;
}

bool all_done(const Group& g, int slot) {
  const size_t n = g.in.size();
  const CgState* s = g.h_st + slot * n;
  for (size_t j = 0; j < n; ++j)
    if (!s[j].done) return false;
  return true;
}

// uhat, FP64 residuals, recovery, error norms, and the read-back of states,
// residuals and the error flag
int enqueue_finish(hdg_batch* b, Group& g) {
  const int n = (int)g.in.size();
  const Grouped gr = rows(g);
  BLAUNCH(b, launch_cg_finish(g.T, g.cg, g.uhat, g.s));
  BLAUNCH(b, launch_true_residual_t(g.pat, g.T, g.uhat, g.cg.part, gr, g.nbt, g.s));
  BLAUNCH(b, launch_reduce_segs(g.cg.part, g.segs, n, g.res, g.s));
  // a group on the Schur form holds only the u rows of K_e and b_e
  if (g.d_spd)
    BLAUNCH(b, launch_recover_spd(g.k, g.mesh, g.d_spd, g.Ke, g.be, g.uhat, g.qu, g.s));
  else
    BLAUNCH(b, launch_recover(g.k, g.mesh, g.Ke, g.be, g.uhat, g.qu, g.s));
  bool want = false;
  for (const Inst& I : g.in) want = want || b->reqs[I.req].uex_kind >= 0;
  if (want) {
    const double p3[3] = {0.0, 0.0, 0.0};
    BLAUNCH(b, launch_errornorm_t(g.mesh, g.d_tab, g.lay, g.k, g.qu, HDG_UEX_SINSIN, p3, g.part2,
                                  Grouped{g.csegs, g.cbinst}, g.nbc, g.s));
    BLAUNCH(b, launch_reduce_segs(g.part2, g.csegs, n, g.res + n, g.s));
  } else {
    BHIP(b, hipMemsetAsync(g.res + n, 0, sizeof(double) * n, g.s));
  }
  BHIP(b, hipMemcpyAsync(g.h_st + 2 * n, g.st, sizeof(CgState) * n, hipMemcpyDeviceToHost, g.s));
  BHIP(b, hipMemcpyAsync(g.h_res, g.res, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, g.s));
  BHIP(b, hipMemcpyAsync(g.h_res + 2 * n, g.err, sizeof(int) * n, hipMemcpyDeviceToHost, g.s));
  return HDG_OK;
}

}  // namespace

extern "C" {

int hdg_batch_create(hdg_batch** out, int device, size_t hbm_budget, int n, const hdg_request* reqs) {
  return hdg_batch_create_in_arena(out, device, hbm_budget, n, reqs, nullptr);
}

}  // extern "C"

int hdg_batch_create_in_arena(hdg_batch** out, int device, size_t hbm_budget, int n, const hdg_request* reqs,
                              hdg_arena* arena) {
  if (!out) return HDG_ERR_ARG;
  *out = nullptr;
  if (n < 1 || !reqs) return HDG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HDG_ERR_HIP;
  if (device < 0 || device >= ndev) return HDG_ERR_ARG;
  if (arena && arena->device != device) return HDG_ERR_ARG;
  hdg_batch* b = new hdg_batch();
  b->device = device;
  b->budget = hbm_budget;
  b->arena = arena;
  b->reqs.assign(reqs, reqs + n);
  b->where.assign(n, {0, 0});
  auto bail = [&](int rc) {
    *out = b;  // keeps the message readable through hdg_batch_last_error
    return rc;
  };
  for (int i = 0; i < n; ++i) {
    hdg_request& q = b->reqs[i];
    if (q.nx < 1 || q.ny < 1) return bail(fail(b, HDG_ERR_ARG, "request %d: nx, ny must be >= 1", i));
    if (q.opt.k < 1 || q.opt.k > 4) return bail(fail(b, HDG_ERR_UNSUPPORTED, "request %d: k=%d (1..4)", i, q.opt.k));
    if (q.opt.quad_degree <= 0) q.opt.quad_degree = q.opt.k + 1;
    if (q.opt.precision != HDG_FP64 && q.opt.precision != HDG_MIXED)
      return bail(fail(b, HDG_ERR_ARG, "request %d: unknown precision %d", i, q.opt.precision));
    if (q.opt.precond < HDG_PC_AUTO || q.opt.precond > HDG_PC_MULTIGRID)
      return bail(fail(b, HDG_ERR_ARG, "request %d: unknown preconditioner %d", i, q.opt.precond));
    if (q.opt.precision == HDG_MIXED && q.opt.precond == HDG_PC_BLOCK_JACOBI)
      return bail(fail(b, HDG_ERR_UNSUPPORTED, "request %d: mixed precision runs on the multigrid path", i));
    if (q.src.kind < 0 || q.src.kind > 2)
      return bail(fail(b, HDG_ERR_UNSUPPORTED, "request %d: source kind %d (batches take analytic sources 0..2)",
                       i, q.src.kind));
    if (q.uex_kind != -1 && q.uex_kind != HDG_UEX_SINSIN)
      return bail(fail(b, HDG_ERR_ARG, "request %d: batches report HDG_UEX_SINSIN error norms only", i));
    // groups never mix sources with and without kappa: without it a group
    // may condense through the Schur complement, as one context would
    const int kap = q.src.kappa_on ? 1 : 0;
    int gi = -1;
    for (size_t t = 0; t < b->groups.size(); ++t)
      if (same_group(b->groups[t].opt, q.opt) && b->groups[t].kappa == kap) gi = (int)t;
    if (gi < 0) {
      b->groups.emplace_back();
      Group& g = b->groups.back();
      g.opt = q.opt;
      g.kappa = kap;
      g.mg = q.opt.precond != HDG_PC_BLOCK_JACOBI;
      g.mixed = g.mg && q.opt.precision == HDG_MIXED;
      g.eager = (q.opt.schedule & HDG_SCHEDULE_EAGER) != 0;
      gi = (int)b->groups.size() - 1;
    }
    Inst I;
    I.req = i;
    I.nx = q.nx;
    I.ny = q.ny;
    b->where[i] = {gi, (int)b->groups[gi].in.size()};
    b->groups[gi].in.push_back(I);
  }
  if (hipSetDevice(device) != hipSuccess) return bail(fail(b, HDG_ERR_HIP, "hipSetDevice failed"));
  {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  }
  for (Group& g : b->groups) {
    const int rc = build_group(b, g);
    if (rc != HDG_OK) return bail(rc);
  }
  *out = b;
  return HDG_OK;
}

namespace {

// One group from condensation to its results on the host, on the group's
// stream: one chunk in flight while the previous chunk's states are read.
int run_group(hdg_batch* b, Group& g, std::chrono::steady_clock::time_point t0) {
  if (hipSetDevice(b->device) != hipSuccess) return fail(b, HDG_ERR_HIP, "hipSetDevice failed");
  BTRY(enqueue_setup(b, g));
  // as for one problem: nothing is queued behind a chunk in which every
  // instance is forecast to converge (chunk_converges)
  const size_t n = g.in.size();
  std::vector<CgState> prev(n), last(n);
  int nread = 0;
  bool behind = false;
  int slot = 0;
  BTRY(launch_chunk(b, g, slot));
  for (;;) {
    const int nxt = slot ^ 1;
    behind = nread < 2;
    for (size_t j = 0; j < n && !behind; ++j)
      behind = !chunk_converges(prev[j], last[j], chunk_iters(g) + (g.mg ? 1 : 0), rtol2_of(g));
    if (behind) BTRY(launch_chunk(b, g, nxt));
    BHIP(b, hipEventSynchronize(g.ev[slot]));
    prev.swap(last);
    std::copy(g.h_st + slot * n, g.h_st + (slot + 1) * n, last.begin());
    ++nread;
    if (all_done(g, slot)) break;
    if (!behind) {  // the forecast missed
      BTRY(launch_chunk(b, g, nxt));
      behind = true;
    }
    slot = nxt;
  }
  if (behind) BHIP(b, hipEventSynchronize(g.ev[slot ^ 1]));
  BTRY(enqueue_finish(b, g));
  BHIP(b, hipStreamSynchronize(g.s));
  g.t_done = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  ++g.runs;
  return HDG_OK;
}

}  // namespace

extern "C" {

int hdg_batch_run(hdg_batch* b, hdg_result* res) {
  if (!b || !res) return HDG_ERR_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return fail(b, HDG_ERR_HIP, "hipSetDevice failed");
  for (Group& g : b->groups)
    if (!g.s) return fail(b, HDG_ERR_STATE, "batch was not built");
  const auto t0 = std::chrono::steady_clock::now();
  // groups run side by side, each driven by a host thread of its own
  std::vector<int> rc(b->groups.size(), HDG_OK);
  {
    std::vector<std::thread> th;
    for (size_t gi = 0; gi < b->groups.size(); ++gi)
      th.emplace_back([&, gi] { rc[gi] = run_group(b, b->groups[gi], t0); });
    for (auto& t : th) t.join();
  }
  for (int r : rc)
    if (r != HDG_OK) return r;
  int first_bad = HDG_OK;
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    Group& g = b->groups[gi];
    const int n = (int)g.in.size();
    const int* flags = reinterpret_cast<const int*>(g.h_res + 2 * n);
    for (int j = 0; j < n; ++j) {
      const Inst& I = g.in[j];
      const int gerr = (flags[j] & 1) ? HDG_ERR_NEG_JACOBIAN : ((flags[j] & 2) ? HDG_ERR_LOCAL_SINGULAR : HDG_OK);
      const CgState& s = g.h_st[2 * n + j];
      hdg_result& r = res[I.req];
      std::memset(&r, 0, sizeof(r));
      r.id = I.req;
      r.iters = s.it;
      r.relres = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
      r.true_relres = s.bb > 0 ? std::sqrt(g.h_res[j] / s.bb) : 0.0;
      r.errornorm = b->reqs[I.req].uex_kind >= 0 ? g.h_res[n + j] : 0.0;
      r.n_cells = I.nc;
      r.n_dofs = I.N;
      r.t_submit = 0.0;
      r.t_start = 0.0;
      r.t_done = g.t_done;
      r.device_bytes = (size_t)((double)g.bytes * I.nc / std::max<int64_t>(g.nc, 1));
      r.slot = (int)gi;
      r.status = gerr != HDG_OK ? gerr : (s.done == 1 ? HDG_OK : HDG_ERR_NOT_CONVERGED);
      if (r.status != HDG_OK && first_bad == HDG_OK) {
        first_bad = r.status;
        fail(b, r.status, "request %d: %s", I.req,
             gerr == HDG_ERR_NEG_JACOBIAN   ? "det(J) is not positive for a cell"
             : gerr == HDG_ERR_LOCAL_SINGULAR ? "singular local (q,u) matrix; raise quad_degree"
             : s.done == 2                    ? "trace solve reached maxit"
                                              : "trace solve broke down");
      }
    }
  }
  b->ran = true;
  return first_bad;
}

int hdg_batch_trace(hdg_batch* b, int i, double* uhat, int64_t cap) {
  if (!b || !uhat || i < 0 || i >= (int)b->reqs.size()) return HDG_ERR_ARG;
  if (!b->ran) return fail(b, HDG_ERR_STATE, "hdg_batch_run has not been called");
  if (hipSetDevice(b->device) != hipSuccess) return fail(b, HDG_ERR_HIP, "hipSetDevice failed");
  Group& g = b->groups[b->where[i].first];
  const Inst& I = g.in[b->where[i].second];
  if (cap < I.N)
    return fail(b, HDG_ERR_ARG, "the trace of instance %d has %lld doubles, more than the buffer's %lld", i,
                (long long)I.N, (long long)cap);
  BHIP(b, hipMemcpyAsync(uhat, g.uhat + I.row0, sizeof(double) * I.N, hipMemcpyDeviceToHost, g.s));
  BHIP(b, hipStreamSynchronize(g.s));
  return HDG_OK;
}

int hdg_batch_info(const hdg_batch* b, int* ngroups, size_t* device_bytes) {
  if (!b) return HDG_ERR_ARG;
  if (ngroups) *ngroups = (int)b->groups.size();
  if (device_bytes) *device_bytes = b->used;
  return HDG_OK;
}

void hdg_batch_destroy(hdg_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  for (Group& g : b->groups) free_group(b, g);
  delete b;
}

const char* hdg_batch_last_error(const hdg_batch* b) { return b ? b->err.c_str() : "null batch"; }

}  // extern "C"
