/*
 * hdg.h -- C ABI of libhdg, the MI355X HDG Poisson pipeline.
 *
 * The library replaces the hot path of HDiscontinuousGalerkin.jl's HDG
 * driver: the per-cell loop of `doassemble` (examples/poisson2D_HDG.jl:58-186),
 * `apply!` + `K \ b` (:189-195, src/boundary.jl:121-175), `get_uσ!`
 * (:197-212) and `errornorm` (src/DiscreteFunctions.jl:97-120).  A foreign
 * caller (Julia `ccall`, Python `ctypes`) binds these symbols directly: all
 * arguments are plain C scalars, pointers and sizes; no C++ types cross the
 * boundary and nothing allocated by the library is returned to the caller.
 *
 * Layouts are the reference's in-memory layouts, so a Julia caller passes
 * `pointer(mesh.nodes)` etc. without copies:
 *   nodes  xy[2*n + c]                       == Vector{Node{2,Float64}}
 *   cells  cells[6*c + 0..2] node ids, cells[6*c + 3..5] face ids, 1-based
 *                                            == Vector{Cell{2,3,3}}
 *   faces  faces[f + nf*j], j = v1 v2 cell1 cell2 (0 = none), 1-based,
 *          column-major                      == PolygonalMesh.faces
 *   q, u   ncells x nq / ncells x nu column-major == TrialFunction.m_values
 *
 * Ownership: the caller owns every host buffer; the library reads or writes
 * them only during a call.  The context owns all device memory until
 * hdg_destroy.  Calls on one context are serialised by the caller; distinct
 * contexts may be used from distinct host threads.  Every entry point
 * returns an hdg_status (0 = HDG_OK); hdg_last_error gives the message.
 */
#ifndef HDG_H
#define HDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HDG_OK = 0,
  HDG_ERR_ARG = 1,            /* ArgumentError: bad size, order, pointer */
  HDG_ERR_NEG_JACOBIAN = 2,   /* det(J) <= 0, src/ScalarFunctionSpaces.jl:110 */
  HDG_ERR_QUAD = 3,           /* quadrature order unavailable, src/quadrature.jl:24 */
  HDG_ERR_LOCAL_SINGULAR = 4, /* local (q,u) matrix singular (the reference
                                 returns garbage silently, e.g. k=3 with the
                                 default quad_degree) */
  HDG_ERR_NOT_CONVERGED = 5,  /* trace solve hit maxit */
  HDG_ERR_HBM_BUDGET = 6,     /* device allocation beyond the context budget */
  HDG_ERR_HIP = 7,            /* HIP runtime failure */
  HDG_ERR_STATE = 8,          /* entry point called out of order */
  HDG_ERR_UNSUPPORTED = 9,    /* e.g. k outside 1..4 */
  HDG_ERR_BOUNDARY = 10       /* face on one cell but not on the box,
                                 src/generate_mesh.jl:80 */
} hdg_status;

typedef struct hdg_ctx hdg_ctx;

/* Precision modes of the trace solve.  HDG_MIXED keeps condensation, the
   trace matrix, the CG recurrences and the iterate in FP64 and runs the
   multigrid preconditioner's trace-level data in float (structured meshes
   only: HDG_ERR_UNSUPPORTED otherwise). */
enum { HDG_FP64 = 0, HDG_MIXED = 1 };

/* Preconditioners of the trace solve.  HDG_PC_AUTO picks multigrid when the
   mesh is a structured rectangle_mesh and block-Jacobi otherwise. */
enum { HDG_PC_AUTO = 0, HDG_PC_BLOCK_JACOBI = 1, HDG_PC_MULTIGRID = 2 };

/* Schedule flags.  By default the first solve after hdg_setup launches its
   kernels from the host, and from the second solve on each chunk of CG
   iterations is captured once as a HIP graph and replayed (capture costs
   more than one solve saves); HDG_SCHEDULE_EAGER never captures.  Same
   results, bit for bit.  Where kappa = 1 and the cell rule leaves the
   reference mass SPD (k = 1 at quad_degree 2, k = 3 at 6, k <= 2 at 4),
   condensation eliminates the q block analytically (A = detJ diag(Mref,
   Mref)) and factors the Schur complement on u; HDG_SCHEDULE_LU eliminates
   the full local matrix as the reference's factorize does instead (same
   operators to rounding, about cond(Me) eps). */
enum { HDG_SCHEDULE_EAGER = 1, HDG_SCHEDULE_AUTOTUNE = 2, HDG_SCHEDULE_LU = 4 };

/* Sources.  HDG_SRC_REFERENCE is the reference's f = 2 pi^2 sin(pi x) sin(pi y)
   (examples/poisson2D_HDG.jl:55).  HDG_SRC_VALUES takes f already evaluated
   at the cell quadrature points (for arbitrary Julia closures). */
enum {
  HDG_SRC_REFERENCE = 0,
  HDG_SRC_KAPPA_MANUFACTURED = 1, /* f = -div(kappa grad(sin(pi x) sin(pi y))) */
  HDG_SRC_SINE_SERIES = 2,        /* f = sum c_pq sin(p pi x) sin(q pi y), p,q=1..4 */
  HDG_SRC_VALUES = 3              /* f_qp: ncells x nqp, column-major */
};

/* Exact solutions known to hdg_errornorm. */
enum { HDG_UEX_SINSIN = 0, HDG_UEX_LINEAR = 1 };

typedef struct {
  int32_t k;           /* Dubiner order of q, u and Legendre order of the trace, 1..4 */
  int32_t quad_degree; /* quad_degree keyword; 0 selects the reference default k+1 */
  double tau;          /* stabilisation (poisson2D_HDG.jl:58) */
  int32_t precision;   /* HDG_FP64 or HDG_MIXED */
  double rtol;         /* trace solve: stop at ||r|| <= rtol ||b|| (0 -> 1e-12) */
  int32_t maxit;       /* trace solve iteration cap (0 -> 200000) */
  int32_t schedule;    /* HDG_SCHEDULE_* flags, 0 = default */
  int32_t precond;     /* HDG_PC_*: block-Jacobi per face, or a V-cycle of
                          block-Jacobi on the traces + geometric multigrid on
                          the P1 vertex space of a structured mesh */
} hdg_options;

typedef struct {
  int32_t kind;        /* HDG_SRC_* */
  int32_t kappa_on;    /* 1: kappa = 1 + 0.5 sin(ka x) cos(kb y) weights (q,v) */
  double ka, kb;
  double coef[16];     /* sine-series coefficients c_pq, p-major */
  const double* f_qp;  /* HDG_SRC_VALUES only; read during hdg_set_source */
} hdg_source;

typedef struct {
  int64_t n_nodes, n_cells, n_faces;
  int64_t n_dofs;      /* trace dofs = n_faces * nt */
  int64_t nnz;         /* stored entries of K (sparse(I,J,V) semantics) */
  int32_t nu, nq, nt, m, ntr, nqp, nfq;
  int32_t sym_blocks;  /* 1: the trace product reads K as symmetric face blocks
                          (nt = 4: K_ff and one owned block per cell and face,
                          3 nt^2 values per face); 0: the face-ELL (nnz values) */
} hdg_sizes;

typedef struct {
  int32_t iters;       /* PCG iterations */
  int32_t refinements; /* 0: the FP64 CG recurrence itself carries the accuracy */
  double relres;       /* final recursive ||r|| / ||b|| */
  double true_relres;  /* ||b - S x|| / ||b|| recomputed in FP64 */
  double seconds;      /* device time of the solve */
} hdg_solve_info;

/* ---- context ---------------------------------------------------------- */
/* A context on HIP device `device`; hbm_budget bytes caps its device
   memory (0 = no cap). */
int hdg_create(hdg_ctx** ctx, int device, size_t hbm_budget);
void hdg_destroy(hdg_ctx* ctx);
const char* hdg_last_error(const hdg_ctx* ctx);
/* Bytes of device memory the context holds now. */
int hdg_memory_used(const hdg_ctx* ctx, size_t* bytes);

/* ---- mesh: replaces rectangle_mesh (src/generate_mesh.jl:96-143) ------- */
/* Face count of rectangle_mesh(TriangleCell, (nx, ny), ...). */
int64_t hdg_rectangle_nfaces(int64_t nx, int64_t ny);
/* rectangle_mesh(TriangleCell, (nx, ny), LL, UR) into caller buffers
   xy[2*(nx+1)*(ny+1)], cells[6*2*nx*ny], faces[4*nf] (column-major), and
   face_side[nf] (0 interior, 1 bottom, 2 right, 3 top, 4 left: the
   facesets of generate_mesh.jl:60-89).  jitter > 0 moves interior nodes by
   up to jitter*h per axis after generation (splitmix64 stream `seed`). */
int hdg_rectangle_mesh(int64_t nx, int64_t ny, double x0, double y0, double x1,
                       double y1, double jitter, uint64_t seed, double* xy,
                       int64_t* cells, int64_t* faces, int8_t* face_side);

/* hdg_rectangle_mesh + hdg_set_mesh in one step, generated on the device:
   the same mesh, numbering and node coordinates bit for bit (so the same
   solution), without host arrays or their upload. */
int hdg_set_rectangle_mesh(hdg_ctx* ctx, int64_t nx, int64_t ny, double x0, double y0, double x1, double y1,
                           double jitter, uint64_t seed);

/* ---- problem set-up ----------------------------------------------------- */
/* Uploads a PolygonalMesh (src/mesh.jl:43-49) and builds the face-level
   sparsity pattern of the trace matrix on the device. */
int hdg_set_mesh(hdg_ctx* ctx, int64_t n_nodes, const double* xy,
                 int64_t n_cells, const int64_t* cells, int64_t n_faces,
                 const int64_t* faces);
/* ScalarFunctionSpace / VectorFunctionSpace / ScalarTraceFunctionSpace
   construction (src/ScalarFunctionSpaces.jl:24-29, VectorFunctionSpaces.jl:
   10-18, TraceFunctionSpaces.jl:11-28): reference tables for (k, quad
   degree) go to the device. */
int hdg_setup(hdg_ctx* ctx, const hdg_options* opt);
int hdg_set_source(hdg_ctx* ctx, const hdg_source* src);
/* Dirichlet(u_hat, mesh, faceset, g) (src/boundary.jl:11-42): prescribed
   trace dofs (1-based) and values.  n = 0 clears. */
int hdg_set_dirichlet(hdg_ctx* ctx, int64_t n, const int64_t* dofs,
                      const double* vals);
/* u = value on every trace dof of every boundary face (a face with one
   cell), free elsewhere; the device-side form of hdg_set_dirichlet for a
   constant on the whole boundary. */
int hdg_set_dirichlet_boundary(hdg_ctx* ctx, double value);
/* Dirichlet(u_hat, mesh, "boundary", g) (src/boundary.jl:11-42) computed on
   the device for g given in closed form: every trace dof of every face with
   one cell is prescribed (ascending by face, mode), the others are freed.
   g_kind HDG_G_LINEAR: g = par[0] + par[1] x + par[2] y (a constant with
   par[1] = par[2] = 0); HDG_G_SINSIN: g = sin(pi x) sin(pi y) (par unused).
   flags 0: the L2 projection onto the orthonormal trace basis;
   HDG_DIRICHLET_REFERENCE: the reference's computation, which weights face
   point q with the cell rule's weight q (fs.fs.qr_weights, boundary.jl:33)
   and does not reset its accumulator between modes (:27-37); both give 0
   for g = 0 and differ otherwise (HDG_ERR_ARG when the cell rule has fewer
   points than the face rule, where the reference's loop leaves its
   weights). */
enum { HDG_G_LINEAR = 0, HDG_G_SINSIN = 1 };
enum { HDG_DIRICHLET_REFERENCE = 1 };
int hdg_set_dirichlet_function(hdg_ctx* ctx, int g_kind, const double* par, int flags);
int hdg_query_sizes(const hdg_ctx* ctx, hdg_sizes* sizes);

/* ---- the pipeline -------------------------------------------------------- */
/* doassemble (poisson2D_HDG.jl:58-186) on the device: element blocks, static
   condensation, trace matrix K and rhs. */
int hdg_assemble_condense(hdg_ctx* ctx);
/* end_assemble's SparseMatrixCSC{Float64,Int} of K (1-based colptr[n+1],
   rowval[nnz], nzval[nnz]) and the assembled rhs[n]; any pointer may be
   NULL. */
int hdg_get_system(hdg_ctx* ctx, int64_t* colptr, int64_t* rowval,
                   double* nzval, double* rhs);
/* apply!(K, b, dbc) with APPLY_TRANSPOSE (src/boundary.jl:121-175) on the
   device, exported as hdg_get_system exports K and b: the rhs lifted by the
   prescribed values (b_i -= v_d K(i, d) over the prescribed dofs d in
   ascending order, skipping v_d = 0), prescribed rows and columns zeroed
   (stored entries stay stored), meandiag(K) = sum_i |K(i,i)| / n
   (:169-175, summed in a fixed blocked order) on their diagonal and
   b_d = v_d meandiag.  The device's own trace system is left unchanged;
   any pointer may be NULL; meandiag_out receives meandiag(K) if not NULL. */
int hdg_get_applied_system(hdg_ctx* ctx, int64_t* colptr, int64_t* rowval, double* nzval, double* rhs,
                           double* meandiag_out);
/* K_element / b_element of doassemble: Ke[c] m x ntr column-major, stored
   cell after cell; be[c] of length m.  Either pointer may be NULL. */
int hdg_get_local_ops(hdg_ctx* ctx, double* Ke, double* be);
/* Condensed element blocks Ate[c] (ntr x ntr column-major) and bte[c]. */
int hdg_get_condensed(hdg_ctx* ctx, double* Ate, double* bte);
/* Element blocks A, B, C, E, F, H, be of cells [first, first+count) as
   built inside doassemble (poisson2D_HDG.jl:77-153), each column-major per
   cell, cell after cell.  Any pointer may be NULL. */
int hdg_get_element_blocks(hdg_ctx* ctx, int64_t first, int64_t count,
                           double* A, double* B, double* C, double* E,
                           double* F, double* H, double* be);
/* apply!(K, b, dbc) + K \ b (poisson2D_HDG.jl:194-195): the trace solve with
   Dirichlet dofs eliminated.  uhat_out[n_dofs] may be NULL (the solution
   stays on the device for hdg_recover). */
int hdg_solve(hdg_ctx* ctx, double* uhat_out, hdg_solve_info* info);
/* Replaces the device trace solution by a caller's uhat[n_dofs] (e.g. the
   result of the caller's own K \ b), so that hdg_recover applies get_uσ! to
   it.  Requires hdg_assemble_condense. */
int hdg_set_trace(hdg_ctx* ctx, const double* uhat);
/* The trace solution û of the last hdg_solve / hdg_run / hdg_set_trace
   (the `K \ b` of poisson2D_HDG.jl:195), n_dofs doubles. */
int hdg_get_trace(hdg_ctx* ctx, double* uhat);
/* get_uσ! (poisson2D_HDG.jl:197-212): q (ncells x nq), u (ncells x nu),
   uhat_cell (ncells x nt x 3), all column-major; any may be NULL. */
int hdg_recover(hdg_ctx* ctx, double* q, double* u, double* uhat_cell);
/* errornorm(u_h, u_ex) (DiscreteFunctions.jl:97-120): squared L2 error of
   the recovered u against HDG_UEX_SINSIN or HDG_UEX_LINEAR (par c0,c1,c2). */
int hdg_errornorm(hdg_ctx* ctx, int uex_kind, const double* uex_par,
                  double* sq_l2);
/* Fused assemble + condense + solve + recover with everything kept on the
   device; the timed unit of the benchmark. */
int hdg_run(hdg_ctx* ctx, hdg_solve_info* info);

/* hdg_errornorm for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops); no counterpart in the reference.
   Allocates nothing: the kernels' block partial sums (at most 2048 doubles)
   and the reference tables of k = 1..4 at quad degrees 1..6 live in static
   device arrays of the library, one copy per GPU, set up at the first call
   on the current device (that call copies the tables and synchronises; later
   calls only enqueue two kernels on `stream`).  Inputs are device arrays:
   xy (2 doubles per node), cells (3 node ids per cell, int32, 0-based) and
   qu (the recovered [q; u] of each cell, 3 nu doubles per cell, cell after
   cell); the squared L2 error of u against HDG_UEX_SINSIN or HDG_UEX_LINEAR
   (uex_par c0, c1, c2, read on the host) goes to sq_l2, one double in device
   memory, asynchronously on stream (a hipStream_t; NULL = the default
   stream).  HDG_ERR_ARG for n_cells outside 1..HDG_NOALLOC_MAX_CELLS or a
   null pointer, HDG_ERR_UNSUPPORTED for k outside 1..4, HDG_ERR_QUAD for a
   quad degree above 6 (0 = k+1).  Every call on one GPU shares that GPU's
   partial sums, which each call writes in full before reading them: calls
   that overlap on one GPU race. */
#define HDG_NOALLOC_MAX_CELLS 2147483647LL
int hdg_errornorm_noalloc(int k, int quad_degree, int64_t n_cells, const double* xy, const int32_t* cells,
                          const double* qu, int uex_kind, const double* uex_par, double* sq_l2, void* stream);

/* ---- reference-element numerics (host) ----------------------------------- */
/* value / gradient of the orthonormal Dubiner basis j = 1..nu at (r, s)
   (src/basis.jl:65-123): val[nu], grad[2*nu]. */
int hdg_dubiner(int k, double r, double s, double* val, double* grad);
/* QuadratureRule{2,RefTetrahedron}(DefaultQuad(), order) (quadrature.jl:
   17-26) and the 1-D rule on [0,1] (:29-39).  Returns the point count, or
   -HDG_ERR_QUAD; pts holds 2 (or 1) coordinates per point; cap = capacity
   in points. */
int hdg_quadrature(int dim, int order, int cap, double* pts, double* w);

/* ---- autotuning ----------------------------------------------------------- */
/* No counterpart in the reference.  The condensation takes one of two forms:
   HDG_CONDENSE_ELIMINATION, the elimination of the bordered local system
   with one element per segment of 16, 32 or 64 lanes (any width that holds
   the element's rows), or, where it applies (see HDG_SCHEDULE_LU),
   HDG_CONDENSE_SCHUR, one element per 16 lanes with a register budget cut
   for 2 or 3 waves per SIMD.  Both stride a grid of blocks over the cells.
   hdg_autotune times every (width or register budget, grid cap) pair of the
   form the current problem takes, with HIP events, and keeps the fastest for
   this context until the next hdg_setup (HDG_SCHEDULE_AUTOTUNE in
   hdg_options.schedule does the same at the first assembly after hdg_setup,
   and for every group of a batch at hdg_batch_create).  Results do not
   depend on the schedule, bit for bit.  Needs hdg_setup and hdg_set_source;
   overwrites the assembled operators (assemble again before solving). */
enum { HDG_CONDENSE_ELIMINATION = 0, HDG_CONDENSE_SCHUR = 1 };
typedef struct {
  int32_t condense_seg;   /* lanes per element (16 for the Schur form) */
  int32_t condense_grid;  /* grid cap, blocks of 256 threads */
  double condense_us;     /* measured duration of the chosen schedule */
  double default_us;      /* of the default schedule: the narrowest width and
                             8192 blocks, or (Schur form) 2 waves per SIMD
                             and 16384 blocks */
  int32_t condense_form;  /* out: HDG_CONDENSE_* of the problem tuned */
  int32_t condense_waves; /* Schur form: 2 or 3 waves per SIMD (0 = 2) */
} hdg_schedule;
int hdg_autotune(hdg_ctx* ctx, hdg_schedule* out);
/* Sets the condensation schedule until the next hdg_setup, e.g. one tuned
   before: condense_seg 0 = narrowest width (the Schur form ignores it),
   condense_grid 0 = the form's default cap, condense_waves 0, 2 or 3 (the
   elimination form ignores it); condense_form is not read.  HDG_ERR_ARG for
   a width not 16/32/64, a width below the element's rows where the problem
   takes the elimination form (where it takes the Schur form such a width,
   e.g. the 16 hdg_autotune reports, leaves the elimination form its
   narrowest), or another waves value. */
int hdg_set_schedule(hdg_ctx* ctx, const hdg_schedule* sched);

/* ---- measurement --------------------------------------------------------- */
/* Average device duration (microseconds) of one launch of an internal
   kernel, timed with HIP events around `reps` back-to-back launches on the
   context stream with the data currently resident.  which: 0 = trace SpMV
   (the dominant PCG kernel), 1 = PCG update, 2 = condensation, 3 = a plain
   copy of the trace-matrix values (5 nt^2 doubles per face), the reference
   point for hardware traffic counters.  Timing runs overwrite the solver
   state; solve again before reading results. */
int hdg_time_kernel(hdg_ctx* ctx, int which, int reps, double* avg_us);

/* No counterpart in the reference: a measurement of what an FP32 operator
   costs in accuracy.  FP64 iterative refinement of the assembled system
   K uhat = rhs (Dirichlet rows fixed) whose corrections solve with K rounded
   entry by entry to float, the operator an FP32 condensation and assembly
   would produce at best: from u_0 = the Dirichlet values, step j forms the
   FP64 residual r_j = rhs - K u_j on the free rows, solves K32 d = r_j with
   the context's preconditioned CG (FP64 arithmetic, stopping at
   ||r|| <= inner_rtol ||r_j||; 0 keeps the context's rtol) and sets
   u_{j+1} = u_j + d.  relres[j] = ||r_j|| / ||r_0|| for j = 0..outer
   (outer + 1 values, outer <= 64); inner_iters[j] (may be null) the CG
   iterations of step j.  Leaves u_outer as the trace (hdg_recover and
   hdg_errornorm then measure it) and the assembled system as it was.
   Allocates a second copy of the trace-matrix values while it runs.
   HDG_ERR_NOT_CONVERGED if an inner solve fails (entries after it are
   NaN). */
int hdg_refine_probe(hdg_ctx* ctx, int outer, double inner_rtol, double* relres, int32_t* inner_iters);

/* ---- serving: a request loop over one device ---------------------------- */
/* No counterpart in the reference (it solves one problem per script run).
   A server owns `slots` worker threads, each with its own context and HIP
   stream, so solves in different slots overlap on the device.  A request is
   admitted when its estimated device footprint fits the HBM budget next to
   the footprints the other slots hold (continuous admission: a slot takes
   the first of the next few queued requests that fits as soon as it is
   free).  A slot keeps its mesh and set-up while consecutive requests share
   them, so repeated meshes stay device-resident.  Slots launch every kernel
   from the host (HDG_SCHEDULE_EAGER is set on every submitted request).  Results are bitwise the
   same whatever else runs in the other slots (no atomics, fixed reduction
   order). */
typedef struct hdg_server hdg_server;

typedef struct {
  int64_t nx, ny;      /* rectangle_mesh(TriangleCell, (nx, ny), (0,0), (1,1)) */
  double jitter;       /* interior nodes moved by up to jitter*h (0: none) */
  uint64_t seed;       /* jitter stream */
  hdg_options opt;     /* k, quad_degree, tau, precision, rtol, maxit, precond */
  hdg_source src;      /* f (and kappa); f_qp must stay valid until completion */
  int32_t uex_kind;    /* -1: no error norm; else HDG_UEX_* for hdg_errornorm */
  int32_t want_trace;  /* 1: keep a host copy of the trace solution for hdg_server_trace */
} hdg_request;

typedef struct {
  int64_t id;
  int32_t status;      /* hdg_status of the solve */
  int32_t iters;
  double relres, true_relres;
  double errornorm;    /* squared L2 error (uex_kind >= 0), else 0 */
  int64_t n_cells, n_dofs;
  double t_submit, t_start, t_done;  /* seconds since hdg_server_create */
  size_t device_bytes; /* device memory the slot held for it */
  int32_t slot;
} hdg_result;

typedef struct {
  int64_t completed, failed;
  int64_t cells;             /* cells of the completed requests */
  double p50_latency, p99_latency, mean_latency;  /* seconds, submit -> done */
  double mean_queue_wait;    /* seconds, submit -> start */
  double busy_seconds;       /* first start -> last done */
  double cells_per_second;   /* cells / busy_seconds */
  size_t peak_reserved;      /* largest sum of admitted footprints */
} hdg_server_summary;

/* hbm_budget 0 = no cap; slots >= 1.  With a cap, the server allocates
   hbm_budget bytes of device memory at creation, as one piece held until
   hdg_server_destroy, and the slots' contexts sub-allocate from it (if that
   allocation fails they use the device's pool). */
int hdg_server_create(hdg_server** srv, int device, size_t hbm_budget, int slots);
/* Continuous batching: a free slot that takes a request also takes up to
   max_requests - 1 more queued requests with the same options (order,
   quadrature, tau, precision, tolerances, preconditioner) while their cells
   stay within max_cells (0 = no cap) and their footprints within the
   budget, and solves them together in grouped launches (hdg_batch_*).
   Results are bitwise those of solving each request alone.  Requests a
   batch does not accept (HDG_SRC_VALUES, HDG_UEX_LINEAR) run alone.
   max_requests = 1 (the default) turns batching off.  Call before
   submitting. */
int hdg_server_set_batching(hdg_server* srv, int max_requests, int64_t max_cells);
/* Queues a request (copied); returns at once with its id. */
int hdg_server_submit(hdg_server* srv, const hdg_request* req, int64_t* id);
/* Blocks until request `id` completes (timeout_s < 0: no timeout), fills
   `out` and forgets the request unless it kept a trace.  HDG_ERR_STATE on
   an unknown id, HDG_ERR_NOT_CONVERGED on timeout. */
int hdg_server_wait(hdg_server* srv, int64_t id, double timeout_s, hdg_result* out);
/* Copies the kept trace solution of a completed request (its result's
   n_dofs doubles) into uhat, a buffer of cap doubles, and releases it;
   HDG_ERR_ARG (the trace kept) when cap is smaller. */
int hdg_server_trace(hdg_server* srv, int64_t id, double* uhat, int64_t cap);
/* Device bytes a request would hold (mesh, per-cell operators, trace system,
   solver and multigrid data). */
size_t hdg_server_footprint(const hdg_request* req);
int hdg_server_stats(hdg_server* srv, hdg_server_summary* out);
/* Finishes the queued requests, joins the slots, frees everything. */
void hdg_server_destroy(hdg_server* srv);
const char* hdg_server_last_error(const hdg_server* srv);

/* ---- batches: many instances in grouped launches ------------------------- */
/* No counterpart in the reference (one problem per script run).  A batch
   holds n independent requests (same fields as the serving loop's; u = 0 on
   the boundary, f and kappa from src, HDG_SRC_VALUES not accepted).
   Requests of the same order and options (and with or without kappa) form
   a group: their rectangle
   meshes are packed into one mesh (a disjoint union), so element build,
   condensation, assembly, recovery and every elementwise kernel of the
   solve run once per group, and the PCG scalars, reductions and multigrid
   vertex levels of the instances are selected per block from launch tables.
   One launch per kernel serves the whole group; groups of different orders
   run side by side on their own streams.  Each instance's results are
   bitwise those of solving it alone with hdg_run (same reductions over the
   same block partition).  Meshes, operators and solutions stay resident
   between runs. */
typedef struct hdg_batch hdg_batch;

/* Builds and uploads the packed meshes of every group and allocates every
   device array (hbm_budget 0 = no cap; HDG_ERR_HBM_BUDGET when the batch
   does not fit).  reqs is copied.  On failure *out may still receive the
   batch, so that hdg_batch_last_error explains it; destroy it then. */
int hdg_batch_create(hdg_batch** out, int device, size_t hbm_budget, int n, const hdg_request* reqs);
/* Element build + condensation + assembly + trace solve + recovery (and the
   error norm where uex_kind >= 0) of every instance; res[i] for request i
   (id = i, slot = its group, times in seconds since the start of this
   call).  Returns HDG_OK when every instance converged, else the first
   failing status (each res[i].status holds its own). */
int hdg_batch_run(hdg_batch* b, hdg_result* res);
/* Trace solution of request i after hdg_batch_run (res[i].n_dofs doubles)
   into uhat, a buffer of cap doubles; HDG_ERR_ARG when cap is smaller. */
int hdg_batch_trace(hdg_batch* b, int i, double* uhat, int64_t cap);
/* Number of groups and device bytes held. */
int hdg_batch_info(const hdg_batch* b, int* ngroups, size_t* device_bytes);
void hdg_batch_destroy(hdg_batch* b);
const char* hdg_batch_last_error(const hdg_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* HDG_H */
