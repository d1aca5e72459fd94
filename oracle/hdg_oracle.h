/*
 * hdg_oracle.h -- CPU restatement of the HDG Poisson pipeline of
 * HDiscontinuousGalerkin.jl (examples/poisson2D_HDG.jl:58-218 and the
 * src/ files it calls).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * library (libhdg) is compared against, and the timed CPU baseline.  It is
 * loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  Nothing in the product links or calls it.
 *
 * Conventions follow the reference's in-memory layouts:
 *   - nodes:  xy[2*n + c]                       (Vector{Node{2,Float64}})
 *   - cells:  cells[6*c + {0,1,2}] node ids, cells[6*c + {3,4,5}] face ids,
 *             all 1-based                       (Vector{Cell{2,3,3}})
 *   - faces:  faces[f + nf*col], col 0..3 = v1 v2 cell1 cell2(0 = none),
 *             1-based, column-major             (PolygonalMesh.faces)
 *   - dense element matrices are column-major (Julia Matrix{Float64}).
 */
#ifndef HDG_ORACLE_H
#define HDG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- reference-element numerics (src/basis.jl, quadrature) --- */
double orc_jacobi(double x, int p, double alpha, double beta);
double orc_dubiner_value(int j, double r, double s);
void orc_dubiner_grad(int j, double r, double s, double g[2]);
double orc_dubiner_basis(double x, double y, int j);
void orc_dubiner_basis_grad(double x, double y, int j, double g[2]);
double orc_legendre_value(int k, double x);
void orc_lagrange2d_p1(double r, double s, double val[3], double grad[6]);
void orc_lagrange1d_p1(double x, double val[2]);
int orc_strang(int order, double* pts, double* w);
int orc_gm(int s, double* pts, double* w);
int orc_gauss_legendre(int n, double* x, double* w);
int orc_quad2d(int order, double* pts, double* w);
int orc_quad1d(int order, double* x, double* w);

/* ---------------- mesh (src/generate_mesh.jl, src/mesh.jl) ---------------- */
int64_t orc_rect_nfaces(int64_t nx, int64_t ny);
int orc_rectangle_mesh(int64_t nx, int64_t ny, double x0, double y0,
                       double x1, double y1, double* xy, int64_t* cells,
                       int64_t* faces, int8_t* face_side);
void orc_jitter(int64_t nx, int64_t ny, double x0, double y0, double x1,
                double y1, double amp, uint64_t seed, double* xy);
int orc_face_orientation(const int64_t* cells, int64_t cell, int face);

/* ---------------- HDG pipeline (examples/poisson2D_HDG.jl) ----------------- */
typedef struct {
  int32_t k;          /* Dubiner order of q and u, Legendre order of the trace */
  int32_t qdeg;       /* quad_degree keyword (reference default k+1) */
  double tau;         /* stabilisation */
  int32_t src_kind;   /* 0: 2pi^2 sin sin; 1: -div(kappa grad sin sin);
                         2: 4x4 sine series; 3: values at quadrature points */
  int32_t kappa_on;   /* 1: kappa = 1 + 0.5 sin(a x) cos(b y), else 1 */
  double ka, kb;      /* a, b of kappa */
  double coef[16];    /* c_pq of the sine series, p-major */
  const double* f_qp; /* kind 3: ncells x nqp, column-major */
} orc_problem;

int orc_sizes(int k, int qdeg, int* nu, int* nt, int* nqp, int* nfq);
int orc_geometry(const double* xy, const int64_t* cells, int64_t cell,
                 double* detJ, double* Jinv, double* normals, double* detJf);
int orc_element_blocks(const orc_problem* P, const double* xy,
                       const int64_t* cells, int64_t ncells, int64_t cell,
                       double* A, double* B, double* C, double* E, double* F,
                       double* H, double* be);
int orc_condense(const orc_problem* P, const double* xy, const int64_t* cells,
                 int64_t ncells, double* Ate, double* bte, double* Ke,
                 double* be, int nthreads);
int orc_assemble(int64_t ncells, const int64_t* cells, int64_t nfaces, int nt,
                 const double* Ate, const double* bte, int64_t* colptr,
                 int64_t* rowval, double* nzval, double* rhs, int64_t* nnz);
int orc_dirichlet_trace(const orc_problem* P, const double* xy,
                        const int64_t* cells, const int64_t* faces,
                        int64_t nfaces, const int8_t* in_set, int g_kind,
                        const double* g_par, int quirk, int64_t* dofs,
                        double* vals, int64_t* ndofs);
int orc_apply(int64_t n, const int64_t* colptr, const int64_t* rowval,
              double* nzval, double* rhs, int64_t nd, const int64_t* dofs,
              const double* vals);
int orc_banded_solve(int64_t n, const int64_t* colptr, const int64_t* rowval,
                     const double* nzval, const double* b, double* x);
int orc_pcg_solve(int64_t n, const int64_t* colptr, const int64_t* rowval,
                  const double* nzval, const double* b, int nt,
                  const int64_t* dir_dofs, const double* dir_vals, int64_t nd,
                  double rtol, int maxit, double* x, int* iters,
                  double* relres, int nthreads);
int orc_recover(int64_t ncells, const int64_t* cells, int k, const double* Ke,
                const double* be, const double* uhat, double* q, double* u,
                double* uhat_cell);
double orc_errornorm(const orc_problem* P, const double* xy,
                     const int64_t* cells, int64_t ncells, const double* u,
                     int uex_kind, const double* uex_par);
int orc_residual(int64_t n, const int64_t* colptr, const int64_t* rowval,
                 const double* nzval, const double* x, const double* b,
                 double* r);

#ifdef __cplusplus
}
#endif
#endif
