/* Runs the oracle's whole pipeline on small jittered meshes, k = 1..4, for
 * the AddressSanitizer / UndefinedBehaviorSanitizer build made by
 * tests/test_sanitizers.py: rectangle mesh + jitter, quadrature and basis
 * tables, element blocks, condensation, assembly (count, then fill), both
 * Dirichlet modes, apply!, the banded direct solve, PCG, recovery, the
 * residual and the error norm.  Exits non-zero on a failed status, a
 * non-finite value or an error above its bound. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hdg_oracle.h"

#define CHECK(c)                                                \
  do {                                                          \
    if (!(c)) {                                                 \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      exit(1);                                                  \
    }                                                           \
  } while (0)

static void* xcalloc(size_t n, size_t s) {
  void* p = calloc(n ? n : 1, s);
  CHECK(p != NULL);
  return p;
}

static int finite_all(const double* a, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(a[i])) return 0;
  return 1;
}

static void run(int nx, int ny, int k, int qdeg, int quirk, int g_kind) {
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1), nc = 2LL * nx * ny, nf = orc_rect_nfaces(nx, ny);
  double* xy = xcalloc(2 * nn, sizeof(double));
  int64_t* cells = xcalloc(6 * nc, sizeof(int64_t));
  int64_t* faces = xcalloc(4 * nf, sizeof(int64_t));
  int8_t* side = xcalloc(nf, 1);
  CHECK(orc_rectangle_mesh(nx, ny, 0.0, 0.0, 1.0, 1.0, xy, cells, faces, side) == 0);
  orc_jitter(nx, ny, 0.0, 0.0, 1.0, 1.0, 0.2, 7u + (uint64_t)nx, xy);
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < 3; ++f) CHECK(abs(orc_face_orientation(cells, c, f)) <= 1);

  orc_problem P = {0};
  P.k = k;
  P.qdeg = qdeg;
  P.tau = 1.0;
  P.src_kind = 0;
  int nu, nt, nqp, nfq;
  CHECK(orc_sizes(k, qdeg, &nu, &nt, &nqp, &nfq) == 0);
  const int m = 3 * nu, ntr = 3 * nt;

  double* A = xcalloc(4 * (size_t)nu * nu, sizeof(double));
  double* B = xcalloc(2 * (size_t)nu * nu, sizeof(double));
  double* Cm = xcalloc((size_t)nu * nu, sizeof(double));
  double* E = xcalloc(2 * (size_t)nu * ntr, sizeof(double));
  double* F = xcalloc((size_t)nu * ntr, sizeof(double));
  double* H = xcalloc((size_t)ntr * ntr, sizeof(double));
  double* bel = xcalloc(nu, sizeof(double));
  CHECK(orc_element_blocks(&P, xy, cells, nc, nc / 2, A, B, Cm, E, F, H, bel) == 0);

  double* Ate = xcalloc(nc * ntr * ntr, sizeof(double));
  double* bte = xcalloc(nc * ntr, sizeof(double));
  double* Ke = xcalloc(nc * m * ntr, sizeof(double));
  double* be = xcalloc(nc * m, sizeof(double));
  CHECK(orc_condense(&P, xy, cells, nc, Ate, bte, Ke, be, 2) == 0);
  CHECK(finite_all(Ate, nc * ntr * ntr) && finite_all(Ke, nc * m * ntr));

  const int64_t n = nf * nt;
  int64_t nnz = 0;
  CHECK(orc_assemble(nc, cells, nf, nt, Ate, bte, NULL, NULL, NULL, NULL, &nnz) == 0);
  int64_t* colptr = xcalloc(n + 1, sizeof(int64_t));
  int64_t* rowval = xcalloc(nnz, sizeof(int64_t));
  double* nzval = xcalloc(nnz, sizeof(double));
  double* rhs = xcalloc(n, sizeof(double));
  CHECK(orc_assemble(nc, cells, nf, nt, Ate, bte, colptr, rowval, nzval, rhs, NULL) == 0);
  CHECK(colptr[n] == nnz + 1);

  int64_t nd = 0;
  int64_t* dofs = xcalloc(nf * nt, sizeof(int64_t));
  double* vals = xcalloc(nf * nt, sizeof(double));
  const double g[3] = {0.5, 1.0, -2.0};
  CHECK(orc_dirichlet_trace(&P, xy, cells, faces, nf, side, g_kind, g, quirk, dofs, vals, &nd) == 0);
  CHECK(nd > 0 && finite_all(vals, nd));

  /* PCG on the unmodified system with the Dirichlet values eliminated */
  double* x = xcalloc(n, sizeof(double));
  int it = 0;
  double rr = 0.0;
  CHECK(orc_pcg_solve(n, colptr, rowval, nzval, rhs, nt, dofs, vals, nd, 1e-12, 20000, x, &it, &rr, 2) >= 0);
  /* apply! then the banded direct solve */
  CHECK(orc_apply(n, colptr, rowval, nzval, rhs, nd, dofs, vals) == 0);
  double* xd = xcalloc(n, sizeof(double));
  CHECK(orc_banded_solve(n, colptr, rowval, nzval, rhs, xd) == 0);
  double* r = xcalloc(n, sizeof(double));
  orc_residual(n, colptr, rowval, nzval, xd, rhs, r);
  double rn = 0.0, bn = 0.0, dx = 0.0, xm = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    rn += r[i] * r[i];
    bn += rhs[i] * rhs[i];
    dx = fmax(dx, fabs(x[i] - xd[i]));
    xm = fmax(xm, fabs(xd[i]));
  }
  CHECK(sqrt(rn) <= 1e-9 * sqrt(bn) + 1e-300);
  if (!(dx <= 1e-7 * xm + 1e-300)) fprintf(stderr, "k=%d n=%dx%d quirk=%d g=%d it=%d rr=%g dx=%g xm=%g\n", k, nx, ny, quirk, g_kind, it, rr, dx, xm);
  CHECK(dx <= 1e-7 * xm + 1e-300);

  double* q = xcalloc((size_t)2 * nu * nc, sizeof(double));
  double* u = xcalloc((size_t)nu * nc, sizeof(double));
  double* uc = xcalloc((size_t)3 * nt * nc, sizeof(double));
  CHECK(orc_recover(nc, cells, k, Ke, be, xd, q, u, uc) == 0);
  const double e2 = orc_errornorm(&P, xy, cells, nc, u, 0, NULL);
  CHECK(isfinite(e2) && e2 >= 0.0);

  free(xy), free(cells), free(faces), free(side), free(A), free(B), free(Cm), free(E), free(F), free(H), free(bel);
  free(Ate), free(bte), free(Ke), free(be), free(colptr), free(rowval), free(nzval), free(rhs), free(dofs);
  free(vals), free(x), free(xd), free(r), free(q), free(u), free(uc);
}

int main(void) {
  double pts[3 * 64], w[64];
  for (int o = 1; o <= 8; ++o) (void)orc_quad2d(o, pts, w);
  for (int o = 1; o <= 12; ++o) (void)orc_quad1d(o, pts, w);
  for (int n = 1; n <= 12; ++n) CHECK(orc_gauss_legendre(n, pts, w) >= 0);
  for (int j = 0; j < 15; ++j) {
    double gr[2];
    (void)orc_dubiner_value(j, 0.2, 0.3);
    orc_dubiner_grad(j, 0.2, 0.3, gr);
  }
  const int qd[5] = {0, 2, 3, 6, 6};
  for (int k = 1; k <= 4; ++k) {
    run(5, 4, k, qd[k], 0, 1);
    run(3, 6, k, qd[k], 3, 1);
    run(1, 1, k, qd[k], 0, 0);
    run(16, 12, k, qd[k], 0, 1);
  }
  printf("oracle sanitized run: ok\n");
  return 0;
}
