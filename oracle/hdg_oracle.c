/*
 * hdg_oracle.c -- CPU restatement of the HDG Poisson pipeline of
 * HDiscontinuousGalerkin.jl, in plain C (C11 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for libhdg and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  The product never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.  Pins (checks
 * that would fail on a plausible mistake) live in tests/test_oracle_*.py:
 *   - reference literals: test/test_FunctionSpace.jl:39-72 (A, B, C, E, H,
 *     face measures), test/test_ScalarFuncSp.jl:15-33 (Jinv, detJ, detJf,
 *     normals), test/test_mesh.jl:11-44 (connectivity, orientation, sets),
 *     test/test_quadrature.jl:6-25, test/test_basis.jl:7-93, and the squared
 *     L2 bounds test/test_FunctionSpace.jl:243, examples/poisson2D_HDG.jl:218;
 *   - closed forms: A = detJ I and H = |e| I under exact quadrature, patch
 *     tests (constant and linear traces reproduced exactly), symmetry and
 *     definiteness of the trace operator, nnz and face-count formulas,
 *     convergence rate k+1;
 *   - textbook reductions: Gauss-Legendre vs numpy.polynomial.legendre,
 *     dense LU vs numpy.linalg.solve, banded vs dense global solves;
 *   - the sine-series source by its manufactured solution
 *     u = sum c_pq sin(p pi x) sin(q pi y) / (pi^2 (p^2 + q^2)) at rate k+1;
 *   - the Dirichlet values: the projection reproduces linear g exactly, and
 *     the reference's own computation (cell-rule weights at face points, an
 *     accumulator never reset, src/boundary.jl:27-37) against closed forms
 *     of those weights and sums.
 * The Lagrange-P1 geometry basis via an inverted Dubiner Vandermonde
 * (src/basis.jl:264-299) and LAPACK's exact pivot order are restated up to
 * rounding; they agree with exact arithmetic to ~1e-16, which the pins above
 * bound.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  FMA
 * contraction is off so that mesh coordinates are computed with the same
 * roundings as the reference's Julia code (generate_mesh.jl:1-18).
 */
#include "hdg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979311600 /* Float64(pi) */
#define MAXQ 64
#define MAXNU 64
#define MAXNT 16

/* ========================================================================= */
/* Reference-element numerics                                                */
/* ========================================================================= */

/* src/basis.jl:132-150 -- Jacobi polynomial by the three-term recursion. */
double orc_jacobi(double x, int p, double al, double be) {
  double a = 1.0;
  double b = ((2.0 + al + be) * x + al - be) / 2.0;
  if (p <= 0) return a;
  if (p == 1) return b;
  for (int n = 2; n <= p; ++n) {
    double a1 = 2.0 * n * (n + al + be) * (2.0 * n - 2.0 + al + be);
    double a2 = (2.0 * n - 1.0 + al + be) * (al + be) * (al - be);
    double a3 = (2.0 * n - 2.0 + al + be) * (2.0 * n - 1.0 + al + be) *
                (2.0 * n + al + be);
    double a4 = 2.0 * (n - 1.0 + al) * (n - 1.0 + be) * (2.0 * n + al + be);
    double nb = ((a2 + a3 * x) * b - a4 * a) / a1;
    a = b;
    b = nb;
  }
  return b;
}

/* src/basis.jl:157-162 */
static double djacobi(double x, int n, double al, double be) {
  if (n <= 0) return 0.0;
  return (n + al + be + 1.0) / 2.0 * orc_jacobi(x, n - 1, al + 1.0, be + 1.0);
}

/* src/basis.jl:169-181 */
static double dubiner_nm(double x, double y, int n, int m) {
  double xi = fabs(x) < 2.220446049250313e-16 ? -1.0 : 2.0 * x / (1.0 - y) - 1.0;
  double eta = 2.0 * y - 1.0;
  int k = 2 * n + 1;
  double P = orc_jacobi(xi, n, 0.0, 0.0) * orc_jacobi(eta, m, (double)k, 0.0) *
             pow((1.0 - eta) / 2.0, n);
  double N = sqrt(2.0 / ((2.0 * n + 1.0) * (m + n + 1.0)));
  return (2.0 * P) / N;
}

/* src/basis.jl:188-204 */
static void grad_dubiner_nm(double x, double y, int n, int m, double g[2]) {
  double xi = fabs(x) < 2.220446049250313e-16 ? -1.0 : 2.0 * x / (1.0 - y) - 1.0;
  int k = 2 * n + 1;
  double eta = 2.0 * y - 1.0;
  double Dpn = djacobi(xi, n, 0.0, 0.0);
  double pn = orc_jacobi(xi, n, 0.0, 0.0);
  double pm = orc_jacobi(eta, m, (double)k, 0.0);
  double Dpm = djacobi(eta, m, (double)k, 0.0);
  double Px = 2.0 / (1.0 - eta) * Dpn * pm * pow((1.0 - eta) / 2.0, n);
  double N = sqrt(2.0 / ((2.0 * n + 1.0) * (m + n + 1.0)));
  double Py = (2.0 * x / ((1.0 - y) * (1.0 - y)) * Dpn * pow((1.0 - eta) / 2.0, n) -
               n * pow((1.0 - eta) / 2.0, n - 1) * pn) * pm +
              2.0 * pn * pow((1.0 - eta) / 2.0, n) * Dpm;
  g[0] = (4.0 * Px) / N;
  g[1] = (2.0 * Py) / N;
}

/* src/basis.jl:211-218 and 225-232: index j -> degrees (n, m). */
static void dubiner_degrees(int j, int* n, int* m) {
  double t = -1.5 + 0.5 * sqrt(1.0 + 8.0 * j);
  double ct = ceil(t);
  *n = (int)(((ct + 1.0) * (ct + 2.0)) / 2.0 - j);
  *m = (int)(ct - *n);
}

double orc_dubiner_basis(double x, double y, int j) {
  int n, m;
  dubiner_degrees(j, &n, &m);
  return dubiner_nm(x, y, n, m);
}

void orc_dubiner_basis_grad(double x, double y, int j, double g[2]) {
  int n, m;
  dubiner_degrees(j, &n, &m);
  grad_dubiner_nm(x, y, n, m, g);
}

/* src/basis.jl:65-86 -- orthonormal Dubiner basis, closed forms j <= 15. */
double orc_dubiner_value(int j, double r, double s) {
  const double s2 = sqrt(2.0), s3 = sqrt(3.0), s6 = sqrt(6.0), s10 = sqrt(10.0);
  const double s14 = sqrt(14.0), s30 = sqrt(30.0), s70 = sqrt(70.0);
  switch (j) {
    case 0: return 0.0;
    case 1: return s2;
    case 2: return 2.0 * s3 * (2.0 * r + s - 1.0);
    case 3: return 2.0 * (3.0 * s - 1.0);
    case 4: return s30 * (6.0 * r * r + 6.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
    case 5: return 3.0 * s2 * (5.0 * s - 1.0) * (2.0 * r + s - 1.0);
    case 6: return s6 * (10.0 * s * s - 8.0 * s + 1.0);
    case 7:
      return 2.0 * s14 * (2.0 * r + s - 1.0) *
             (10.0 * r * r + 10.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
    case 8:
      return 2.0 * s10 * (7.0 * s - 1.0) *
             (6.0 * r * r + 6.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
    case 9: return 2.0 * s6 * (21.0 * s * s - 12.0 * s + 1.0) * (2.0 * r + s - 1.0);
    case 10: return 2.0 * s2 * (35.0 * s * s * s - 45.0 * s * s + 15.0 * s - 1.0);
    case 11:
      return 3.0 * s10 *
             (70.0 * pow(r, 4) + 140.0 * r * r * r * (s - 1.0) +
              90.0 * r * r * (s * s - 2.0 * s + 1.0) +
              20.0 * r * (s * s * s - 3.0 * s * s + 3.0 * s - 1.0) + pow(s, 4) -
              4.0 * s * s * s + 6.0 * s * s - 4.0 * s + 1.0);
    case 12:
      return s70 * (9.0 * s - 1.0) * (2.0 * r + s - 1.0) *
             (10.0 * r * r + 10.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
    case 13:
      return 5.0 * s2 * (36.0 * s * s - 16.0 * s + 1.0) *
             (6.0 * r * r + 6.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
    case 14:
      return s30 * (84.0 * s * s * s - 84.0 * s * s + 21.0 * s - 1.0) *
             (2.0 * r + s - 1.0);
    case 15:
      return s10 * (126.0 * pow(s, 4) - 224.0 * s * s * s + 126.0 * s * s - 24.0 * s + 1.0);
    default: return orc_dubiner_basis(r, s, j);
  }
}

/* src/basis.jl:93-123 -- gradients of the closed forms. */
void orc_dubiner_grad(int j, double r, double s, double g[2]) {
  const double s2 = sqrt(2.0), s3 = sqrt(3.0), s6 = sqrt(6.0), s10 = sqrt(10.0);
  const double s14 = sqrt(14.0), s30 = sqrt(30.0), s70 = sqrt(70.0);
  switch (j) {
    case 0:
    case 1: g[0] = 0.0; g[1] = 0.0; return;
    case 2: g[0] = 4.0 * s3; g[1] = 2.0 * s3; return;
    case 3: g[0] = 0.0; g[1] = 6.0; return;
    case 4:
      g[0] = 12.0 * s30 * r + 6.0 * s30 * (s - 1.0);
      g[1] = 6.0 * s30 * r + 2.0 * s30 * s - 2.0 * s30;
      return;
    case 5:
      g[0] = 6.0 * s2 * (5.0 * s - 1.0);
      g[1] = 30.0 * s2 * r + 30.0 * s2 * s - 18.0 * s2;
      return;
    case 6: g[0] = 0.0; g[1] = 20.0 * s6 * s - 8.0 * s6; return;
    case 7:
      g[0] = 24.0 * s14 * (5.0 * r * r + 5.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
      g[1] = 60.0 * s14 * r * r + r * (48.0 * s14 * s - 48.0 * s14) +
             6.0 * s14 * s * s - 12.0 * s14 * s + 6.0 * s14;
      return;
    case 8:
      g[0] = 12.0 * s10 * (7.0 * s - 1.0) * (2.0 * r + s - 1.0);
      g[1] = 84.0 * s10 * r * r + r * (168.0 * s10 * s - 96.0 * s10) +
             42.0 * s10 * s * s - 60.0 * s10 * s + 18.0 * s10;
      return;
    case 9:
      g[0] = 4.0 * s6 * (21.0 * s * s - 12.0 * s + 1.0);
      g[1] = r * (168.0 * s6 * s - 48.0 * s6) + 126.0 * s6 * s * s -
             132.0 * s6 * s + 26.0 * s6;
      return;
    case 10: g[0] = 0.0; g[1] = 210.0 * s2 * s * s - 180.0 * s2 * s + 30.0 * s2; return;
    case 11:
      g[0] = 840.0 * s10 * r * r * r + 1260.0 * s10 * r * r * (s - 1.0) +
             540.0 * s10 * r * (s * s - 2.0 * s + 1.0) +
             60.0 * s10 * (s * s * s - 3.0 * s * s + 3.0 * s - 1.0);
      g[1] = 420.0 * s10 * r * r * r + r * r * (540.0 * s10 * s - 540.0 * s10) +
             r * (180.0 * s10 * s * s - 360.0 * s10 * s + 180.0 * s10) +
             12.0 * s10 * s * s * s - 36.0 * s10 * s * s + 36.0 * s10 * s - 12.0 * s10;
      return;
    case 12:
      g[0] = 12.0 * s70 * (9.0 * s - 1.0) * (5.0 * r * r + 5.0 * r * (s - 1.0) + s * s - 2.0 * s + 1.0);
      g[1] = 180.0 * s70 * r * r * r + r * r * (540.0 * s70 * s - 300.0 * s70) +
             r * (324.0 * s70 * s * s - 456.0 * s70 * s + 132.0 * s70) +
             36.0 * s70 * s * s * s - 84.0 * s70 * s * s + 60.0 * s70 * s - 12.0 * s70;
      return;
    case 13:
      g[0] = 30.0 * s2 * (36.0 * s * s - 16.0 * s + 1.0) * (2.0 * r + s - 1.0);
      g[1] = r * r * (2160.0 * s2 * s - 480.0 * s2) +
             r * (3240.0 * s2 * s * s - 3120.0 * s2 * s + 510.0 * s2) +
             720.0 * s2 * s * s * s - 1320.0 * s2 * s * s + 690.0 * s2 * s - 90.0 * s2;
      return;
    case 14:
      g[0] = 2.0 * s30 * (84.0 * s * s * s - 84.0 * s * s + 21.0 * s - 1.0);
      g[1] = r * (504.0 * s30 * s * s - 336.0 * s30 * s + 42.0 * s30) +
             336.0 * s30 * s * s * s - 504.0 * s30 * s * s + 210.0 * s30 * s - 22.0 * s30;
      return;
    case 15:
      g[0] = 0.0;
      g[1] = 504.0 * s10 * s * s * s - 672.0 * s10 * s * s + 252.0 * s10 * s - 24.0 * s10;
      return;
    default: orc_dubiner_basis_grad(r, s, j, g); return;
  }
}

/* src/basis.jl:351-354 -- orthonormal Legendre on [0,1], k 1-based. */
double orc_legendre_value(int k, double x) {
  return sqrt((double)(2 * (k - 1) + 1)) * orc_jacobi(2.0 * x - 1.0, k - 1, 0.0, 0.0);
}

/* Small dense inverse by Gauss-Jordan with partial pivoting (stands in for
   Julia's inv, used by the Lagrange constructor src/basis.jl:264-273). */
static void small_inverse(int n, const double* V, double* Vi) {
  double a[8][16];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j)
      a[i][j] = j < n ? V[i + n * j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i)
      if (fabs(a[i][c]) > fabs(a[p][c])) p = i;
    if (p != c)
      for (int j = 0; j < 2 * n; ++j) { double t = a[c][j]; a[c][j] = a[p][j]; a[p][j] = t; }
    double d = a[c][c];
    for (int j = 0; j < 2 * n; ++j) a[c][j] /= d;
    for (int i = 0; i < n; ++i)
      if (i != c) {
        double l = a[i][c];
        for (int j = 0; j < 2 * n; ++j) a[i][j] -= l * a[c][j];
      }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Vi[i + n * j] = a[i][n + j];
}

/* src/basis.jl:264-273,289-293 + FiniteElement.jl:41-46: P1 Lagrange on the
   reference triangle built from the Dubiner P1 basis through the inverse
   Vandermonde matrix V[i,j] = phi_j(node_i); gradients follow the same dot
   product (Tensors AD of value, ScalarFunctionSpaces.jl:82,85). */
void orc_lagrange2d_p1(double r, double s, double val[3], double grad[6]) {
  static const double nodes[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
  double V[9], Vi[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) V[i + 3 * j] = orc_dubiner_value(j + 1, nodes[i][0], nodes[i][1]);
  small_inverse(3, V, Vi);
  double phi[3], dphi[3][2];
  for (int j = 0; j < 3; ++j) {
    phi[j] = orc_dubiner_value(j + 1, r, s);
    orc_dubiner_grad(j + 1, r, s, dphi[j]);
  }
  for (int k = 0; k < 3; ++k) {
    double v = 0.0, gx = 0.0, gy = 0.0;
    for (int j = 0; j < 3; ++j) {
      v += Vi[j + 3 * k] * phi[j];
      gx += Vi[j + 3 * k] * dphi[j][0];
      gy += Vi[j + 3 * k] * dphi[j][1];
    }
    val[k] = v;
    if (grad) { grad[2 * k] = gx; grad[2 * k + 1] = gy; }
  }
}

/* src/basis.jl:295-299: P1 Lagrange on [0,1] from the Legendre basis. */
void orc_lagrange1d_p1(double x, double val[2]) {
  double V[4], Vi[4];
  const double nodes[2] = {0.0, 1.0};
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) V[i + 2 * j] = orc_legendre_value(j + 1, nodes[i]);
  small_inverse(2, V, Vi);
  double phi[2] = {orc_legendre_value(1, x), orc_legendre_value(2, x)};
  for (int k = 0; k < 2; ++k) val[k] = Vi[0 + 2 * k] * phi[0] + Vi[1 + 2 * k] * phi[1];
}

/* src/StrangQuad.jl:1-61 -- Strang-Fix rules on the reference triangle
   (points and weights are the published table values). */
int orc_strang(int order, double* p, double* w) {
  int n = 0;
#define PT(a, b) do { p[2 * n] = (a); p[2 * n + 1] = (b); ++n; } while (0)
  if (order == 0 || order == 1) {
    PT(1.0 / 3.0, 1.0 / 3.0);
    w[0] = 0.5;
  } else if (order == 2) {
    PT(1.0 / 6.0, 1.0 / 6.0); PT(1.0 / 6.0, 2.0 / 3.0); PT(2.0 / 3.0, 1.0 / 6.0);
    for (int i = 0; i < 3; ++i) w[i] = 1.0 / 6.0;
  } else if (order == 3) {
    PT(0.659027622374092, 0.231933368553031); PT(0.659027622374092, 0.109039009072877);
    PT(0.231933368553031, 0.659027622374092); PT(0.231933368553031, 0.109039009072877);
    PT(0.109039009072877, 0.659027622374092); PT(0.109039009072877, 0.231933368553031);
    for (int i = 0; i < 6; ++i) w[i] = 1.0 / 12.0 * 1.0;
  } else if (order == 4) {
    PT(0.816847572980459, 0.091576213509771); PT(0.091576213509771, 0.816847572980459);
    PT(0.091576213509771, 0.091576213509771); PT(0.108103018168070, 0.445948490915965);
    PT(0.445948490915965, 0.108103018168070); PT(0.445948490915965, 0.445948490915965);
    for (int i = 0; i < 3; ++i) w[i] = 0.109951743655322 * (1.0 / 2.0);
    for (int i = 3; i < 6; ++i) w[i] = 0.223381589678011 * (1.0 / 2.0);
  } else if (order == 5) {
    PT(0.33333333333333333, 0.33333333333333333); PT(0.79742698535308720, 0.10128650732345633);
    PT(0.10128650732345633, 0.79742698535308720); PT(0.10128650732345633, 0.10128650732345633);
    PT(0.05971587178976981, 0.47014206410511505); PT(0.47014206410511505, 0.05971587178976981);
    PT(0.47014206410511505, 0.47014206410511505);
    w[0] = 0.22500000000000000 * (1.0 / 2.0);
    for (int i = 1; i < 4; ++i) w[i] = 0.12593918054482717 * (1.0 / 2.0);
    for (int i = 4; i < 7; ++i) w[i] = 0.13239415278850616 * (1.0 / 2.0);
  } else if (order == 6) {
    PT(0.873821971016996, 0.063089014491502); PT(0.063089014491502, 0.873821971016996);
    PT(0.063089014491502, 0.063089014491502); PT(0.501426509658179, 0.249286745170910);
    PT(0.249286745170910, 0.501426509658179); PT(0.249286745170910, 0.249286745170910);
    PT(0.636502499121399, 0.310352451033785); PT(0.636502499121399, 0.053145049844816);
    PT(0.310352451033785, 0.636502499121399); PT(0.310352451033785, 0.053145049844816);
    PT(0.053145049844816, 0.636502499121399); PT(0.053145049844816, 0.310352451033785);
    for (int i = 0; i < 3; ++i) w[i] = 0.050844906370207 * (1.0 / 2.0);
    for (int i = 3; i < 6; ++i) w[i] = 0.116786275726379 * (1.0 / 2.0);
    for (int i = 6; i < 12; ++i) w[i] = 0.082851075618374 * (1.0 / 2.0);
  } else {
    return -1;
  }
#undef PT
  return n;
}

static int64_t ifact(int n) { int64_t f = 1; for (int i = 2; i <= n; ++i) f *= i; return f; }
static int64_t ipow_i(int64_t b, int e) { int64_t r = 1; for (int i = 0; i < e; ++i) r *= b; return r; }

/* src/GrundmannMoellerQuad.jl:29-49 -- compositions of n into k parts in the
   order of the "next composition" enumeration.  Writes rows of k ints. */
static int all_exponentials(int n, int k, int* out) {
  int a[8];
  int t = n, h = 0, cnt = 0;
  a[0] = n;
  for (int i = 1; i < k; ++i) a[i] = 0;
  memcpy(out + k * cnt++, a, sizeof(int) * k);
  while (a[k - 1] != n) {
    if (1 < t) h = 0;
    h = h + 1;
    t = a[h - 1];
    a[h - 1] = 0;
    a[0] = t - 1;
    a[h] = a[h] + 1;
    memcpy(out + k * cnt++, a, sizeof(int) * k);
  }
  return cnt;
}

/* src/GrundmannMoellerQuad.jl:10-27 -- Grundmann-Moeller rule of index s on
   the 2-simplex; weights normalised to sum 1/2. */
int orc_gm(int s, double* p, double* w) {
  const int n = 2;
  int d = 2 * s + 1, k = 0;
  int ex[4096 * 3];
  for (int i = 0; i <= s; ++i) {
    double wi = ((i % 2 ? -1.0 : 1.0) * pow(2.0, -2 * s) *
                 (double)ipow_i(d + n - 2 * i, d)) /
                (double)(ifact(i) * ifact(d + n - i));
    int cnt = all_exponentials(s - i, n + 1, ex);
    for (int c = 0; c < cnt; ++c) {
      p[2 * k] = (2.0 * ex[3 * c + 1] + 1.0) / (double)(d + n - 2 * i);
      p[2 * k + 1] = (2.0 * ex[3 * c + 2] + 1.0) / (double)(d + n - 2 * i);
      w[k] = wi;
      ++k;
    }
  }
  double sum = 0.0;
  for (int i = 0; i < k; ++i) sum += 2.0 * w[i];
  for (int i = 0; i < k; ++i) w[i] = w[i] / sum;
  return k;
}

/* Gauss-Legendre nodes and weights on [-1,1], ascending.  Restates
   FastGaussQuadrature.gausslegendre (called at src/quadrature.jl:30) by the
   textbook Newton iteration on the three-term recurrence. */
int orc_gauss_legendre(int n, double* x, double* w) {
  if (n < 1) return -1;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int kk = 2; kk <= n; ++kk) {
        double p2 = ((2.0 * kk - 1.0) * z * p1 - (kk - 1.0) * p0) / kk;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int kk = 2; kk <= n; ++kk) {
        double p2 = ((2.0 * kk - 1.0) * z * p1 - (kk - 1.0) * p0) / kk;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
    }
    if (2 * i + 1 == n) z = 0.0;
    double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  return n;
}

/* src/quadrature.jl:17-26 -- DefaultQuad on the triangle. */
int orc_quad2d(int order, double* p, double* w) {
  if (order <= 6) return orc_strang(order, p, w);
  if (order % 2 == 1) return orc_gm((order - 1) / 2, p, w);
  return -1;
}

/* src/quadrature.jl:29-39 -- DefaultQuad on [0,1]: Gauss-Legendre with
   `order` points, mapped from [-1,1]. */
int orc_quad1d(int order, double* x, double* w) {
  int n = orc_gauss_legendre(order, x, w);
  if (n < 0) return n;
  for (int i = 0; i < n; ++i) {
    w[i] *= 0.5;
    x[i] = (x[i] + 1.0) / 2.0;
  }
  return n;
}

/* ========================================================================= */
/* Reference element tables (src/ScalarFunctionSpaces.jl:31-99,              */
/* src/TraceFunctionSpaces.jl:11-28, src/quadrature.jl:52-75)              */
/* ========================================================================= */
typedef struct {
  int k, qdeg, nu, nt, nqp, nfq;
  double qp[2 * MAXQ], qw[MAXQ];
  double fx[MAXQ], fw[MAXQ];
  double N[MAXNU * MAXQ];      /* N[i*nqp+q] */
  double dN[MAXNU * MAXQ * 2]; /* dN[(i*nqp+q)*2+d] */
  double M[3 * MAXQ];          /* geometric P1 at cell qps: M[i*nqp+q] */
  double dM[6];                /* geometric gradients (constant), at qp 1 */
  double E[3 * MAXNU * MAXQ];  /* E[(f*nu+i)*nfq+q] cell basis at face qps */
  double L[3 * 3 * MAXQ];      /* L[(f*3+i)*nfq+q] geometric P1 at face qps */
  double Nt[MAXNT * MAXQ];     /* Nt[i*nfq+q] trace basis at 1-D qps */
} refel;

/* src/shapes.jl:19-24 -- reference edges, CCW, edge i opposite vertex i. */
static const double ref_edges[3][2][2] = {
    {{1.0, 0.0}, {0.0, 1.0}}, {{0.0, 1.0}, {0.0, 0.0}}, {{0.0, 0.0}, {1.0, 0.0}}};

static int refel_init(refel* R, int k, int qdeg) {
  memset(R, 0, sizeof(*R));
  R->k = k;
  R->qdeg = qdeg;
  R->nu = (k + 1) * (k + 2) / 2;
  R->nt = k + 1;
  if (R->nu > MAXNU || R->nt > MAXNT) return -1;
  R->nqp = orc_quad2d(qdeg, R->qp, R->qw);
  if (R->nqp <= 0 || R->nqp > MAXQ) return -2;
  R->nfq = orc_quad1d(qdeg, R->fx, R->fw);
  if (R->nfq <= 0 || R->nfq > MAXQ) return -2;
  for (int q = 0; q < R->nqp; ++q) {
    double r = R->qp[2 * q], s = R->qp[2 * q + 1];
    for (int i = 0; i < R->nu; ++i) {
      R->N[i * R->nqp + q] = orc_dubiner_value(i + 1, r, s);
      orc_dubiner_grad(i + 1, r, s, &R->dN[(i * R->nqp + q) * 2]);
    }
    double mv[3], mg[6];
    orc_lagrange2d_p1(r, s, mv, mg);
    for (int i = 0; i < 3; ++i) R->M[i * R->nqp + q] = mv[i];
    if (q == 0) memcpy(R->dM, mg, sizeof(mg));
  }
  /* face quadrature points on the three reference edges (quadrature.jl:60-75) */
  for (int f = 0; f < 3; ++f)
    for (int q = 0; q < R->nfq; ++q) {
      double lv[2];
      orc_lagrange1d_p1(R->fx[q], lv);
      double ex = 0.0, ey = 0.0;
      for (int c = 0; c < 2; ++c) {
        ex += lv[c] * ref_edges[f][c][0];
        ey += lv[c] * ref_edges[f][c][1];
      }
      for (int i = 0; i < R->nu; ++i) R->E[(f * R->nu + i) * R->nfq + q] = orc_dubiner_value(i + 1, ex, ey);
      double mv[3];
      orc_lagrange2d_p1(ex, ey, mv, NULL);
      for (int i = 0; i < 3; ++i) R->L[(f * 3 + i) * R->nfq + q] = mv[i];
    }
  for (int i = 0; i < R->nt; ++i)
    for (int q = 0; q < R->nfq; ++q) R->Nt[i * R->nfq + q] = orc_legendre_value(i + 1, R->fx[q]);
  return 0;
}

int orc_sizes(int k, int qdeg, int* nu, int* nt, int* nqp, int* nfq) {
  refel R;
  int e = refel_init(&R, k, qdeg);
  if (e) return e;
  *nu = R.nu; *nt = R.nt; *nqp = R.nqp; *nfq = R.nfq;
  return 0;
}

/* ========================================================================= */
/* Mesh                                                                       */
/* ========================================================================= */

/* generate_mesh.jl:121 */
int64_t orc_rect_nfaces(int64_t nx, int64_t ny) { return nx + ny + 3 * nx * ny; }

/* open-addressing map (min,max) node pair -> face id, standing in for the
   Dict used by _build_cells (generate_mesh.jl:20-46). */
typedef struct { int64_t* keys; int64_t* vals; int64_t cap; } pairmap;
static uint64_t mix64(uint64_t z) {
  z ^= z >> 33; z *= 0xff51afd7ed558ccdULL; z ^= z >> 33;
  z *= 0xc4ceb9fe1a85ec53ULL; z ^= z >> 33; return z;
}
static int64_t* pm_slot(pairmap* m, int64_t a, int64_t b) {
  uint64_t key = (uint64_t)a * 0x100000000ULL + (uint64_t)b;
  uint64_t h = mix64(key) & (uint64_t)(m->cap - 1);
  for (;;) {
    if (m->keys[2 * h] == 0) { m->keys[2 * h] = a; m->keys[2 * h + 1] = b; m->vals[h] = 0; return &m->vals[h]; }
    if (m->keys[2 * h] == a && m->keys[2 * h + 1] == b) return &m->vals[h];
    h = (h + 1) & (uint64_t)(m->cap - 1);
  }
}

/* generate_mesh.jl:49-57 */
static void check_node_data(const double* xy, int64_t* n) {
  double ax = xy[2 * (n[1] - 1)] - xy[2 * (n[0] - 1)];
  double ay = xy[2 * (n[1] - 1) + 1] - xy[2 * (n[0] - 1) + 1];
  double bx = xy[2 * (n[2] - 1)] - xy[2 * (n[0] - 1)];
  double by = xy[2 * (n[2] - 1) + 1] - xy[2 * (n[0] - 1) + 1];
  if ((ax * by - ay * bx) < 0) { int64_t t = n[1]; n[1] = n[2]; n[2] = t; }
}

/* generate_mesh.jl:20-46 */
static int64_t build_cell(int64_t face_idx, int64_t n_el, const int64_t* el_nodes,
                          int64_t* cells, int64_t* faces, int64_t nf, pairmap* dict) {
  static const int ren[3][2] = {{1, 2}, {2, 0}, {0, 1}};
  for (int i = 0; i < 3; ++i) {
    int64_t v1 = el_nodes[ren[i][0]], v2 = el_nodes[ren[i][1]];
    int64_t a = v1 < v2 ? v1 : v2, b = v1 < v2 ? v2 : v1;
    int64_t* slot = pm_slot(dict, a, b);
    if (*slot > 0) {
      cells[6 * (n_el - 1) + 3 + i] = *slot;
      if (n_el != faces[(*slot - 1) + 3 * nf]) faces[(*slot - 1) + 3 * nf] = n_el;
    } else {
      face_idx = face_idx + 1;
      *slot = face_idx;
      faces[(face_idx - 1) + 0 * nf] = v1;
      faces[(face_idx - 1) + 1 * nf] = v2;
      faces[(face_idx - 1) + 2 * nf] = n_el;
      faces[(face_idx - 1) + 3 * nf] = 0;
      cells[6 * (n_el - 1) + 3 + i] = face_idx;
    }
  }
  for (int i = 0; i < 3; ++i) cells[6 * (n_el - 1) + i] = el_nodes[i];
  return face_idx;
}

/* generate_mesh.jl:1-18 (nodes), 101-143 (cells, faces), 60-89 (boundary
   sets).  face_side[f]: 0 interior, 1 bottom, 2 right, 3 top, 4 left. */
int orc_rectangle_mesh(int64_t nx, int64_t ny, double x0, double y0, double x1,
                       double y1, double* xy, int64_t* cells, int64_t* faces,
                       int8_t* face_side) {
  const double LL[2] = {x0, y0}, UR[2] = {x1, y1};
  const double LR[2] = {x1, y0}, UL[2] = {x0, y1};
  int64_t nnx = nx + 1, nny = ny + 1;
  int64_t nf = orc_rect_nfaces(nx, ny);
  int64_t cnt = 0;
  for (int64_t i = 0; i <= nny - 1; ++i) {
    double rb = (double)i / (double)(nny - 1);
    double xa = LL[0] * (1 - rb) + rb * UL[0];
    double xb = LR[0] * (1 - rb) + rb * UR[0];
    double ya = LL[1] * (1 - rb) + rb * UL[1];
    double yb = LR[1] * (1 - rb) + rb * UR[1];
    for (int64_t j = 0; j <= nnx - 1; ++j) {
      double ratio = (double)j / (double)(nnx - 1);
      xy[2 * cnt] = xa * (1 - ratio) + ratio * xb;
      xy[2 * cnt + 1] = ya * (1 - ratio) + ratio * yb;
      ++cnt;
    }
  }
  pairmap dict;
  dict.cap = 1;
  while (dict.cap < 4 * nf) dict.cap <<= 1;
  dict.keys = (int64_t*)calloc((size_t)dict.cap * 2, sizeof(int64_t));
  dict.vals = (int64_t*)calloc((size_t)dict.cap, sizeof(int64_t));
  if (!dict.keys || !dict.vals) { free(dict.keys); free(dict.vals); return -1; }
  int64_t n_el = 0, face_idx = 0, el[3];
#define NODE(i, j) ((i) + ((j) - 1) * nnx) /* node_array[i,j], 1-based */
  for (int64_t j = 1; j <= ny; ++j)
    for (int64_t i = 1; i <= nx; ++i) {
      n_el++;
      el[0] = NODE(i, j); el[1] = NODE(i + 1, j); el[2] = NODE(i, j + 1);
      check_node_data(xy, el);
      face_idx = build_cell(face_idx, n_el, el, cells, faces, nf, &dict);
      n_el++;
      el[0] = NODE(i + 1, j); el[1] = NODE(i + 1, j + 1); el[2] = NODE(i, j + 1);
      check_node_data(xy, el);
      face_idx = build_cell(face_idx, n_el, el, cells, faces, nf, &dict);
    }
#undef NODE
  free(dict.keys);
  free(dict.vals);
  if (face_idx != nf) return -2;
  for (int64_t f = 0; f < nf; ++f) {
    face_side[f] = 0;
    if (faces[f + 3 * nf] != 0) continue;
    const double* p1 = xy + 2 * (faces[f] - 1);
    const double* p2 = xy + 2 * (faces[f + nf] - 1);
    if (p1[1] == LL[1] && p2[1] == LL[1]) face_side[f] = 1;
    else if (p1[0] == UR[0] && p2[0] == UR[0]) face_side[f] = 2;
    else if (p1[1] == UR[1] && p2[1] == UR[1]) face_side[f] = 3;
    else if (p1[0] == LL[0] && p2[0] == LL[0]) face_side[f] = 4;
    else return -3; /* generate_mesh.jl:80 */
  }
  return 0;
}

/* Interior-node jitter (an extension used by the jittered workloads; the
   reference has none): node n (0-based) at grid (i, j) moves by
   amp * h * (2u - 1) per axis, u from splitmix64 of seed + 2n (+1). */
static uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
void orc_jitter(int64_t nx, int64_t ny, double x0, double y0, double x1,
                double y1, double amp, uint64_t seed, double* xy) {
  double hx = (x1 - x0) / (double)nx, hy = (y1 - y0) / (double)ny;
  for (int64_t j = 1; j < ny; ++j)
    for (int64_t i = 1; i < nx; ++i) {
      uint64_t n = (uint64_t)(i + j * (nx + 1));
      double ux = (double)(splitmix64(seed + 2 * n) >> 11) * (1.0 / 9007199254740992.0);
      double uy = (double)(splitmix64(seed + 2 * n + 1) >> 11) * (1.0 / 9007199254740992.0);
      xy[2 * n] = xy[2 * n] + amp * hx * (2.0 * ux - 1.0);
      xy[2 * n + 1] = xy[2 * n + 1] + amp * hy * (2.0 * uy - 1.0);
    }
}

/* src/mesh.jl:51-54 (face 1-based in the reference; 0-based here). */
int orc_face_orientation(const int64_t* cells, int64_t cell, int face) {
  static const int ren[3][2] = {{1, 2}, {2, 0}, {0, 1}};
  const int64_t* c = cells + 6 * cell;
  return c[ren[face][1]] > c[ren[face][0]];
}

/* ========================================================================= */
/* Cell geometry: reinit! (src/ScalarFunctionSpaces.jl:101-132)              */
/* ========================================================================= */
typedef struct {
  double J[4];    /* column-major J[a + 2b] */
  double detJ;
  double Jinv[4]; /* column-major */
  double normals[6];
  double detJf[3];
  double x[6];    /* vertex coordinates */
} cellgeo;

static int reinit(const refel* R, const double* xy, const int64_t* cells,
                  int64_t cell, cellgeo* G) {
  const int64_t* c = cells + 6 * cell;
  for (int v = 0; v < 3; ++v) {
    G->x[2 * v] = xy[2 * (c[v] - 1)];
    G->x[2 * v + 1] = xy[2 * (c[v] - 1) + 1];
  }
  double J[4] = {0, 0, 0, 0};
  for (int j = 0; j < 3; ++j)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) J[a + 2 * b] += G->x[2 * j + a] * R->dM[2 * j + b];
  memcpy(G->J, J, sizeof(J));
  double det = J[0] * J[3] - J[2] * J[1];
  if (!(det > 0.0)) return -10; /* ScalarFunctionSpaces.jl:110 */
  G->detJ = det;
  double dinv = 1.0 / det;
  G->Jinv[0] = J[3] * dinv;
  G->Jinv[1] = -J[1] * dinv;
  G->Jinv[2] = -J[2] * dinv;
  G->Jinv[3] = J[0] * dinv;
  /* shapes.jl:80-87 weighted normals */
  for (int l = 0; l < 3; ++l) {
    double wn[2];
    if (l == 0) { wn[0] = -(J[1] - J[3]); wn[1] = J[0] - J[2]; }
    else if (l == 1) { wn[0] = -J[3]; wn[1] = J[2]; }
    else { wn[0] = J[1]; wn[1] = -J[0]; }
    double dj = sqrt(wn[0] * wn[0] + wn[1] * wn[1]);
    G->normals[2 * l] = wn[0] / dj;
    G->normals[2 * l + 1] = wn[1] / dj;
    if (!(dj > 0.0)) return -10;
    G->detJf[l] = dj;
  }
  return 0;
}

int orc_geometry(const double* xy, const int64_t* cells, int64_t cell,
                 double* detJ, double* Jinv, double* normals, double* detJf) {
  refel R;
  refel_init(&R, 1, 2);
  cellgeo G;
  int e = reinit(&R, xy, cells, cell, &G);
  if (e) return e;
  *detJ = G.detJ;
  memcpy(Jinv, G.Jinv, sizeof(G.Jinv));
  memcpy(normals, G.normals, sizeof(G.normals));
  memcpy(detJf, G.detJf, sizeof(G.detJf));
  return 0;
}

/* ========================================================================= */
/* Sources                                                                    */
/* ========================================================================= */
static double kappa_at(const orc_problem* P, double x, double y) {
  if (!P->kappa_on) return 1.0;
  return 1.0 + 0.5 * sin(P->ka * x) * cos(P->kb * y);
}

/* f at a physical point.  Kind 0 is the reference's source
   (examples/poisson2D_HDG.jl:55); kinds 1-2 are the extension workloads. */
static double source_at(const orc_problem* P, double x, double y) {
  switch (P->src_kind) {
    case 0: return 2 * (PI * PI) * sin(PI * x) * sin(PI * y);
    case 1: {
      double u = sin(PI * x) * sin(PI * y);
      double ux = PI * cos(PI * x) * sin(PI * y), uy = PI * sin(PI * x) * cos(PI * y);
      double k = kappa_at(P, x, y), kx = 0.0, ky = 0.0;
      if (P->kappa_on) {
        kx = 0.5 * P->ka * cos(P->ka * x) * cos(P->kb * y);
        ky = -0.5 * P->kb * sin(P->ka * x) * sin(P->kb * y);
      }
      return k * 2.0 * PI * PI * u - (kx * ux + ky * uy);
    }
    case 2: {
      double f = 0.0;
      for (int p = 1; p <= 4; ++p)
        for (int q = 1; q <= 4; ++q)
          f += P->coef[(p - 1) * 4 + (q - 1)] * sin(p * PI * x) * sin(q * PI * y);
      return f;
    }
    default: return 0.0;
  }
}

/* ========================================================================= */
/* Element blocks: the cell loop body of doassemble                          */
/* (examples/poisson2D_HDG.jl:77-153)                                        */
/* ========================================================================= */
static int element_blocks_impl(const orc_problem* P, const refel* R,
                               const double* xy, const int64_t* cells,
                               int64_t ncells, int64_t cell, double* A,
                               double* B, double* C, double* E, double* F,
                               double* H, double* be) {
  const int nu = R->nu, nq = 2 * nu, nt = R->nt, ntr = 3 * nt;
  const int nqp = R->nqp, nfq = R->nfq;
  const double tau = P->tau;
  cellgeo G;
  int e = reinit(R, xy, cells, cell, &G);
  if (e) return e;
  memset(A, 0, sizeof(double) * nq * nq);
  memset(B, 0, sizeof(double) * nq * nu);
  memset(C, 0, sizeof(double) * nu * nu);
  memset(E, 0, sizeof(double) * nq * ntr);
  memset(F, 0, sizeof(double) * nu * ntr);
  memset(H, 0, sizeof(double) * ntr * ntr);
  memset(be, 0, sizeof(double) * nu);
  /* dNdx = dNdxi . Jinv (ScalarFunctionSpaces.jl:114-116) */
  double dNdx[MAXNU * MAXQ * 2];
  for (int i = 0; i < nu; ++i)
    for (int q = 0; q < nqp; ++q)
      for (int c = 0; c < 2; ++c) {
        const double* g = &R->dN[(i * nqp + q) * 2];
        dNdx[(i * nqp + q) * 2 + c] = g[0] * G.Jinv[0 + 2 * c] + g[1] * G.Jinv[1 + 2 * c];
      }
  /* cell integrals (poisson2D_HDG.jl:88-104) */
  for (int q = 0; q < nqp; ++q) {
    double dO = G.detJ * R->qw[q];
    double xq = 0.0, yq = 0.0; /* spatial_coordinate, DiscreteFunctions.jl:16-24 */
    for (int i = 0; i < 3; ++i) {
      xq += R->M[i * nqp + q] * G.x[2 * i];
      yq += R->M[i * nqp + q] * G.x[2 * i + 1];
    }
    double kinv = P->kappa_on ? 1.0 / kappa_at(P, xq, yq) : 1.0;
    for (int i = 0; i < nq; ++i) {
      int ci = i / nu, ii = i % nu;  /* VectorFunctionSpaces.jl:28-57 */
      double vh = R->N[ii * nqp + q];
      double div_vh = dNdx[(ii * nqp + q) * 2 + ci];
      for (int j = 0; j < nq; ++j) {
        int cj = j / nu, jj = j % nu;
        double sv = (ci == cj) ? R->N[jj * nqp + q] * vh : 0.0;
        if (P->kappa_on) A[i + nq * j] += sv * dO * kinv;
        else A[i + nq * j] += sv * dO;
      }
      for (int j = 0; j < nu; ++j) B[i + nq * j] += (R->N[j * nqp + q] * div_vh) * dO;
    }
  }
  /* RHS (poisson2D_HDG.jl:106-114) */
  for (int q = 0; q < nqp; ++q) {
    double dO = G.detJ * R->qw[q];
    double fh;
    if (P->src_kind == 3) {
      fh = P->f_qp[cell + ncells * q];
    } else {
      double xq = 0.0, yq = 0.0;
      for (int i = 0; i < 3; ++i) {
        xq += R->M[i * nqp + q] * G.x[2 * i];
        yq += R->M[i * nqp + q] * G.x[2 * i + 1];
      }
      fh = source_at(P, xq, yq);
    }
    for (int i = 0; i < nu; ++i) be[i] += fh * R->N[i * nqp + q] * dO;
  }
  /* face integrals (poisson2D_HDG.jl:116-153) */
  for (int f = 0; f < 3; ++f) {
    int o = orc_face_orientation(cells, cell, f);
    for (int q = 0; q < nfq; ++q) {
      double dS = G.detJf[f] * R->fw[q];
      int qo = o ? q : nfq - 1 - q; /* ScalarFunctionSpaces.jl:153 */
      for (int i = 0; i < nu; ++i) {
        double w = R->E[(f * nu + i) * nfq + q];
        for (int j = 0; j < nu; ++j) {
          double u = R->E[(f * nu + j) * nfq + q];
          C[i + nu * j] += tau * (u * w) * dS;
        }
        w = R->E[(f * nu + i) * nfq + qo];
        for (int j = 0; j < nt; ++j) {
          double uh = R->Nt[j * nfq + q];
          F[i + nu * (nt * f + j)] += (tau * (uh * w)) * dS;
        }
      }
      for (int i = 0; i < nq; ++i) {
        int ci = i / nu, ii = i % nu;
        double vn = R->E[(f * nu + ii) * nfq + qo] * G.normals[2 * f + ci];
        for (int j = 0; j < nt; ++j) {
          double uh = R->Nt[j * nfq + q];
          E[i + nq * (nt * f + j)] += (uh * vn) * dS;
        }
      }
      for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j) {
          double mu = R->Nt[i * nfq + q], uh = R->Nt[j * nfq + q];
          H[(nt * f + i) + ntr * (nt * f + j)] += (uh * mu) * dS;
        }
    }
  }
  return 0;
}

int orc_element_blocks(const orc_problem* P, const double* xy,
                       const int64_t* cells, int64_t ncells, int64_t cell,
                       double* A, double* B, double* C, double* E, double* F,
                       double* H, double* be) {
  refel R;
  int e = refel_init(&R, P->k, P->qdeg);
  if (e) return e;
  return element_blocks_impl(P, &R, xy, cells, ncells, cell, A, B, C, E, F, H, be);
}

/* Dense LU with partial pivoting (LAPACK getf2 semantics; stands in for
   factorize(Array(Me)), poisson2D_HDG.jl:160) and the matching solve
   (Mei \ X, :171,173).  a is n x n column-major, overwritten. */
static int lu_factor(int n, double* a, int* piv) {
  for (int j = 0; j < n; ++j) {
    int p = j;
    double mx = fabs(a[j + n * j]);
    for (int i = j + 1; i < n; ++i)
      if (fabs(a[i + n * j]) > mx) { mx = fabs(a[i + n * j]); p = i; }
    piv[j] = p;
    if (a[p + n * j] == 0.0) return -20;
    if (p != j)
      for (int c = 0; c < n; ++c) { double t = a[j + n * c]; a[j + n * c] = a[p + n * c]; a[p + n * c] = t; }
    double inv = 1.0 / a[j + n * j];
    for (int i = j + 1; i < n; ++i) a[i + n * j] *= inv;
    for (int c = j + 1; c < n; ++c) {
      double t = a[j + n * c];
      if (t != 0.0)
        for (int i = j + 1; i < n; ++i) a[i + n * c] -= a[i + n * j] * t;
    }
  }
  return 0;
}
static void lu_solve(int n, const double* a, const int* piv, double* b, int nrhs) {
  for (int r = 0; r < nrhs; ++r) {
    double* x = b + n * r;
    for (int j = 0; j < n; ++j)
      if (piv[j] != j) { double t = x[j]; x[j] = x[piv[j]]; x[piv[j]] = t; }
    for (int j = 0; j < n; ++j)
      for (int i = j + 1; i < n; ++i) x[i] -= a[i + n * j] * x[j];
    for (int j = n - 1; j >= 0; --j) {
      x[j] /= a[j + n * j];
      for (int i = 0; i < j; ++i) x[i] -= a[i + n * j] * x[j];
    }
  }
}

/* Static condensation (poisson2D_HDG.jl:155-174): Me = [A -B; B' C],
   EFe = [-E; F], Ge = [E; F]', K_e = Me \ EFe, Ate = Ge K_e - H,
   b_e = Me \ [0; be], bte = -Ge b_e.  Per cell outputs (column-major):
   Ate ntr x ntr, bte ntr, Ke m x ntr, be m. */
int orc_condense(const orc_problem* P, const double* xy, const int64_t* cells,
                 int64_t ncells, double* Ate, double* bte, double* Ke,
                 double* beo, int nthreads) {
  refel R;
  int e0 = refel_init(&R, P->k, P->qdeg);
  if (e0) return e0;
  const int nu = R.nu, nq = 2 * nu, m = 3 * nu, ntr = 3 * R.nt;
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double* buf = (double*)malloc(sizeof(double) * (size_t)(nq * nq + nq * nu + nu * nu + nq * ntr +
                                                           nu * ntr + ntr * ntr + nu + m * m +
                                                           m * (ntr + 1)));
    int* piv = (int*)malloc(sizeof(int) * m);
    double *A = buf, *B = A + nq * nq, *C = B + nq * nu, *E = C + nu * nu;
    double *F = E + nq * ntr, *H = F + nu * ntr, *be = H + ntr * ntr;
    double *Me = be + nu, *X = Me + m * m;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t c = 0; c < ncells; ++c) {
      int e = element_blocks_impl(P, &R, xy, cells, ncells, c, A, B, C, E, F, H, be);
      if (e) { err = e; continue; }
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
          double v;
          if (i < nq && j < nq) v = A[i + nq * j];
          else if (i < nq) v = -B[i + nq * (j - nq)];
          else if (j < nq) v = B[j + nq * (i - nq)];
          else v = C[(i - nq) + nu * (j - nq)];
          Me[i + m * j] = v;
        }
      for (int j = 0; j < ntr; ++j)
        for (int i = 0; i < m; ++i) X[i + m * j] = i < nq ? -E[i + nq * j] : F[(i - nq) + nu * j];
      for (int i = 0; i < m; ++i) X[i + m * ntr] = i < nq ? 0.0 : be[i - nq];
      e = lu_factor(m, Me, piv);
      if (e) { err = e; continue; }
      lu_solve(m, Me, piv, X, ntr + 1);
      double* ke = Ke + (size_t)c * m * ntr;
      double* bo = beo + (size_t)c * m;
      memcpy(ke, X, sizeof(double) * m * ntr);
      memcpy(bo, X + m * ntr, sizeof(double) * m);
      double* at = Ate + (size_t)c * ntr * ntr;
      double* bt = bte + (size_t)c * ntr;
      for (int b = 0; b < ntr; ++b)
        for (int a = 0; a < ntr; ++a) {
          double s = 0.0;
          for (int j = 0; j < m; ++j) {
            double g = j < nq ? E[j + nq * a] : F[(j - nq) + nu * a];
            s += g * ke[j + m * b];
          }
          at[a + ntr * b] = s - H[a + ntr * b];
        }
      for (int a = 0; a < ntr; ++a) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) {
          double g = j < nq ? E[j + nq * a] : F[(j - nq) + nu * a];
          s += g * bo[j];
        }
        bt[a] = -s;
      }
    }
    free(buf);
    free(piv);
  }
  return err;
}

/* ========================================================================= */
/* Global assembly: COO triplets + sparse(I,J,V) (src/assembler.jl:1-60,     */
/* poisson2D_HDG.jl:176-185).  Output is Julia's SparseMatrixCSC of K,       */
/* 1-based, with duplicates summed in encounter order and stored zeros kept. */
/* Pass colptr == NULL to query nnz only.                                    */
/* ========================================================================= */
int orc_assemble(int64_t ncells, const int64_t* cells, int64_t nfaces, int nt,
                 const double* Ate, const double* bte, int64_t* colptr,
                 int64_t* rowval, double* nzval, double* rhs, int64_t* nnz) {
  const int ntr = 3 * nt;
  const int64_t n = nfaces * nt;
  const int64_t ntrip = ncells * ntr * ntr;
  /* count triplets per column (gdof, poisson2D_HDG.jl:176-181) */
  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  if (!cnt) return -1;
  for (int64_t c = 0; c < ncells; ++c)
    for (int lf = 0; lf < 3; ++lf)
      for (int j = 0; j < nt; ++j) {
        int64_t col = (cells[6 * c + 3 + lf] - 1) * nt + j;
        cnt[col + 1] += ntr;
      }
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  int64_t* trow = (int64_t*)malloc(sizeof(int64_t) * (size_t)ntrip);
  double* tval = (double*)malloc(sizeof(double) * (size_t)ntrip);
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!trow || !tval || !pos) { free(cnt); free(trow); free(tval); free(pos); return -1; }
  memcpy(pos, cnt, sizeof(int64_t) * n);
  /* triplets in encounter order: cell, then column j, then row i
     (assembler.jl:31-40: V = vec(Ate), I repeats edof, J = edof[j]). */
  for (int64_t c = 0; c < ncells; ++c) {
    int64_t gdof[48];
    for (int lf = 0; lf < 3; ++lf)
      for (int j = 0; j < nt; ++j) gdof[lf * nt + j] = (cells[6 * c + 3 + lf] - 1) * nt + j;
    const double* at = Ate + (size_t)c * ntr * ntr;
    for (int j = 0; j < ntr; ++j)
      for (int i = 0; i < ntr; ++i) {
        int64_t p = pos[gdof[j]]++;
        trow[p] = gdof[i];
        tval[p] = at[i + ntr * j];
      }
  }
  /* per column: stable sort by row, combine duplicates in order */
  int64_t total = 0;
  for (int64_t col = 0; col < n; ++col) {
    int64_t b = cnt[col], e = cnt[col + 1];
    for (int64_t p = b + 1; p < e; ++p) { /* stable insertion sort */
      int64_t r = trow[p];
      double v = tval[p];
      int64_t q = p - 1;
      while (q >= b && trow[q] > r) { trow[q + 1] = trow[q]; tval[q + 1] = tval[q]; --q; }
      trow[q + 1] = r;
      tval[q + 1] = v;
    }
    int64_t u = 0;
    for (int64_t p = b; p < e; ++p)
      if (p == b || trow[p] != trow[p - 1]) u++;
    total += u;
  }
  if (nnz) *nnz = total;
  if (colptr) {
    int64_t w = 0;
    colptr[0] = 1;
    for (int64_t col = 0; col < n; ++col) {
      int64_t b = cnt[col], e = cnt[col + 1];
      for (int64_t p = b; p < e; ++p) {
        if (p == b || trow[p] != trow[p - 1]) {
          rowval[w] = trow[p] + 1;
          nzval[w] = tval[p];
          ++w;
        } else {
          nzval[w - 1] = nzval[w - 1] + tval[p];
        }
      }
      colptr[col + 1] = w + 1;
    }
  }
  if (rhs) { /* assemble!(rhs, gdof, bte), assembler.jl:55-60 */
    for (int64_t i = 0; i < n; ++i) rhs[i] = 0.0;
    for (int64_t c = 0; c < ncells; ++c)
      for (int lf = 0; lf < 3; ++lf)
        for (int j = 0; j < nt; ++j)
          rhs[(cells[6 * c + 3 + lf] - 1) * nt + j] += bte[(size_t)c * ntr + lf * nt + j];
  }
  free(cnt); free(trow); free(tval); free(pos);
  return 0;
}

/* ========================================================================= */
/* Trace Dirichlet data (src/boundary.jl:11-42)                              */
/* g_kind 0: g = 0; 1: g = g0 + g1 x + g2 y.  quirk 0 is the L2 projection   */
/* onto the orthonormal trace basis.  quirk bit 1 keeps the reference's      */
/* accumulator, which is not reset between modes (boundary.jl:27-37); bit 2  */
/* weights the face points with the cell rule's weights fs.fs.qr_weights[q]  */
/* as the reference does (boundary.jl:33) instead of the face rule's.        */
/* quirk 3 is the reference's computation.                                   */
/* ========================================================================= */
int orc_dirichlet_trace(const orc_problem* P, const double* xy,
                        const int64_t* cells, const int64_t* faces,
                        int64_t nfaces, const int8_t* in_set, int g_kind,
                        const double* g_par, int quirk, int64_t* dofs,
                        double* vals, int64_t* ndofs) {
  refel R;
  int e0 = refel_init(&R, P->k, P->qdeg);
  if (e0) return e0;
  const int nt = R.nt, nfq = R.nfq;
  int64_t kk = 0;
  for (int64_t f = 0; f < nfaces; ++f) {
    if (!in_set[f]) continue;
    if (faces[f + 3 * nfaces] != 0) return -30; /* boundary.jl:22 */
    int64_t cell = faces[f + 2 * nfaces] - 1;
    int lf = -1;
    for (int i = 0; i < 3; ++i)
      if (cells[6 * cell + 3 + i] == f + 1) { lf = i; break; }
    if (lf < 0) return -31;
    int o = orc_face_orientation(cells, cell, lf);
    double N = 0.0;
    for (int i = 0; i < nt; ++i) {
      if (!(quirk & 1)) N = 0.0;
      for (int q = 0; q < nfq; ++q) {
        int qo = o ? q : nfq - 1 - q; /* TraceFunctionSpaces.jl:47-56 */
        double x = 0.0, y = 0.0;
        for (int v = 0; v < 3; ++v) {
          double lv = R.L[(lf * 3 + v) * nfq + qo];
          x += lv * xy[2 * (cells[6 * cell + v] - 1)];
          y += lv * xy[2 * (cells[6 * cell + v] - 1) + 1];
        }
        double g = g_kind == 1 ? g_par[0] + g_par[1] * x + g_par[2] * y : 0.0;
        const double wq = (quirk & 2) ? R.qw[q] : R.fw[q];  /* boundary.jl:33 */
        N += wq * g * R.Nt[i * nfq + q];
      }
      if (vals) vals[kk] = N;
      if (dofs) dofs[kk] = (f + 1) * nt - nt + (i + 1);
      ++kk;
    }
  }
  if (ndofs) *ndofs = kk;
  return 0;
}

/* src/boundary.jl:121-175 -- apply!(K, f, dbc) with APPLY_TRANSPOSE. */
int orc_apply(int64_t n, const int64_t* colptr, const int64_t* rowval,
              double* nzval, double* rhs, int64_t nd, const int64_t* dofs,
              const double* vals) {
  if (n <= 0) return -1;
  double md = 0.0; /* meandiag */
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t p = colptr[i] - 1; p < colptr[i + 1] - 1; ++p)
      if (rowval[p] - 1 == i) d = nzval[p];
    md += fabs(d);
  }
  md = md / (double)n;
  for (int64_t t = 0; t < nd; ++t) {
    int64_t d = dofs[t] - 1;
    double v = vals[t];
    if (v != 0)
      for (int64_t p = colptr[d] - 1; p < colptr[d + 1] - 1; ++p) rhs[rowval[p] - 1] -= v * nzval[p];
  }
  char* isd = (char*)calloc((size_t)n, 1);
  for (int64_t t = 0; t < nd; ++t) isd[dofs[t] - 1] = 1;
  for (int64_t col = 0; col < n; ++col)
    for (int64_t p = colptr[col] - 1; p < colptr[col + 1] - 1; ++p)
      if (isd[col] || isd[rowval[p] - 1]) nzval[p] = 0.0;
  for (int64_t t = 0; t < nd; ++t) {
    int64_t d = dofs[t] - 1;
    for (int64_t p = colptr[d] - 1; p < colptr[d + 1] - 1; ++p)
      if (rowval[p] - 1 == d) nzval[p] = md;
    rhs[d] = vals[t] * md;
  }
  free(isd);
  return 0;
}

/* r = b - K x for K in CSC. */
int orc_residual(int64_t n, const int64_t* colptr, const int64_t* rowval,
                 const double* nzval, const double* x, const double* b, double* r) {
  for (int64_t i = 0; i < n; ++i) r[i] = b[i];
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c] - 1; p < colptr[c + 1] - 1; ++p) r[rowval[p] - 1] -= nzval[p] * x[c];
  return 0;
}

/* Direct solve K \ b (poisson2D_HDG.jl:195) by banded LU with partial
   pivoting (LAPACK gbtrf/gbtrs semantics) -- exact up to rounding, used for
   meshes small enough that the band fits in memory. */
int orc_banded_solve(int64_t n, const int64_t* colptr, const int64_t* rowval,
                     const double* nzval, const double* b, double* x) {
  int64_t kl = 0, ku = 0;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c] - 1; p < colptr[c + 1] - 1; ++p) {
      int64_t r = rowval[p] - 1;
      if (r - c > kl) kl = r - c;
      if (c - r > ku) ku = c - r;
    }
  int64_t ldab = 2 * kl + ku + 1;
  if ((double)ldab * (double)n > 4e8) return -40;
  double* ab = (double*)calloc((size_t)(ldab * n), sizeof(double));
  if (!ab) return -1;
  /* band storage: A(i,j) at ab[(kl + ku + i - j) + ldab * j] */
#define AB(i, j) ab[(kl + ku + (i) - (j)) + ldab * (j)]
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c] - 1; p < colptr[c + 1] - 1; ++p) AB(rowval[p] - 1, c) += nzval[p];
  int64_t* piv = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  int64_t kv = ku + kl;
  for (int64_t j = 0; j < n; ++j) {
    int64_t km = kl < n - 1 - j ? kl : n - 1 - j;
    int64_t p = j;
    double mx = fabs(AB(j, j));
    for (int64_t i = j + 1; i <= j + km; ++i)
      if (fabs(AB(i, j)) > mx) { mx = fabs(AB(i, j)); p = i; }
    piv[j] = p;
    if (AB(p, j) == 0.0) { free(ab); free(piv); return -20; }
    int64_t ju = j + kv < n - 1 ? j + kv : n - 1;
    if (p != j)
      for (int64_t c = j; c <= ju; ++c) { double t = AB(j, c); AB(j, c) = AB(p, c); AB(p, c) = t; }
    double inv = 1.0 / AB(j, j);
    for (int64_t i = j + 1; i <= j + km; ++i) AB(i, j) *= inv;
    for (int64_t c = j + 1; c <= ju; ++c) {
      double t = AB(j, c);
      if (t != 0.0)
        for (int64_t i = j + 1; i <= j + km; ++i) AB(i, c) -= AB(i, j) * t;
    }
  }
  for (int64_t i = 0; i < n; ++i) x[i] = b[i];
  for (int64_t j = 0; j < n; ++j) {
    if (piv[j] != j) { double t = x[j]; x[j] = x[piv[j]]; x[piv[j]] = t; }
    int64_t km = kl < n - 1 - j ? kl : n - 1 - j;
    for (int64_t i = j + 1; i <= j + km; ++i) x[i] -= AB(i, j) * x[j];
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    x[j] /= AB(j, j);
    int64_t i0 = j - kv > 0 ? j - kv : 0;
    for (int64_t i = i0; i < j; ++i) x[i] -= AB(i, j) * x[j];
  }
#undef AB
  free(ab);
  free(piv);
  return 0;
}

/* Preconditioned CG on S = -K restricted to the non-Dirichlet dofs, with a
   block-Jacobi preconditioner made of the inverted nt x nt diagonal block of
   each face.  This is the iterative stand-in for K \ b used beyond the sizes
   the banded solve handles; Dirichlet values are lifted into the rhs. */
int orc_pcg_solve(int64_t n, const int64_t* colptr, const int64_t* rowval,
                  const double* nzval, const double* b, int nt,
                  const int64_t* dir_dofs, const double* dir_vals, int64_t nd,
                  double rtol, int maxit, double* x, int* iters,
                  double* relres, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t nnz = colptr[n] - 1;
  /* CSR of K by transposing the CSC arrays */
  int64_t* rp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t* ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)nnz);
  double* va = (double*)malloc(sizeof(double) * (size_t)nnz);
  char* isd = (char*)calloc((size_t)n, 1);
  double* bb = (double*)malloc(sizeof(double) * (size_t)n);
  double *r = (double*)malloc(sizeof(double) * (size_t)n), *z = (double*)malloc(sizeof(double) * (size_t)n);
  double *p = (double*)malloc(sizeof(double) * (size_t)n), *q = (double*)malloc(sizeof(double) * (size_t)n);
  int64_t nf = n / nt;
  double* Dinv = (double*)malloc(sizeof(double) * (size_t)(nf * nt * nt));
  if (!rp || !ci || !va || !isd || !bb || !r || !z || !p || !q || !Dinv) return -1;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t k = colptr[c] - 1; k < colptr[c + 1] - 1; ++k) rp[rowval[k]]++;
  for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
  {
    int64_t* w = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    memcpy(w, rp, sizeof(int64_t) * n);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t k = colptr[c] - 1; k < colptr[c + 1] - 1; ++k) {
        int64_t row = rowval[k] - 1, d = w[row]++;
        ci[d] = c;
        va[d] = nzval[k];
      }
    free(w);
  }
  for (int64_t t = 0; t < nd; ++t) isd[dir_dofs[t] - 1] = 1;
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  for (int64_t t = 0; t < nd; ++t) x[dir_dofs[t] - 1] = dir_vals[t];
  /* rhs of S x_I = -(b_I - K_ID x_D) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) {
    if (isd[i]) { bb[i] = 0.0; continue; }
    double s = b[i];
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      if (isd[ci[k]]) s -= va[k] * x[ci[k]];
    bb[i] = -s;
  }
  /* block-Jacobi blocks: inverse of -K[f,f] */
  for (int64_t f = 0; f < nf; ++f) {
    double D[256];
    for (int a = 0; a < nt; ++a)
      for (int c = 0; c < nt; ++c) D[a + nt * c] = 0.0;
    for (int a = 0; a < nt; ++a) {
      int64_t row = f * nt + a;
      for (int64_t k = rp[row]; k < rp[row + 1]; ++k)
        if (ci[k] / nt == f) D[a + nt * (ci[k] % nt)] = -va[k];
    }
    small_inverse(nt, D, Dinv + f * nt * nt);
  }
  double bn = 0.0;
  for (int64_t i = 0; i < n; ++i) bn += bb[i] * bb[i];
  bn = sqrt(bn);
  for (int64_t i = 0; i < n; ++i) { r[i] = bb[i]; }
  double rz = 0.0, rr = 0.0;
  int it = 0;
  if (bn == 0.0) { *iters = 0; *relres = 0.0; goto done; }
#define PRECOND()                                                            \
  _Pragma("omp parallel for schedule(static)")                               \
  for (int64_t f = 0; f < nf; ++f)                                           \
    for (int a = 0; a < nt; ++a) {                                           \
      int64_t row = f * nt + a;                                              \
      if (isd[row]) { z[row] = 0.0; continue; }                              \
      double s = 0.0;                                                        \
      for (int c = 0; c < nt; ++c) s += Dinv[f * nt * nt + a + nt * c] * r[f * nt + c]; \
      z[row] = s;                                                            \
    }
  PRECOND();
  for (int64_t i = 0; i < n; ++i) { p[i] = z[i]; rz += r[i] * z[i]; }
  for (it = 1; it <= maxit; ++it) {
    double pq = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : pq)
#endif
    for (int64_t i = 0; i < n; ++i) {
      if (isd[i]) { q[i] = 0.0; continue; }
      double s = 0.0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
        if (!isd[ci[k]]) s -= va[k] * p[ci[k]];
      q[i] = s;
      pq += p[i] * s;
    }
    double alpha = rz / pq;
    rr = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : rr)
#endif
    for (int64_t i = 0; i < n; ++i) {
      if (isd[i]) continue;
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      rr += r[i] * r[i];
    }
    if (sqrt(rr) <= rtol * bn) break;
    PRECOND();
    double rzn = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : rzn)
#endif
    for (int64_t i = 0; i < n; ++i) rzn += r[i] * z[i];
    double beta = rzn / rz;
    rz = rzn;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
#undef PRECOND
  *iters = it > maxit ? maxit : it;
  *relres = sqrt(rr) / bn;
done:
  free(rp); free(ci); free(va); free(isd); free(bb); free(r); free(z); free(p); free(q); free(Dinv);
  return (*relres <= rtol) ? 0 : 1;
}

/* get_u_sigma! (poisson2D_HDG.jl:197-212) with the trace gather written for
   any nt: dofs (f-1)nt+1 .. f nt.  q: ncells x nq, u: ncells x nu,
   uhat_cell: ncells x nt x 3, all column-major as TrialFunction.m_values. */
int orc_recover(int64_t ncells, const int64_t* cells, int k, const double* Ke,
                const double* be, const double* uhat, double* q, double* u,
                double* uhat_cell) {
  const int nu = (k + 1) * (k + 2) / 2, nq = 2 * nu, m = 3 * nu, nt = k + 1, ntr = 3 * nt;
  for (int64_t c = 0; c < ncells; ++c) {
    double ue[48];
    for (int lf = 0; lf < 3; ++lf) {
      int64_t f = cells[6 * c + 3 + lf];
      for (int j = 0; j < nt; ++j) {
        ue[lf * nt + j] = uhat[(f - 1) * nt + j];
        if (uhat_cell) uhat_cell[c + ncells * (j + nt * lf)] = ue[lf * nt + j];
      }
    }
    const double* ke = Ke + (size_t)c * m * ntr;
    const double* bo = be + (size_t)c * m;
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int j = 0; j < ntr; ++j) s += ke[i + m * j] * ue[j];
      s += bo[i];
      if (i < nq) q[c + ncells * i] = s;
      else u[c + ncells * (i - nq)] = s;
    }
  }
  return 0;
}

/* errornorm(u_h, u_ex) -- squared L2 (src/DiscreteFunctions.jl:97-120).
   uex_kind 0: sin(pi x) sin(pi y); 1: c0 + c1 x + c2 y. */
double orc_errornorm(const orc_problem* P, const double* xy,
                     const int64_t* cells, int64_t ncells, const double* u,
                     int uex_kind, const double* uex_par) {
  refel R;
  if (refel_init(&R, P->k, P->qdeg)) return NAN;
  const int nu = R.nu, nqp = R.nqp;
  double Et = 0.0;
  for (int64_t c = 0; c < ncells; ++c) {
    cellgeo G;
    if (reinit(&R, xy, cells, c, &G)) return NAN;
    double El = 0.0;
    for (int q = 0; q < nqp; ++q) {
      double dO = G.detJ * R.qw[q];
      double uh = 0.0;
      for (int i = 0; i < nu; ++i) uh += u[c + ncells * i] * R.N[i * nqp + q];
      double x = 0.0, y = 0.0;
      for (int i = 0; i < 3; ++i) {
        x += R.M[i * nqp + q] * G.x[2 * i];
        y += R.M[i * nqp + q] * G.x[2 * i + 1];
      }
      double ue = uex_kind == 1 ? uex_par[0] + uex_par[1] * x + uex_par[2] * y
                                : sin(PI * x) * sin(PI * y);
      double d = uh - ue;
      El += d * d * dO;
    }
    Et += El;
  }
  return Et;
}
