// Reference-element tables (host).  See hdg_tables.h.
#include "hdg_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace hdg {

namespace {

// Jacobi polynomials P_n^{(a,0)}(x), n = 0..nmax, and d/dx, by the
// three-term recurrence (Abramowitz & Stegun 22.7.1 with beta = 0).
void jacobi_a0(int nmax, double a, double x, double* P, double* dP) {
  P[0] = 1.0;
  dP[0] = 0.0;
  if (nmax >= 1) {
    P[1] = 0.5 * ((a + 2.0) * x + a);
    dP[1] = 0.5 * (a + 2.0);
  }
  for (int n = 2; n <= nmax; ++n) {
    const double c = 2.0 * n + a;
    const double a1 = 2.0 * n * (n + a) * (c - 2.0);
    const double a2 = (c - 1.0) * a * a;
    const double a3 = (c - 2.0) * (c - 1.0) * c;
    const double a4 = 2.0 * (n + a - 1.0) * (n - 1.0) * c;
    P[n] = ((a2 + a3 * x) * P[n - 1] - a4 * P[n - 2]) / a1;
    dP[n] = ((a2 + a3 * x) * dP[n - 1] + a3 * P[n - 1] - a4 * dP[n - 2]) / a1;
  }
}

}  // namespace

// psi_{p,q}(r,s) = sqrt(2(2p+1)(p+q+1)) g_p(r,s) P_q^{(2p+1,0)}(2s-1), where
// g_p = P_p(xi) (1-s)^p is generated without the collapsed-coordinate
// singularity by the homogeneous Legendre recurrence in a = 2r+s-1, b = 1-s:
// (p+1) g_{p+1} = (2p+1) a g_p - p b^2 g_{p-1}.  Index order: total degree
// d = 0,1,..,k and within a degree p = d, d-1, .., 0 (src/basis.jl:211-218).
void dubiner_eval(int k, double r, double s, double* val, double* grad) {
  const double a = 2.0 * r + s - 1.0, b = 1.0 - s;
  double g[16], gr[16], gs[16];
  g[0] = 1.0; gr[0] = 0.0; gs[0] = 0.0;
  if (k >= 1) { g[1] = a; gr[1] = 2.0; gs[1] = 1.0; }
  for (int p = 1; p < k; ++p) {
    const double c1 = (2.0 * p + 1.0) / (p + 1.0), c2 = double(p) / (p + 1.0);
    g[p + 1] = c1 * a * g[p] - c2 * b * b * g[p - 1];
    gr[p + 1] = c1 * (2.0 * g[p] + a * gr[p]) - c2 * b * b * gr[p - 1];
    gs[p + 1] = c1 * (g[p] + a * gs[p]) - c2 * (-2.0 * b * g[p - 1] + b * b * gs[p - 1]);
  }
  int j = 0;
  for (int d = 0; d <= k; ++d)
    for (int p = d; p >= 0; --p) {
      const int q = d - p;
      double P[16], dP[16];
      jacobi_a0(q, 2.0 * p + 1.0, 2.0 * s - 1.0, P, dP);
      const double nrm = std::sqrt(2.0 * (2.0 * p + 1.0) * (p + q + 1.0));
      val[j] = nrm * g[p] * P[q];
      if (grad) {
        grad[2 * j] = nrm * gr[p] * P[q];
        grad[2 * j + 1] = nrm * (gs[p] * P[q] + g[p] * 2.0 * dP[q]);
      }
      ++j;
    }
}

void legendre01_eval(int nt, double x, double* val) {
  const double t = 2.0 * x - 1.0;
  double p0 = 1.0, p1 = t;
  for (int i = 0; i < nt; ++i) {
    double pi;
    if (i == 0) pi = 1.0;
    else if (i == 1) pi = t;
    else {
      pi = ((2.0 * i - 1.0) * t * p1 - (i - 1.0) * p0) / i;
      p0 = p1;
      p1 = pi;
    }
    val[i] = std::sqrt(2.0 * i + 1.0) * pi;
  }
}

void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= n; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int j = 2; j <= n; ++j) {
      const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) p0 = 1.0;
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    if (2 * i + 1 == n) z = 0.0;
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    // ascending on [-1,1], then mapped to [0,1]
    x[i] = (1.0 - z) / 2.0;
    x[n - 1 - i] = (z + 1.0) / 2.0;
    w[i] = w[n - 1 - i] = 0.5 * wi;
  }
}

namespace {

// Strang-Fix rules: (point, weight/2) for orders 1..6, the published table
// values the reference uses (src/StrangQuad.jl:1-61).
struct StrangPt { double r, s, w; };
const StrangPt kStrang1[] = {{1.0 / 3.0, 1.0 / 3.0, 1.0}};
const StrangPt kStrang2[] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0},
                             {1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0},
                             {2.0 / 3.0, 1.0 / 6.0, 1.0 / 3.0}};
const StrangPt kStrang3[] = {{0.659027622374092, 0.231933368553031, 1.0 / 6.0},
                             {0.659027622374092, 0.109039009072877, 1.0 / 6.0},
                             {0.231933368553031, 0.659027622374092, 1.0 / 6.0},
                             {0.231933368553031, 0.109039009072877, 1.0 / 6.0},
                             {0.109039009072877, 0.659027622374092, 1.0 / 6.0},
                             {0.109039009072877, 0.231933368553031, 1.0 / 6.0}};
const StrangPt kStrang4[] = {{0.816847572980459, 0.091576213509771, 0.109951743655322},
                             {0.091576213509771, 0.816847572980459, 0.109951743655322},
                             {0.091576213509771, 0.091576213509771, 0.109951743655322},
                             {0.108103018168070, 0.445948490915965, 0.223381589678011},
                             {0.445948490915965, 0.108103018168070, 0.223381589678011},
                             {0.445948490915965, 0.445948490915965, 0.223381589678011}};
const StrangPt kStrang5[] = {{0.33333333333333333, 0.33333333333333333, 0.22500000000000000},
                             {0.79742698535308720, 0.10128650732345633, 0.12593918054482717},
                             {0.10128650732345633, 0.79742698535308720, 0.12593918054482717},
                             {0.10128650732345633, 0.10128650732345633, 0.12593918054482717},
                             {0.05971587178976981, 0.47014206410511505, 0.13239415278850616},
                             {0.47014206410511505, 0.05971587178976981, 0.13239415278850616},
                             {0.47014206410511505, 0.47014206410511505, 0.13239415278850616}};
const StrangPt kStrang6[] = {{0.873821971016996, 0.063089014491502, 0.050844906370207},
                             {0.063089014491502, 0.873821971016996, 0.050844906370207},
                             {0.063089014491502, 0.063089014491502, 0.050844906370207},
                             {0.501426509658179, 0.249286745170910, 0.116786275726379},
                             {0.249286745170910, 0.501426509658179, 0.116786275726379},
                             {0.249286745170910, 0.249286745170910, 0.116786275726379},
                             {0.636502499121399, 0.310352451033785, 0.082851075618374},
                             {0.636502499121399, 0.053145049844816, 0.082851075618374},
                             {0.310352451033785, 0.636502499121399, 0.082851075618374},
                             {0.310352451033785, 0.053145049844816, 0.082851075618374},
                             {0.053145049844816, 0.636502499121399, 0.082851075618374},
                             {0.053145049844816, 0.310352451033785, 0.082851075618374}};

template <size_t N>
void take(const StrangPt (&t)[N], std::vector<double>& p, std::vector<double>& w) {
  p.clear();
  w.clear();
  for (const auto& q : t) {
    p.push_back(q.r);
    p.push_back(q.s);
    w.push_back(q.w * 0.5);
  }
}

double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

}  // namespace

// Grundmann & Moeller, SIAM J. Numer. Anal. 15 (1978), Theorem 4, on the
// 2-simplex: degree 2s+1, points (2 beta + 1)/(d + 2 - 2i) over all
// compositions |beta| = s - i, weights (-1)^i 2^-2s (d+2-2i)^d /
// (i! (d+2-i)!), normalised to the reference area 1/2.
bool triangle_rule(int order, std::vector<double>& p, std::vector<double>& w) {
  switch (order) {
    case 0:
    case 1: take(kStrang1, p, w); return true;
    case 2: take(kStrang2, p, w); return true;
    case 3: take(kStrang3, p, w); return true;
    case 4: take(kStrang4, p, w); return true;
    case 5: take(kStrang5, p, w); return true;
    case 6: take(kStrang6, p, w); return true;
    default: break;
  }
  if (order < 0 || order % 2 == 0) return false;
  const int s = (order - 1) / 2, d = 2 * s + 1;
  p.clear();
  w.clear();
  double sum = 0.0;
  for (int i = 0; i <= s; ++i) {
    const int den = d + 2 - 2 * i;
    const double wi = ((i % 2) ? -1.0 : 1.0) * std::ldexp(1.0, -2 * s) * std::pow(double(den), d) /
                      (factorial(i) * factorial(d + 2 - i));
    const int t = s - i;
    for (int b1 = 0; b1 <= t; ++b1)
      for (int b2 = 0; b1 + b2 <= t; ++b2) {
        p.push_back((2.0 * b1 + 1.0) / den);
        p.push_back((2.0 * b2 + 1.0) / den);
        w.push_back(wi);
        sum += wi;
      }
  }
  for (auto& x : w) x = 0.5 * x / sum;
  return true;
}

bool RefTables::build(int k_, int qdeg_) {
  k = k_;
  qdeg = qdeg_;
  if (!triangle_rule(qdeg, qp, qw)) return false;
  gauss_legendre01(qdeg, fx, fw);
  const int nu = (k + 1) * (k + 2) / 2, nt = k + 1;
  const int nqp = (int)qw.size(), nfq = (int)fw.size();
  TableLayout& L = lay;
  L.nu = nu; L.nt = nt; L.nqp = nqp; L.nfq = nfq;
  int o = 0;
  L.Mref = o; o += nu * nu;
  L.Dref = o; o += 2 * nu * nu;
  L.Cref = o; o += 3 * nu * nu;
  L.G = o; o += 6 * nu * nt;
  L.Href = o; o += nt * nt;
  L.Nw = o; o += nu * nqp;
  L.N = o; o += nu * nqp;
  L.Mg = o; o += 3 * nqp;
  L.W = o; o += nqp;
  L.total = o;
  data.assign(o, 0.0);

  std::vector<double> N(nu * nqp), dN(2 * nu * nqp), v(nu), g(2 * nu);
  for (int q = 0; q < nqp; ++q) {
    dubiner_eval(k, qp[2 * q], qp[2 * q + 1], v.data(), g.data());
    for (int i = 0; i < nu; ++i) {
      N[i * nqp + q] = v[i];
      dN[(i * nqp + q) * 2] = g[2 * i];
      dN[(i * nqp + q) * 2 + 1] = g[2 * i + 1];
    }
  }
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nu; ++j) {
      double m = 0.0, d0 = 0.0, d1 = 0.0;
      for (int q = 0; q < nqp; ++q) {
        m += qw[q] * N[i * nqp + q] * N[j * nqp + q];
        d0 += qw[q] * dN[(i * nqp + q) * 2] * N[j * nqp + q];
        d1 += qw[q] * dN[(i * nqp + q) * 2 + 1] * N[j * nqp + q];
      }
      data[L.Mref + i * nu + j] = m;
      data[L.Dref + i * nu + j] = d0;
      data[L.Dref + nu * nu + i * nu + j] = d1;
    }
  for (int i = 0; i < nu; ++i)
    for (int q = 0; q < nqp; ++q) {
      data[L.Nw + i * nqp + q] = qw[q] * N[i * nqp + q];
      data[L.N + i * nqp + q] = N[i * nqp + q];
    }
  for (int q = 0; q < nqp; ++q) {
    const double r = qp[2 * q], s = qp[2 * q + 1];
    data[L.Mg + 0 * nqp + q] = 1.0 - r - s;
    data[L.Mg + 1 * nqp + q] = r;
    data[L.Mg + 2 * nqp + q] = s;
    data[L.W + q] = qw[q];
  }
  // face tables: reference edges in CCW order, edge f opposite vertex f,
  // traversed from its first to its second local vertex
  // ((2,3),(3,1),(1,2) of src/shapes.jl:19-24).
  const double ev[3][2][2] = {{{1.0, 0.0}, {0.0, 1.0}}, {{0.0, 1.0}, {0.0, 0.0}},
                              {{0.0, 0.0}, {1.0, 0.0}}};
  std::vector<double> E(3 * nu * nfq), Nt(nt * nfq), lt(nt);
  for (int q = 0; q < nfq; ++q) {
    legendre01_eval(nt, fx[q], lt.data());
    for (int t = 0; t < nt; ++t) Nt[t * nfq + q] = lt[t];
    for (int f = 0; f < 3; ++f) {
      const double r = (1.0 - fx[q]) * ev[f][0][0] + fx[q] * ev[f][1][0];
      const double s = (1.0 - fx[q]) * ev[f][0][1] + fx[q] * ev[f][1][1];
      dubiner_eval(k, r, s, v.data(), nullptr);
      for (int i = 0; i < nu; ++i) E[(f * nu + i) * nfq + q] = v[i];
    }
  }
  for (int f = 0; f < 3; ++f) {
    for (int i = 0; i < nu; ++i)
      for (int j = 0; j < nu; ++j) {
        double c = 0.0;
        for (int q = 0; q < nfq; ++q) c += fw[q] * E[(f * nu + i) * nfq + q] * E[(f * nu + j) * nfq + q];
        data[L.Cref + (f * nu + i) * nu + j] = c;
      }
    // orientation o = 1: the trace parameter runs with the local edge;
    // o = 0: against it, so the cell basis is read at the mirrored point
    // (src/ScalarFunctionSpaces.jl:153).
    for (int o = 0; o < 2; ++o)
      for (int i = 0; i < nu; ++i)
        for (int t = 0; t < nt; ++t) {
          double gsum = 0.0;
          for (int q = 0; q < nfq; ++q) {
            const int qo = o ? q : nfq - 1 - q;
            gsum += fw[q] * E[(f * nu + i) * nfq + qo] * Nt[t * nfq + q];
          }
          data[L.G + ((f * 2 + o) * nu + i) * nt + t] = gsum;
        }
  }
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      double h = 0.0;
      for (int q = 0; q < nfq; ++q) h += fw[q] * Nt[i * nfq + q] * Nt[j * nfq + q];
      data[L.Href + i * nt + j] = h;
    }
  return true;
}

bool RefTables::mass_spd() const {
  for (double w : qw)
    if (!(w > 0.0)) return false;
  const int n = lay.nu;
  std::vector<double> a(data.begin() + lay.Mref, data.begin() + lay.Mref + (size_t)n * n);
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, a[i * n + i]);
  if (!(dmax > 0.0)) return false;
  for (int j = 0; j < n; ++j) {  // Cholesky, lower triangle in place
    double d = a[j * n + j];
    for (int t = 0; t < j; ++t) d -= a[j * n + t] * a[j * n + t];
    if (!(d > 1e-10 * dmax)) return false;
    const double l = std::sqrt(d);
    a[j * n + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double v = a[i * n + j];
      for (int t = 0; t < j; ++t) v -= a[i * n + t] * a[j * n + t];
      a[i * n + j] = v / l;
    }
  }
  return true;
}

bool build_spd_tables(const RefTables& rt, SpdLayout& L, std::vector<double>& data) {
  if (!rt.mass_spd()) return false;
  const TableLayout& T = rt.lay;
  const int nu = T.nu, nt = T.nt, nqp = T.nqp;
  const double* d = rt.data.data();
  // The derived tables enter every cell alike, so their rounding errors do
  // not average out over the mesh: they perturb the constant-trace identity
  // Ate 1 = 0 of all cells the same way, an error the trace solve amplifies
  // by cond(K) ~ h^-2.  They are therefore formed in extended precision
  // from the base tables and rounded once.
  using X = long double;
  // Mi = Mref^-1 by Gauss-Jordan with partial pivoting (Mref is SPD and
  // close to the identity for an orthonormal basis)
  std::vector<X> a(d + T.Mref, d + T.Mref + nu * nu), mi(nu * nu, 0.0L);
  for (int i = 0; i < nu; ++i) mi[i * nu + i] = 1.0L;
  for (int j = 0; j < nu; ++j) {
    int p = j;
    for (int i = j + 1; i < nu; ++i)
      if (std::fabs(a[i * nu + j]) > std::fabs(a[p * nu + j])) p = i;
    for (int c = 0; c < nu; ++c) {
      std::swap(a[j * nu + c], a[p * nu + c]);
      std::swap(mi[j * nu + c], mi[p * nu + c]);
    }
    const X piv = a[j * nu + j];
    for (int c = 0; c < nu; ++c) {
      a[j * nu + c] /= piv;
      mi[j * nu + c] /= piv;
    }
    for (int i = 0; i < nu; ++i) {
      if (i == j) continue;
      const X f = a[i * nu + j];
      for (int c = 0; c < nu; ++c) {
        a[i * nu + c] -= f * a[j * nu + c];
        mi[i * nu + c] -= f * mi[j * nu + c];
      }
    }
  }
  auto D = [&](int dd, int i, int j) { return (X)d[T.Dref + dd * nu * nu + i * nu + j]; };
  auto Gm = [&](int fo, int i, int t) { return (X)d[T.G + (fo * nu + i) * nt + t]; };
  L.nu = nu;
  L.nt = nt;
  L.nqp = nqp;
  int o = 0;
  L.R = o; o += 3 * nu * nu;
  L.C = o; o += 3 * nu * nu;
  L.Gam = o; o += 6 * nu * nt;
  L.Z = o; o += 12 * nu * nt;
  L.V = o; o += 6 * nu * nt;
  L.P = o; o += 2 * nu * nu;
  L.Om = o; o += 36 * nt * nt;
  L.H = o; o += nt * nt;
  L.Nw = o; o += nu * nqp;
  L.Mg = o; o += 3 * nqp;
  L.total = o;
  data.assign(o, 0.0);
  // P_d = Mi D_d, V_fo = Mi Gam_fo
  std::vector<X> P(2 * nu * nu), V(6 * nu * nt);
  for (int dd = 0; dd < 2; ++dd)
    for (int i = 0; i < nu; ++i)
      for (int j = 0; j < nu; ++j) {
        X v = 0.0L;
        for (int t = 0; t < nu; ++t) v += mi[i * nu + t] * D(dd, t, j);
        P[(dd * nu + i) * nu + j] = v;
        data[L.P + (dd * nu + i) * nu + j] = (double)v;
      }
  for (int fo = 0; fo < 6; ++fo)
    for (int i = 0; i < nu; ++i)
      for (int t = 0; t < nt; ++t) {
        X v = 0.0L;
        for (int j = 0; j < nu; ++j) v += mi[i * nu + j] * Gm(fo, j, t);
        V[(fo * nu + i) * nt + t] = v;
        data[L.V + (fo * nu + i) * nt + t] = (double)v;
      }
  // R_de = D_d' Mi D_e = D_d' P_e (symmetric up to rounding: symmetrised)
  for (int i = 0; i < nu; ++i)
    for (int j = i; j < nu; ++j) {
      X r[2][2][2];  // [d][e][(i,j) or (j,i)]
      for (int dd = 0; dd < 2; ++dd)
        for (int e = 0; e < 2; ++e) {
          X v = 0.0L, w = 0.0L;
          for (int t = 0; t < nu; ++t) {
            v += D(dd, t, i) * P[(e * nu + t) * nu + j];
            w += D(dd, t, j) * P[(e * nu + t) * nu + i];
          }
          r[dd][e][0] = v;
          r[dd][e][1] = w;
        }
      const X m0 = 0.5L * (r[0][0][0] + r[0][0][1]);
      const X m1 = 0.5L * (r[0][1][0] + r[1][0][0] + r[0][1][1] + r[1][0][1]);
      const X m2 = 0.5L * (r[1][1][0] + r[1][1][1]);
      const X mm[3] = {m0, m1, m2};
      for (int m = 0; m < 3; ++m)
        data[L.R + (m * nu + i) * nu + j] = data[L.R + (m * nu + j) * nu + i] = (double)mm[m];
    }
  for (int x = 0; x < 3 * nu * nu; ++x) data[L.C + x] = d[T.Cref + x];
  for (int x = 0; x < 6 * nu * nt; ++x) data[L.Gam + x] = d[T.G + x];
  // Z_dfo = D_d' V_fo
  for (int dd = 0; dd < 2; ++dd)
    for (int fo = 0; fo < 6; ++fo)
      for (int i = 0; i < nu; ++i)
        for (int t = 0; t < nt; ++t) {
          X v = 0.0L;
          for (int j = 0; j < nu; ++j) v += D(dd, j, i) * V[(fo * nu + j) * nt + t];
          data[L.Z + ((dd * 6 + fo) * nu + i) * nt + t] = (double)v;
        }
  // Om_fo,gp = Gam_fo' Mi Gam_gp, exactly symmetric
  for (int fo = 0; fo < 6; ++fo)
    for (int gp = fo; gp < 6; ++gp)
      for (int t = 0; t < nt; ++t)
        for (int u = 0; u < nt; ++u) {
          if (gp == fo && u < t) continue;
          X v = 0.0L, w = 0.0L;
          for (int i = 0; i < nu; ++i) {
            v += Gm(fo, i, t) * V[(gp * nu + i) * nt + u];
            w += Gm(gp, i, u) * V[(fo * nu + i) * nt + t];
          }
          const double s = (double)(0.5L * (v + w));
          data[L.Om + ((fo * 6 + gp) * nt + t) * nt + u] = s;
          data[L.Om + ((gp * 6 + fo) * nt + u) * nt + t] = s;
        }
  for (int x = 0; x < nt * nt; ++x) data[L.H + x] = d[T.Href + x];
  for (int x = 0; x < nu * nqp; ++x) data[L.Nw + x] = d[T.Nw + x];
  for (int x = 0; x < 3 * nqp; ++x) data[L.Mg + x] = d[T.Mg + x];
  return true;
}

}  // namespace hdg
