// Reference-element tables for the HDG element kernel (host side).
//
// Everything the element kernel needs that does not depend on the cell is
// folded into small reference matrices here, once per (k, quad degree):
// the per-cell blocks of doassemble (examples/poisson2D_HDG.jl:88-153) are
// affine-invariant combinations of them with the cell's detJ, J^-1, edge
// lengths and normals.
#pragma once
#include <vector>

namespace hdg {

// Orthonormal Dubiner basis on the reference triangle (0,0),(1,0),(0,1)
// (the basis of src/basis.jl:65-86), values val[nu] and gradients
// grad[2*nu] at (r, s), through the collapsed-coordinate recursion.
void dubiner_eval(int k, double r, double s, double* val, double* grad);
// Orthonormal Legendre basis on [0,1] (src/basis.jl:351-354), nt functions.
void legendre01_eval(int nt, double x, double* val);
// DefaultQuad on the triangle (src/quadrature.jl:17-26): Strang-Fix for
// order <= 6, Grundmann-Moeller for odd order > 6.  Returns false if the
// order is unavailable.
bool triangle_rule(int order, std::vector<double>& pts, std::vector<double>& w);
// Gauss-Legendre with n points on [0,1] (src/quadrature.jl:29-39).
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w);

// Offsets of the packed table buffer uploaded to the device (doubles).
struct TableLayout {
  int nu, nt, nqp, nfq;
  int Mref;  // [nu][nu]       sum_q w N_i N_j                     (A / detJ)
  int Dref;  // [2][nu][nu]    sum_q w dN_i/dxi_d N_j               (B)
  int Cref;  // [3][nu][nu]    sum_q w E^f_i E^f_j                 (C / (tau |e_f|))
  int G;     // [3][2][nu][nt] sum_q w E^f_i(q or flipped) Nt_t(q) (E, F)
  int Href;  // [nt][nt]       sum_q w Nt_i Nt_j                   (H / |e_f|)
  int Nw;    // [nu][nqp]      w_q N_i(q)                          (be, kappa A)
  int N;     // [nu][nqp]      N_i(q)
  int Mg;    // [3][nqp]       P1 geometry basis at cell points    (x_q)
  int W;     // [nqp]          cell quadrature weights
  int total;
};

struct RefTables {
  int k = 0, qdeg = 0;
  TableLayout lay{};
  std::vector<double> data;  // packed per TableLayout
  std::vector<double> qp, qw, fx, fw;
  // builds the tables; returns false when the quadrature order is
  // unavailable (src/quadrature.jl:24)
  bool build(int k, int qdeg);
  // True when every cell weight is positive and the reference mass
  // sum_q w N_i N_j is positive definite (Cholesky pivots above 1e-10 of
  // the largest diagonal entry): then the q block A = (kappa^-1 v_j, v_i)
  // of every cell with det J > 0 and kappa > 0 is symmetric positive
  // definite, and so is the Schur complement C + B^T A^-1 B of a
  // nonsingular local matrix.  False for the default rules at k = 2
  // (6 points on a conic) and k = 4 (12 points for 15 functions).
  bool mass_spd() const;
};

// Tables of the condensation for kappa = 1 and an SPD reference mass
// (RefTables::mass_spd).  The q block of an affine cell is then
// A = detJ diag(Mref, Mref), so every product through A^-1 in the Schur
// complement S = C + B' A^-1 B, in R = F + B' A^-1 E, in E' A^-1 E and in
// the recovered q is a combination of constant matrices with a few scalars
// of the cell's geometry (Mi = Mref^-1, D_d = Dref[d], Gam = G):
struct SpdLayout {
  int nu, nt, nqp;
  int R;    // [3][nu][nu]        D0' Mi D0, D0' Mi D1 + D1' Mi D0, D1' Mi D1
  int C;    // [3][nu][nu]        Cref
  int Gam;  // [3][2][nu][nt]     G (face f, orientation o)
  int Z;    // [2][3][2][nu][nt]  D_d' Mi Gam_fo
  int V;    // [3][2][nu][nt]     Mi Gam_fo
  int P;    // [2][nu][nu]        Mi D_d
  int Om;   // [6][6][nt][nt]     Gam_fo' Mi Gam_gp, exactly symmetric under (fo,t) <-> (gp,s)
  int H;    // [nt][nt]           Href
  int Nw;   // [nu][nqp]          w_q N_i(q)
  int Mg;   // [3][nqp]           P1 geometry basis at the cell points
  int total;
};
// Builds them from rt; false when the reference mass is not SPD.
bool build_spd_tables(const RefTables& rt, SpdLayout& L, std::vector<double>& data);

}  // namespace hdg
