// Structured triangle mesh generator (host).
//
// Produces exactly the arrays of rectangle_mesh(TriangleCell, (nx,ny), LL, UR)
// (src/generate_mesh.jl:1-18,101-143): row-major nodes, two triangles per
// grid cell (lower-left then upper-right), and faces numbered in order of
// first encounter over local edges (2,3),(3,1),(1,2).  The reference finds
// face ids through a hash map; here the first-encounter order of the sweep
// is written in closed form, so generation is a single O(cells) pass with no
// map.  Node coordinates use the reference's interpolation formula, and the
// file is compiled without FMA contraction so they round identically.
#include <cstdint>

#include "../../include/hdg.h"

extern "C" int64_t hdg_rectangle_nfaces(int64_t nx, int64_t ny) {
  return nx + ny + 3 * nx * ny;
}

namespace {

uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Sweep {
  int64_t nx, ny;
  // ids (1-based) of the faces first created by grid cell (i, j): diagonal,
  // left, bottom of the lower triangle; top, right of the upper triangle.
  int64_t base(int64_t j) const { return j == 1 ? 0 : (4 * nx + 1) + (j - 2) * (3 * nx + 1); }
  int64_t D(int64_t i, int64_t j) const {
    if (i == 1) return base(j) + 1;
    return base(j) + (j == 1 ? 5 + 4 * (i - 2) : 4 + 3 * (i - 2)) + 1;
  }
  int64_t T(int64_t i, int64_t j) const {
    if (j == 1) return i == 1 ? base(j) + 4 : base(j) + 5 + 4 * (i - 2) + 3;
    return i == 1 ? base(j) + 3 : base(j) + 4 + 3 * (i - 2) + 2;
  }
  int64_t R(int64_t i, int64_t j) const {
    if (j == 1) return i == 1 ? base(j) + 5 : base(j) + 5 + 4 * (i - 2) + 4;
    return i == 1 ? base(j) + 4 : base(j) + 4 + 3 * (i - 2) + 3;
  }
  int64_t L(int64_t i, int64_t j) const { return i == 1 ? base(j) + 2 : R(i - 1, j); }
  int64_t B(int64_t i, int64_t j) const {
    if (j > 1) return T(i, j - 1);
    return i == 1 ? base(j) + 3 : base(j) + 5 + 4 * (i - 2) + 2;
  }
};

}  // namespace

extern "C" int hdg_rectangle_mesh(int64_t nx, int64_t ny, double x0, double y0,
                                  double x1, double y1, double jitter,
                                  uint64_t seed, double* xy, int64_t* cells,
                                  int64_t* faces, int8_t* face_side) {
  if (nx < 1 || ny < 1 || !xy || !cells || !faces || !face_side) return HDG_ERR_ARG;
  const int64_t nnx = nx + 1, nny = ny + 1, nf = hdg_rectangle_nfaces(nx, ny);
  const double LL[2] = {x0, y0}, UR[2] = {x1, y1}, LR[2] = {x1, y0}, UL[2] = {x0, y1};
  for (int64_t i = 0; i < nny; ++i) {
    const double rb = double(i) / double(nny - 1);
    const double xa = LL[0] * (1 - rb) + rb * UL[0];
    const double xb = LR[0] * (1 - rb) + rb * UR[0];
    const double ya = LL[1] * (1 - rb) + rb * UL[1];
    const double yb = LR[1] * (1 - rb) + rb * UR[1];
    for (int64_t j = 0; j < nnx; ++j) {
      const double ratio = double(j) / double(nnx - 1);
      xy[2 * (i * nnx + j)] = xa * (1 - ratio) + ratio * xb;
      xy[2 * (i * nnx + j) + 1] = ya * (1 - ratio) + ratio * yb;
    }
  }
  const Sweep S{nx, ny};
  auto node = [&](int64_t i, int64_t j) { return i + (j - 1) * nnx; };
  // _check_node_data (generate_mesh.jl:49-57): a triangle whose vertices
  // run clockwise (a box with one side given in decreasing order) has its
  // vertices 2 and 3 swapped.  The sign is evaluated as the reference
  // evaluates it, so that degenerate boxes decide the same way.
  auto clockwise = [&](int64_t n1, int64_t n2, int64_t n3) {
    const double* p1 = xy + 2 * (n1 - 1);
    const double* p2 = xy + 2 * (n2 - 1);
    const double* p3 = xy + 2 * (n3 - 1);
    const double ax = p2[0] - p1[0], ay = p2[1] - p1[1];
    const double bx = p3[0] - p1[0], by = p3[1] - p1[1];
    return (ax * by - ay * bx) < 0;
  };
  auto setf = [&](int64_t f, int64_t v1, int64_t v2, int64_t c1, int64_t c2, int8_t side) {
    faces[(f - 1)] = v1;
    faces[(f - 1) + nf] = v2;
    faces[(f - 1) + 2 * nf] = c1;
    faces[(f - 1) + 3 * nf] = c2;
    face_side[f - 1] = side;
  };
  for (int64_t j = 1; j <= ny; ++j)
    for (int64_t i = 1; i <= nx; ++i) {
      const int64_t lo = 2 * ((j - 1) * nx + (i - 1)) + 1, up = lo + 1;
      const int64_t a = node(i, j), b = node(i + 1, j), c = node(i, j + 1), d = node(i + 1, j + 1);
      const bool sl = clockwise(a, b, c), su = clockwise(b, d, c);
      // With vertices 2 and 3 swapped the local edges (2,3),(3,1),(1,2)
      // meet the faces in another order: lower (D, B, L) instead of
      // (D, L, B), upper (T, R, D) instead of (T, D, R); every face is
      // stored in the reverse traversal.  The ids only change where the
      // lower triangle of grid cell (1, 1) creates both L and B.
      int64_t fL = S.L(i, j), fB = S.B(i, j);
      if (sl && i == 1 && j == 1) {
        const int64_t t = fL;
        fL = fB;
        fB = t;
      }
      const int64_t fD = S.D(i, j), fT = S.T(i, j), fR = S.R(i, j);
      int64_t* cl = cells + 6 * (lo - 1);
      int64_t* cu = cells + 6 * (up - 1);
      if (!sl) {
        cl[0] = a; cl[1] = b; cl[2] = c;
        cl[3] = fD; cl[4] = fL; cl[5] = fB;
      } else {
        cl[0] = a; cl[1] = c; cl[2] = b;
        cl[3] = fD; cl[4] = fB; cl[5] = fL;
      }
      if (!su) {
        cu[0] = b; cu[1] = d; cu[2] = c;
        cu[3] = fT; cu[4] = fD; cu[5] = fR;
      } else {
        cu[0] = b; cu[1] = c; cu[2] = d;
        cu[3] = fT; cu[4] = fR; cu[5] = fD;
      }
      // faces first met by this grid cell, with the first cell's traversal
      if (!sl) setf(fD, b, c, lo, up, 0); else setf(fD, c, b, lo, up, 0);
      if (i == 1) {
        if (!sl) setf(fL, c, a, lo, 0, 4); else setf(fL, a, c, lo, 0, 4);
      }
      if (j == 1) {
        if (!sl) setf(fB, a, b, lo, 0, 1); else setf(fB, b, a, lo, 0, 1);
      }
      const int64_t cT = j < ny ? 2 * (j * nx + (i - 1)) + 1 : 0;
      const int64_t cR = i < nx ? 2 * ((j - 1) * nx + i) + 1 : 0;
      if (!su) setf(fT, d, c, up, cT, j < ny ? 0 : 3); else setf(fT, c, d, up, cT, j < ny ? 0 : 3);
      if (!su) setf(fR, b, d, up, cR, i < nx ? 0 : 2); else setf(fR, d, b, up, cR, i < nx ? 0 : 2);
    }
  // boundary sets are found as the reference finds them: by exact equality
  // of both end points with a side of the box, bottom, right, top, left in
  // that order; a one-cell face on none of them is an error
  // (_get_rectangular_boundary_sets, generate_mesh.jl:60-89)
  for (int64_t f = 0; f < nf; ++f) {
    if (faces[f + 3 * nf] != 0) {
      face_side[f] = 0;
      continue;
    }
    const double* p1 = xy + 2 * (faces[f] - 1);
    const double* p2 = xy + 2 * (faces[f + nf] - 1);
    int8_t side = 0;
    if (p1[1] == LL[1] && p2[1] == LL[1]) side = 1;
    else if (p1[0] == UR[0] && p2[0] == UR[0]) side = 2;
    else if (p1[1] == UR[1] && p2[1] == UR[1]) side = 3;
    else if (p1[0] == LL[0] && p2[0] == LL[0]) side = 4;
    else return HDG_ERR_BOUNDARY;
    face_side[f] = side;
  }
  if (jitter > 0.0) {
    const double hx = (x1 - x0) / double(nx), hy = (y1 - y0) / double(ny);
    for (int64_t j = 1; j < ny; ++j)
      for (int64_t i = 1; i < nx; ++i) {
        const uint64_t n = uint64_t(i + j * nnx);
        const double ux = double(splitmix64(seed + 2 * n) >> 11) * (1.0 / 9007199254740992.0);
        const double uy = double(splitmix64(seed + 2 * n + 1) >> 11) * (1.0 / 9007199254740992.0);
        xy[2 * n] = xy[2 * n] + jitter * hx * (2.0 * ux - 1.0);
        xy[2 * n + 1] = xy[2 * n + 1] + jitter * hy * (2.0 * uy - 1.0);
      }
  }
  return HDG_OK;
}
