// Device helpers shared by the element kernels: the reference's sources
// and the affine geometry of a triangle.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "hdg_internal.h"

namespace hdg {

constexpr double kPi = 3.14159265358979311600;

// ------------------------------------------------------------- sources
struct Src {
  __device__ __forceinline__ static double kappa(const SrcParams& S, double x, double y) {
    return S.kappa_on ? 1.0 + 0.5 * sin(S.ka * x) * cos(S.kb * y) : 1.0;
  }
  __device__ __forceinline__ static double f(const SrcParams& S, double x, double y) {
    switch (S.kind) {
      case 0: return 2.0 * (kPi * kPi) * sin(kPi * x) * sin(kPi * y);
      case 1: {
        double sx, cx, sy, cy;
        sincos(kPi * x, &sx, &cx);
        sincos(kPi * y, &sy, &cy);
        const double u = sx * sy, ux = kPi * cx * sy, uy = kPi * sx * cy;
        double k = 1.0, kx = 0.0, ky = 0.0;
        if (S.kappa_on) {
          double sa, ca, sb, cb;
          sincos(S.ka * x, &sa, &ca);
          sincos(S.kb * y, &sb, &cb);
          k = 1.0 + 0.5 * sa * cb;
          kx = 0.5 * S.ka * ca * cb;
          ky = -0.5 * S.kb * sa * sb;
        }
        return k * 2.0 * kPi * kPi * u - (kx * ux + ky * uy);
      }
      case 2: {
        double sxp[4], syq[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          sxp[p] = sin((p + 1) * kPi * x);
          syq[p] = sin((p + 1) * kPi * y);
        }
        double v = 0.0;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) v += S.coef[p * 4 + q] * sxp[p] * syq[q];
        return v;
      }
      default: return 0.0;
    }
  }
};

// ------------------------------------------------------------- geometry
// reinit! (src/ScalarFunctionSpaces.jl:101-132) for an affine triangle.
struct Geo {
  double x[6];
  double detJ;
  double Ji[4];   // J^-1 column-major: Ji[a + 2b]
  double nrm[6];  // unit outward normals of local faces 0..2
  double len[3];  // edge lengths (detJf)
  int ori[3];     // face_orientation (src/mesh.jl:51-54)
};

__device__ __forceinline__ void cell_geometry(const DevMesh& m, int64_t c, Geo& g) {
  int nd[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    nd[v] = __ldg(m.cn + 3 * c + v);
    g.x[2 * v] = __ldg(m.xy + 2 * (int64_t)nd[v]);
    g.x[2 * v + 1] = __ldg(m.xy + 2 * (int64_t)nd[v] + 1);
  }
  const double J0 = g.x[2] - g.x[0], J1 = g.x[3] - g.x[1];  // column 1
  const double J2 = g.x[4] - g.x[0], J3 = g.x[5] - g.x[1];  // column 2
  g.detJ = J0 * J3 - J2 * J1;
  const double di = 1.0 / g.detJ;
  g.Ji[0] = J3 * di;
  g.Ji[1] = -J1 * di;
  g.Ji[2] = -J2 * di;
  g.Ji[3] = J0 * di;
  const double wn[6] = {-(J1 - J3), J0 - J2, -J3, J2, J1, -J0};  // src/shapes.jl:80-87
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const double l = sqrt(wn[2 * f] * wn[2 * f] + wn[2 * f + 1] * wn[2 * f + 1]);
    g.len[f] = l;
    g.nrm[2 * f] = wn[2 * f] / l;
    g.nrm[2 * f + 1] = wn[2 * f + 1] / l;
  }
  g.ori[0] = nd[2] > nd[1];
  g.ori[1] = nd[0] > nd[2];
  g.ori[2] = nd[1] > nd[0];
}

}  // namespace hdg
