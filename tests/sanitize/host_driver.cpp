// Runs libhdg's host-only code for the AddressSanitizer /
// UndefinedBehaviorSanitizer build made by tests/test_sanitizers.py: the
// rectangle mesh generator (regular, one-cell, thin and inverted boxes, the
// argument checks), the reference tables of every (k, quad degree) with the
// SPD test and the Schur-form tables, the bases and rules, and the serving
// arena's free list under a random take/give sequence checked against a
// shadow map.  Exits non-zero on a failed check.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "hdg.h"
#include "hdg_arena.h"
#include "hdg_tables.h"

#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      std::exit(1);                                                           \
    }                                                                         \
  } while (0)

static void mesh(int64_t nx, int64_t ny, double x0, double y0, double x1, double y1) {
  const int64_t nf = hdg_rectangle_nfaces(nx, ny), nn = (nx + 1) * (ny + 1), nc = 2 * nx * ny;
  CHECK(nf == nx + ny + 3 * nx * ny);
  std::vector<double> xy(2 * nn);
  std::vector<int64_t> cells(6 * nc), faces(4 * nf);
  std::vector<int8_t> side(nf);
  const int st = hdg_rectangle_mesh(nx, ny, x0, y0, x1, y1, 0.0, 1, xy.data(), cells.data(), faces.data(),
                                    side.data());
  CHECK(st == HDG_OK || st == HDG_ERR_BOUNDARY);
  if (st != HDG_OK) return;
  for (int64_t c = 0; c < nc; ++c)
    for (int j = 0; j < 3; ++j) {
      CHECK(cells[6 * c + j] >= 1 && cells[6 * c + j] <= nn);
      CHECK(cells[6 * c + 3 + j] >= 1 && cells[6 * c + 3 + j] <= nf);
    }
  int64_t nb = 0;
  for (int64_t f = 0; f < nf; ++f) nb += side[f] != 0;
  CHECK(nb == 2 * (nx + ny));
}

static void tables() {
  for (int k = 1; k <= 4; ++k) {
    std::vector<double> v((k + 1) * (k + 2) / 2), g(2 * v.size()), lt(k + 1);
    hdg::dubiner_eval(k, 0.25, 0.5, v.data(), g.data());
    hdg::legendre01_eval(k + 1, 0.3, lt.data());
    for (int q = 1; q <= 9; ++q) {
      hdg::RefTables rt;
      if (!rt.build(k, q)) continue;
      CHECK((int)rt.data.size() >= rt.lay.total);
      for (double x : rt.data) CHECK(std::isfinite(x));
      if (k <= 3 && rt.mass_spd()) {
        hdg::SpdLayout L{};
        std::vector<double> d;
        CHECK(hdg::build_spd_tables(rt, L, d));
        CHECK((int)d.size() >= L.total);
        for (double x : d) CHECK(std::isfinite(x));
      }
    }
  }
  for (int o = 1; o <= 13; ++o) {
    std::vector<double> p, w;
    if (hdg::triangle_rule(o, p, w)) {
      double s = 0.0;
      for (double x : w) s += x;
      CHECK(std::fabs(s - 0.5) < 1e-13);
    }
  }
  for (int n = 1; n <= 12; ++n) {
    std::vector<double> x, w;
    hdg::gauss_legendre01(n, x, w);
    double s = 0.0;
    for (double y : w) s += y;
    CHECK((int)x.size() == n && std::fabs(s - 1.0) < 1e-14);
  }
}

// The arena's free list on a host buffer (it never dereferences base).
static void arena() {
  std::vector<char> buf(1 << 20);
  hdg_arena a;
  a.base = buf.data();
  a.size = buf.size();
  a.free_[0] = buf.size();
  std::mt19937 rng(3);
  std::map<char*, size_t> live;
  for (int it = 0; it < 20000; ++it) {
    if (live.empty() || (rng() % 3 != 0 && live.size() < 200)) {
      const size_t b = 1 + rng() % 20000;
      char* p = static_cast<char*>(a.take(b));
      if (!p) continue;
      CHECK(a.owns(p) && p + hdg_arena::round(b) <= a.base + a.size);
      auto nx = live.lower_bound(p);
      if (nx != live.end()) CHECK(p + hdg_arena::round(b) <= nx->first);
      if (nx != live.begin()) {
        auto pv = std::prev(nx);
        CHECK(pv->first + hdg_arena::round(pv->second) <= p);
      }
      live[p] = b;
    } else {
      auto itl = live.begin();
      std::advance(itl, rng() % live.size());
      a.give(itl->first, itl->second);
      live.erase(itl);
    }
  }
  for (auto& kv : live) a.give(kv.first, kv.second);
  CHECK(a.free_.size() == 1 && a.free_.begin()->first == 0 && a.free_.begin()->second == buf.size());
}

int main() {
  mesh(1, 1, 0.0, 0.0, 1.0, 1.0);
  mesh(7, 5, 0.0, 0.0, 1.0, 1.0);
  mesh(64, 3, -1.0, 0.5, 2.0, 3.25);
  mesh(5, 6, 1.0, 1.0, 0.0, 0.0);  // inverted box
  mesh(40, 40, 0.1, 0.2, 1.3, 0.9);
  std::vector<double> xy(8);
  std::vector<int64_t> c(12), f(20);
  std::vector<int8_t> s(5);
  CHECK(hdg_rectangle_mesh(0, 1, 0, 0, 1, 1, 0, 1, xy.data(), c.data(), f.data(), s.data()) == HDG_ERR_ARG);
  CHECK(hdg_rectangle_mesh(1, 1, 0, 0, 1, 1, 0, 1, nullptr, c.data(), f.data(), s.data()) == HDG_ERR_ARG);
  tables();
  arena();
  std::printf("host sanitized run: ok\n");
  return 0;
}
