// hdg_errornorm_noalloc: the squared L2 error of a recovered solution for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  It runs the kernels of hdg_errornorm on device
// arrays the caller passes; the block partial sums live in `scratch`, a
// file-scope device array, and the reference tables of every (k, quad
// degree) it accepts in `tables`, both filled at the first call on a device.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/hdg.h"
#include "hdg_internal.h"
#include "hdg_tables.h"

using namespace hdg;

namespace {

// errornorm_blocks caps the grid of the error-norm kernel at 2048 blocks,
// one partial sum each
constexpr int kPartials = 2048;
constexpr int kMaxK = 4, kMaxQ = 6;
// per (k, quad degree): N [nu][nqp], Mg [3][nqp], W [nqp]; nu <= 15, nqp <= 12
constexpr int kTabStride = 15 * 12 + 3 * 12 + 12;
constexpr int kTables = kMaxK * kMaxQ * kTabStride;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kPartials];
static __device__ __align__(256) double tables[kTables];

struct DeviceArrays {
  std::once_flag once;
  int status = HDG_OK;
  double* scratch = nullptr;  // device addresses of the two arrays on this device
  double* tables = nullptr;
};
DeviceArrays g_dev[kMaxDevices];
TabLayoutDev g_lay[kMaxK][kMaxQ];  // offsets inside one (k, q) slot; nqp = 0 where the rule is missing

// Once per device, with the device current: the arrays' addresses and the
// reference tables of k = 1..4 and quad degrees 1..6.
void init_device(DeviceArrays& d) {
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    d.status = HDG_ERR_HIP;
    return;
  }
  d.scratch = static_cast<double*>(p);
  p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(tables)) != hipSuccess || !p) {
    d.status = HDG_ERR_HIP;
    return;
  }
  d.tables = static_cast<double*>(p);
  std::vector<double> host(kTables, 0.0);
  for (int k = 1; k <= kMaxK; ++k)
    for (int q = 1; q <= kMaxQ; ++q) {
      RefTables rt;
      TabLayoutDev& L = g_lay[k - 1][q - 1];
      L = TabLayoutDev{};
      if (!rt.build(k, q)) continue;
      const TableLayout& T = rt.lay;
      double* slot = host.data() + ((k - 1) * kMaxQ + (q - 1)) * kTabStride;
      const int nu = T.nu, nqp = T.nqp;
      L.nu = nu;
      L.nt = T.nt;
      L.nqp = nqp;
      L.nfq = T.nfq;
      L.N = 0;
      L.Mg = nu * nqp;
      L.W = L.Mg + 3 * nqp;
      L.total = L.W + nqp;
      for (int x = 0; x < nu * nqp; ++x) slot[L.N + x] = rt.data[T.N + x];
      for (int x = 0; x < 3 * nqp; ++x) slot[L.Mg + x] = rt.data[T.Mg + x];
      for (int x = 0; x < nqp; ++x) slot[L.W + x] = rt.data[T.W + x];
    }
  if (hipMemcpy(d.tables, host.data(), sizeof(double) * kTables, hipMemcpyHostToDevice) != hipSuccess)
    d.status = HDG_ERR_HIP;
}

}  // namespace

extern "C" int hdg_errornorm_noalloc(int k, int quad_degree, int64_t n_cells, const double* xy,
                                     const int32_t* cells, const double* qu, int uex_kind, const double* uex_par,
                                     double* sq_l2, void* stream) {
  if (k < 1 || k > kMaxK) return HDG_ERR_UNSUPPORTED;
  const int q = quad_degree > 0 ? quad_degree : k + 1;
  if (q > kMaxQ) return HDG_ERR_QUAD;
  if (n_cells < 1 || n_cells > HDG_NOALLOC_MAX_CELLS) return HDG_ERR_ARG;
  if (!xy || !cells || !qu || !sq_l2) return HDG_ERR_ARG;
  if (uex_kind != HDG_UEX_SINSIN && uex_kind != HDG_UEX_LINEAR) return HDG_ERR_ARG;
  if (uex_kind == HDG_UEX_LINEAR && !uex_par) return HDG_ERR_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return HDG_ERR_HIP;
  DeviceArrays& d = g_dev[dev];
  std::call_once(d.once, init_device, std::ref(d));
  if (d.status != HDG_OK) return d.status;
  const TabLayoutDev& L = g_lay[k - 1][q - 1];
  if (L.nqp == 0) return HDG_ERR_QUAD;
  const double* tab = d.tables + ((k - 1) * kMaxQ + (q - 1)) * kTabStride;
  const double p3[3] = {uex_par ? uex_par[0] : 0.0, uex_par ? uex_par[1] : 0.0, uex_par ? uex_par[2] : 0.0};
  DevMesh m;
  m.nc = n_cells;
  m.xy = xy;
  m.cn = cells;
  // partials: errornorm_blocks(n_cells) <= kPartials, every one written before it is read
  if (errornorm_blocks(n_cells) > kPartials) return HDG_ERR_ARG;
  const int e = launch_errornorm(m, tab, L, k, qu, uex_kind, p3, d.scratch, sq_l2, static_cast<hipStream_t>(stream));
  return e ? HDG_ERR_HIP : HDG_OK;
}
