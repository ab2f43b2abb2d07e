// Setup work on the device: the matrix-free operators of a material, forcing
// tables from signal parameters, and the source/receiver grid points of many
// locations at once.
//
// The reference evaluates the material and assembles J_vp / J_sp on the host
// at setup (shwavepp.hpp:246-466), evaluates each forcing on the host every
// step (forcing_rank_{one,two}.hpp, signal.hpp:86-128) and maps each source
// and receiver with a host scan (map_point_source_to_velocity_grid_point.hpp:
// 61-133, map_nominal_location_to_velocity_grid_point.hpp:57-117). Here the
// same formulas (formulas.hpp) run one thread per grid point, time step or
// location, so a caller that changes the material (inversion loops) or admits
// many sources (serving) does not assemble on the host and upload operators or
// tables over PCIe.
//
// The structured mesh enters as per-row and per-column tables (GridTables):
// every per-point coordinate, cotangent, stencil multiplier and label of the
// mesh is checked once against them, so the device reads the same doubles the
// host assembly reads. Operators are bit-identical to Operators::build;
// forcing values differ from the host table only where the device's exp/sin
// round differently; mapping uses the host's distance formula and tie rule.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "formulas.hpp"
#include "rom_internal.hpp"
#include "shaw_app.hpp"
#include "shaw_common.hpp"

namespace shaw {
namespace {

template <class T>
struct Buf {
  T* p = nullptr;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { if (p) (void)hipFree(p); }
  void upload(const T* h, size_t n) {
    SHAW_HIP(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    if (n) SHAW_HIP(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
  }
  void alloc(size_t n) { SHAW_HIP(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T))); }
};

struct DevGrid {
  const double *rInvRow, *rInvSrp, *thVp, *thStp, *cotV, *cotS;
  uint64_t nr, nth;
  double dthInv, drrInv, rSurfM;
};

// build_graph_impl.py:57-120 as in Mesh::structured: W, N, E, S multipliers,
// all zero on the axis columns
__host__ __device__ inline void stencilAt(uint64_t i, uint64_t j, uint64_t nr, uint64_t nth, double (&c)[4]) {
  const bool axis = i == 0 || i == nth - 1;
  c[0] = axis ? 0.0 : (i >= 1 ? 1.0 : 0.0);
  c[1] = axis ? 0.0 : (j + 1 < nr ? 1.0 : -1.0);
  c[2] = axis ? 0.0 : (i + 1 < nth ? 1.0 : 0.0);
  c[3] = axis ? 0.0 : (j >= 1 ? 1.0 : -1.0);
}

// J_vp rows (shwavepp.hpp:319-400) and 1/rho, one velocity point per thread:
// the expressions of Operators::build in its order
__global__ void k_assemble_vp(DevGrid g, int kind, const double* __restrict__ mp, double* __restrict__ coefVp,
                              double* __restrict__ rhoInv) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.nr * g.nth) return;
  const uint64_t i = p % g.nth, j = p / g.nth;
  const double rInv = g.rInvRow[j];
  double rho, vs;
  materialAt(kind, mp, g.rSurfM, 1.0 / rInv, rho, vs);
  const double ri = 1.0 / rho;
  if (rhoInv) rhoInv[p] = ri;
  if (!coefVp) return;
  double c[4];
  stencilAt(i, j, g.nr, g.nth, c);
  const double cot = g.cotV[i];
  const double cW = (-rInv * g.dthInv + rInv * cot) * c[0] * ri;
  const double cN = (g.drrInv + 1.5 * rInv) * c[1] * ri;
  const double cE = (rInv * g.dthInv + rInv * cot) * c[2] * ri;
  const double cS = (-g.drrInv + 1.5 * rInv) * c[3] * ri;
  double* o = coefVp + 4 * p;
  if (j == 0 || j + 1 == g.nr) { o[0] = cN + cS; o[1] = 0.0; }
  else { o[0] = cN; o[1] = cS; }
  if (i == 0 || i + 1 == g.nth) { o[2] = cW + cE; o[3] = 0.0; }
  else { o[2] = cW; o[3] = cE; }
}

// J_sp rows (shwavepp.hpp:402-466), one stress point per thread
__global__ void k_assemble_sp(DevGrid g, int kind, const double* __restrict__ mp, double* __restrict__ coefSp) {
  const uint64_t nth = g.nth, rs = 2 * nth - 1, last = (g.nr - 1) * rs;
  const uint64_t nSp = last + nth - 1;
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nSp) return;
  bool srp;
  uint64_t i, j;
  if (q < last) {
    j = q / rs;
    const uint64_t kk = q % rs;
    srp = (kk & 1) == 0;
    i = srp ? kk / 2 : (kk - 1) / 2;
  } else {
    srp = false;
    j = g.nr - 1;
    i = q - last;
  }
  const double rInv = srp ? g.rInvSrp[j] : g.rInvRow[j];
  double rho, vs;
  materialAt(kind, mp, g.rSurfM, 1.0 / rInv, rho, vs);
  const double G = rho * vs * vs;
  double* o = coefSp + 2 * q;
  if (srp) {
    o[0] = (g.drrInv - rInv * 0.5) * G;
    o[1] = (-g.drrInv - rInv * 0.5) * G;
  } else {
    const double cot = g.cotS[i];
    o[0] = (-g.dthInv - 0.5 * cot) * rInv * G;
    o[1] = (g.dthInv - 0.5 * cot) * rInv * G;
  }
}

// out[s * m + k] (ldOut 0) or out[k * ldOut + s] = signal_k(s * dt)
// (forcingTable's time: the product s dt); signal k: kind (as a double),
// delay, period at par[3 k .. 3 k + 2]. Consecutive threads take consecutive
// entries of out in either layout, so the stores coalesce.
__global__ void k_forcing_table(uint64_t m, const double* __restrict__ par, double dt, uint64_t nsteps,
                                double* __restrict__ out, uint64_t ldOut) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * nsteps) return;
  const uint64_t s = ldOut ? e % nsteps : e / m, k = ldOut ? e / nsteps : e % m;
  const double period = par[3 * k + 2];
  const double freq = 1.0 / period;
  const double t = (double)s * dt;
  out[ldOut ? k * ldOut + s : e] = signalAt((int)par[3 * k], par[3 * k + 1], period, freq * freq, t);
}

// nearest vp point of each target (r [m], theta [rad]) by polar distance,
// ties to the last gid, over the rows/columns window around the bracketing
// row and column (the host's windowed search, shaw_host.cpp)
__global__ void k_map_points(DevGrid g, const double* __restrict__ rRow, uint64_t n, const double* __restrict__ tr,
                             const double* __restrict__ tth, uint64_t* __restrict__ gids) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double r = tr[e], th = tth[e];
  uint64_t lo = 0, hi = g.nr - 1;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) / 2;
    if (rRow[mid] <= r) lo = mid;
    else hi = mid;
  }
  const uint64_t jc = lo;
  lo = 0;
  hi = g.nth - 1;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) / 2;
    if (g.thVp[mid] <= th) lo = mid;
    else hi = mid;
  }
  const uint64_t ic = lo, W = 4;
  const uint64_t j0 = jc > W ? jc - W : 0, j1 = jc + W + 1 < g.nr - 1 ? jc + W + 1 : g.nr - 1;
  const uint64_t i0 = ic > W ? ic - W : 0, i1 = ic + W + 1 < g.nth - 1 ? ic + W + 1 : g.nth - 1;
  double trial = std::numeric_limits<double>::max();
  uint64_t out = 0;
  for (uint64_t j = j0; j <= j1; ++j)
    for (uint64_t i = i0; i <= i1; ++i) {
      const double r1 = 1.0 / g.rInvRow[j], th1 = g.thVp[i];
      const double d = std::sqrt((r1 * r1) + (r * r) - 2.0 * r1 * r * std::cos(th - th1));
      if (d <= trial) {
        trial = d;
        out = j * g.nth + i;
      }
    }
  gids[e] = out;
}

// the mesh's per-row / per-column tables, checked against every point
const GridTables& gridTables(shaw_app& app) {
  GridTables& t = app.tables;
  if (t.checked) {
    SHAW_REQUIRE(t.ok, SHAW_ERR_ARG, "the mesh is not separable into per-row and per-column tables");
    return t;
  }
  t.checked = true;
  const Mesh& m = app.mesh;
  const Grid& g = m.grid;
  const uint64_t nr = g.nr, nth = g.nth;
  t.rInvRow.resize(nr);
  t.rInvSrp.resize(nr - 1);
  t.thVp.resize(nth);
  t.thStp.resize(nth - 1);
  t.cotV.resize(nth);
  t.cotS.resize(nth - 1);
  for (uint64_t j = 0; j < nr; ++j) t.rInvRow[j] = m.coordsVp[2 * g.vp(0, j) + 1];
  for (uint64_t j = 0; j + 1 < nr; ++j) t.rInvSrp[j] = m.coordsSp[2 * g.srp(0, j) + 1];
  for (uint64_t i = 0; i < nth; ++i) {
    t.thVp[i] = m.coordsVp[2 * g.vp(i, 0)];
    t.cotV[i] = m.cotVp[g.vp(i, 0)];
  }
  for (uint64_t i = 0; i + 1 < nth; ++i) {
    t.thStp[i] = m.coordsSp[2 * g.stp(i, 0)];
    t.cotS[i] = m.cotSp[g.stp(i, 0)];
  }
  bool ok = true;
  for (uint64_t j = 0; j < nr && ok; ++j)
    for (uint64_t i = 0; i < nth && ok; ++i) {
      const uint64_t p = g.vp(i, j);
      double c[4];
      stencilAt(i, j, nr, nth, c);
      ok = m.coordsVp[2 * p] == t.thVp[i] && m.coordsVp[2 * p + 1] == t.rInvRow[j] && m.cotVp[p] == t.cotV[i];
      for (int k = 0; k < 4; ++k) ok = ok && m.stencilVp[4 * p + k] == c[k];
      if (ok && j + 1 < nr) {
        const uint64_t q = g.srp(i, j);
        ok = m.labelsSp[q] == 1 && m.coordsSp[2 * q + 1] == t.rInvSrp[j];
      }
      if (ok && i + 1 < nth) {
        const uint64_t q = g.stp(i, j);
        ok = m.labelsSp[q] == 2 && m.coordsSp[2 * q] == t.thStp[i] && m.coordsSp[2 * q + 1] == t.rInvRow[j] &&
             m.cotSp[q] == t.cotS[i];
      }
    }
  t.ok = ok;
  SHAW_REQUIRE(ok, SHAW_ERR_ARG, "the mesh is not separable into per-row and per-column tables");
  return t;
}

struct UploadedGrid {
  Buf<double> rInvRow, rInvSrp, thVp, thStp, cotV, cotS;
  DevGrid d{};
  explicit UploadedGrid(shaw_app& app) {
    const GridTables& t = gridTables(app);
    rInvRow.upload(t.rInvRow.data(), t.rInvRow.size());
    rInvSrp.upload(t.rInvSrp.data(), t.rInvSrp.size());
    thVp.upload(t.thVp.data(), t.thVp.size());
    thStp.upload(t.thStp.data(), t.thStp.size());
    cotV.upload(t.cotV.data(), t.cotV.size());
    cotS.upload(t.cotS.data(), t.cotS.size());
    d = DevGrid{rInvRow.p, rInvSrp.p, thVp.p, thStp.p, cotV.p, cotS.p, app.mesh.grid.nr, app.mesh.grid.nth,
                app.mesh.info.dthInv, app.mesh.info.drrInv, app.mesh.info.bounds[3]};
  }
};

}  // namespace

void assembleOperatorsDevice(shaw_app& app, const Material& mat, double* coefVp, double* coefSp, double* rhoInv,
                             hipStream_t stream) {
  SHAW_REQUIRE(mat.kind >= Material::Unilayer && mat.kind <= Material::Layered, SHAW_ERR_ARG,
               "unknown material kind");
  UploadedGrid ug(app);
  Buf<double> mp;
  mp.upload(mat.params.data(), mat.params.size());
  DevGrid d = ug.d;
  d.rSurfM = mat.rSurfM;
  const uint64_t nVp = app.mesh.grid.nVp(), nSp = app.mesh.grid.nSp();
  if (coefVp || rhoInv)
    hipLaunchKernelGGL(k_assemble_vp, dim3((unsigned)((nVp + 255) / 256)), dim3(256), 0, stream, d, mat.kind, mp.p,
                       coefVp, rhoInv);
  if (coefSp)
    hipLaunchKernelGGL(k_assemble_sp, dim3((unsigned)((nSp + 255) / 256)), dim3(256), 0, stream, d, mat.kind, mp.p,
                       coefSp);
  SHAW_HIP(hipGetLastError());
  SHAW_HIP(hipStreamSynchronize(stream));  // the uploaded tables are freed on return
}

std::vector<double> signalParams(uint64_t m, const int* kinds, const double* delays, const double* periods) {
  std::vector<double> par(3 * m);
  for (uint64_t k = 0; k < m; ++k) {
    SHAW_REQUIRE(kinds[k] >= 0 && kinds[k] <= 2, SHAW_ERR_ARG, "unknown signal kind");
    SHAW_REQUIRE(periods[k] > 0, SHAW_ERR_ARG, "Period for signal must be positive.");
    par[3 * k] = (double)kinds[k];
    par[3 * k + 1] = delays[k];
    par[3 * k + 2] = periods[k];
  }
  return par;
}

void forcingTableFromParams(uint64_t m, const double* parDevice, double dt, uint64_t nsteps, double* out,
                            uint64_t ldOut, hipStream_t stream) {
  const uint64_t n = m * nsteps;
  if (n == 0) return;
  hipLaunchKernelGGL(k_forcing_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, m, parDevice, dt,
                     nsteps, out, ldOut);
  SHAW_HIP(hipGetLastError());
}

void forcingTableDevice(uint64_t m, const int* kinds, const double* delays, const double* periods, double dt,
                        uint64_t nsteps, double* out, uint64_t ldOut, hipStream_t stream) {
  const std::vector<double> par = signalParams(m, kinds, delays, periods);
  if (m == 0 || nsteps == 0) return;
  Buf<double> dp;
  dp.upload(par.data(), par.size());
  forcingTableFromParams(m, dp.p, dt, nsteps, out, ldOut, stream);
  SHAW_HIP(hipStreamSynchronize(stream));  // the parameter array is freed on return
}

}  // namespace shaw

using shaw::guarded;

extern "C" {

int shaw_app_assemble_device(shaw_app_t app, const shaw_material* mat, double* coef_vp, double* coef_sp,
                             double* rho_inv, void* stream) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(app && mat, SHAW_ERR_ARG, "null argument");
    Material m = materialFromC(mat);
    m.rSurfM = app->mesh.info.bounds[3];
    m.validate(app->mesh.info);
    assembleOperatorsDevice(*app, m, coef_vp, coef_sp, rho_inv, static_cast<hipStream_t>(stream));
  });
}

int shaw_forcing_table_device(uint64_t m, const int* kinds, const double* delays, const double* periods, double dt,
                              uint64_t nsteps, double* table, void* stream) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE((m == 0 || (kinds && delays && periods)) && (m * nsteps == 0 || table), SHAW_ERR_ARG,
                 "null argument");
    forcingTableDevice(m, kinds, delays, periods, dt, nsteps, table, 0, static_cast<hipStream_t>(stream));
  });
}

int shaw_map_points_device(shaw_app_t app, uint64_t n, const double* depths_km, const double* angles_deg,
                           uint64_t* gids) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(app && (n == 0 || (angles_deg && gids)), SHAW_ERR_ARG, "null argument");
    if (n == 0) return;
    const MeshInfo& mi = app->mesh.info;
    constexpr double toRad = M_PI / 180.0;
    std::vector<double> tr(n), tth(n);
    for (uint64_t e = 0; e < n; ++e) {
      const double a = angles_deg[e];
      if (!depths_km) {  // map_nominal_location_to_velocity_grid_point.hpp:57-87
        SHAW_REQUIRE(a >= mi.bounds[0] && a <= mi.bounds[1], SHAW_ERR_ARG, "Target angle is outside theta axis");
        tr[e] = mi.bounds[3];
        tth[e] = a * toRad;
        continue;
      }
      // map_point_source_to_velocity_grid_point.hpp:61-105
      const double radiusKm = mi.maxRadiusKm() - depths_km[e];
      SHAW_REQUIRE(a >= mi.bounds[0] && a <= mi.bounds[1], SHAW_ERR_ARG, "Source angle is outside theta axis");
      SHAW_REQUIRE(radiusKm * 1000. >= mi.bounds[2] && radiusKm * 1000. <= mi.bounds[3], SHAW_ERR_ARG,
                   "Source angle is outside radius axis");
      double th = a * toRad;
      auto eq = [](double x, double y) {
        const double eps = std::numeric_limits<double>::epsilon();
        return std::abs(x - y) <= ((std::abs(x) > std::abs(y) ? std::abs(y) : std::abs(x)) * eps);
      };
      if (eq(a, mi.bounds[0])) th = mi.bounds[0] * toRad + mi.dth;
      if (eq(a, mi.bounds[1])) th = mi.bounds[1] * toRad - mi.dth;
      tr[e] = radiusKm * 1000.;
      tth[e] = th;
    }
    UploadedGrid ug(*app);
    const GridTables& t = app->tables;
    std::vector<double> rRow(t.rInvRow.size());
    for (size_t j = 0; j < rRow.size(); ++j) rRow[j] = 1.0 / t.rInvRow[j];
    Buf<double> dRow, dr, dth;
    Buf<uint64_t> dg;
    dRow.upload(rRow.data(), rRow.size());
    dr.upload(tr.data(), n);
    dth.upload(tth.data(), n);
    dg.alloc(n);
    hipLaunchKernelGGL(k_map_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, ug.d, dRow.p, n, dr.p, dth.p,
                       dg.p);
    SHAW_HIP(hipGetLastError());
    SHAW_HIP(hipMemcpy(gids, dg.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
