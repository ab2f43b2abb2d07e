// FOM energy weights of the energy-consistent Galerkin ROM.
//
// The leapfrog of the reference (fom_update_kernels.hpp) conserves a discrete
// energy E = 1/2 v^T Mv v + 1/2 s^T Ms s with diagonal Mv, Ms: on every
// coupling that is not merged at the CMB/surface rows or on the axis,
// Mv[p] J_vp[p, q] = -Ms[q] J_sp[q, p] (both coefficients are those of
// ShWavePP, shwavepp.hpp:319-466). The ratios Ms[q] / Mv[p] of the unmerged
// couplings therefore fix the weights up to one scale: Mv follows by chained
// products along grid row 1 and then up every grid column, the stress weights
// from the unmerged coupling of each stress point. Points whose couplings are
// all merged take the weight their merged coupling implies (CMB and surface
// rows) or that of the adjacent point (the two axis columns, where the field
// vanishes). The log-domain chaining keeps the products free of overflow on
// large grids; the weights are normalised to 1 at the mesh centre.
#include <cmath>
#include <cstring>
#include <vector>

#include "shaw_app.hpp"
#include "shaw_common.hpp"

namespace shaw {

void energyWeights(const Operators& ops, std::vector<double>& mv, std::vector<double>& ms) {
  const Grid& g = ops.grid;
  const uint64_t nr = g.nr, nth = g.nth, rs = g.rs();
  SHAW_REQUIRE(nr >= 4 && nth >= 4, SHAW_ERR_ARG, "energy weights need at least 4 x 4 grid points");
  const double* cv = ops.coefVp.data();
  const double* cs = ops.coefSp.data();
  // J_vp slots of point (i, j): N, S, W, E
  auto cN = [&](uint64_t j, uint64_t i) { return cv[4 * (j * nth + i) + 0]; };
  auto cS = [&](uint64_t j, uint64_t i) { return cv[4 * (j * nth + i) + 1]; };
  auto cW = [&](uint64_t j, uint64_t i) { return cv[4 * (j * nth + i) + 2]; };
  auto cE = [&](uint64_t j, uint64_t i) { return cv[4 * (j * nth + i) + 3]; };
  // J_sp pairs: srp(i, j) couples v(i, j+1) (slot 0, north) and v(i, j)
  // (slot 1, south); stp(i, j) couples v(i, j) (slot 0, west) and v(i+1, j)
  auto srp0 = [&](uint64_t j, uint64_t i) { return cs[2 * g.srp(i, j) + 0]; };
  auto srp1 = [&](uint64_t j, uint64_t i) { return cs[2 * g.srp(i, j) + 1]; };
  auto stp0 = [&](uint64_t j, uint64_t i) { return cs[2 * g.stp(i, j) + 0]; };
  auto stp1 = [&](uint64_t j, uint64_t i) { return cs[2 * g.stp(i, j) + 1]; };
  // Mv[j+1][i] / Mv[j][i] through srp(i, j); Mv[j][i+1] / Mv[j][i] through stp(i, j)
  auto up = [&](uint64_t j, uint64_t i) { return (cN(j, i) / srp1(j, i)) / (cS(j + 1, i) / srp0(j, i)); };
  auto east = [&](uint64_t j, uint64_t i) { return (cE(j, i) / stp0(j, i)) / (cW(j, i + 1) / stp1(j, i)); };

  std::vector<double> lv(nr * nth, 0.0);
  auto L = [&](uint64_t j, uint64_t i) -> double& { return lv[j * nth + i]; };
  // along row 1 from column 1, then up each interior column from row 1
  double acc = 0.0;
  for (uint64_t i = 2; i + 1 < nth; ++i) {
    acc += std::log(east(1, i - 1));
    L(1, i) = acc;
  }
  for (uint64_t i = 1; i + 1 < nth; ++i) {
    double a = 0.0;
    for (uint64_t j = 2; j + 1 < nr; ++j) {
      a += std::log(up(j - 1, i));
      L(j, i) = L(1, i) + a;
    }
  }
  const double centre = L(nr / 2, nth / 2);
  mv.assign(nr * nth, 0.0);
  for (uint64_t e = 0; e < nr * nth; ++e) mv[e] = std::exp(lv[e] - centre);
  auto M = [&](uint64_t j, uint64_t i) -> double& { return mv[j * nth + i]; };

  // radial stress weights from the unmerged side: the south velocity point
  // for rows j >= 1, the north one for row 0
  std::vector<double> srp((nr - 1) * nth), stp(nr * (nth - 1));
  auto R = [&](uint64_t j, uint64_t i) -> double& { return srp[j * nth + i]; };
  auto T = [&](uint64_t j, uint64_t i) -> double& { return stp[j * (nth - 1) + i]; };
  for (uint64_t j = 0; j + 1 < nr; ++j)
    for (uint64_t i = 0; i < nth; ++i)
      R(j, i) = j >= 1 ? (-M(j, i) * cN(j, i)) / srp1(j, i) : (-M(j + 1, i) * cS(j + 1, i)) / srp0(j, i);
  // CMB and surface rows: N and S of a point there are one stress point, so
  // the merged coefficient (in the N slot) and that stress point's fix it
  for (uint64_t i = 0; i < nth; ++i) {
    M(0, i) = (-R(0, i) * srp1(0, i)) / cN(0, i);
    M(nr - 1, i) = (-R(nr - 2, i) * srp0(nr - 2, i)) / cN(nr - 1, i);
  }
  // axis columns: every coefficient vanishes there; the neighbour's weight
  for (uint64_t j = 0; j < nr; ++j) {
    M(j, 0) = M(j, 1);
    M(j, nth - 1) = M(j, nth - 2);
  }
  for (uint64_t j = 0; j + 1 < nr; ++j) {
    R(j, 0) = R(j, 1);
    R(j, nth - 1) = R(j, nth - 2);
  }
  // angular stress weights from the point off the axis
  for (uint64_t j = 0; j < nr; ++j)
    for (uint64_t i = 0; i + 1 < nth; ++i)
      T(j, i) = i >= 1 ? (-M(j, i) * cE(j, i)) / stp0(j, i) : (-M(j, i + 1) * cW(j, i + 1)) / stp1(j, i);
  ms.assign(g.nSp(), 0.0);
  for (uint64_t j = 0; j + 1 < nr; ++j)
    for (uint64_t i = 0; i < nth; ++i) {
      ms[g.srp(i, j)] = R(j, i);
      if (i + 1 < nth) ms[g.stp(i, j)] = T(j, i);
    }
  for (uint64_t i = 0; i + 1 < nth; ++i) ms[g.stp(i, nr - 1)] = T(nr - 1, i);
  uint64_t bad = 0;
  for (double x : mv) bad += !(std::isfinite(x) && x > 0.0);
  for (double x : ms) bad += !(std::isfinite(x) && x > 0.0);
  SHAW_REQUIRE(bad == 0, SHAW_ERR_ARG, "energy weights: " + std::to_string(bad) + " points without a usable coupling");
}

}  // namespace shaw

extern "C" int shaw_app_energy_weights(shaw_app_t app, double* mv, double* ms) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(app && (mv || ms), SHAW_ERR_ARG, "null argument");
    std::vector<double> a, b;
    energyWeights(app->ops, a, b);
    if (mv) std::memcpy(mv, a.data(), a.size() * sizeof(double));
    if (ms) std::memcpy(ms, b.data(), b.size() * sizeof(double));
  });
}
