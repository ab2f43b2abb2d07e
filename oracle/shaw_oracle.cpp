// SHAW time-stepping oracle: a plain C++ CPU restatement of the reference's
// FP64 FOM (CSR operators, leapfrog, forcing, receivers, snapshots) and
// Galerkin ROM, used ONLY as test infrastructure (parity checker, CPU
// baseline). Nothing in the product links or calls this file.
//
// Every function cites the reference file:line it restates. Arithmetic is
// written in the same association order as the reference and the library is
// compiled with -ffp-contract=off (no FMA), which is what a default x86-64
// gcc build of the reference produces.
//
// Parity status of each function (what pins it):
//   orc_app_*           CSR values/ptr/ind: tests/jacobian_vp, jacobian_sp golds
//                       (4x5 mesh); graph/coords/labels golds; mesh-info golds.
//   orc_map_source      tests/forcing_rank1 gold gid 636; FOM golds.
//   orc_map_receiver    tests/seismogram golds 1027/1034/1053/1064.
//   orc_signal          Ricker closed form + FOM golds (ricker, sinusoid).
//   orc_fom_run         FOM state/seismogram golds, rank-1 and rank-2.
//   orc_rom_project     brute-force dense Phi^T J Phi; Phi = I reproduces CSR.
//   orc_rom_run         Phi = I ROM == FOM (invariant); rank-2 column j ==
//                       rank-1 run j. The reference's KokkosBlas gemv/gemm
//                       summation order itself is parity unpinned (no active
//                       reference test covers the ROM).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

// src/shared/various/equality.hpp:54-58
bool essentiallyEqual(double a, double b) {
  const double eps = std::numeric_limits<double>::epsilon();
  return std::abs(a - b) <= ((std::abs(a) > std::abs(b) ? std::abs(b) : std::abs(a)) * eps);
}

// src/shared/various/angular_helpers.hpp:52-91
double computeCotangent(double theta) {
  const double c = std::cos(theta), s = std::sin(theta);
  if (essentiallyEqual(s, 0.0)) return 0.0;
  return c / s;
}
double degToRad(double v) { constexpr double f = M_PI / 180.0; return v * f; }
double computePolarDistance(double r1, double th1, double r2, double th2) {
  return std::sqrt((r1 * r1) + (r2 * r2) - 2.0 * r1 * r2 * std::cos(th2 - th1));
}

struct Material {
  // kind: 0 unilayer, 1 bilayer, 2 prem, 3 layered (piecewise linear in r)
  int kind = 0;
  std::vector<double> p;
  double rSurfM = 0;
  // src/shared/material_models/material_model_unilayer.hpp:72-83,
  // material_model_bilayer.hpp:74-101, material_model_prem.hpp:85-176
  void computeAt(double r, double& rho, double& vs) const {
    if (kind == 0) {
      const double dM = rSurfM - r;
      rho = p[0] + p[1] * dM + p[2] * dM * dM;
      vs = p[3] + p[4] * dM + p[5] * dM * dM;
    } else if (kind == 1) {
      // p = [depth2Km, l1 rho c0..c2, l1 vs c0..c2, l2 rho c0..c2, l2 vs c0..c2]
      const double thisPtDepthM = rSurfM - r;
      if (thisPtDepthM < p[0] * 1000.0) {
        const double dM = thisPtDepthM;
        rho = p[1] + p[2] * dM + p[3] * dM * dM;
        vs = p[4] + p[5] * dM + p[6] * dM * dM;
      } else {
        const double dM = thisPtDepthM - (p[0] * 1000.0);
        rho = p[7] + p[8] * dM + p[9] * dM * dM;
        vs = p[10] + p[11] * dM + p[12] * dM * dM;
      }
    } else if (kind == 2) {
      const double rKm = r / 1000.0;
      const double x = rKm / 6371.0, xSq = x * x, xCu = x * x * x;
      if (rKm >= 6356.0) { rho = 2.6; vs = 3.2; }
      else if (rKm >= 6346.6 && rKm < 6356.0) { rho = 2.9; vs = 3.9; }
      else if (rKm >= 6291.0 && rKm < 6346.6) { rho = 2.691 + 0.6924 * x; vs = 2.1519 + 2.3481 * x; }
      else if (rKm >= 6151.0 && rKm < 6291.0) { rho = 2.691 + 0.6924 * x; vs = 2.1519 + 2.3481 * x; }
      else if (rKm >= 5971.0 && rKm < 6151.0) { rho = 7.1089 - 3.8045 * x; vs = 8.9496 - 4.4597 * x; }
      else if (rKm >= 5771.0 && rKm < 5971.0) { rho = 11.2494 - 8.0298 * x; vs = 22.3512 - 18.5856 * x; }
      else if (rKm >= 5701.0 && rKm < 5771.0) { rho = 5.3197 - 1.4836 * x; vs = 9.9839 - 4.9324 * x; }
      else if (rKm >= 5600.0 && rKm < 5701.0) {
        rho = 7.9565 - 6.4761 * x + 5.5283 * xSq - 3.0807 * xCu;
        vs = 22.3459 - 17.2473 * x - 2.0834 * xSq + 0.9783 * xCu;
      } else if (rKm >= 3630.0 && rKm < 5600.0) {
        rho = 7.9565 - 6.4761 * x + 5.5283 * xSq - 3.0807 * xCu;
        vs = 11.1671 - 13.7818 * x + 17.4575 * xSq - 9.2777 * xCu;
      } else if (rKm >= 3480.0 && rKm < 3630.0) {
        rho = 7.9565 - 6.4761 * x + 5.5283 * xSq - 3.0807 * xCu;
        vs = 6.9254 + 1.4672 * x - 2.0834 * xSq + 0.9783 * xCu;
      } else if (rKm >= 1221.5 && rKm < 3480.0) {
        rho = 12.5815 - 1.2638 * x - 3.6426 * xSq - 5.5281 * xCu; vs = 0.0;
      } else { rho = 13.0885 - 8.8381 * xSq; vs = 3.6678 - 4.4475 * xSq; }
      vs *= 1000.0;
      rho *= 1000.0;
    } else {
      // layered: p = [n, rKm_0..rKm_{n-1}, rho_0.., vs_0..]; nodes ascending in
      // radius, rho/vs linear between nodes, held constant outside.
      const int n = (int)p[0];
      const double* rk = &p[1];
      const double* ro = &p[1 + n];
      const double* vv = &p[1 + 2 * n];
      const double rKm = r / 1000.0;
      int k = 0;
      while (k < n - 2 && rKm >= rk[k + 1]) ++k;
      double w = (rKm - rk[k]) / (rk[k + 1] - rk[k]);
      if (w < 0.0) w = 0.0;
      if (w > 1.0) w = 1.0;
      rho = ro[k] + (ro[k + 1] - ro[k]) * w;
      vs = vv[k] + (vv[k + 1] - vv[k]) * w;
    }
  }
};

struct Csr {
  std::vector<size_t> ptr, ind;
  std::vector<double> val;
};

struct App {
  // mesh info, src/shared/mesh_helpers/mesh_info.hpp:116-214
  double bounds[4] = {0, 0, 0, 0};  // thL deg, thR deg, rCmb m, rSurf m
  double dth = 0, drr = 0, dthInv = 0, drrInv = 0;
  size_t nVp = 0, nSp = 0, nth = 0, nr = 0;
  // graphs/coords, src/shared/mesh_helpers/read_graph_file.hpp:54-124
  std::vector<size_t> gVp, gSp;          // [n*5], [n*3]
  std::vector<double> cVp, cSp;          // [n*2]: theta, 1/r
  std::vector<double> cotVp, cotSp;
  std::vector<int> labels;
  std::vector<double> coeffVp;           // [n*4]
  std::vector<double> rhoInv, shearMod;
  double minVs = std::numeric_limits<double>::max();
  double maxVs = std::numeric_limits<double>::min();
  Csr Jvp, Jsp;
};

void readMeshInfo(App& a, const std::string& dir) {
  std::ifstream f(dir + "/mesh_info.dat");
  if (!f) throw std::runtime_error("File named: mesh_info.dat not found");
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string k, v;
    ss >> k;
    if (k == "thL") { ss >> v; a.bounds[0] = std::stod(v); }
    if (k == "thR") { ss >> v; a.bounds[1] = std::stod(v); }
    if (k == "rCmb") { ss >> v; a.bounds[2] = std::stod(v) * 1000.0; }
    if (k == "rSurf") { ss >> v; a.bounds[3] = std::stod(v) * 1000.0; }
    if (k == "dth") { ss >> v; a.dth = std::stod(v); a.dthInv = 1.0 / a.dth; }
    if (k == "dr") { ss >> v; a.drr = std::stod(v) * 1000.0; a.drrInv = 1.0 / a.drr; }
    if (k == "numPtsVp") { ss >> v; a.nVp = std::stoi(v); }
    if (k == "numPtsSp") { ss >> v; a.nSp = std::stoi(v); }
    if (k == "nth") { ss >> v; a.nth = std::stoi(v); }
    if (k == "nr") { ss >> v; a.nr = std::stoi(v); }
  }
}

void readGraph(App& a, const std::string& dir, bool sp) {
  const std::string path = dir + (sp ? "/graph_sp.dat" : "/graph_vp.dat");
  std::ifstream f(path);
  if (!f) throw std::runtime_error("graph file: " + path + " not found");
  const size_t n = sp ? a.nSp : a.nVp, w = sp ? 3 : 5;
  auto& g = sp ? a.gSp : a.gVp;
  auto& c = sp ? a.cSp : a.cVp;
  auto& cot = sp ? a.cotSp : a.cotVp;
  g.assign(n * w, 0); c.assign(n * 2, 0); cot.assign(n, 0);
  if (sp) a.labels.assign(n, 0);
  std::string line, v;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    if (!(ss >> v)) continue;
    const size_t gid = std::stoi(v);
    g[gid * w] = gid;
    if (sp) { ss >> v; a.labels[gid] = std::stoi(v); }
    ss >> v; const int onSym = std::stoi(v);
    ss >> v; c[gid * 2] = std::stod(v);
    ss >> v; c[gid * 2 + 1] = 1.0 / (std::stod(v) * 1000.0);
    cot[gid] = onSym == 1 ? 0.0 : computeCotangent(c[gid * 2]);
    for (size_t i = 1; i < w; ++i) { ss >> v; g[gid * w + i] = std::stoi(v); }
  }
}

void readCoeff(App& a, const std::string& dir) {
  std::ifstream f(dir + "/coeff_vp.dat");
  if (!f) throw std::runtime_error("coeff file not found");
  a.coeffVp.assign(a.nVp * 4, 0);
  std::string line, v;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    if (!(ss >> v)) continue;
    const size_t gid = std::stoi(v);
    for (int i = 0; i < 4; ++i) { ss >> v; a.coeffVp[gid * 4 + i] = std::stod(v); }
  }
}

// src/kokkos/shwavepp.hpp:246-286
void setMaterial(App& a, const Material& m) {
  a.rhoInv.assign(a.nVp, 0); a.shearMod.assign(a.nSp, 0);
  double rho, vs;
  for (size_t i = 0; i < a.nVp; ++i) {
    const size_t g = a.gVp[i * 5];
    m.computeAt(1.0 / a.cVp[g * 2 + 1], rho, vs);
    a.rhoInv[i] = 1.0 / rho;
  }
  for (size_t i = 0; i < a.nSp; ++i) {
    const size_t g = a.gSp[i * 3];
    m.computeAt(1.0 / a.cSp[g * 2 + 1], rho, vs);
    a.shearMod[i] = rho * vs * vs;
    if (vs != 0.) a.minVs = std::min(a.minVs, vs);
    a.maxVs = std::max(a.maxVs, vs);
  }
}

// src/kokkos/shwavepp.hpp:319-400 (row order N, S, W, E with N==S / W==E merges)
void fillVp(App& a) {
  Csr& J = a.Jvp;
  J.ptr.assign(a.nVp + 1, 0); J.ind.clear(); J.val.clear();
  for (size_t i = 0; i < a.nVp; ++i) {
    const size_t* g = &a.gVp[i * 5];
    const size_t gid = g[0], gw = g[1], gn = g[2], ge = g[3], gs = g[4];
    const double rInv = a.cVp[gid * 2 + 1], rhoInv = a.rhoInv[i];
    const double* c = &a.coeffVp[i * 4];
    const double cot = a.cotVp[gid];
    const double cW = (-rInv * a.dthInv + rInv * cot) * c[0] * rhoInv;
    const double cN = (a.drrInv + 1.5 * rInv) * c[1] * rhoInv;
    const double cE = (rInv * a.dthInv + rInv * cot) * c[2] * rhoInv;
    const double cS = (-a.drrInv + 1.5 * rInv) * c[3] * rhoInv;
    size_t shift = 0;
    if (gn == gs) { shift += 1; J.ind.push_back(gn); J.val.push_back(cN + cS); }
    else { shift += 2; J.ind.push_back(gn); J.val.push_back(cN); J.ind.push_back(gs); J.val.push_back(cS); }
    if (gw == ge) { shift += 1; J.ind.push_back(gw); J.val.push_back(cW + cE); }
    else { shift += 2; J.ind.push_back(gw); J.val.push_back(cW); J.ind.push_back(ge); J.val.push_back(cE); }
    J.ptr[i + 1] = J.ptr[i] + shift;
  }
}

// src/kokkos/shwavepp.hpp:402-466
void fillSp(App& a) {
  Csr& J = a.Jsp;
  J.ptr.assign(a.nSp + 1, 0); J.ind.assign(a.nSp * 2, 0); J.val.assign(a.nSp * 2, 0);
  for (size_t r = 0; r < a.nSp; ++r) J.ptr[r + 1] = J.ptr[r] + 2;
  size_t k = 0;
  for (size_t i = 0; i < a.nSp; ++i) {
    const size_t gid = a.gSp[i * 3];
    const double rInv = a.cSp[gid * 2 + 1];
    const double G = a.shearMod[i];
    if (a.labels[i] == 1) {
      J.ind[k] = a.gSp[i * 3 + 1]; J.val[k++] = (a.drrInv - rInv * 0.5) * G;
      J.ind[k] = a.gSp[i * 3 + 2]; J.val[k++] = (-a.drrInv - rInv * 0.5) * G;
    }
    if (a.labels[i] == 2) {
      const double cot = a.cotSp[gid];
      J.ind[k] = a.gSp[i * 3 + 1]; J.val[k++] = (-a.dthInv - 0.5 * cot) * rInv * G;
      J.ind[k] = a.gSp[i * 3 + 2]; J.val[k++] = (a.dthInv - 0.5 * cot) * rInv * G;
    }
  }
}

// src/shared/various/signal.hpp:86-128 ; kind 0 ricker, 1 sinusoid, 2 gaussDer
double signalAt(int kind, double delay, double period, double t) {
  const double one = 1.0, two = 2.0, myPISq = M_PI * M_PI;
  const double f = one / period, fSq = f * f;
  if (kind == 0) {
    const double tDiffSq = (t - delay) * (t - delay);
    const double expTerm = std::exp(-myPISq * fSq * tDiffSq);
    return (one - two * myPISq * tDiffSq * fSq) * expTerm;
  } else if (kind == 1) {
    constexpr double rn = 0.001;
    constexpr double term2 = rn / (rn + 2.0);
    if (t < period) {
      const double term1 = std::sin(rn * M_PI * t / period);
      const double term3 = std::sin((rn + two) * M_PI * t / period);
      return term1 - term2 * term3;
    }
    const double term1 = std::sin(rn * M_PI);
    const double term3 = std::sin((rn + two) * M_PI);
    return term1 - term2 * term3;
  }
  const double tDiff = t - delay, tDiffSq = tDiff * tDiff;
  const double expTerm = std::exp(-fSq * tDiffSq);
  return -two * tDiff * fSq * expTerm;
}

// src/shared/nominal_to_grid_mappers/map_point_source_to_velocity_grid_point.hpp:61-133
size_t mapSource(const App& a, double angleDeg, double depthKm) {
  const double rSurfKm = a.bounds[3] / 1000.0;
  const double radiusKm = rSurfKm - depthKm;
  const double thLRad = degToRad(a.bounds[0]), thRRad = degToRad(a.bounds[1]);
  double srcRad = degToRad(angleDeg);
  if (angleDeg < a.bounds[0] || angleDeg > a.bounds[1])
    throw std::runtime_error("Source angle is outside theta axis");
  if (radiusKm * 1000.0 < a.bounds[2] || radiusKm * 1000.0 > a.bounds[3])
    throw std::runtime_error("Source angle is outside radius axis");
  if (essentiallyEqual(angleDeg, a.bounds[0])) srcRad = thLRad + a.dth;
  if (essentiallyEqual(angleDeg, a.bounds[1])) srcRad = thRRad - a.dth;
  double trial = std::numeric_limits<double>::max();
  size_t out = 0;
  for (size_t i = 0; i < a.nVp; ++i) {
    const size_t g = a.gVp[i * 5];
    const double d = computePolarDistance(1.0 / a.cVp[g * 2 + 1], a.cVp[g * 2], radiusKm * 1000.0, srcRad);
    if (d <= trial) { trial = d; out = g; }
  }
  return out;
}

// src/shared/nominal_to_grid_mappers/map_nominal_location_to_velocity_grid_point.hpp:57-117
size_t mapReceiver(const App& a, double angleDeg) {
  const double rM = a.bounds[3];
  if (angleDeg < a.bounds[0] || angleDeg > a.bounds[1])
    throw std::runtime_error("Target angle is outside theta axis");
  const double th = degToRad(angleDeg);
  double trial = std::numeric_limits<double>::max();
  size_t out = 0;
  for (size_t i = 0; i < a.nVp; ++i) {
    const size_t g = a.gVp[i * 5];
    const double d = computePolarDistance(1.0 / a.cVp[g * 2 + 1], a.cVp[g * 2], rM, th);
    if (d <= trial) { trial = d; out = g; }
  }
  return out;
}

// KokkosSparse::spmv(N, alpha, A, x, 1, y) restated per column of a
// row-major [n x M] multivector: y(r,:) += alpha * sum_k val[k] x(ind[k],:).
// src/kokkos/fom_update_kernels.hpp:84,103,154,180
void spmmAdd(const Csr& J, size_t nrows, size_t M, double alpha, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)nrows; ++r) {
    for (size_t m = 0; m < M; ++m) {
      double s = 0.0;
      for (size_t k = J.ptr[r]; k < J.ptr[r + 1]; ++k) s += J.val[k] * x[J.ind[k] * M + m];
      y[r * M + m] += alpha * s;
    }
  }
}

void setErr(const std::exception& e) { g_err = e.what(); }

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// material kind/params as in Material above. Returns an App* or null.
void* orc_app_create(const char* meshDir, int matKind, const double* params, int nParams) {
  try {
    App* a = new App();
    const std::string dir(meshDir);
    readMeshInfo(*a, dir);
    readGraph(*a, dir, false);
    readCoeff(*a, dir);
    readGraph(*a, dir, true);
    Material m;
    m.kind = matKind;
    m.p.assign(params, params + nParams);
    m.rSurfM = a->bounds[3];
    setMaterial(*a, m);
    fillVp(*a);
    fillSp(*a);
    return a;
  } catch (const std::exception& e) { setErr(e); return nullptr; }
}

// The same App for a mesh generated in memory by restating the Python mesh
// generator cell by cell: meshing/helpers/full_mesh_utils.py:8-89 (cells in
// natural order, dof gids vp/srp/stp), unit_cell.py:58-66 (gid counters),
// build_graph_impl.py:7-167 (neighbours and +-1/0 multipliers),
// write_files.py (values printed with %.20f and parsed back).
void* orc_app_create_structured(uint64_t nr, uint64_t nth, double thLDeg, double thRDeg, double rCmbKm,
                                double rSurfKm, int matKind, const double* params, int nParams) {
  try {
    App* a = new App();
    auto rt = [](double x) {  // %.20f print + strtod parse of the mesh files
      // 20 decimals carry >= 17 significant digits for |x| >= 1e-3: exact round trip
      if (x == 0.0 || std::abs(x) >= 1e-3) return x;
      char buf[64];
      std::snprintf(buf, sizeof buf, "%.20f", x);
      return std::strtod(buf, nullptr);
    };
    const double thL = thLDeg * M_PI / 180., thR = thRDeg * M_PI / 180.;
    const double dth = (thR - thL) / (double)(nth - 1), dr = (rSurfKm - rCmbKm) / (double)(nr - 1);
    a->bounds[0] = thLDeg; a->bounds[1] = thRDeg;
    a->bounds[2] = rCmbKm * 1000.0; a->bounds[3] = rSurfKm * 1000.0;
    a->dth = rt(dth); a->dthInv = 1.0 / a->dth;
    a->drr = rt(dr) * 1000.0; a->drrInv = 1.0 / a->drr;
    a->nr = nr; a->nth = nth;
    // pass 1: gids of every cell's dofs (vp, then srp if j<nr-1, then stp if i<nth-1)
    const size_t ncell = nr * nth;
    std::vector<long> gvp(ncell), gsrp(ncell, -1), gstp(ncell, -1);
    long cv = -1, cs = -1;
    for (size_t j = 0; j < nr; ++j)
      for (size_t i = 0; i < nth; ++i) {
        const size_t k = j * nth + i;
        gvp[k] = ++cv;
        if (j < nr - 1) gsrp[k] = ++cs;
        if (i < nth - 1) gstp[k] = ++cs;
      }
    a->nVp = cv + 1; a->nSp = cs + 1;
    a->gVp.assign(a->nVp * 5, 0); a->cVp.assign(a->nVp * 2, 0); a->cotVp.assign(a->nVp, 0);
    a->coeffVp.assign(a->nVp * 4, 0);
    a->gSp.assign(a->nSp * 3, 0); a->cSp.assign(a->nSp * 2, 0); a->cotSp.assign(a->nSp, 0);
    a->labels.assign(a->nSp, 0);
    auto setPt = [&](bool sp, size_t gid, double th, double rKm, bool onSym) {
      auto& c = sp ? a->cSp : a->cVp;
      auto& cot = sp ? a->cotSp : a->cotVp;
      c[gid * 2] = rt(th);
      c[gid * 2 + 1] = 1.0 / (rt(rKm) * 1000.0);
      cot[gid] = onSym ? 0.0 : computeCotangent(c[gid * 2]);
    };
    for (size_t j = 0; j < nr; ++j)
      for (size_t i = 0; i < nth; ++i) {
        const size_t k = j * nth + i;
        const double th = thL + i * dth, r = rCmbKm + j * dr;
        const bool sym = (i == 0 || i == nth - 1);
        // vp: W, N, E, S with multipliers (build_graph_impl.py:57-120)
        long W, N, E, S;
        double c[4];
        if (i >= 1) { W = gstp[k - 1]; c[0] = 1.; } else { W = gstp[k]; c[0] = 0.; }
        if (i < nth - 1) { E = gstp[k]; c[2] = 1.; } else { E = gstp[k - 1]; c[2] = 0.; }
        if (j < nr - 1) { N = gsrp[k]; c[1] = 1.; } else { N = gsrp[k - nth]; c[1] = -1.; }
        if (j >= 1) { S = gsrp[k - nth]; c[3] = 1.; } else { S = gsrp[k]; c[3] = -1.; }
        if (sym) c[0] = c[1] = c[2] = c[3] = 0.;
        const size_t g = gvp[k];
        size_t* gr = &a->gVp[g * 5];
        gr[0] = g; gr[1] = W; gr[2] = N; gr[3] = E; gr[4] = S;
        for (int q = 0; q < 4; ++q) a->coeffVp[g * 4 + q] = c[q];
        setPt(false, g, th, r, sym);
        if (j < nr - 1) {  // srp: north vp, south vp
          const size_t q = gsrp[k];
          a->gSp[q * 3] = q; a->gSp[q * 3 + 1] = gvp[k + nth]; a->gSp[q * 3 + 2] = gvp[k];
          a->labels[q] = 1;
          setPt(true, q, th, r + dr * 0.5, sym);
        }
        if (i < nth - 1) {  // stp: west vp, east vp
          const size_t q = gstp[k];
          a->gSp[q * 3] = q; a->gSp[q * 3 + 1] = gvp[k]; a->gSp[q * 3 + 2] = gvp[k + 1];
          a->labels[q] = 2;
          setPt(true, q, th + dth * 0.5, r, false);
        }
      }
    Material m;
    m.kind = matKind;
    m.p.assign(params, params + nParams);
    m.rSurfM = a->bounds[3];
    setMaterial(*a, m);
    fillVp(*a);
    fillSp(*a);
    return a;
  } catch (const std::exception& e) { setErr(e); return nullptr; }
}

void orc_app_destroy(void* h) { delete (App*)h; }

// info: [nVp, nSp, nth, nr, nnzVp, nnzSp]; dinfo: [thL, thR, rCmb, rSurf, dth, drr, minVs, maxVs]
void orc_app_info(void* h, uint64_t* info, double* dinfo) {
  const App& a = *(App*)h;
  info[0] = a.nVp; info[1] = a.nSp; info[2] = a.nth; info[3] = a.nr;
  info[4] = a.Jvp.val.size(); info[5] = a.Jsp.val.size();
  for (int i = 0; i < 4; ++i) dinfo[i] = a.bounds[i];
  dinfo[4] = a.dth; dinfo[5] = a.drr; dinfo[6] = a.minVs; dinfo[7] = a.maxVs;
}

void orc_app_csr(void* h, int dof, uint64_t* ptr, uint64_t* ind, double* val) {
  const App& a = *(App*)h;
  const Csr& J = dof == 0 ? a.Jvp : a.Jsp;
  for (size_t i = 0; i < J.ptr.size(); ++i) ptr[i] = J.ptr[i];
  for (size_t i = 0; i < J.ind.size(); ++i) { ind[i] = J.ind[i]; val[i] = J.val[i]; }
}

// graph [n*5 | n*3], coords [n*2] (theta, 1/r), labels (sp only)
void orc_app_graph(void* h, int dof, uint64_t* graph, double* coords, int32_t* labels) {
  const App& a = *(App*)h;
  const auto& g = dof == 0 ? a.gVp : a.gSp;
  const auto& c = dof == 0 ? a.cVp : a.cSp;
  for (size_t i = 0; i < g.size(); ++i) graph[i] = g[i];
  for (size_t i = 0; i < c.size(); ++i) coords[i] = c[i];
  if (dof == 1 && labels) for (size_t i = 0; i < a.labels.size(); ++i) labels[i] = a.labels[i];
}

void orc_app_material(void* h, double* rhoInv, double* shearMod) {
  const App& a = *(App*)h;
  std::copy(a.rhoInv.begin(), a.rhoInv.end(), rhoInv);
  std::copy(a.shearMod.begin(), a.shearMod.end(), shearMod);
}

int64_t orc_map_source(void* h, double depthKm, double angleDeg) {
  try { return (int64_t)mapSource(*(App*)h, angleDeg, depthKm); }
  catch (const std::exception& e) { setErr(e); return -1; }
}

int64_t orc_map_receiver(void* h, double angleDeg) {
  try { return (int64_t)mapReceiver(*(App*)h, angleDeg); }
  catch (const std::exception& e) { setErr(e); return -1; }
}

// Material::computeAt at n radii [m] for a domain whose surface lies at
// r_surf_m (the depth origin of unilayer and bilayer); rho [kg/m^3], vs [m/s].
void orc_material_at(int kind, const double* params, int nparams, double rSurfM, uint64_t n, const double* r,
                     double* rho, double* vs) {
  Material m;
  m.kind = kind;
  m.p.assign(params, params + nparams);
  m.rSurfM = rSurfM;
  for (uint64_t i = 0; i < n; ++i) m.computeAt(r[i], rho[i], vs[i]);
}

double orc_signal(int kind, double delay, double period, double t) {
  return signalAt(kind, delay, period, t);
}

// One FOM run over M forcing columns (rank2 = 0 requires M == 1).
// src/kokkos/fom_run.hpp:57-161 ; forcing: forcing_rank_one.hpp:150-178,
// forcing_rank_two.hpp:143-151 ; rank-1 add KokkosBlas::mult
// (fom_update_kernels.hpp:89), rank-2 add AddForcingRank2 (:106-129).
// Seismogram [R x T x M] LayoutLeft (seismogram.hpp:52-79, 192-214);
// snapshots [n x T x M] LayoutLeft (state_observer.hpp:52-78, 174-203).
// Any output pointer may be null. Returns 0 or -1 (orc_last_error).
int orc_fom_run(void* h, double dt, uint64_t nsteps, uint64_t M, int rank2, int sigKind,
                const double* depths, const double* angles, const double* periods,
                const double* delays, uint64_t R, const double* recvAngles, uint64_t seisFreq,
                double* seis, uint64_t snapFreqVp, double* snapVp, uint64_t snapFreqSp,
                double* snapSp, double* finalVp, double* finalSp) {
  try {
    const App& a = *(App*)h;
    if (!rank2 && M != 1) throw std::runtime_error("rank-1 run needs M == 1");
    std::vector<size_t> src(M), rec(R);
    for (size_t m = 0; m < M; ++m) src[m] = mapSource(a, angles[m], depths[m]);
    for (size_t r = 0; r < R; ++r) rec[r] = mapReceiver(a, recvAngles[r]);
    std::vector<double> xv(a.nVp * M, 0.0), xs(a.nSp * M, 0.0), f(M, 0.0);
    const size_t Ts = seisFreq ? nsteps / seisFreq : 0;
    const size_t Tv = snapFreqVp ? nsteps / snapFreqVp : 0;
    const size_t Tsp = snapFreqSp ? nsteps / snapFreqSp : 0;
    size_t cs = 0, cv = 0, csp = 0;
    double timeVp = 0.0;
    for (size_t step = 1; step <= nsteps; ++step) {
      // forcing value at time (step-1)*dt computed as a product, not a running sum
      for (size_t m = 0; m < M; ++m) f[m] = signalAt(sigKind, delays[m], periods[m], timeVp);
      spmmAdd(a.Jvp, a.nVp, M, dt, xs.data(), xv.data());
      if (!rank2) {
        const size_t g = src[0];
        xv[g] = xv[g] + dt * a.rhoInv[g] * f[0];
      } else {
        for (size_t m = 0; m < M; ++m) {
          const size_t g = src[m];
          xv[g * M + m] += a.rhoInv[g] * f[m] * dt;
        }
      }
      if (snapVp && snapFreqVp && step % snapFreqVp == 0) {
        for (size_t m = 0; m < M; ++m)
          for (size_t i = 0; i < a.nVp; ++i) snapVp[i + a.nVp * (cv + Tv * m)] = xv[i * M + m];
        ++cv;
      }
      if (seis && seisFreq && step % seisFreq == 0) {
        for (size_t m = 0; m < M; ++m)
          for (size_t r = 0; r < R; ++r) seis[r + R * (cs + Ts * m)] = xv[rec[r] * M + m];
        ++cs;
      }
      timeVp = step * dt;
      spmmAdd(a.Jsp, a.nSp, M, dt, xv.data(), xs.data());
      if (snapSp && snapFreqSp && step % snapFreqSp == 0) {
        for (size_t m = 0; m < M; ++m)
          for (size_t i = 0; i < a.nSp; ++i) snapSp[i + a.nSp * (csp + Tsp * m)] = xs[i * M + m];
        ++csp;
      }
    }
    if (finalVp) std::copy(xv.begin(), xv.end(), finalVp);
    if (finalSp) std::copy(xs.begin(), xs.end(), finalSp);
    return 0;
  } catch (const std::exception& e) { setErr(e); return -1; }
}

// Galerkin projection: ArV = PhiV^T (Jvp PhiS) [Kv x Ks], ArS = PhiS^T (Jsp PhiV)
// [Ks x Kv]; Phi column-major (LayoutLeft), Ar row-major.
// src/kokkos/rom_compute_jacobians.hpp:58-109
int orc_rom_project(void* h, const double* phiV, uint64_t Kv, const double* phiS, uint64_t Ks,
                    double* ArV, double* ArS) {
  const App& a = *(App*)h;
  // tmp = J * phiR (J: nrows x ncols, phiR: ncols x KR col-major), out = phiL^T tmp
  auto proj = [&](const Csr& J, size_t nrows, size_t ncols, const double* phiL, size_t KL,
                  const double* phiR, size_t KR, double* out) {
    std::vector<double> tmp(nrows * KR, 0.0);  // column-major
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)KR; ++c)
      for (size_t r = 0; r < nrows; ++r) {
        double s = 0.0;
        for (size_t k = J.ptr[r]; k < J.ptr[r + 1]; ++k) s += J.val[k] * phiR[J.ind[k] + ncols * c];
        tmp[r + nrows * c] = s;
      }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)KL; ++i)
      for (size_t j = 0; j < KR; ++j) {
        double s = 0.0;
        for (size_t r = 0; r < nrows; ++r) s += phiL[r + nrows * i] * tmp[r + nrows * j];
        out[i * KR + j] = s;
      }
  };
  proj(a.Jvp, a.nVp, a.nSp, phiV, Kv, phiS, Ks, ArV);
  proj(a.Jsp, a.nSp, a.nVp, phiS, Ks, phiV, Kv, ArS);
  return 0;
}

// Rank-2 reduced leapfrog in FP64 (rom_run_rank_one.hpp:112-151 per column;
// torevise/run_rom_rank_two_forcing.hpp:149-188 batched form):
//   qv += dt * ArV qs ; qv += (f*dt) * B[:,m] ; y = C qv (every seisFreq) ;
//   qs += dt * ArS qv.
// ArV [Kv x Ks] row-major, ArS [Ks x Kv] row-major, B [Kv x M] row-major,
// ftab [nsteps x M] (value for step i at row i-1), C [R x Kv] row-major.
// seis [R x T x M] LayoutLeft; qvOut [Kv x M], qsOut [Ks x M] row-major.
int orc_rom_run(uint64_t Kv, uint64_t Ks, const double* ArV, const double* ArS, const double* B,
                const double* ftab, double dt, uint64_t nsteps, uint64_t M, uint64_t R,
                const double* C, uint64_t seisFreq, double* seis, double* qvOut, double* qsOut) {
  std::vector<double> qv(Kv * M, 0.0), qs(Ks * M, 0.0);
  const size_t T = seisFreq ? nsteps / seisFreq : 0;
  size_t cnt = 0;
  for (size_t step = 1; step <= nsteps; ++step) {
#pragma omp parallel for schedule(static)
    for (long m = 0; m < (long)M; ++m) {
      const double fdt = ftab[(step - 1) * M + m] * dt;
      for (size_t i = 0; i < Kv; ++i) {
        double s = 0.0;
        for (size_t k = 0; k < Ks; ++k) s += ArV[i * Ks + k] * qs[k * M + m];
        qv[i * M + m] += dt * s;
      }
      for (size_t i = 0; i < Kv; ++i) qv[i * M + m] += fdt * B[i * M + m];
    }
    if (seis && seisFreq && step % seisFreq == 0) {
      for (size_t m = 0; m < M; ++m)
        for (size_t r = 0; r < R; ++r) {
          double s = 0.0;
          for (size_t k = 0; k < Kv; ++k) s += C[r * Kv + k] * qv[k * M + m];
          seis[r + R * (cnt + T * m)] = s;
        }
      ++cnt;
    }
#pragma omp parallel for schedule(static)
    for (long m = 0; m < (long)M; ++m)
      for (size_t i = 0; i < Ks; ++i) {
        double s = 0.0;
        for (size_t k = 0; k < Kv; ++k) s += ArS[i * Kv + k] * qv[k * M + m];
        qs[i * M + m] += dt * s;
      }
  }
  if (qvOut) std::copy(qv.begin(), qv.end(), qvOut);
  if (qsOut) std::copy(qs.begin(), qs.end(), qsOut);
  return 0;
}

// The reconstruction loop of src/tools/main_reconstruct_seismo.cc:28-38:
// out(i, j) = 0, then out(i, j) += phi(i, k) * q(k, j) for k = 0 .. K-1.
// reconstructFomState's KokkosBlas::gemv (main_reconstruct_fom_state.cc:
// 158-160) computes the same sums in an order that library leaves open.
// phi LayoutLeft [n x K], q LayoutLeft [K x ncols], out LayoutLeft [n x ncols].
int orc_reconstruct(const double* phi, uint64_t n, uint64_t K, const double* q, uint64_t ncols, double* out) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i)
    for (uint64_t j = 0; j < ncols; ++j) {
      double s = 0.0;
      for (uint64_t k = 0; k < K; ++k) s += phi[i + n * k] * q[k + K * j];
      out[i + n * j] = s;
    }
  return 0;
}

}  // extern "C"
