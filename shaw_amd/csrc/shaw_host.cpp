#include "shaw_host.hpp"
#include "formulas.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace shaw {

namespace {

// src/shared/various/equality.hpp:54-58
bool essentiallyEqual(double a, double b) {
  const double eps = std::numeric_limits<double>::epsilon();
  return std::abs(a - b) <= ((std::abs(a) > std::abs(b) ? std::abs(b) : std::abs(a)) * eps);
}
// src/shared/various/angular_helpers.hpp:52-91
double cotangent(double theta) {
  const double c = std::cos(theta), s = std::sin(theta);
  return essentiallyEqual(s, 0.0) ? 0.0 : c / s;
}
double deg2rad(double v) { constexpr double f = M_PI / 180.0; return v * f; }
double rad2deg(double v) { constexpr double f = 180.0 / M_PI; return v * f; }
double polarDistance(double r1, double th1, double r2, double th2) {
  return std::sqrt((r1 * r1) + (r2 * r2) - 2.0 * r1 * r2 * std::cos(th2 - th1));
}

// Read-only memory map of a text file with a tiny token scanner.
class TextFile {
 public:
  explicit TextFile(const std::string& path) {
    fd_ = ::open(path.c_str(), O_RDONLY);
    if (fd_ < 0) throw std::runtime_error("file: " + path + " not found");
    struct stat st;
    if (fstat(fd_, &st) != 0) throw std::runtime_error("cannot stat " + path);
    n_ = (size_t)st.st_size;
    if (n_) {
      p_ = (const char*)mmap(nullptr, n_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p_ == MAP_FAILED) throw std::runtime_error("cannot map " + path);
      madvise((void*)p_, n_, MADV_SEQUENTIAL);
    }
    path_ = path;
  }
  ~TextFile() {
    if (p_ && p_ != MAP_FAILED) munmap((void*)p_, n_);
    if (fd_ >= 0) ::close(fd_);
  }
  bool eof() { skipBlank(); return i_ >= n_; }
  // skip spaces and newlines
  void skipBlank() { while (i_ < n_ && (p_[i_] == ' ' || p_[i_] == '\n' || p_[i_] == '\r' || p_[i_] == '\t')) ++i_; }
  void skipSpaces() { while (i_ < n_ && (p_[i_] == ' ' || p_[i_] == '\t' || p_[i_] == '\r')) ++i_; }
  long long integer() {
    skipSpaces();
    bool neg = false;
    if (i_ < n_ && (p_[i_] == '-' || p_[i_] == '+')) neg = p_[i_++] == '-';
    if (i_ >= n_ || p_[i_] < '0' || p_[i_] > '9') bad("integer");
    long long v = 0;
    while (i_ < n_ && p_[i_] >= '0' && p_[i_] <= '9') v = v * 10 + (p_[i_++] - '0');
    // tolerate a trailing fractional part of zeros ("1.0" read with stoi)
    if (i_ < n_ && p_[i_] == '.') { ++i_; while (i_ < n_ && p_[i_] >= '0' && p_[i_] <= '9') ++i_; }
    return neg ? -v : v;
  }
  double real() {
    skipSpaces();
    char buf[64];
    size_t k = 0;
    while (i_ < n_ && k < 63 && p_[i_] != ' ' && p_[i_] != '\n' && p_[i_] != '\t' && p_[i_] != '\r') buf[k++] = p_[i_++];
    buf[k] = 0;
    char* end = nullptr;
    const double v = std::strtod(buf, &end);
    if (k == 0 || *end) bad("number");
    return v;
  }
  // a number anywhere ahead (line breaks are separators too)
  double realAcrossLines() {
    skipBlank();
    return real();
  }
  void endLine() {
    skipSpaces();
    if (i_ < n_ && p_[i_] != '\n') bad("end of line");
    if (i_ < n_) ++i_;
  }
  [[noreturn]] void bad(const char* what) {
    size_t line = 1 + std::count(p_, p_ + std::min(i_, n_), '\n');
    throw std::runtime_error(path_ + ":" + std::to_string(line) + ": expected " + what);
  }

 private:
  int fd_ = -1;
  const char* p_ = nullptr;
  size_t n_ = 0, i_ = 0;
  std::string path_;
};

MeshInfo readMeshInfo(const std::string& dir) {
  MeshInfo mi;
  mi.meshDir = dir;
  std::ifstream f(dir + "/mesh_info.dat");
  if (!f) throw std::runtime_error("File named: mesh_info.dat not found in " + dir);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string k, v;
    ss >> k;
    if (!(ss >> v)) continue;
    if (k == "thL") mi.bounds[0] = std::stod(v);
    if (k == "thR") mi.bounds[1] = std::stod(v);
    if (k == "rCmb") mi.bounds[2] = std::stod(v) * 1000.0;
    if (k == "rSurf") mi.bounds[3] = std::stod(v) * 1000.0;
    if (k == "dth") { mi.dth = std::stod(v); mi.dthInv = 1.0 / mi.dth; }
    if (k == "dr") { mi.drr = std::stod(v) * 1000.0; mi.drrInv = 1.0 / mi.drr; }
    if (k == "numPtsVp") mi.nVp = std::stoull(v);
    if (k == "numPtsSp") mi.nSp = std::stoull(v);
    if (k == "nth") mi.nth = std::stoull(v);
    if (k == "nr") mi.nr = std::stoull(v);
  }
  return mi;
}

// %.20f print/parse round trip of the Python mesh writer (write_files.py:7)
double roundTrip20(double x) {
  if (x == 0.0 || std::abs(x) >= 1e-3) return x;  // >= 17 significant digits: exact
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.20f", x);
  return std::strtod(buf, nullptr);
}

}  // namespace

// ---------------------------------------------------------------- mesh
Mesh Mesh::load(const std::string& dir) {
  Mesh m;
  m.info = readMeshInfo(dir);
  const MeshInfo& mi = m.info;
  m.grid.nr = mi.nr;
  m.grid.nth = mi.nth;
  const Grid& g = m.grid;
  if (mi.nr < 2 || mi.nth < 3) throw std::runtime_error("mesh needs nr >= 2 and nth >= 3");
  if (g.nVp() != mi.nVp || g.nSp() != mi.nSp)
    throw std::runtime_error("mesh_info.dat point counts do not match a structured full mesh");
  m.coordsVp.assign(mi.nVp * 2, 0.0);
  m.cotVp.assign(mi.nVp, 0.0);
  m.stencilVp.assign(mi.nVp * 4, 0.0);
  m.coordsSp.assign(mi.nSp * 2, 0.0);
  m.cotSp.assign(mi.nSp, 0.0);
  m.labelsSp.assign(mi.nSp, 0);
  std::vector<uint8_t> seen;

  {  // graph_vp.dat: gid onSym theta r W N E S   (read_graph_file.hpp:54-124)
    TextFile f(dir + "/graph_vp.dat");
    seen.assign(mi.nVp, 0);
    while (!f.eof()) {
      const long long gid = f.integer();
      if (gid < 0 || (uint64_t)gid >= mi.nVp || seen[gid]) f.bad("a new vp gid in range");
      seen[gid] = 1;
      const long long onSym = f.integer();
      const double th = f.real();
      const double r = f.real();
      long long nb[4];
      for (auto& x : nb) x = f.integer();
      f.endLine();
      const uint64_t i = gid % g.nth, j = gid / g.nth;
      const uint64_t W = i >= 1 ? g.stp(i - 1, j) : g.stp(i, j);
      const uint64_t E = i + 1 < g.nth ? g.stp(i, j) : g.stp(i - 1, j);
      const uint64_t N = j + 1 < g.nr ? g.srp(i, j) : g.srp(i, j - 1);
      const uint64_t S = j >= 1 ? g.srp(i, j - 1) : g.srp(i, j);
      if ((uint64_t)nb[0] != W || (uint64_t)nb[1] != N || (uint64_t)nb[2] != E || (uint64_t)nb[3] != S)
        throw std::runtime_error("graph_vp.dat: gid " + std::to_string(gid) +
                                 " is not connected as in a structured full mesh");
      m.coordsVp[2 * gid] = th;
      m.coordsVp[2 * gid + 1] = 1.0 / (r * 1000.0);
      m.cotVp[gid] = onSym == 1 ? 0.0 : cotangent(th);
    }
    if (std::count(seen.begin(), seen.end(), 0)) throw std::runtime_error("graph_vp.dat: missing vp points");
  }
  {  // coeff_vp.dat: gid c0 c1 c2 c3 (read_vpcoeff_file.hpp:52-81)
    TextFile f(dir + "/coeff_vp.dat");
    seen.assign(mi.nVp, 0);
    while (!f.eof()) {
      const long long gid = f.integer();
      if (gid < 0 || (uint64_t)gid >= mi.nVp || seen[gid]) f.bad("a new vp gid in range");
      seen[gid] = 1;
      for (int k = 0; k < 4; ++k) m.stencilVp[4 * gid + k] = f.real();
      f.endLine();
    }
    if (std::count(seen.begin(), seen.end(), 0)) throw std::runtime_error("coeff_vp.dat: missing vp points");
  }
  {  // graph_sp.dat: gid label onSym theta r n1 n2
    TextFile f(dir + "/graph_sp.dat");
    seen.assign(mi.nSp, 0);
    const uint64_t lastRow = (g.nr - 1) * g.rs();
    while (!f.eof()) {
      const long long gid = f.integer();
      if (gid < 0 || (uint64_t)gid >= mi.nSp || seen[gid]) f.bad("a new sp gid in range");
      seen[gid] = 1;
      const long long label = f.integer();
      const long long onSym = f.integer();
      const double th = f.real();
      const double r = f.real();
      const long long n1 = f.integer(), n2 = f.integer();
      f.endLine();
      uint64_t e1, e2;
      int elab;
      if ((uint64_t)gid < lastRow) {
        const uint64_t j = gid / g.rs(), k = gid % g.rs();
        if (k % 2 == 0) { elab = 1; e1 = g.vp(k / 2, j + 1); e2 = g.vp(k / 2, j); }
        else { elab = 2; e1 = g.vp((k - 1) / 2, j); e2 = e1 + 1; }
      } else {
        elab = 2;
        e1 = g.vp(gid - lastRow, g.nr - 1);
        e2 = e1 + 1;
      }
      if (label != elab || (uint64_t)n1 != e1 || (uint64_t)n2 != e2)
        throw std::runtime_error("graph_sp.dat: gid " + std::to_string(gid) +
                                 " is not connected as in a structured full mesh");
      m.labelsSp[gid] = (int32_t)label;
      m.coordsSp[2 * gid] = th;
      m.coordsSp[2 * gid + 1] = 1.0 / (r * 1000.0);
      m.cotSp[gid] = onSym == 1 ? 0.0 : cotangent(th);
    }
    if (std::count(seen.begin(), seen.end(), 0)) throw std::runtime_error("graph_sp.dat: missing sp points");
  }
  return m;
}

Mesh Mesh::structured(uint64_t nr, uint64_t nth, double thLDeg, double thRDeg, double rCmbKm,
                      double rSurfKm) {
  if (nr < 2 || nth < 3) throw std::runtime_error("mesh needs nr >= 2 and nth >= 3");
  if (rCmbKm >= rSurfKm || rCmbKm <= 0 || thLDeg >= thRDeg)
    throw std::runtime_error("invalid mesh bounds");
  Mesh m;
  m.grid.nr = nr;
  m.grid.nth = nth;
  const Grid& g = m.grid;
  // create_single_mesh.py:28-37,183-184 ; full_mesh_utils.py:41-67
  const double thL = thLDeg * M_PI / 180., thR = thRDeg * M_PI / 180.;
  const double dth = (thR - thL) / (double)(nth - 1);
  const double dr = (rSurfKm - rCmbKm) / (double)(nr - 1);
  MeshInfo& mi = m.info;
  mi.meshDir = "<structured>";
  mi.bounds = {thLDeg, thRDeg, rCmbKm * 1000.0, rSurfKm * 1000.0};
  mi.dth = roundTrip20(dth);
  mi.dthInv = 1.0 / mi.dth;
  mi.drr = roundTrip20(dr) * 1000.0;
  mi.drrInv = 1.0 / mi.drr;
  mi.nr = nr;
  mi.nth = nth;
  mi.nVp = g.nVp();
  mi.nSp = g.nSp();
  m.coordsVp.resize(mi.nVp * 2);
  m.cotVp.resize(mi.nVp);
  m.stencilVp.resize(mi.nVp * 4);
  m.coordsSp.resize(mi.nSp * 2);
  m.cotSp.resize(mi.nSp);
  m.labelsSp.resize(mi.nSp);
  std::vector<double> thVp(nth), thStp(nth), cotV(nth), cotS(nth);
  for (uint64_t i = 0; i < nth; ++i) {
    const double t = thL + (double)i * dth;
    thVp[i] = roundTrip20(t);
    thStp[i] = roundTrip20(t + dth * 0.5);
    cotV[i] = (i == 0 || i == nth - 1) ? 0.0 : cotangent(thVp[i]);
    cotS[i] = cotangent(thStp[i]);
  }
#pragma omp parallel for schedule(static)
  for (long long jj = 0; jj < (long long)nr; ++jj) {
    const uint64_t j = (uint64_t)jj;
    const double r = rCmbKm + (double)j * dr;
    const double rInv = 1.0 / (r * 1000.0);
    const double rInvSrp = 1.0 / ((r + dr * 0.5) * 1000.0);
    for (uint64_t i = 0; i < nth; ++i) {
      const uint64_t p = g.vp(i, j);
      m.coordsVp[2 * p] = thVp[i];
      m.coordsVp[2 * p + 1] = rInv;
      m.cotVp[p] = cotV[i];
      // build_graph_impl.py:57-120: W, N, E, S multipliers; axis rows all zero
      double c[4] = {i >= 1 ? 1.0 : 0.0, j + 1 < nr ? 1.0 : -1.0, i + 1 < nth ? 1.0 : 0.0, j >= 1 ? 1.0 : -1.0};
      if (i == 0 || i == nth - 1) c[0] = c[1] = c[2] = c[3] = 0.0;
      for (int k = 0; k < 4; ++k) m.stencilVp[4 * p + k] = c[k];
      if (j + 1 < nr) {
        const uint64_t q = g.srp(i, j);
        m.coordsSp[2 * q] = thVp[i];
        m.coordsSp[2 * q + 1] = rInvSrp;
        m.cotSp[q] = cotV[i];
        m.labelsSp[q] = 1;
      }
      if (i + 1 < nth) {
        const uint64_t q = g.stp(i, j);
        m.coordsSp[2 * q] = thStp[i];
        m.coordsSp[2 * q + 1] = rInv;
        m.cotSp[q] = cotS[i];
        m.labelsSp[q] = 2;
      }
    }
  }
  return m;
}

void Mesh::write(const std::string& dir, uint64_t nr, uint64_t nth, double thLDeg, double thRDeg,
                 double rCmbKm, double rSurfKm) {
  if (nr < 2 || nth < 3) throw std::runtime_error("mesh needs nr >= 2 and nth >= 3");
  Grid g{nr, nth};
  const double thL = thLDeg * M_PI / 180., thR = thRDeg * M_PI / 180.;
  const double dth = (thR - thL) / (double)(nth - 1);
  const double dr = (rSurfKm - rCmbKm) / (double)(nr - 1);
  auto open = [&](const char* name) {
    FILE* f = std::fopen((dir + "/" + name).c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + dir + "/" + name);
    return f;
  };
  // write_files.py:12-18, 34-42
  FILE* f = open("mesh_info.dat");
  std::fprintf(f, "thL   %.20f\nthR   %.20f\nrCmb  %.20f\nrSurf %.20f\ndth   %.20f\ndr    %.20f\n",
               thLDeg, thRDeg, rCmbKm, rSurfKm, dth, dr);
  std::fprintf(f, "numPtsVp %llu\nnumPtsSp %llu\nnth %llu\nnr %llu\n", (unsigned long long)g.nVp(),
               (unsigned long long)g.nSp(), (unsigned long long)nth, (unsigned long long)nr);
  std::fclose(f);
  // write_files.py:85-98 (graph_vp), 100-110 (coeff_vp), 112-128 (graph_sp)
  FILE* fv = open("graph_vp.dat");
  FILE* fc = open("coeff_vp.dat");
  for (uint64_t j = 0; j < nr; ++j)
    for (uint64_t i = 0; i < nth; ++i) {
      const double th = thL + (double)i * dth, r = rCmbKm + (double)j * dr;
      const uint64_t W = i >= 1 ? g.stp(i - 1, j) : g.stp(i, j);
      const uint64_t E = i + 1 < nth ? g.stp(i, j) : g.stp(i - 1, j);
      const uint64_t N = j + 1 < nr ? g.srp(i, j) : g.srp(i, j - 1);
      const uint64_t S = j >= 1 ? g.srp(i, j - 1) : g.srp(i, j);
      const int onSym = (i == 0 || i == nth - 1) ? 1 : 0;
      std::fprintf(fv, "%llu %d %.20f %.20f %llu %llu %llu %llu \n", (unsigned long long)g.vp(i, j), onSym,
                   th, r, (unsigned long long)W, (unsigned long long)N, (unsigned long long)E,
                   (unsigned long long)S);
      double c[4] = {i >= 1 ? 1.0 : 0.0, j + 1 < nr ? 1.0 : -1.0, i + 1 < nth ? 1.0 : 0.0, j >= 1 ? 1.0 : -1.0};
      if (onSym) c[0] = c[1] = c[2] = c[3] = 0.0;
      std::fprintf(fc, "%llu %.1f %.1f %.1f %.1f \n", (unsigned long long)g.vp(i, j), c[0], c[1], c[2], c[3]);
    }
  std::fclose(fv);
  std::fclose(fc);
  FILE* fs = open("graph_sp.dat");
  for (uint64_t j = 0; j < nr; ++j)
    for (uint64_t i = 0; i < nth; ++i) {
      const double th = thL + (double)i * dth, r = rCmbKm + (double)j * dr;
      if (j + 1 < nr)
        std::fprintf(fs, "%llu %1d %d %.20f %.20f %llu %llu \n", (unsigned long long)g.srp(i, j), 1,
                     (i == 0 || i == nth - 1) ? 1 : 0, th, r + dr * 0.5, (unsigned long long)g.vp(i, j + 1),
                     (unsigned long long)g.vp(i, j));
      if (i + 1 < nth)
        std::fprintf(fs, "%llu %1d %d %.20f %.20f %llu %llu \n", (unsigned long long)g.stp(i, j), 2, 0,
                     th + dth * 0.5, r, (unsigned long long)g.vp(i, j), (unsigned long long)g.vp(i + 1, j));
    }
  std::fclose(fs);
}

// ---------------------------------------------------------------- material
void Material::computeAt(double r, double /*theta*/, double& rho, double& vs) const {
  if (kind < Unilayer || kind > Layered) throw std::runtime_error("unknown material kind");
  materialAt(kind, params.data(), rSurfM, r, rho, vs);
}

void Material::validate(const MeshInfo& mi) const {
  const size_t need[] = {6, 13, 0, 0};
  if (kind < 0 || kind > 3) throw std::runtime_error("Unknown material model kind");
  if (kind < 2 && params.size() < need[kind]) throw std::runtime_error("material: too few coefficients");
  if (kind == Prem) {  // material_model_prem.hpp:64-80
    if (std::abs(mi.maxRadiusKm() - 6371.) > 1e-13 || std::abs(mi.minRadiusKm() - 3480.) > 1e-13)
      throw std::runtime_error("To use the PREM model, your Earth domain must have r_cmb=3480 km and r_surface=6371 km.");
  }
  if (kind == Layered) {
    if (params.empty()) throw std::runtime_error("layered material: empty");
    const int n = (int)params[0];
    if (n < 2 || params.size() != (size_t)(1 + 3 * n))
      throw std::runtime_error("layered material: need radius/density/velocity lists of equal length >= 2");
    for (int k = 0; k + 1 < n; ++k)
      if (!(params[1 + k + 1] > params[1 + k])) throw std::runtime_error("layered material: radii must increase");
  }
}

// ---------------------------------------------------------------- operators
Operators Operators::build(const Mesh& mesh, const Material& mat) {
  Operators op;
  const Grid& g = mesh.grid;
  const MeshInfo& mi = mesh.info;
  op.grid = g;
  const uint64_t nVp = mi.nVp, nSp = mi.nSp;
  op.rhoInv.resize(nVp);
  op.shearMod.resize(nSp);
  op.coefVp.resize(nVp * 4);
  op.coefSp.resize(nSp * 2);
  const double dthInv = mi.dthInv, drrInv = mi.drrInv;
  // shwavepp.hpp:246-286 (material) and :319-400 (J_vp rows)
#pragma omp parallel for schedule(static)
  for (long long pp = 0; pp < (long long)nVp; ++pp) {
    const uint64_t p = (uint64_t)pp, i = p % g.nth, j = p / g.nth;
    double rho, vs;
    const double rInv = mesh.coordsVp[2 * p + 1];
    mat.computeAt(1.0 / rInv, mesh.coordsVp[2 * p], rho, vs);
    const double rhoInv = 1.0 / rho;
    op.rhoInv[p] = rhoInv;
    const double* c = &mesh.stencilVp[4 * p];
    const double cot = mesh.cotVp[p];
    const double cW = (-rInv * dthInv + rInv * cot) * c[0] * rhoInv;
    const double cN = (drrInv + 1.5 * rInv) * c[1] * rhoInv;
    const double cE = (rInv * dthInv + rInv * cot) * c[2] * rhoInv;
    const double cS = (-drrInv + 1.5 * rInv) * c[3] * rhoInv;
    double* o = &op.coefVp[4 * p];
    if (j == 0 || j + 1 == g.nr) { o[0] = cN + cS; o[1] = 0.0; }
    else { o[0] = cN; o[1] = cS; }
    if (i == 0 || i + 1 == g.nth) { o[2] = cW + cE; o[3] = 0.0; }
    else { o[2] = cW; o[3] = cE; }
  }
  // shwavepp.hpp:262-283 (shear modulus, vs range) and :402-466 (J_sp rows)
  double vmin = std::numeric_limits<double>::max(), vmax = std::numeric_limits<double>::min();
#pragma omp parallel for schedule(static) reduction(min : vmin) reduction(max : vmax)
  for (long long qq = 0; qq < (long long)nSp; ++qq) {
    const uint64_t q = (uint64_t)qq;
    double rho, vs;
    const double rInv = mesh.coordsSp[2 * q + 1];
    mat.computeAt(1.0 / rInv, mesh.coordsSp[2 * q], rho, vs);
    const double G = rho * vs * vs;
    op.shearMod[q] = G;
    if (vs != 0.) vmin = std::min(vmin, vs);
    vmax = std::max(vmax, vs);
    double* o = &op.coefSp[2 * q];
    if (mesh.labelsSp[q] == 1) {
      o[0] = (drrInv - rInv * 0.5) * G;
      o[1] = (-drrInv - rInv * 0.5) * G;
    } else {
      const double cot = mesh.cotSp[q];
      o[0] = (-dthInv - 0.5 * cot) * rInv * G;
      o[1] = (dthInv - 0.5 * cot) * rInv * G;
    }
  }
  op.minVs = vmin;
  op.maxVs = vmax;
  return op;
}

// ---------------------------------------------------------------- signals
SignalKind signalKindFromString(const std::string& s) {
  if (s == "ricker" || s == "Ricker") return SignalKind::Ricker;
  if (s == "sinusoid" || s == "Sinusoid") return SignalKind::Sinusoid;
  if (s == "gaussDer" || s == "gaussder" || s == "GaussDer") return SignalKind::GaussDer;
  return SignalKind::Unknown;
}
std::string signalKindToString(SignalKind k) {
  switch (k) {
    case SignalKind::Ricker: return "ricker";
    case SignalKind::Sinusoid: return "sinusoid";
    case SignalKind::GaussDer: return "gaussDer";
    default: return "unknown";
  }
}

double Signal::operator()(double t) const {
  if (kind != SignalKind::Ricker && kind != SignalKind::Sinusoid && kind != SignalKind::GaussDer)
    throw std::runtime_error("unknown signal kind");
  return signalAt((int)kind, delay, period, freqSq, t);
}

std::vector<double> forcingTable(const std::vector<Signal>& sig, double dt, uint64_t nsteps) {
  const size_t M = sig.size();
  std::vector<double> tab(nsteps * M);
#pragma omp parallel for schedule(static)
  for (long long s = 0; s < (long long)nsteps; ++s) {
    const double t = (double)s * dt;
    for (size_t m = 0; m < M; ++m) tab[(size_t)s * M + m] = sig[m](t);
  }
  return tab;
}

// ---------------------------------------------------------------- mappers
namespace {

struct Target { double r, th; };

uint64_t nearestWindow(const Mesh& m, Target t) {
  const Grid& g = m.grid;
  // rows are ordered by radius, columns by angle: bracket the target analytically
  auto rowR = [&](uint64_t j) { return 1.0 / m.coordsVp[2 * g.vp(0, j) + 1]; };
  auto colT = [&](uint64_t i) { return m.coordsVp[2 * g.vp(i, 0)]; };
  uint64_t lo = 0, hi = g.nr - 1;
  while (hi - lo > 1) { const uint64_t mid = (lo + hi) / 2; (rowR(mid) <= t.r ? lo : hi) = mid; }
  const uint64_t jc = lo;
  lo = 0; hi = g.nth - 1;
  while (hi - lo > 1) { const uint64_t mid = (lo + hi) / 2; (colT(mid) <= t.th ? lo : hi) = mid; }
  const uint64_t ic = lo;
  const uint64_t W = 4;
  const uint64_t j0 = jc > W ? jc - W : 0, j1 = std::min(g.nr - 1, jc + W + 1);
  const uint64_t i0 = ic > W ? ic - W : 0, i1 = std::min(g.nth - 1, ic + W + 1);
  double trial = std::numeric_limits<double>::max();
  uint64_t out = 0;
  for (uint64_t j = j0; j <= j1; ++j)
    for (uint64_t i = i0; i <= i1; ++i) {
      const uint64_t p = g.vp(i, j);
      const double d = polarDistance(1.0 / m.coordsVp[2 * p + 1], m.coordsVp[2 * p], t.r, t.th);
      if (d <= trial) { trial = d; out = p; }
    }
  return out;
}

uint64_t nearestFull(const Mesh& m, Target t) {
  double trial = std::numeric_limits<double>::max();
  uint64_t out = 0;
  for (uint64_t p = 0; p < m.info.nVp; ++p) {
    const double d = polarDistance(1.0 / m.coordsVp[2 * p + 1], m.coordsVp[2 * p], t.r, t.th);
    if (d <= trial) { trial = d; out = p; }
  }
  return out;
}

// map_point_source_to_velocity_grid_point.hpp:61-105
Target sourceTarget(const Mesh& m, double depthKm, double angleDeg) {
  const MeshInfo& mi = m.info;
  const double radiusKm = mi.maxRadiusKm() - depthKm;
  double th = deg2rad(angleDeg);
  if (angleDeg < mi.bounds[0] || angleDeg > mi.bounds[1]) throw std::runtime_error("Source angle is outside theta axis");
  if (radiusKm * 1000. < mi.bounds[2] || radiusKm * 1000. > mi.bounds[3])
    throw std::runtime_error("Source angle is outside radius axis");
  if (essentiallyEqual(angleDeg, mi.bounds[0])) th = deg2rad(mi.bounds[0]) + mi.dth;
  if (essentiallyEqual(angleDeg, mi.bounds[1])) th = deg2rad(mi.bounds[1]) - mi.dth;
  return {radiusKm * 1000., th};
}

// map_nominal_location_to_velocity_grid_point.hpp:57-87
Target receiverTarget(const Mesh& m, double angleDeg) {
  const MeshInfo& mi = m.info;
  if (angleDeg < mi.bounds[0] || angleDeg > mi.bounds[1]) throw std::runtime_error("Target angle is outside theta axis");
  return {mi.bounds[3], deg2rad(angleDeg)};
}

void report(const Mesh& m, const char* what, uint64_t gid) {
  const double th = m.coordsVp[2 * gid], rr = 1.0 / m.coordsVp[2 * gid + 1];
  std::cout << what << " mapped to GID = " << gid << " with depth (km) = " << std::setprecision(17)
            << (m.info.bounds[3] - rr) / 1000. << " with angle (rad) = " << th
            << " angle (deg) = " << rad2deg(th) << std::endl;
}

}  // namespace

uint64_t mapSource(const Mesh& m, double depthKm, double angleDeg, bool verbose) {
  const uint64_t g = nearestWindow(m, sourceTarget(m, depthKm, angleDeg));
  if (verbose) report(m, "Source", g);
  return g;
}
uint64_t mapReceiver(const Mesh& m, double angleDeg, bool verbose) {
  const uint64_t g = nearestWindow(m, receiverTarget(m, angleDeg));
  if (verbose) report(m, "Point", g);
  return g;
}
uint64_t mapSourceFullScan(const Mesh& m, double depthKm, double angleDeg) {
  return nearestFull(m, sourceTarget(m, depthKm, angleDeg));
}
uint64_t mapReceiverFullScan(const Mesh& m, double angleDeg) {
  return nearestFull(m, receiverTarget(m, angleDeg));
}

// ---------------------------------------------------------------- checkers
void checkCfl(const MeshInfo& mi, double dt, double maxVel) {
  const double hRef = std::min(mi.drr, std::min(mi.maxArc(), mi.minArc()));
  const double value = (dt * std::sqrt(2.0) * maxVel) / hRef;
  if (value <= 0.28) {
    std::cout << "CFL: OK! " << std::endl << "cfl = " << value << std::endl;
  } else {
    std::cout << "CFL = " << value << std::endl;
    throw std::runtime_error("CFL error");
  }
}

void checkDispersion(const MeshInfo& mi, double maxFreq, double minVel) {
  const double ratio = minVel / ((double)10 * maxFreq);
  const double f0 = minVel / ((double)10 * std::max(mi.drr, mi.maxArc()));
  std::cout << "targetPeriod = " << 1. / maxFreq << std::endl
            << "centerFreq = " << f0 << std::endl
            << "minPeriod  = " << 1. / f0 << std::endl;
  std::cout << "drr = " << mi.drr << ", ratio = " << ratio << std::endl;
  if (!(mi.drr <= ratio)) throw std::runtime_error("Numerical dispersion criterion violated by radial spacing");
  std::cout << "maxArc = " << mi.maxArc() << ", ratio = " << ratio << std::endl;
  if (!(mi.maxArc() <= ratio)) throw std::runtime_error("Numerical dispersion criterion violated along theta");
}

// ---------------------------------------------------------------- files
void writeMatrix(const std::string& file, const double* A, uint64_t m, uint64_t n, bool binary, bool writeSize) {
  if (binary) {
    std::ofstream out(file, std::ios::out | std::ios::binary | std::ios::trunc);
    if (writeSize) { out.write((const char*)&m, 8); out.write((const char*)&n, 8); }
    out.write((const char*)A, (std::streamsize)(m * n * sizeof(double)));
    return;
  }
  std::ofstream f(file);
  if (writeSize) f << m << " " << n << "\n";
  f << std::setprecision(std::numeric_limits<double>::max_digits10);
  for (uint64_t i = 0; i < m; ++i) {
    for (uint64_t j = 0; j < n; ++j) f << A[i + m * j] << " ";
    f << "\n";
  }
}

void write3d(const std::string& file, const double* A, uint64_t s0, uint64_t s1, uint64_t s2, bool binary,
             bool writeSize) {
  if (binary) {
    std::ofstream out(file, std::ios::out | std::ios::binary | std::ios::trunc);
    if (writeSize) { out.write((const char*)&s0, 8); out.write((const char*)&s1, 8); out.write((const char*)&s2, 8); }
    out.write((const char*)A, (std::streamsize)(s0 * s1 * s2 * sizeof(double)));
    return;
  }
  std::ofstream f(file);
  if (writeSize) f << s0 << " " << s1 << " " << s2 << "\n";
  f << std::setprecision(std::numeric_limits<double>::max_digits10);
  for (uint64_t k = 0; k < s2; ++k)
    for (uint64_t i = 0; i < s0; ++i) {
      for (uint64_t j = 0; j < s1; ++j) f << A[i + s0 * (j + s1 * k)] << " ";
      f << "\n";
    }
}

std::vector<double> readMatrix(const std::string& file, bool binary, uint64_t& m, uint64_t& n) {
  std::vector<double> A;
  if (binary) {  // matrix_read.hpp: [u64 rows][u64 cols] + column-major data
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + file);
    in.read((char*)&m, 8);
    in.read((char*)&n, 8);
    A.resize(m * n);
    in.read((char*)A.data(), (std::streamsize)(m * n * sizeof(double)));
    if (!in) throw std::runtime_error("short read in " + file);
    return A;
  }
  TextFile f(file);
  m = (uint64_t)f.integer();
  n = (uint64_t)f.integer();
  A.resize(m * n);
  for (uint64_t i = 0; i < m; ++i)
    for (uint64_t j = 0; j < n; ++j) A[i + m * j] = f.realAcrossLines();
  return A;
}

void writeCoordinates(const Mesh& mesh) {
  for (int d = 0; d < 2; ++d) {
    const std::vector<double>& c = d == 0 ? mesh.coordsVp : mesh.coordsSp;
    std::ofstream f(d == 0 ? "coords_vp.txt" : "coords_sp.txt");
    f << std::setprecision(std::numeric_limits<double>::max_digits10);
    for (size_t i = 0; i < c.size() / 2; ++i) f << c[2 * i] << " " << 1.0 / c[2 * i + 1] << "\n";
  }
}

}  // namespace shaw
