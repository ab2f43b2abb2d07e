// Host-side domain layer for the SHAW time-stepping path: mesh files,
// materials, matrix-free operator coefficients, signals, source/receiver
// mapping, checkers and the snapshot/seismogram file formats.
//
// The mesh is the reference's structured full mesh (meshing/helpers/
// full_mesh_utils.py:8-89, unit_cell.py:58-66, build_graph_impl.py:7-167):
//   vp(i,j)  gid = j*nth + i                                  (i: theta, j: r)
//   srp(i,j) gid = j*(2nth-1) + 2i           for j <= nr-2
//   stp(i,j) gid = j*(2nth-1) + 2i + 1       for j <= nr-2, i <= nth-2
//            gid = (nr-1)*(2nth-1) + i       for j == nr-1 (surface row)
// Graph files are read and checked against this closed form; the kernels
// then address neighbours arithmetically instead of through CSR indices.
#pragma once

#include <array>
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

namespace shaw {

// src/shared/mesh_helpers/mesh_info.hpp:52-214
struct MeshInfo {
  std::string meshDir;
  std::array<double, 4> bounds{};  // thL [deg], thR [deg], rCmb [m], rSurf [m]
  double dth = 0, drr = 0, dthInv = 0, drrInv = 0;
  uint64_t nVp = 0, nSp = 0, nth = 0, nr = 0;
  double maxRadiusKm() const { return bounds[3] / 1000.; }
  double minRadiusKm() const { return bounds[2] / 1000.; }
  double maxArc() const { return dth * bounds[3]; }
  double minArc() const { return dth * bounds[2]; }
};

// Structured index helpers (see header comment).
struct Grid {
  uint64_t nr = 0, nth = 0;
  uint64_t rs() const { return 2 * nth - 1; }
  uint64_t vp(uint64_t i, uint64_t j) const { return j * nth + i; }
  uint64_t srp(uint64_t i, uint64_t j) const { return j * rs() + 2 * i; }
  uint64_t stp(uint64_t i, uint64_t j) const { return j + 1 < nr ? j * rs() + 2 * i + 1 : (nr - 1) * rs() + i; }
  uint64_t nVp() const { return nr * nth; }
  uint64_t nSp() const { return (nr - 1) * nth + nr * (nth - 1); }
};

// A full mesh: coordinates stored as (theta, 1/r[m]) like
// read_graph_file.hpp:95-106, cotangent per point (0 on the symmetry axis),
// and the four vp stencil multipliers (W, N, E, S) of coeff_vp.dat.
struct Mesh {
  MeshInfo info;
  Grid grid;
  std::vector<double> coordsVp, coordsSp;  // [n][2]
  std::vector<double> cotVp, cotSp;
  std::vector<int32_t> labelsSp;           // 1 = srp, 2 = stp
  std::vector<double> stencilVp;           // [nVp][4] in file order W, N, E, S

  // Read mesh_info.dat, graph_vp.dat, coeff_vp.dat, graph_sp.dat from `dir`.
  // Throws std::runtime_error if a file is missing or the graph is not the
  // structured full mesh.
  static Mesh load(const std::string& dir);
  // Build the same mesh in memory (identical doubles to what the Python
  // generator writes and load() parses back). Angles in degrees, radii in km.
  static Mesh structured(uint64_t nr, uint64_t nth, double thLDeg = 0., double thRDeg = 180.,
                         double rCmbKm = 3480., double rSurfKm = 6371.);
  // Write the four mesh files byte-identically to meshing/create_single_mesh.py.
  static void write(const std::string& dir, uint64_t nr, uint64_t nth, double thLDeg = 0.,
                    double thRDeg = 180., double rCmbKm = 3480., double rSurfKm = 6371.);
};

// Material models. src/shared/material_models/material_model_*.hpp;
// kind "layered" (piecewise-linear in r between radius nodes) is added for
// radially layered synthetic Earth models.
struct Material {
  enum Kind { Unilayer = 0, Bilayer = 1, Prem = 2, Layered = 3 };
  int kind = Unilayer;
  std::vector<double> params;  // same flattening as the C-ABI (include/shaw_amd.h)
  double rSurfM = 0;
  void computeAt(double radiusM, double thetaRad, double& rho, double& vs) const;
  void validate(const MeshInfo& mi) const;
};

// Matrix-free FOM operators. coefVp[p][4] holds the CSR row of J_vp in slot
// order N, S, W, E where a merged N==S (CMB/surface rows) or W==E (axis
// columns) entry sits in the N (resp. W) slot and the S (resp. E) slot is 0;
// coefSp[q][2] holds the CSR row of J_sp (srp: N, S; stp: W, E).
// Values are computed exactly as src/kokkos/shwavepp.hpp:319-466.
struct Operators {
  Grid grid;
  std::vector<double> coefVp, coefSp, rhoInv, shearMod;
  double minVs = std::numeric_limits<double>::max();
  double maxVs = std::numeric_limits<double>::min();
  static Operators build(const Mesh& mesh, const Material& mat);
};

// src/shared/various/signal.hpp:52-128
enum class SignalKind { Ricker = 0, Sinusoid = 1, GaussDer = 2, Unknown = 3 };
SignalKind signalKindFromString(const std::string& s);
std::string signalKindToString(SignalKind k);
struct Signal {
  SignalKind kind = SignalKind::Ricker;
  double delay = 0, period = 1, freq = 1, freqSq = 1;
  Signal() = default;
  Signal(SignalKind k, double delayS, double periodS)
      : kind(k), delay(delayS), period(periodS), freq(1.0 / periodS), freqSq(freq * freq) {}
  double operator()(double t) const;
};
// table[(step-1)*M + m] = signals[m]((step-1)*dt), step = 1..nsteps; the time is
// the product (step-1)*dt as in forcing_rank_one.hpp:168-178 and fom_run.hpp:123.
std::vector<double> forcingTable(const std::vector<Signal>& signals, double dt, uint64_t nsteps);

// src/shared/nominal_to_grid_mappers/*.hpp. Nearest vp point by polar
// distance with ties to the last gid, evaluated over a window around the
// analytic nearest row/column (identical result to the full scan, which
// mapSourceFullScan performs for checking).
uint64_t mapSource(const Mesh& m, double depthKm, double angleDeg, bool verbose = false);
uint64_t mapReceiver(const Mesh& m, double angleDeg, bool verbose = false);
uint64_t mapSourceFullScan(const Mesh& m, double depthKm, double angleDeg);
uint64_t mapReceiverFullScan(const Mesh& m, double angleDeg);

// src/shared/checkers/check_cfl.hpp:52-75, check_dispersion_criterion.hpp:52-84
void checkCfl(const MeshInfo& mi, double dt, double maxVel);
void checkDispersion(const MeshInfo& mi, double maxFreq, double minVel);

// File formats, src/shared/io/matrix_write.hpp:60-191, matrix_read.hpp:57-157.
// `A` is column-major (LayoutLeft) [m x n] or [s0 x s1 x s2].
void writeMatrix(const std::string& file, const double* A, uint64_t m, uint64_t n, bool binary,
                 bool writeSize);
void write3d(const std::string& file, const double* A, uint64_t s0, uint64_t s1, uint64_t s2,
             bool binary, bool writeSize);
// Reads a 2D matrix with its size header; returns column-major data.
std::vector<double> readMatrix(const std::string& file, bool binary, uint64_t& m, uint64_t& n);
void writeCoordinates(const Mesh& mesh);  // coords_vp.txt, coords_sp.txt (shwavepp.hpp:231-243)

}  // namespace shaw
