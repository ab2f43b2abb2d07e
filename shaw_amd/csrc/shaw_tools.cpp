// The reference's offline tools (src/tools/) with the same command lines and
// file formats, their products computed on the GPU through the C ABI:
//   computeThinSVD         POD basis of snapshot files (main_eigen_svd.cc)
//   extractStateFromSnaps  states at given steps from a snapshot file
//                          (main_extract_state_from_snaps.cc)
//   reconstructFomState    Phi q at given steps (main_reconstruct_fom_state.cc)
//   reconstructSeismogram  Phi[gids, :] q for every sample
//                          (main_reconstruct_seismo.cc)
// One program; the name it is invoked under selects the tool. Rank-2
// (three-dimensional) snapshot files are read by all four; the reference's
// POD tool reads two-dimensional files only and its two reconstructors stop
// with "reconstruction for rank2 to renable" for them.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "shaw_amd.h"
#include "shaw_host.hpp"

namespace {

// ---------------------------------------------------------------- options
// The subset of CLI11 the tools use: `--name value...` or `--name=value
// value...`; values run to the next `--` token.
struct Options {
  std::map<std::string, std::vector<std::string>> v;
  Options(int argc, char** argv) {
    std::string cur;
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
        const size_t eq = a.find('=');
        cur = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
        v[cur];
        if (eq != std::string::npos) v[cur].push_back(a.substr(eq + 1));
      } else {
        if (cur.empty()) throw std::runtime_error("unexpected argument " + a);
        v[cur].push_back(a);
      }
    }
  }
  bool has(const std::string& k) const { return v.count(k) && !v.at(k).empty(); }
  const std::vector<std::string>& list(const std::string& k, size_t minCount = 1) const {
    auto it = v.find(k);
    if (it == v.end() || it->second.size() < minCount) throw std::runtime_error("--" + k + " is required");
    return it->second;
  }
  std::string str(const std::string& k) const { return list(k)[0]; }
  uint64_t u64(const std::string& k) const { return toU64(str(k), k); }
  static uint64_t toU64(const std::string& s, const std::string& k) {
    size_t pos = 0;
    long long x = 0;
    try { x = std::stoll(s, &pos); } catch (...) { pos = 0; }
    if (pos != s.size() || x < 0) throw std::runtime_error("--" + k + ": not a non-negative integer: " + s);
    return (uint64_t)x;
  }
  std::vector<uint64_t> u64s(const std::string& k) const {
    std::vector<uint64_t> out;
    for (const auto& s : list(k)) out.push_back(toU64(s, k));
    return out;
  }
  // `--name file format` pairs (CLI11 std::pair options)
  std::pair<std::string, bool> fileFormat(const std::string& k) const {
    const auto& l = list(k, 2);
    if (l[1] != "binary" && l[1] != "ascii") throw std::runtime_error("--" + k + ": format must be binary or ascii");
    return {l[0], l[1] == "binary"};
  }
};

void check(int rc) {
  if (rc != SHAW_OK) throw std::runtime_error(shaw_last_error());
}

// ---------------------------------------------------------------- files
uint64_t fileSize(const std::string& f) {
  std::ifstream in(f, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error("cannot open " + f);
  return (uint64_t)in.tellg();
}

// A binary snapshot file with its size header, two- or three-dimensional
// (matrix_write.hpp:60-91); the extents are told apart by the file size.
// Returned as column-major [n x cols] with cols = s1 (2D) or s1*s2 (3D),
// which is the LayoutLeft order of both.
struct Snapshots {
  uint64_t n = 0, s1 = 0, s2 = 1;
  bool threeD = false;
  std::vector<double> data;
  uint64_t cols() const { return s1 * s2; }
};

Snapshots readSnapshots(const std::string& file, bool binary) {
  Snapshots s;
  if (!binary) {
    s.data = shaw::readMatrix(file, false, s.n, s.s1);
    return s;
  }
  const uint64_t bytes = fileSize(file);
  std::ifstream in(file, std::ios::binary);
  uint64_t h[3] = {0, 0, 0};
  in.read((char*)h, bytes >= 24 ? 24 : 16);
  if (!in) throw std::runtime_error("short read in " + file);
  if (bytes == 16 + 8 * h[0] * h[1]) {
    s.n = h[0];
    s.s1 = h[1];
  } else if (bytes >= 24 && bytes == 24 + 8 * h[0] * h[1] * h[2]) {
    s.n = h[0];
    s.s1 = h[1];
    s.s2 = h[2];
    s.threeD = true;
  } else {
    throw std::runtime_error(file + ": size does not match a 2D or 3D binary header");
  }
  in.seekg(s.threeD ? 24 : 16);
  s.data.resize(s.n * s.cols());
  in.read((char*)s.data.data(), (std::streamsize)(s.data.size() * 8));
  if (!in) throw std::runtime_error("short read in " + file);
  return s;
}

// vector_write.hpp:60-90
void writeVector(const std::string& file, const double* x, uint64_t n, bool binary) {
  if (binary) {
    std::ofstream out(file, std::ios::out | std::ios::binary | std::ios::trunc);
    out.write((const char*)&n, 8);
    out.write((const char*)x, (std::streamsize)(n * 8));
    return;
  }
  std::ofstream f(file);
  f << n << std::endl;
  f << std::setprecision(std::numeric_limits<double>::max_digits10);
  for (uint64_t i = 0; i < n; ++i) f << x[i] << " \n";
}

// A column vector as Eigen prints it: one value per line at the stream's
// precision, right-aligned to the widest value (main_eigen_svd.cc:137-139).
void writeAlignedColumn(const std::string& file, const double* x, uint64_t n) {
  std::vector<std::string> s(n);
  size_t width = 0;
  for (uint64_t i = 0; i < n; ++i) {
    std::ostringstream o;
    o << std::setprecision(std::numeric_limits<double>::max_digits10) << x[i];
    s[i] = o.str();
    width = std::max(width, s[i].size());
  }
  std::ofstream f(file);
  for (uint64_t i = 0; i < n; ++i) {
    if (i) f << "\n";
    f << std::string(width - s[i].size(), ' ') << s[i];
  }
  f << "\n";
}

std::string suffix(const std::string& append) { return append.empty() ? "" : "_" + append; }

double secondsSince(std::chrono::high_resolution_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
}

// ---------------------------------------------------------------- tools
// main_eigen_svd.cc:199-310. For each dof: the snapshots snaps_<dof>_0 of
// every directory side by side, their POD on the GPU, lsv_<dof> (binary,
// [u64 n][u64 k] + column-major U) and sva_<dof> (singular values).
int computeThinSVD(const Options& o) {
  if (!o.has("method")) throw std::runtime_error("Invalid method");
  const uint64_t method = o.u64("method");
  if (method != 1 && method != 2) throw std::runtime_error("Invalid method");
  const auto& dirs = o.list("dirs");
  o.list("informat");
  o.list("outformat");
  for (const std::string dof : {"vp", "sp"}) {
    const auto t0 = std::chrono::high_resolution_clock::now();
    std::vector<double> all;
    uint64_t n = 0, ns = 0;
    for (size_t d = 0; d < dirs.size(); ++d) {
      const std::string file = dirs[d] + "/snaps_" + dof + "_0";
      std::cout << "\nReading snapshots: " << file << std::endl;
      Snapshots s = readSnapshots(file, true);
      std::cout << "Current snapshots size: " << s.n << " " << s.cols() << std::endl;
      if (d == 0) n = s.n;
      else if (s.n != n) throw std::runtime_error("Mismatching # rows of current data with previous");
      all.insert(all.end(), s.data.begin(), s.data.end());
      ns += s.cols();
    }
    std::cout << "Final snapshot matrix size: " << n << " " << ns << std::endl;
    std::cout << "svd:read:" << dof << ": " << std::fixed << std::setprecision(10) << secondsSince(t0) << std::endl;
    std::cout << std::defaultfloat;
    if (n == 0 || ns == 0) throw std::runtime_error("empty snapshot matrix");
    const uint64_t kmax = std::min(n, ns);
    std::vector<double> phi(n * kmax), sigma(ns);
    uint64_t k = 0;
    std::cout << "Computing SVD method " << method << " (GPU)" << std::endl;
    check(shaw_pod_basis_upto(all.data(), n, ns, kmax, phi.data(), sigma.data(), &k));
    std::cout << "svd_matrix_" << dof << "_rank = " << k << std::endl;
    writeAlignedColumn("sva_" + dof, sigma.data(), kmax);
    shaw::writeMatrix("lsv_" + dof, phi.data(), n, k, true, true);
    std::cout << "Done with SVD" << std::endl << "----------------" << std::endl;
    std::cout << "svd:time" << dof << ": " << std::fixed << std::setprecision(10) << secondsSince(t0) << std::endl;
    std::cout << std::defaultfloat;
  }
  return 0;
}

// column index of a target step: step / samplingFreq - 1
// (main_extract_state_from_snaps.cc:65-68, 91)
uint64_t sampleColumn(uint64_t step, uint64_t freq, uint64_t cols) {
  if (freq == 0) throw std::runtime_error("--samplingfreq must be > 0");
  const uint64_t idx = step / freq;
  if (idx == 0 || idx > cols)
    throw std::runtime_error("time step " + std::to_string(step) + " is not in the snapshot matrix (" +
                             std::to_string(cols) + " samples every " + std::to_string(freq) + " steps)");
  return idx - 1;
}

void printSteps(const std::vector<uint64_t>& steps) {
  std::cout << "Target time steps = ";
  for (auto s : steps) std::cout << s << " ";
  std::cout << std::endl;
}

int extractStateFromSnaps(const Options& o) {
  const auto snaps = o.fileFormat("snaps");
  const uint64_t freq = o.u64("samplingfreq"), fSize = o.u64("fsize");
  const auto steps = o.u64s("timesteps");
  const std::string outFmt = o.str("outformat");
  const std::string app = o.has("outfileappend") ? o.str("outfileappend") : "";
  std::cout << "snaps file = " << snaps.first << std::endl
            << "snaps format = " << (snaps.second ? "binary" : "ascii") << std::endl
            << "Size of f = " << fSize << std::endl;
  printSteps(steps);
  std::cout << "Sampling freq = " << freq << std::endl << "Output format = " << outFmt << std::endl;
  const bool outBin = outFmt == "binary";
  if (fSize == 0) throw std::runtime_error("--fsize must be >= 1");
  if (fSize == 1) {
    Snapshots s = readSnapshots(snaps.first, snaps.second);
    if (s.threeD) throw std::runtime_error("--fsize 1 needs a two-dimensional snapshot file");
    std::cout << "snap size: " << s.n << " " << s.s1 << std::endl;
    for (auto step : steps) {
      const uint64_t c = sampleColumn(step, freq, s.s1);
      writeVector("state_timestep_" + std::to_string(step) + suffix(app), s.data.data() + s.n * c, s.n, outBin);
    }
  } else {
    if (!snaps.second) throw std::runtime_error("Rank-2  snaps ascii not supported yet");
    Snapshots s = readSnapshots(snaps.first, true);
    if (!s.threeD) throw std::runtime_error("--fsize >= 2 needs a three-dimensional snapshot file");
    std::cout << "snap size: " << s.n << " " << s.s1 << " " << s.s2 << std::endl;
    for (uint64_t f = 0; f < s.s2; ++f)
      for (auto step : steps) {
        const uint64_t c = sampleColumn(step, freq, s.s1);
        writeVector("state_timestep_" + std::to_string(step) + "_f_" + std::to_string(f) + suffix(app),
                    s.data.data() + s.n * (c + s.s1 * f), s.n, outBin);
      }
  }
  std::cout << "success" << std::endl;
  return 0;
}

// readBasis (read_basis.hpp:56-101): the first `romSize` columns of a basis
// file with its size header
std::vector<double> loadModes(const std::pair<std::string, bool>& file, uint64_t romSize, uint64_t& n) {
  uint64_t m = 0, cols = 0;
  std::vector<double> all = shaw::readMatrix(file.first, file.second, m, cols);
  if (romSize == 0 || romSize > cols)
    throw std::runtime_error("--romsize " + std::to_string(romSize) + " exceeds the " + std::to_string(cols) +
                             " modes in " + file.first);
  n = m;
  all.resize(n * romSize);
  double mn = 0.0, mx = 0.0;  // the reference starts both at 0
  for (double x : all) { mn = std::min(mn, x); mx = std::max(mx, x); }
  std::cout << "POD modes size = " << n << " " << romSize << std::endl
            << "POD modes MIN/MAX = " << mn << " " << mx << std::endl;
  return all;
}

// the leading k entries of columns `cols` of a column-major [rows x *] matrix
std::vector<double> gatherColumns(const std::vector<double>& a, uint64_t rows, uint64_t k,
                                  const std::vector<uint64_t>& cols) {
  std::vector<double> q(k * cols.size());
  for (size_t c = 0; c < cols.size(); ++c)
    std::copy(a.begin() + rows * cols[c], a.begin() + rows * cols[c] + k, q.begin() + k * c);
  return q;
}

int reconstructFomState(const Options& o) {
  const uint64_t romSize = o.u64("romsize");
  const auto modes = o.fileFormat("podmodes");
  const auto rsnaps = o.fileFormat("romsnaps");
  const uint64_t freq = o.u64("samplingfreq"), fSize = o.u64("fsize");
  const auto steps = o.u64s("timesteps");
  const std::string outFmt = o.str("outformat");
  const std::string app = o.has("outfileappend") ? o.str("outfileappend") : "";
  std::cout << "POD modes = " << modes.first << " " << (modes.second ? "binary" : "ascii") << std::endl
            << "ROM snaps = " << rsnaps.first << " " << (rsnaps.second ? "binary" : "ascii") << std::endl
            << "ROM size = " << romSize << std::endl
            << "Size of f = " << fSize << std::endl;
  printSteps(steps);
  std::cout << "Sampling freq = " << freq << std::endl << "Output format = " << outFmt << std::endl;
  if (fSize == 0) throw std::runtime_error("--fsize must be >= 1");
  uint64_t n = 0;
  const std::vector<double> phi = loadModes(modes, romSize, n);
  if (fSize >= 2 && !rsnaps.second) throw std::runtime_error("Rank-2 ROM snaps ascii not supported yet");
  Snapshots s = readSnapshots(rsnaps.first, rsnaps.second);
  if (s.threeD != (fSize >= 2))
    throw std::runtime_error(fSize >= 2 ? "--fsize >= 2 needs a three-dimensional ROM snapshot file"
                                        : "--fsize 1 needs a two-dimensional ROM snapshot file");
  std::cout << "ROM snap size: " << s.n << " " << s.s1 << (s.threeD ? " " + std::to_string(s.s2) : "") << std::endl;
  if (s.n != romSize)
    throw std::runtime_error("ROM snapshot length " + std::to_string(s.n) + " does not match --romsize " +
                             std::to_string(romSize));
  // every requested (forcing, step) column in one product on the GPU
  std::vector<uint64_t> cols;
  for (uint64_t f = 0; f < s.s2; ++f)
    for (auto step : steps) cols.push_back(sampleColumn(step, freq, s.s1) + s.s1 * f);
  const std::vector<double> q = gatherColumns(s.data, s.n, romSize, cols);
  std::vector<double> out(n * cols.size());
  check(shaw_reconstruct(phi.data(), n, romSize, q.data(), cols.size(), out.data()));
  size_t c = 0;
  for (uint64_t f = 0; f < s.s2; ++f)
    for (auto step : steps) {
      std::string file = "fomReconstructedState_timestep_" + std::to_string(step);
      if (s.threeD) file += "_f_" + std::to_string(f);
      if (!s.threeD) writeVector("ROMSTATE.txt", q.data() + romSize * c, romSize, false);
      writeVector(file + suffix(app), out.data() + n * c, n, outFmt == "binary");
      ++c;
    }
  std::cout << "success" << std::endl;
  return 0;
}

// rows `gids` of the first k columns of a binary basis file, read in place
// (main_reconstruct_seismo.cc:42-87); returned column-major [gids x k]
std::vector<double> readModeRows(const std::string& file, const std::vector<uint64_t>& gids, uint64_t k) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + file);
  uint64_t rows = 0, cols = 0;
  in.read((char*)&rows, 8);
  in.read((char*)&cols, 8);
  if (!in || fileSize(file) != 16 + 8 * rows * cols) throw std::runtime_error(file + ": not a binary basis file");
  if (k == 0 || k > cols)
    throw std::runtime_error("--romsize " + std::to_string(k) + " exceeds the " + std::to_string(cols) +
                             " modes in " + file);
  std::vector<double> out(gids.size() * k);
  for (size_t i = 0; i < gids.size(); ++i) {
    if (gids[i] >= rows) throw std::runtime_error("grid id " + std::to_string(gids[i]) + " out of range");
    for (uint64_t j = 0; j < k; ++j) {
      in.seekg((std::streamoff)(16 + 8 * (gids[i] + rows * j)));
      in.read((char*)&out[i + gids.size() * j], 8);
    }
  }
  if (!in) throw std::runtime_error("short read in " + file);
  return out;
}

int reconstructSeismogram(const Options& o) {
  const uint64_t romSize = o.u64("romsize");
  const auto modes = o.fileFormat("podmodes");
  const auto rsnaps = o.fileFormat("romsnaps");
  const auto gids = o.u64s("gridids");
  const uint64_t fSize = o.u64("fsize");
  const std::string outFmt = o.str("outformat");
  const std::string app = o.has("outfileappend") ? o.str("outfileappend") : "";
  std::cout << "POD modes = " << modes.first << " " << (modes.second ? "binary" : "ascii") << std::endl
            << "ROM snaps = " << rsnaps.first << " " << (rsnaps.second ? "binary" : "ascii") << std::endl
            << "ROM size = " << romSize << std::endl
            << "Size of f = " << fSize << std::endl
            << "Output format = " << outFmt << std::endl
            << "Target gids = ";
  for (auto g : gids) std::cout << g << " ";
  std::cout << std::endl;
  if (fSize == 0) throw std::runtime_error("--fsize must be >= 1");
  const std::vector<double> phi = readModeRows(modes.first, gids, romSize);
  if (!rsnaps.second)
    throw std::runtime_error(fSize == 1 ? "Rank-1 ROM snaps ascii not supported yet"
                                        : "Rank-2 ROM snaps ascii not supported yet");
  Snapshots s = readSnapshots(rsnaps.first, true);
  if (s.threeD != (fSize >= 2))
    throw std::runtime_error(fSize >= 2 ? "--fsize >= 2 needs a three-dimensional ROM snapshot file"
                                        : "--fsize 1 needs a two-dimensional ROM snapshot file");
  std::cout << "ROM snap size: " << s.n << " " << s.s1 << (s.threeD ? " " + std::to_string(s.s2) : "") << std::endl;
  // the sum runs over the first romSize coordinates (main_reconstruct_seismo.cc:35)
  if (s.n < romSize)
    throw std::runtime_error("ROM snapshots hold " + std::to_string(s.n) + " coordinates, fewer than --romsize");
  std::vector<uint64_t> cols(s.cols());
  for (uint64_t c = 0; c < cols.size(); ++c) cols[c] = c;
  const std::vector<double> q = gatherColumns(s.data, s.n, romSize, cols);
  const uint64_t np = gids.size();
  std::vector<double> S(np * cols.size());
  check(shaw_reconstruct(phi.data(), np, romSize, q.data(), cols.size(), S.data()));
  if (!s.threeD) {
    shaw::writeMatrix("rom_seismo.txt", S.data(), np, s.s1, false, false);
  } else {
    for (uint64_t f = 0; f < s.s2; ++f)
      shaw::writeMatrix("seismo_f_" + std::to_string(f) + suffix(app), S.data() + np * s.s1 * f, np, s.s1,
                        outFmt == "binary", true);
  }
  std::cout << "success" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::string name = argv[0];
  const size_t slash = name.find_last_of('/');
  if (slash != std::string::npos) name = name.substr(slash + 1);
  try {
    const Options o(argc, argv);
    if (name == "computeThinSVD") return computeThinSVD(o);
    if (name == "extractStateFromSnaps") return extractStateFromSnaps(o);
    if (name == "reconstructFomState") return reconstructFomState(o);
    if (name == "reconstructSeismogram") return reconstructSeismogram(o);
    std::cerr << "invoke as computeThinSVD, extractStateFromSnaps, reconstructFomState or reconstructSeismogram"
              << std::endl;
    return 2;
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
}
