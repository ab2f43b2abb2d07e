// shawExe: drop-in replacement of the reference executable
// (src/kokkos/main.cc:41-94) -- `shawExe <input.yaml>` reads the same input
// file and mesh files, writes the same output files (snaps_vp_<run>,
// snaps_sp_<run>, seismogram_<run>, coords_vp.txt, coords_sp.txt) in the same
// formats and prints the same stdout keys, with the time stepping on the GPU
// through the C ABI (include/shaw_amd.h).
//
// Drivers (L3 of the reference):
//   FOM rank-1: fom_problem_rank_one.hpp:105-191 (single and multi-forcing,
//               realizations in the nested order depth -> period -> angle ->
//               delay; realizations are batched as independent columns of one
//               GPU run, each column computed exactly as a rank-1 run)
//   FOM rank-2: fom_problem_rank_two.hpp:109-193 (sets of forcingSize)
//   ROM:        rom_problem_rank_one.hpp:90-159 (rank-1; with forcingSize the
//               batched rank-2 ROM of torevise/rom_problem_rank_two_forcing.hpp)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "shaw_amd.h"
#include "shaw_host.hpp"
#include "shaw_parser.hpp"

namespace {

using shaw::InputParser;

void check(int rc) {
  if (rc != SHAW_OK) throw std::runtime_error(shaw_last_error());
}

struct Realization {
  double depth, angle, period, delay;
};

// fom_problem_rank_two.hpp:141-162 (same nested order for rank-1 multi runs)
std::vector<Realization> realizations(const InputParser& p) {
  std::vector<Realization> r;
  for (double d : p.depths)
    for (double t : p.periods)
      for (double a : p.angles)
        for (double dl : p.delays) r.push_back({d, a, t, dl});
  return r;
}

// src/shared/various/print_perf.hpp:57-84; perfTimes = {min, max, total}
// seconds per step (fom_run.hpp:144-148), the byte/flop model from
// shaw_fom_complexity
void printPerf(uint64_t loops, const double (&perf)[3], double memMB, double flops) {
  const double minSec = perf[0], maxSec = perf[1], aveSec = perf[2] / (double)loops;
  const double memGB = memMB / 1024.;
  std::cout << "flops = " << flops << std::endl;
  std::cout << "memMB = " << memMB << std::endl;
  std::cout << "flops/bytes = " << flops / (memMB * 1024. * 1024.) << std::endl;
  std::printf("aveBandwidth(GB/s) = %8.2lf \n", memGB / aveSec);
  std::printf("minBandwidth(GB/s) = %8.2lf \n", memGB / maxSec);
  std::printf("maxBandwidth(GB/s) = %8.2lf \n", memGB / minSec);
  std::printf("aveGFlop = %7.4lf \n", flops / aveSec / 1e9);
  std::printf("minGFlop = %7.4lf \n", flops / maxSec / 1e9);
  std::printf("maxGFlop = %7.4lf \n", flops / minSec / 1e9);
  std::printf("totTime(se) = %7.4lf \n", perf[2]);
  std::printf("aveTime(ms) = %7.4lf \n", aveSec * 1000);
  std::printf("maxTime(ms) = %7.4lf \n", maxSec * 1000);
  std::printf("minTime(ms) = %7.4lf \n", minSec * 1000);
  std::fflush(stdout);
}

// steps per timed chunk of the FOM run (two graph replays at six steps per
// launch)
constexpr uint64_t kTimingChunk = 144;

// per-step times of a run timed in chunks: chunk i's duration over its steps
struct PerfTimes {
  double v[3] = {1e300, 0.0, 0.0};  // min, max, total [s]
  void add(double chunkMs, uint64_t steps) {
    const double perStep = chunkMs / 1000.0 / (double)steps;
    v[0] = std::min(v[0], perStep);
    v[1] = std::max(v[1], perStep);
    v[2] += chunkMs / 1000.0;
  }
};

class Driver {
 public:
  explicit Driver(const InputParser& p) : p_(p) {
    shaw_material mat{p.material.kind, (int)p.material.params.size(), p.material.params.data()};
    std::cout << std::endl << "*** Reading meshfile info ***" << std::endl;
    check(shaw_app_create_from_mesh(p.meshDir.c_str(), &mat, &app_));
    check(shaw_app_info_get(app_, &info_));
    std::cout << "thetaLeft (deg) = " << std::setprecision(17) << info_.th_left_deg << std::endl
              << "thetaRight (deg) = " << info_.th_right_deg << std::endl
              << "minimum radius (km) = " << info_.r_cmb_m / 1000. << std::endl
              << "surface radius (km) = " << info_.r_surf_m / 1000. << std::endl
              << "dth [rad] = " << info_.dth << std::endl
              << "drr [km] = " << info_.drr / 1000. << std::endl
              << "numGptVp = " << info_.nvp << std::endl
              << "numGptSp = " << info_.nsp << std::endl
              << "nth = " << info_.nth << std::endl
              << "nr = " << info_.nr << std::endl;
    std::cout << "minMaxVs = " << info_.min_vs << " " << info_.max_vs << std::endl;
    std::cout << "jacVp:  nnz = " << 4 * info_.nvp - 2 * info_.nr - 2 * info_.nth << " nrows = " << info_.nvp
              << " ncols = " << info_.nsp << std::endl;
    std::cout << "jacSp:  nnz = " << 2 * info_.nsp << " nrows = " << info_.nsp << " ncols = " << info_.nvp << std::endl;
    mi_.bounds = {info_.th_left_deg, info_.th_right_deg, info_.r_cmb_m, info_.r_surf_m};
    mi_.dth = info_.dth;
    mi_.drr = info_.drr;
  }
  ~Driver() { shaw_app_destroy(app_); }

  void run() {
    if (p_.romOn) romRun();
    else if (p_.rank2) fomRank2();
    else fomRank1();
  }

 private:
  // ------------------------------------------------------------- common
  void checkCfl() {
    if (p_.checkCflFlag) shaw::checkCfl(mi_, p_.dt, info_.max_vs);
  }
  void checkDispersion(double freq) {
    if (p_.checkNumDispersion) shaw::checkDispersion(mi_, freq, info_.min_vs);
  }
  std::vector<uint64_t> receiverGids() {
    std::vector<uint64_t> g;
    if (!p_.enableSeismo) return g;
    std::cout << std::endl << "*** Constructing seismogram ***" << std::endl;
    for (double a : p_.receivers) {
      uint64_t gid;
      check(shaw_map_receiver(app_, a, &gid));
      g.push_back(gid);
    }
    if (p_.nSteps % p_.seismoFreq != 0)
      throw std::runtime_error("Seismogram sampling frequency not a divisor of steps");
    std::cout << "Done Mapping receivers to grid " << std::endl;
    return g;
  }
  void checkObserver() {
    if (!p_.enableSnapMatrix) return;
    if (p_.nSteps % p_.vpSnapFreq != 0) throw std::runtime_error("Snapshot Vp frequency not a divisor of steps");
    if (p_.nSteps % p_.spSnapFreq != 0) throw std::runtime_error("Snapshot Sp frequency not a divisor of steps");
  }
  uint64_t sourceGid(const Realization& r) {
    uint64_t gid;
    check(shaw_map_source(app_, r.depth, r.angle, &gid));
    return gid;
  }
  std::vector<double> table(const std::vector<Realization>& rs) {
    std::vector<int> kinds(rs.size(), (int)p_.signalKind);
    std::vector<double> del, per;
    for (const auto& r : rs) { del.push_back(r.delay); per.push_back(r.period); }
    std::vector<double> t(p_.nSteps * rs.size());
    check(shaw_forcing_table((int)rs.size(), kinds.data(), del.data(), per.data(), p_.dt, p_.nSteps, t.data()));
    return t;
  }
  void writeCoordinates() {
    if (!p_.enableSnapMatrix) return;
    for (int d = 0; d < 2; ++d) {
      const uint64_t n = d == 0 ? info_.nvp : info_.nsp;
      std::vector<double> c(2 * n);
      check(shaw_app_coords(app_, d, c.data()));
      FILE* f = std::fopen(d == 0 ? "coords_vp.txt" : "coords_sp.txt", "w");
      if (!f) throw std::runtime_error("cannot write coordinates");
      for (uint64_t i = 0; i < n; ++i) std::fprintf(f, "%.17g %.17g\n", c[2 * i], c[2 * i + 1]);
      std::fclose(f);
    }
  }

  // One GPU FOM run over `rs` columns; writes per-run files. rank1Runs: each
  // column is an independent rank-1 realization with its own run id.
  void fomBatch(const std::vector<Realization>& rs, bool rank1Runs, const std::vector<uint64_t>& runIds,
                const std::vector<uint64_t>& recv) {
    const uint64_t M = rs.size(), N = p_.nSteps;
    // forcing objects first (forcing_rank_{one,two}.hpp ctors map the sources)
    std::vector<uint64_t> src;
    for (const auto& r : rs) src.push_back(sourceGid(r));
    const std::vector<double> tab = table(rs);
    shaw_fom_options o{};
    o.precision = SHAW_FP64;
    o.kernel = SHAW_KERNEL_AUTO;
    o.rank1_forcing = rank1Runs ? 1 : 0;
    shaw_fom_t fom;
    check(shaw_fom_create(app_, M, p_.dt, &o, &fom));
    check(shaw_fom_set_sources(fom, src.data()));
    check(shaw_fom_set_forcing(fom, N, tab.data()));
    if (!recv.empty()) check(shaw_fom_set_receivers(fom, recv.size(), recv.data(), p_.seismoFreq, N / p_.seismoFreq));
    std::vector<uint64_t> cols(M);
    for (uint64_t m = 0; m < M; ++m) cols[m] = m;
    if (p_.enableSnapMatrix) {
      check(shaw_fom_set_snapshots(fom, 0, p_.vpSnapFreq, M, cols.data(), N / p_.vpSnapFreq));
      check(shaw_fom_set_snapshots(fom, 1, p_.spSnapFreq, M, cols.data(), N / p_.spSnapFreq));
    }
    const auto t0 = std::chrono::high_resolution_clock::now();
    // blocks of 2000 steps (the reference's progress lines), each timed in
    // chunks of kTimingChunk steps for the per-step minimum and maximum
    const uint64_t block = 2000;
    PerfTimes perf;
    std::vector<double> chunkMs(block / kTimingChunk + 1);
    for (uint64_t s = 0; s < N; s += block) {
      const uint64_t n = std::min(block, N - s);
      check(shaw_fom_time_chunks(fom, n, kTimingChunk, chunkMs.data()));
      for (uint64_t c = 0; c * kTimingChunk < n; ++c) perf.add(chunkMs[c], std::min(kTimingChunk, n - c * kTimingChunk));
      if ((s + n) % 2000 == 0) std::cout << "Doing step = " << s + n << std::endl;
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    std::cout << "\nloopTime = " << std::fixed << std::setprecision(10)
              << std::chrono::duration<double>(t1 - t0).count();
    // data collection (device -> host, files) is reported separately
    const auto t2 = std::chrono::high_resolution_clock::now();
    std::vector<double> seis, sv, ss;
    uint64_t Ts = 0, Tv = 0, Tsp = 0;
    if (!recv.empty()) { seis.resize(recv.size() * (N / p_.seismoFreq) * M); check(shaw_fom_get_seismogram(fom, seis.data(), &Ts)); }
    if (p_.enableSnapMatrix) {
      sv.resize(info_.nvp * (N / p_.vpSnapFreq) * M);
      ss.resize(info_.nsp * (N / p_.spSnapFreq) * M);
      check(shaw_fom_get_snapshots(fom, 0, sv.data(), &Tv));
      check(shaw_fom_get_snapshots(fom, 1, ss.data(), &Tsp));
    }
    shaw_fom_destroy(fom);
    std::cout << "\ndataCollectionTime = " << std::fixed << std::setprecision(10)
              << std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t2).count() << std::endl;
    std::cout.unsetf(std::ios::fixed);
    std::cout << std::setprecision(6);
    double memMB, flops;
    // nnz(J_vp) = 4 nvp - 2 nr - 2 nth = 2 nsp and nnz(J_sp) = 2 nsp on the full mesh
    check(shaw_fom_complexity(info_.nvp, info_.nsp, 2 * info_.nsp, 2 * info_.nsp, M, rank1Runs ? 0 : 1, &memMB,
                              &flops));
    printPerf(N, perf.v, memMB * (rank1Runs ? (double)M : 1.0), flops * (rank1Runs ? (double)M : 1.0));

    const auto w0 = std::chrono::high_resolution_clock::now();
    const uint64_t R = recv.size();
    if (rank1Runs) {  // one file set per realization, 2D
      for (uint64_t m = 0; m < M; ++m) {
        const std::string id = "_" + std::to_string(runIds[m]);
        if (p_.enableSnapMatrix) {
          shaw::writeMatrix(p_.vpSnapFileName + id, sv.data() + info_.nvp * Tv * m, info_.nvp, Tv, p_.snapBinary, true);
          shaw::writeMatrix(p_.spSnapFileName + id, ss.data() + info_.nsp * Tsp * m, info_.nsp, Tsp, p_.snapBinary, true);
        }
        if (R) shaw::writeMatrix(p_.seismoFileName + id, seis.data() + R * Ts * m, R, Ts, p_.seismoBinary, false);
      }
    } else {  // one file set per set of forcingSize columns (3D unless M == 1)
      const std::string id = "_" + std::to_string(runIds[0]);
      if (p_.enableSnapMatrix) {
        if (M == 1) {
          shaw::writeMatrix(p_.vpSnapFileName + id, sv.data(), info_.nvp, Tv, p_.snapBinary, true);
          shaw::writeMatrix(p_.spSnapFileName + id, ss.data(), info_.nsp, Tsp, p_.snapBinary, true);
        } else {
          shaw::write3d(p_.vpSnapFileName + id, sv.data(), info_.nvp, Tv, M, p_.snapBinary, true);
          shaw::write3d(p_.spSnapFileName + id, ss.data(), info_.nsp, Tsp, M, p_.snapBinary, true);
        }
      }
      if (R) {
        if (M == 1) shaw::writeMatrix(p_.seismoFileName + id, seis.data(), R, Ts, p_.seismoBinary, false);
        else shaw::write3d(p_.seismoFileName + id, seis.data(), R, Ts, M, p_.seismoBinary, false);
      }
    }
    std::cout << "\nfinalProcessTime = " << std::fixed << std::setprecision(10)
              << std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - w0).count() << "\n";
    std::cout.unsetf(std::ios::fixed);
  }

  // ------------------------------------------------------------- FOM rank-1
  void fomRank1() {
    const std::vector<uint64_t> recv = receiverGids();
    checkObserver();
    const std::vector<Realization> rs = realizations(p_);
    checkCfl();
    if (!p_.multiForcing) {
      checkDispersion(1.0 / p_.periods[0]);
      fomBatch(rs, true, {0}, recv);
      writeCoordinates();
      return;
    }
    std::cout << "Doing FOM with sampling" << std::endl;
    for (double T : p_.periods) checkDispersion(1.0 / T);
    // batch independent realizations as columns, bounded by memory
    const uint64_t perCol = 16 * (info_.nvp + info_.nsp) +
                            (p_.enableSnapMatrix ? 8 * (info_.nvp * (p_.nSteps / p_.vpSnapFreq) +
                                                        info_.nsp * (p_.nSteps / p_.spSnapFreq)) : 0);
    const uint64_t budget = 32ull << 30;
    const uint64_t B = std::max<uint64_t>(1, std::min<uint64_t>(32, budget / std::max<uint64_t>(perCol, 1)));
    for (uint64_t s = 0; s < rs.size(); s += B) {
      const uint64_t e = std::min<uint64_t>(rs.size(), s + B);
      std::vector<Realization> part(rs.begin() + s, rs.begin() + e);
      std::vector<uint64_t> ids;
      for (uint64_t i = s; i < e; ++i) ids.push_back(i);
      fomBatch(part, true, ids, recv);
    }
    writeCoordinates();
  }

  // ------------------------------------------------------------- FOM rank-2
  void fomRank2() {
    const std::vector<uint64_t> recv = receiverGids();
    checkObserver();
    checkCfl();
    for (double T : p_.periods) checkDispersion(1.0 / T);
    const std::vector<Realization> rs = realizations(p_);
    const uint64_t fs = (uint64_t)p_.forcingSize;
    if (rs.size() % fs != 0)
      throw std::runtime_error(
          "totFRealizations & fSize != 0: tot num of forcing realizations must be divisible by fSize");
    const uint64_t numSets = rs.size() / fs;
    std::cout << "Doing rank-2 FOM" << std::endl
              << "Total number of samples " << rs.size() << std::endl
              << "Total number of set " << numSets << std::endl;
    for (uint64_t i = 0; i < numSets; ++i)
      fomBatch(std::vector<Realization>(rs.begin() + i * fs, rs.begin() + (i + 1) * fs), false, {i}, recv);
    writeCoordinates();
  }

  // ------------------------------------------------------------- ROM
  // rom_load_basis.hpp:56-101 / read_basis.hpp:54-96: first K columns of the
  // lsv file, or a seeded random dummy basis of amplitude 0.001
  std::vector<double> loadBasis(const std::string& file, bool binary, uint64_t n, uint64_t K, uint64_t seed) {
    std::vector<double> phi(n * K);
    if (p_.dummyBasis) {
      std::mt19937_64 g(seed);
      std::uniform_real_distribution<double> u(0.0, 0.001);
      for (auto& x : phi) x = u(g);
      return phi;
    }
    std::cout << "Loading basis... \n";
    uint64_t m, c;
    const std::vector<double> A = shaw::readMatrix(file, binary, m, c);
    if (m != n) throw std::runtime_error("basis file " + file + " has " + std::to_string(m) + " rows, mesh has " + std::to_string(n));
    if (c < K) throw std::runtime_error("basis file " + file + " has fewer columns than numModes");
    std::copy(A.begin(), A.begin() + n * K, phi.begin());
    return phi;
  }

  // q <- R^-1 q for the upper-triangular R [k][k] (back substitution)
  static void backSolve(const std::vector<double>& R, uint64_t k, double* q) {
    for (uint64_t i = k; i-- > 0;) {
      double x = q[i];
      for (uint64_t j = i + 1; j < k; ++j) x -= R[i * k + j] * q[j];
      q[i] = x / R[i * k + i];
    }
  }

  void romRun() {
    const uint64_t Kv = p_.vpRomSize, Ks = p_.spRomSize, N = p_.nSteps;
    std::vector<double> phiV = loadBasis(p_.vpBasisFile, p_.vpBasisBinary, info_.nvp, Kv, 13718);
    std::vector<double> phiS = loadBasis(p_.spBasisFile, p_.spBasisBinary, info_.nsp, Ks, 13719);
    std::vector<double> arV(Kv * Ks, 0.0), arS(Ks * Kv, 0.0);
    // energy projection: the bases made orthonormal in the FOM energy weights
    // (same spans), Ar_v = Phi_v^T Mv J_vp Phi_s, Ar_s = -Ar_v^T, reduced
    // forcing Mv[g] rho^-1[g] Phi_v[g, :]; reduced coordinates are written
    // back in the files' bases (q = R^-1 q_energy)
    const bool energy = p_.romEnergy && !p_.dummyBasis;
    std::vector<double> mv, Rv, Rs;
    const auto t0 = std::chrono::high_resolution_clock::now();
    if (energy) {
      std::vector<double> ms(info_.nsp), tmpV(phiV.size()), tmpS(phiS.size());
      mv.resize(info_.nvp);
      Rv.resize(Kv * Kv);
      Rs.resize(Ks * Ks);
      check(shaw_app_energy_weights(app_, mv.data(), ms.data()));
      check(shaw_basis_orthonormalize(phiV.data(), info_.nvp, Kv, mv.data(), tmpV.data(), Rv.data()));
      check(shaw_basis_orthonormalize(phiS.data(), info_.nsp, Ks, ms.data(), tmpS.data(), Rs.data()));
      phiV.swap(tmpV);
      phiS.swap(tmpS);
      std::cout << "rom projection: energy-consistent" << std::endl;
    }
    if (p_.disableCompRomJacs) std::cout << "The computation of the rom jacobians is disabled" << std::endl;
    else if (energy) check(shaw_rom_project_energy(app_, phiV.data(), Kv, phiS.data(), Ks, mv.data(), arV.data(), arS.data()));
    else check(shaw_rom_project(app_, phiV.data(), Kv, phiS.data(), Ks, arV.data(), arS.data()));
    std::cout << "romJac time: " << std::fixed << std::setprecision(10)
              << std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count() << std::endl;
    std::cout.unsetf(std::ios::fixed);
    checkObserver();
    std::vector<Realization> rs = realizations(p_);
    if (!p_.rank2) rs.resize(1);  // rank-1 ROM runs the first realization (rom_problem_rank_one.hpp:127)
    const uint64_t fs = p_.rank2 ? (uint64_t)p_.forcingSize : 1;
    if (rs.size() % fs != 0) throw std::runtime_error("tot num of forcing realizations must be divisible by fSize");
    checkCfl();
    for (const auto& r : rs) checkDispersion(1.0 / r.period);
    // receiver rows of Phi_v (main_reconstruct_seismo.cc:13-41, fused)
    std::vector<uint64_t> recv;
    if (p_.enableSeismo) recv = receiverGids();
    const uint64_t R = recv.size();
    std::vector<double> C(R * Kv);
    for (uint64_t r = 0; r < R; ++r)
      for (uint64_t k = 0; k < Kv; ++k) C[r * Kv + k] = phiV[recv[r] + info_.nvp * k];
    std::vector<double> rhoInv(info_.nvp);
    check(shaw_app_operators(app_, nullptr, nullptr, rhoInv.data(), nullptr));
    // reduced snapshots (the reference's observer on the reduced states)
    const uint64_t every = p_.enableSeismo ? p_.seismoFreq : 0;
    shaw_rom_desc d{};
    d.kv = Kv; d.ks = Ks; d.dt = p_.dt; d.precision = p_.romPrecision;
    d.ar_v = arV.data(); d.ar_s = arS.data(); d.r = R; d.c = R ? C.data() : nullptr;
    d.max_cols = fs; d.max_samples = every ? N / every : 0;
    shaw_rom_t rom;
    check(shaw_rom_create(&d, &rom));
    for (uint64_t set = 0; set < rs.size() / fs; ++set) {
      std::vector<Realization> part(rs.begin() + set * fs, rs.begin() + (set + 1) * fs);
      std::vector<double> b(Kv * fs);
      for (uint64_t j = 0; j < fs; ++j) {
        const uint64_t g = sourceGid(part[j]);
        for (uint64_t k = 0; k < Kv; ++k)
          b[k * fs + j] = energy ? mv[g] * rhoInv[g] * phiV[g + info_.nvp * k] : rhoInv[g] * phiV[g + info_.nvp * k];
      }
      const std::vector<double> tab = table(part);
      std::vector<uint64_t> hz(fs, N), slots(fs);
      check(shaw_rom_admit(rom, fs, b.data(), N, tab.data(), hz.data(), every, slots.data()));
      // advance, collecting reduced snapshots at their sampling steps
      const bool snaps = p_.enableSnapMatrix;
      std::vector<double> av, as;
      const uint64_t Tv = snaps ? N / p_.vpSnapFreq : 0, Tsp = snaps ? N / p_.spSnapFreq : 0;
      av.resize(Kv * Tv * fs);
      as.resize(Ks * Tsp * fs);
      const auto l0 = std::chrono::high_resolution_clock::now();
      PerfTimes perf;
      uint64_t step = 0;
      while (step < N) {
        uint64_t next = N;
        if (snaps) {
          next = std::min(next, (step / p_.vpSnapFreq + 1) * p_.vpSnapFreq);
          next = std::min(next, (step / p_.spSnapFreq + 1) * p_.spSnapFreq);
        }
        double tot;
        check(shaw_rom_time(rom, next - step, &tot));
        perf.add(tot, next - step);
        step = next;
        if (snaps) {
          std::vector<double> qv(Kv), qs(Ks);
          for (uint64_t j = 0; j < fs; ++j) {
            check(shaw_rom_state(rom, slots[j], qv.data(), qs.data(), nullptr));
            if (energy) {
              backSolve(Rv, Kv, qv.data());
              backSolve(Rs, Ks, qs.data());
            }
            if (step % p_.vpSnapFreq == 0)
              std::copy(qv.begin(), qv.end(), av.begin() + Kv * (step / p_.vpSnapFreq - 1 + Tv * j));
            if (step % p_.spSnapFreq == 0)
              std::copy(qs.begin(), qs.end(), as.begin() + Ks * (step / p_.spSnapFreq - 1 + Tsp * j));
          }
        }
      }
      std::cout << "\nloopTime = " << std::fixed << std::setprecision(10)
                << std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - l0).count()
                << "\ndataCollectionTime = 0.0000000000" << std::endl;
      std::cout.unsetf(std::ios::fixed);
      std::cout << std::setprecision(6);
      const double memMB = (2.0 * Kv * Ks * 8 * 2 + 3.0 * (Kv + Ks) * 8 * fs) / 1024. / 1024.;
      const double flops = 4.0 * Kv * Ks * fs + 3.0 * Kv * fs;
      printPerf(N, perf.v, memMB, flops);
      std::vector<double> seis(R * (every ? N / every : 0) * fs);
      for (uint64_t j = 0; j < fs; ++j) {
        uint64_t T = 0;
        check(shaw_rom_retire(rom, slots[j], R ? seis.data() + R * (N / std::max<uint64_t>(every, 1)) * j : nullptr,
                              nullptr, nullptr, &T));
      }
      if (!p_.dummyBasis) {  // rom_problem_rank_one.hpp:169-185
        const std::string id = "_" + std::to_string(set);
        if (snaps) {
          if (fs == 1) {
            shaw::writeMatrix(p_.vpSnapFileName + id, av.data(), Kv, Tv, p_.snapBinary, true);
            shaw::writeMatrix(p_.spSnapFileName + id, as.data(), Ks, Tsp, p_.snapBinary, true);
          } else {
            shaw::write3d(p_.vpSnapFileName + id, av.data(), Kv, Tv, fs, p_.snapBinary, true);
            shaw::write3d(p_.spSnapFileName + id, as.data(), Ks, Tsp, fs, p_.snapBinary, true);
          }
        }
        if (R) {
          const uint64_t T = N / every;
          if (fs == 1) shaw::writeMatrix(p_.seismoFileName + id, seis.data(), R, T, p_.seismoBinary, false);
          else shaw::write3d(p_.seismoFileName + id, seis.data(), R, T, fs, p_.seismoBinary, false);
        }
      }
    }
    shaw_rom_destroy(rom);
  }

  const InputParser& p_;
  shaw_app_t app_ = nullptr;
  shaw_app_info info_{};
  shaw::MeshInfo mi_;
};

}  // namespace

int main(int argc, char** argv) {
  try {
    InputParser parser(argc, argv);
    Driver drv(parser);
    drv.run();
  } catch (const std::exception& e) {
    std::cerr << "shawExe: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
