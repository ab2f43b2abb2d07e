// The macro-step operator of the reduced leapfrog (host, OpenMP). With
// x = [q_v; q_s] one unforced step is x' = F x, F = [[I, dt Ar_v],
// [dt Ar_s, I + dt^2 Ar_s Ar_v]] (rom_run_rank_one.hpp:112-151: q_v first,
// then q_s from the new q_v). m steps are x' = F^m x = x + dt G x with
// G = (F^m - I) / dt, built here by stepping the identity's columns m times
// with the leapfrog's own updates (q += dt (A q')), so the BF16 kernels can
// advance a tile of quiet columns m steps with one product per block.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "rom_internal.hpp"

namespace shaw {

namespace {
// Y[r][:] += dt * (A[r][:] . X), A rows x k row-major, X k x n, Y rows x n
void leapfrogHalf(const double* A, uint64_t rows, uint64_t k, const std::vector<double>& X, uint64_t n, double dt,
                  std::vector<double>& Y) {
#pragma omp parallel
  {
    std::vector<double> acc(n);
#pragma omp for schedule(static)
    for (long long rr = 0; rr < (long long)rows; ++rr) {
      const uint64_t r = (uint64_t)rr;
      std::fill(acc.begin(), acc.end(), 0.0);
      for (uint64_t j = 0; j < k; ++j) {
        const double a = A[r * k + j];
        if (a == 0.0) continue;
        const double* x = &X[j * n];
        for (uint64_t c = 0; c < n; ++c) acc[c] += a * x[c];
      }
      double* y = &Y[r * n];
      for (uint64_t c = 0; c < n; ++c) y[c] += dt * acc[c];
    }
  }
}
}  // namespace

void macroOperator(const double* arV, const double* arS, uint64_t kv, uint64_t ks, double dt, uint32_t m,
                   std::vector<double>& gvv, std::vector<double>& gvs, std::vector<double>& gsv,
                   std::vector<double>& gss) {
  const uint64_t n = kv + ks;
  std::vector<double> V(kv * n, 0.0), S(ks * n, 0.0);
  for (uint64_t r = 0; r < kv; ++r) V[r * n + r] = 1.0;
  for (uint64_t r = 0; r < ks; ++r) S[r * n + kv + r] = 1.0;
  for (uint32_t step = 0; step < m; ++step) {
    leapfrogHalf(arV, kv, ks, S, n, dt, V);  // q_v += dt Ar_v q_s
    leapfrogHalf(arS, ks, kv, V, n, dt, S);  // q_s += dt Ar_s q_v
  }
  gvv.assign(kv * kv, 0.0);
  gvs.assign(kv * ks, 0.0);
  gsv.assign(ks * kv, 0.0);
  gss.assign(ks * ks, 0.0);
  for (uint64_t r = 0; r < kv; ++r) {
    for (uint64_t c = 0; c < kv; ++c) gvv[r * kv + c] = (V[r * n + c] - (r == c ? 1.0 : 0.0)) / dt;
    for (uint64_t c = 0; c < ks; ++c) gvs[r * ks + c] = V[r * n + kv + c] / dt;
  }
  for (uint64_t r = 0; r < ks; ++r) {
    for (uint64_t c = 0; c < kv; ++c) gsv[r * kv + c] = S[r * n + c] / dt;
    for (uint64_t c = 0; c < ks; ++c) gss[r * ks + c] = (S[r * n + kv + c] - (r == c ? 1.0 : 0.0)) / dt;
  }
}

}  // namespace shaw
