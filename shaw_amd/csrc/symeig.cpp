// Symmetric eigen-decomposition on the host, used by the POD (the Gram matrix
// of the snapshots, at most a few thousand rows).
//
// Method (G. H. Golub and C. F. Van Loan, Matrix Computations, 4th ed.,
// Johns Hopkins University Press, 2013):
//  1. Householder tridiagonalization, section 8.3.1 (Algorithm 8.3.1), with
//     the Householder vector of Algorithm 5.1.1 and the rank-2 update
//     A <- A - v w^T - w v^T, w = p - (beta p^T v / 2) v, p = beta A v.
//     The orthogonal factor Q = P_0 P_1 ... P_{n-3} is formed by backward
//     accumulation (section 5.1.6).
//  2. The symmetric QR algorithm, section 8.3.3 (Algorithm 8.3.3): implicit
//     QR steps with the Wilkinson shift (Algorithm 8.3.2) on the unreduced
//     trailing block, each step one bulge chase of Givens rotations from the
//     top of the block to its bottom, with deflation when a subdiagonal
//     entry is negligible.
// The rotations of each QR step are recorded and applied to the columns of Q
// afterwards, in parallel over blocks of Q's rows (each row of Q is rotated
// independently of the others).
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "rom_internal.hpp"
#include "shaw_common.hpp"

namespace shaw {
namespace {

// Householder vector (Algorithm 5.1.1): for x of length m, v with v[0] = 1
// and beta such that (I - beta v v^T) x = mu e_0, mu = +-||x||. Returns mu.
double house(const double* x, int m, double* v, double& beta) {
  double sigma = 0.0;
  for (int i = 1; i < m; ++i) sigma += x[i] * x[i];
  v[0] = 1.0;
  for (int i = 1; i < m; ++i) v[i] = x[i];
  if (sigma == 0.0) {  // x is already a multiple of e_0: no reflection
    beta = 0.0;
    return x[0];
  }
  const double mu = std::sqrt(x[0] * x[0] + sigma);
  // v0 = x0 - mu, evaluated without cancellation when x0 > 0
  const double v0 = x[0] <= 0.0 ? x[0] - mu : -sigma / (x[0] + mu);
  beta = 2.0 * v0 * v0 / (sigma + v0 * v0);
  for (int i = 1; i < m; ++i) v[i] /= v0;
  return mu;
}

struct Rotation {
  int k;
  double c, s;
};

}  // namespace

void symmetricEigen(std::vector<double>& A, int n, std::vector<double>& w, std::vector<double>& V) {
  SHAW_REQUIRE(n > 0 && A.size() >= (size_t)n * n, SHAW_ERR_ARG, "symmetricEigen: bad size");
  auto at = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
  std::vector<double> d(n), e(n > 1 ? n - 1 : 1, 0.0);
  // Householder vectors, v_k of length n-k-1 stored in rows k of H (v_k[0] = 1)
  std::vector<double> H((size_t)n * n, 0.0), betas(n, 0.0), p(n), wv(n);

  // 1. Tridiagonalization: at step k, the column x = A[k+1:n, k] is reduced
  // to mu e_0 and the trailing block becomes P A P.
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double* v = &H[(size_t)k * n];
    std::vector<double> x(m);
    for (int i = 0; i < m; ++i) x[i] = at(k + 1 + i, k);
    double beta = 0.0;
    const double mu = house(x.data(), m, v, beta);
    betas[k] = beta;
    d[k] = at(k, k);
    e[k] = mu;
    if (beta == 0.0) continue;
    // p = beta * B v, B = A[k+1:n, k+1:n]
#pragma omp parallel for schedule(static) if (m > 256)
    for (int i = 0; i < m; ++i) {
      const double* row = &A[(size_t)(k + 1 + i) * n + (k + 1)];
      double sum = 0.0;
      for (int j = 0; j < m; ++j) sum += row[j] * v[j];
      p[i] = beta * sum;
    }
    double pv = 0.0;
    for (int i = 0; i < m; ++i) pv += p[i] * v[i];
    const double half = 0.5 * beta * pv;
    for (int i = 0; i < m; ++i) wv[i] = p[i] - half * v[i];
    // B <- B - v w^T - w v^T
#pragma omp parallel for schedule(static) if (m > 256)
    for (int i = 0; i < m; ++i) {
      double* row = &A[(size_t)(k + 1 + i) * n + (k + 1)];
      const double vi = v[i], wi = wv[i];
      for (int j = 0; j < m; ++j) row[j] -= vi * wv[j] + wi * v[j];
    }
  }
  if (n >= 2) {
    d[n - 2] = at(n - 2, n - 2);
    e[n - 2] = at(n - 1, n - 2);
  }
  d[n - 1] = at(n - 1, n - 1);

  // Q by backward accumulation: Q = I; for k = n-3..0: Q[k+1:, k+1:] <- P_k Q[k+1:, k+1:].
  // Stored transposed (Zt = Q^T, row-major), so that a Givens rotation of two
  // columns of Q below touches two contiguous rows of Zt; P_k is symmetric,
  // so Zt[k+1:, k+1:] <- Zt[k+1:, k+1:] P_k.
  std::vector<double> Zt((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Zt[(size_t)i * n + i] = 1.0;
  for (int k = n - 3; k >= 0; --k) {
    const double beta = betas[k];
    if (beta == 0.0) continue;
    const int m = n - k - 1;
    const double* v = &H[(size_t)k * n];
    // rows k+1.. of Zt: row <- row - beta (row . v) v^T
#pragma omp parallel for schedule(static) if (m > 256)
    for (int i = 0; i < m; ++i) {
      double* row = &Zt[(size_t)(k + 1 + i) * n + (k + 1)];
      double dot = 0.0;
      for (int j = 0; j < m; ++j) dot += row[j] * v[j];
      const double f = beta * dot;
      for (int j = 0; j < m; ++j) row[j] -= f * v[j];
    }
  }
  // Row c of Zt = Q^T is column c of Q.
  std::vector<double>& Qt = Zt;

  // 2. Symmetric QR on (d, e). The rotation R = [[c, s], [-s, c]] of plane
  // (k, k+1) maps T to R T R^T and Q to Q R^T, i.e. row k and row k+1 of Qt
  // (= columns k, k+1 of Q) become c q_k + s q_{k+1} and -s q_k + c q_{k+1}.
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) tnorm = std::max(tnorm, std::abs(d[i]) + (i + 1 < n ? std::abs(e[i]) : 0.0) +
                                                          (i > 0 ? std::abs(e[i - 1]) : 0.0));
  const double eps = std::numeric_limits<double>::epsilon();
  // An off-diagonal entry is negligible relative to its diagonal neighbours
  // (the test of Algorithm 8.3.3), or below 1e-3 eps ||T||: an absolute
  // eigenvalue change far below the O(eps ||T||) error of any backward-stable
  // method, which ends the iteration on clusters of noise-level eigenvalues
  // (the trailing eigenvalues of a rank-deficient Gram matrix).
  const double floorAbs = 1e-3 * eps * tnorm;
  std::vector<Rotation> rot;
  rot.reserve(n);
  const int maxSteps = 60 * std::max(n, 1);
  int steps = 0;
  int hi = n - 1;
  while (hi > 0) {
    // deflate the trailing 1x1 blocks
    auto negligible = [&](int i) {
      return std::abs(e[i]) <= eps * (std::abs(d[i]) + std::abs(d[i + 1])) || std::abs(e[i]) <= floorAbs;
    };
    if (negligible(hi - 1)) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    // the unreduced block [lo, hi]
    int lo = hi - 1;
    while (lo > 0 && !negligible(lo - 1)) --lo;
    if (lo > 0) e[lo - 1] = 0.0;
    SHAW_REQUIRE(++steps <= maxSteps, SHAW_ERR_ARG, "symmetric QR did not converge");
    // Wilkinson shift: the eigenvalue of the trailing 2x2 block closer to d[hi]
    const double t = 0.5 * (d[hi - 1] - d[hi]);
    const double b2 = e[hi - 1] * e[hi - 1];
    const double mu = d[hi] - b2 / (t + (t >= 0.0 ? 1.0 : -1.0) * std::hypot(t, e[hi - 1]));
    // bulge chase (Algorithm 8.3.2)
    double x = d[lo] - mu, z = e[lo];
    rot.clear();
    for (int k = lo; k < hi; ++k) {
      const double r = std::hypot(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
      if (k > lo) e[k - 1] = r;  // the bulge below e[k-1] is annihilated
      const double a = d[k], b = d[k + 1], f = e[k];
      const double cc = c * c, ss = s * s, cs = c * s;
      d[k] = cc * a + 2.0 * cs * f + ss * b;
      d[k + 1] = ss * a - 2.0 * cs * f + cc * b;
      e[k] = cs * (b - a) + (cc - ss) * f;
      if (k + 1 < hi) {  // the new bulge at (k+2, k)
        z = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
        x = e[k];
      }
      rot.push_back({k, c, s});
    }
    // apply the step's rotations to Q: Qt rows k, k+1 for each rotation in
    // order; independent across the columns of Qt, so split those
    const int nrot = (int)rot.size();
    const int64_t work = (int64_t)nrot * n;
    const int blk = 64;
    const int nblk = (n + blk - 1) / blk;
#pragma omp parallel for schedule(static) if (work > (1 << 18))
    for (int bI = 0; bI < nblk; ++bI) {
      const int j0 = bI * blk, j1 = std::min(n, j0 + blk);
      for (int q = 0; q < nrot; ++q) {
        const Rotation& g = rot[q];
        double* r0 = &Qt[(size_t)g.k * n];
        double* r1 = r0 + n;
        for (int j = j0; j < j1; ++j) {
          const double u = r0[j], v = r1[j];
          r0[j] = g.c * u + g.s * v;
          r1[j] = -g.s * u + g.c * v;
        }
      }
    }
  }
  // outputs: w (unordered) and V row-major with eigenvector c in column c
  w = d;
  V.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c)
    for (int j = 0; j < n; ++j) V[(size_t)j * n + c] = Qt[(size_t)c * n + j];
}

}  // namespace shaw
