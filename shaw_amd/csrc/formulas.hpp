// Material models and source signals as functions the host and the device
// both compile. The host operator assembly (Operators::build), the host
// forcing table and their device counterparts (device_setup.hip) evaluate the
// same expressions in the same order with contraction off on both sides, so
// operators assembled on the device equal the host's bit for bit; forcing
// tables differ only where the device's exp/sin round differently from the
// host's.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace shaw {

// Material kinds, parameter flattening as shaw_material (include/shaw_amd.h).
enum MaterialKind : int { kUnilayer = 0, kBilayer = 1, kPrem = 2, kLayered = 3 };

// density [kg/m^3] and shear velocity [m/s] at radius r [m]
__host__ __device__ inline void materialAt(int kind, const double* p, double rSurfM, double r, double& rho,
                                           double& vs) {
  switch (kind) {
    case kUnilayer: {  // material_model_unilayer.hpp:72-83
      const double dM = rSurfM - r;
      rho = p[0] + p[1] * dM + p[2] * dM * dM;
      vs = p[3] + p[4] * dM + p[5] * dM * dM;
      return;
    }
    case kBilayer: {  // material_model_bilayer.hpp:74-101
      const double depthM = rSurfM - r;
      if (depthM < p[0] * 1000.0) {
        const double dM = depthM;
        rho = p[1] + p[2] * dM + p[3] * dM * dM;
        vs = p[4] + p[5] * dM + p[6] * dM * dM;
      } else {
        const double dM = depthM - (p[0] * 1000.0);
        rho = p[7] + p[8] * dM + p[9] * dM * dM;
        vs = p[10] + p[11] * dM + p[12] * dM * dM;
      }
      return;
    }
    case kPrem: {  // material_model_prem.hpp:85-176 (g/cc, km/s -> SI)
      const double rKm = r / 1000.0;
      const double x = rKm / 6371.0, xSq = x * x, xCu = x * x * x;
      if (rKm >= 6356.0) { rho = 2.6; vs = 3.2; }
      else if (rKm >= 6346.6) { rho = 2.9; vs = 3.9; }
      else if (rKm >= 6151.0) { rho = 2.691 + 0.6924 * x; vs = 2.1519 + 2.3481 * x; }
      else if (rKm >= 5971.0) { rho = 7.1089 - 3.8045 * x; vs = 8.9496 - 4.4597 * x; }
      else if (rKm >= 5771.0) { rho = 11.2494 - 8.0298 * x; vs = 22.3512 - 18.5856 * x; }
      else if (rKm >= 5701.0) { rho = 5.3197 - 1.4836 * x; vs = 9.9839 - 4.9324 * x; }
      else if (rKm >= 5600.0) {
        rho = 7.9565 - 6.4761 * x + 5.5283 * xSq - 3.0807 * xCu;
        vs = 22.3459 - 17.2473 * x - 2.0834 * xSq + 0.9783 * xCu;
      } else if (rKm >= 3630.0) {
        rho = 7.9565 - 6.4761 * x + 5.5283 * xSq - 3.0807 * xCu;
        vs = 11.1671 - 13.7818 * x + 17.4575 * xSq - 9.2777 * xCu;
      } else if (rKm >= 3480.0) {
        rho = 7.9565 - 6.4761 * x + 5.5283 * xSq - 3.0807 * xCu;
        vs = 6.9254 + 1.4672 * x - 2.0834 * xSq + 0.9783 * xCu;
      } else if (rKm >= 1221.5) {
        rho = 12.5815 - 1.2638 * x - 3.6426 * xSq - 5.5281 * xCu;
        vs = 0.0;
      } else {
        rho = 13.0885 - 8.8381 * xSq;
        vs = 3.6678 - 4.4475 * xSq;
      }
      vs *= 1000.0;
      rho *= 1000.0;
      return;
    }
    default: {  // kLayered: nodes ascending in radius [km]; linear between, held outside
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
      return;
    }
  }
}

// Signal kinds (signal.hpp:52-128): 0 Ricker, 1 sinusoid, 2 gaussDer.
// freqSq = (1 / period)^2.
__host__ __device__ inline double signalAt(int kind, double delay, double period, double freqSq, double t) {
  const double one = 1.0, two = 2.0, piSq = M_PI * M_PI;
  if (kind == 0) {
    const double tDiffSq = (t - delay) * (t - delay);
    const double expTerm = std::exp(-piSq * freqSq * tDiffSq);
    return (one - two * piSq * tDiffSq * freqSq) * expTerm;
  }
  if (kind == 1) {
    constexpr double rn = 0.001;
    constexpr double term2 = rn / (rn + 2.0);
    if (t < period) return std::sin(rn * M_PI * t / period) - term2 * std::sin((rn + two) * M_PI * t / period);
    return std::sin(rn * M_PI) - term2 * std::sin((rn + two) * M_PI);
  }
  const double tDiff = t - delay, tDiffSq = tDiff * tDiff;
  return -two * tDiff * freqSq * std::exp(-freqSq * tDiffSq);
}

}  // namespace shaw
