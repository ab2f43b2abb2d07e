// Galerkin ROM on gfx950 (projection + reduced leapfrog). Filled in below.
#include "shaw_app.hpp"
#include "shaw_common.hpp"

using shaw::guarded;

extern "C" {
int shaw_rom_project(shaw_app_t, const double*, uint64_t, const double*, uint64_t, double*, double*) {
  return guarded([&] { throw shaw::Error(SHAW_ERR_STATE, "ROM not built yet"); });
}
}
