// The object behind shaw_app_t: a mesh, its material and the assembled
// matrix-free operators (the ShWavePP replacement, src/kokkos/shwavepp.hpp).
#pragma once

#include "shaw_host.hpp"

struct shaw_app {
  shaw::Mesh mesh;
  shaw::Material mat;
  shaw::Operators ops;
};
