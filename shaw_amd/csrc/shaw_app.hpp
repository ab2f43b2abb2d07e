// The object behind shaw_app_t: a mesh, its material and the assembled
// matrix-free operators (the ShWavePP replacement, src/kokkos/shwavepp.hpp).
#pragma once

#include "shaw_amd.h"
#include "shaw_host.hpp"

namespace shaw {
// The structured mesh as per-row and per-column tables (the device-side
// setup's inputs, device_setup.hip): 1/r of the vp (and stp) points and of
// the srp points of each row, theta and cot of the vp and stp points of each
// column. `ok` once every per-point value of the mesh was checked to be the
// table value of its row or column (and the vp stencil multipliers and sp
// labels the structured closed form).
struct GridTables {
  bool checked = false, ok = false;
  std::vector<double> rInvRow, rInvSrp, thVp, thStp, cotV, cotS;
};
}  // namespace shaw

struct shaw_app {
  shaw::Mesh mesh;
  shaw::Material mat;
  shaw::Operators ops;
  shaw::GridTables tables;  // built on first use by the device-side setup
};

namespace shaw {
// the C-ABI material description as a Material (params copied; throws on
// a null pointer or a bad parameter count); rSurfM is left to the caller
Material materialFromC(const shaw_material* m);
}  // namespace shaw
