// C ABI: last-error plumbing, mesh/operator (shaw_app_*), mapping and
// forcing-table entry points. FOM and ROM entry points live next to their
// kernels (fom.hip, rom.hip).
#include <cstring>
#include <memory>
#include <string>

#include "shaw_app.hpp"
#include "shaw_common.hpp"

namespace shaw {
namespace {
thread_local std::string g_lastError;
}
void setLastError(const std::string& msg) { g_lastError = msg; }
void clearLastError() { g_lastError.clear(); }

Material materialFromC(const shaw_material* m) {
  SHAW_REQUIRE(m != nullptr, SHAW_ERR_ARG, "null material");
  SHAW_REQUIRE(m->nparams >= 0 && (m->nparams == 0 || m->params), SHAW_ERR_ARG, "bad material params");
  Material mat;
  mat.kind = m->kind;
  mat.params.assign(m->params, m->params + m->nparams);
  return mat;
}

static shaw_app* finishApp(std::unique_ptr<shaw_app> a) {
  a->mat.rSurfM = a->mesh.info.bounds[3];
  a->mat.validate(a->mesh.info);
  a->ops = Operators::build(a->mesh, a->mat);
  return a.release();
}

}  // namespace shaw

using shaw::guarded;

extern "C" {

const char* shaw_last_error(void) { return shaw::g_lastError.c_str(); }
const char* shaw_version(void) { return "shaw_amd 0.1 (gfx950)"; }

int shaw_app_create_from_mesh(const char* dir, const shaw_material* mat, shaw_app_t* out) {
  return guarded([&] {
    SHAW_REQUIRE(dir && out, SHAW_ERR_ARG, "null argument");
    auto a = std::make_unique<shaw_app>();
    a->mat = shaw::materialFromC(mat);
    a->mesh = shaw::Mesh::load(dir);
    *out = shaw::finishApp(std::move(a));
  });
}

int shaw_app_create_structured(uint64_t nr, uint64_t nth, double thL, double thR, double rCmb, double rSurf,
                               const shaw_material* mat, shaw_app_t* out) {
  return guarded([&] {
    SHAW_REQUIRE(out, SHAW_ERR_ARG, "null argument");
    auto a = std::make_unique<shaw_app>();
    a->mat = shaw::materialFromC(mat);
    a->mesh = shaw::Mesh::structured(nr, nth, thL, thR, rCmb, rSurf);
    *out = shaw::finishApp(std::move(a));
  });
}

int shaw_mesh_write(const char* dir, uint64_t nr, uint64_t nth, double thL, double thR, double rCmb, double rSurf) {
  return guarded([&] {
    SHAW_REQUIRE(dir, SHAW_ERR_ARG, "null argument");
    shaw::Mesh::write(dir, nr, nth, thL, thR, rCmb, rSurf);
  });
}

int shaw_app_info_get(shaw_app_t app, shaw_app_info* info) {
  return guarded([&] {
    SHAW_REQUIRE(app && info, SHAW_ERR_ARG, "null argument");
    const shaw::MeshInfo& mi = app->mesh.info;
    info->nr = mi.nr;
    info->nth = mi.nth;
    info->nvp = mi.nVp;
    info->nsp = mi.nSp;
    info->th_left_deg = mi.bounds[0];
    info->th_right_deg = mi.bounds[1];
    info->r_cmb_m = mi.bounds[2];
    info->r_surf_m = mi.bounds[3];
    info->dth = mi.dth;
    info->drr = mi.drr;
    info->min_vs = app->ops.minVs;
    info->max_vs = app->ops.maxVs;
  });
}

int shaw_app_operators(shaw_app_t app, double* cv, double* cs, double* rhoInv, double* shear) {
  return guarded([&] {
    SHAW_REQUIRE(app, SHAW_ERR_ARG, "null handle");
    const shaw::Operators& o = app->ops;
    if (cv) std::memcpy(cv, o.coefVp.data(), o.coefVp.size() * sizeof(double));
    if (cs) std::memcpy(cs, o.coefSp.data(), o.coefSp.size() * sizeof(double));
    if (rhoInv) std::memcpy(rhoInv, o.rhoInv.data(), o.rhoInv.size() * sizeof(double));
    if (shear) std::memcpy(shear, o.shearMod.data(), o.shearMod.size() * sizeof(double));
  });
}

int shaw_app_coords(shaw_app_t app, int dof, double* coords) {
  return guarded([&] {
    SHAW_REQUIRE(app && coords && (dof == 0 || dof == 1), SHAW_ERR_ARG, "bad argument");
    const std::vector<double>& c = dof == 0 ? app->mesh.coordsVp : app->mesh.coordsSp;
    for (size_t i = 0; i < c.size() / 2; ++i) {
      coords[2 * i] = c[2 * i];
      coords[2 * i + 1] = 1.0 / c[2 * i + 1];
    }
  });
}

int shaw_map_source(shaw_app_t app, double depthKm, double angleDeg, uint64_t* gid) {
  return guarded([&] {
    SHAW_REQUIRE(app && gid, SHAW_ERR_ARG, "null argument");
    *gid = shaw::mapSource(app->mesh, depthKm, angleDeg);
  });
}

int shaw_map_receiver(shaw_app_t app, double angleDeg, uint64_t* gid) {
  return guarded([&] {
    SHAW_REQUIRE(app && gid, SHAW_ERR_ARG, "null argument");
    *gid = shaw::mapReceiver(app->mesh, angleDeg);
  });
}

void shaw_app_destroy(shaw_app_t app) { delete app; }

int shaw_forcing_table(int m, const int* kinds, const double* delays, const double* periods, double dt,
                       uint64_t nsteps, double* table) {
  return guarded([&] {
    SHAW_REQUIRE(m >= 0 && kinds && delays && periods && table, SHAW_ERR_ARG, "null argument");
    std::vector<shaw::Signal> sig;
    for (int k = 0; k < m; ++k) {
      SHAW_REQUIRE(kinds[k] >= 0 && kinds[k] <= 2, SHAW_ERR_ARG, "unknown signal kind");
      SHAW_REQUIRE(periods[k] > 0, SHAW_ERR_ARG, "Period for signal must be positive.");
      sig.emplace_back((shaw::SignalKind)kinds[k], delays[k], periods[k]);
    }
    const std::vector<double> t = shaw::forcingTable(sig, dt, nsteps);
    std::memcpy(table, t.data(), t.size() * sizeof(double));
  });
}

// complexityFom (src/kokkos/fom_complexities.hpp:57-137) with the counts of
// src/shared/complexities/spmv.hpp, spmm.hpp and mult.hpp (mult_beta_one):
// FP64 values, 8-byte ordinals, each term summed in the same order.
int shaw_fom_complexity(uint64_t nvp, uint64_t nsp, uint64_t nnzVp, uint64_t nnzSp, uint64_t m, int rank2,
                        double* memMB, double* flops) {
  return guarded([&] {
    SHAW_REQUIRE(memMB && flops && m > 0, SHAW_ERR_ARG, "bad argument");
    constexpr uint64_t sc = sizeof(double), ord = sizeof(uint64_t);
    // spmm with k columns (spmv: k = 1, the same counts)
    auto spmm = [&](uint64_t nnz, uint64_t rows, uint64_t k, double& mb, double& fl) {
      const double matsize = 1.0 * (nnz * (sc + ord) + rows * ord);
      const double xr = 1. * nnz * k * sc, cr = 1. * rows * k * sc, cw = 1. * rows * k * sc;
      mb = (matsize + xr + cr + cw) / 1024. / 1024.;
      fl = 2. * nnz * k + rows * k;
    };
    double mb[3], fl[3];
    const uint64_t k = rank2 ? m : 1;
    spmm(nnzVp, nvp, k, mb[0], fl[0]);
    if (rank2) {  // the per-forcing add: about 4 words and 3 flops per forcing
      mb[1] = 4. * m * sc / 1024. / 1024.;
      fl[1] = 3. * m;
    } else {  // mult_beta_one over nvp
      mb[1] = (4.0 * nvp * sc) / 1024. / 1024.;
      fl[1] = 3.0 * nvp;
    }
    spmm(nnzSp, nsp, k, mb[2], fl[2]);
    *memMB = mb[0] + mb[1] + mb[2];
    *flops = fl[0] + fl[1] + fl[2];
  });
}

}  // extern "C"
