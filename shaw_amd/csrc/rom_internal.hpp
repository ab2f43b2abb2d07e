// Entry points shared between the FOM and ROM translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "shaw_app.hpp"

namespace shaw {
// device_setup.hip: the operators of `mat` on app's mesh assembled on the
// device into coef_vp [nVp][4], coef_sp [nSp][2], rho_inv [nVp] (device
// doubles, any may be null), equal to Operators::build bit for bit; and the
// forcing table of m signals, out[s * m + k] (ldOut = 0) or out[k * ldOut + s]
// (device), equal to forcingTable up to the device exp/sin rounding
void assembleOperatorsDevice(shaw_app& app, const Material& mat, double* coefVp, double* coefSp, double* rhoInv,
                             hipStream_t stream);
void forcingTableDevice(uint64_t m, const int* kinds, const double* delays, const double* periods, double dt,
                        uint64_t nsteps, double* out, uint64_t ldOut, hipStream_t stream);
// the same with the parameters already on the device (par [m][3]: kind,
// delay, period, as signalParams packs them); only enqueues
std::vector<double> signalParams(uint64_t m, const int* kinds, const double* delays, const double* periods);
void forcingTableFromParams(uint64_t m, const double* parDevice, double dt, uint64_t nsteps, double* out,
                            uint64_t ldOut, hipStream_t stream);
// POD of device-resident snapshots dS [n x ns] column-major (overwritten);
// phi: host [n x k] column-major, sigma: host [ns] (may be null).
// kOut == nullptr: exactly k modes or an error; else up to k modes, stopping at
// the numerical rank, with the count written to *kOut
void podDevice(double* dS, uint64_t n, uint64_t ns, uint64_t k, double* phi, double* sigma, uint64_t* kOut);
// As podDevice with phi left on the device: dPhi [n x k] column-major,
// allocated by the caller.
void podDeviceTo(double* dS, uint64_t n, uint64_t ns, uint64_t k, double* dPhi, double* sigma, uint64_t* kOut);
// Symmetric eigen-decomposition on the host (Householder tridiagonalization
// and the implicit symmetric QR algorithm): A row-major [n x n] (destroyed),
// w[n] eigenvalues (unordered), V row-major [n x n] with the eigenvector of
// w[c] in column c. symeig.cpp.
void symmetricEigen(std::vector<double>& A, int n, std::vector<double>& w, std::vector<double>& V);
// out (device, row-major [Ka][Kb]) = L^T diag(w) (J X) in FP64 on MFMA
// (proj.hip): op 0 J = J_vp (L: n = nVp rows, X: nX = nSp rows, coef =
// coefVp [nVp][4]), op 1 J = J_sp (n = nSp, nX = nVp, coef = coefSp
// [nSp][2]), op 2 J = I (nX = n, coef unused). L, X column-major on the
// device; w [n] or null. part: device scratch of projPartDoubles(n, Ka, Kb)
// doubles for ops 0 and 1, gramPartDoubles(n, Ka, Kb) for op 2 (split-K
// partial tiles, every entry written before it is read).
void projectTN(int op, uint32_t nr, uint32_t nth, uint64_t n, uint64_t nX, const double* coef, const double* L,
               uint32_t Ka, const double* w, const double* X, uint32_t Kb, double* part, double* out,
               hipStream_t stream);
uint32_t projChunks(uint64_t n, uint64_t Ka, uint64_t Kb);
// out (device, [n x kout] column-major) = base + alpha * X W in FP64 on MFMA
// (proj.hip): X [n x kin] column-major, W [kin x kout] row-major, both on
// the device; base may be null (0) or alias out.
// G = (F^m - I) / dt of the reduced leapfrog's one-step matrix F, in blocks
// (row-major G_vv [kv x kv], G_vs [kv x ks], G_sv [ks x kv], G_ss [ks x ks])
void macroOperator(const double* arV, const double* arS, uint64_t kv, uint64_t ks, double dt, uint32_t m,
                   std::vector<double>& gvv, std::vector<double>& gvs, std::vector<double>& gsv,
                   std::vector<double>& gss);
void gemmNN(const double* X, uint64_t n, uint32_t kin, const double* W, uint32_t kout, double* out,
            const double* base, double alpha, hipStream_t stream);
uint64_t projPartDoubles(uint64_t n, uint64_t Ka, uint64_t Kb);
uint32_t gramChunks(uint64_t n, uint64_t Ka, uint64_t Kb);
uint64_t gramPartDoubles(uint64_t n, uint64_t Ka, uint64_t Kb);
}  // namespace shaw

// POD bases of both dofs kept on the device (shaw_fom_pod_basis): phi[d]
// [n[d] x k[d]] column-major, whose first j columns are the j-mode basis.
struct shaw_basis {
  uint64_t n[2] = {0, 0}, k[2] = {0, 0}, ns[2] = {0, 0};
  double* phi[2] = {nullptr, nullptr};
  std::vector<double> sigma[2];  // ns[d] singular values each
  shaw_basis() = default;
  shaw_basis(const shaw_basis&) = delete;
  shaw_basis& operator=(const shaw_basis&) = delete;
  ~shaw_basis() {
    for (double* p : phi)
      if (p) (void)hipFree(p);
  }
};
