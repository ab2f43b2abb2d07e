// Entry points shared between the FOM and ROM translation units.
#pragma once
#include <cstdint>
#include <vector>

namespace shaw {
// POD of device-resident snapshots dS [n x ns] column-major (overwritten);
// phi: host [n x k] column-major, sigma: host [ns] (may be null).
// kOut == nullptr: exactly k modes or an error; else up to k modes, stopping at
// the numerical rank, with the count written to *kOut
void podDevice(double* dS, uint64_t n, uint64_t ns, uint64_t k, double* phi, double* sigma, uint64_t* kOut);
// Symmetric eigen-decomposition on the host (Householder tridiagonalization
// and the implicit symmetric QR algorithm): A row-major [n x n] (destroyed),
// w[n] eigenvalues (unordered), V row-major [n x n] with the eigenvector of
// w[c] in column c. symeig.cpp.
void symmetricEigen(std::vector<double>& A, int n, std::vector<double>& w, std::vector<double>& V);
}  // namespace shaw
