// Tall-skinny FP64 products on MFMA for the offline ROM steps.
//
//   out[a][b] = sum_r w[r] L[r][a] (J X)[r][b]
//
// L is n x Ka and X is nX x Kb, both column-major (the lsv_<dof> layout of a
// basis); J is the FOM operator applied to X's columns on the fly:
//   OP 0: J_vp (rows = velocity points, 4 slots N, S, W, E over stress X),
//   OP 1: J_sp (rows = stress points, 2 slots over velocity X),
//   OP 2: the identity (X = a second n x Kb matrix; Gram products of POD).
// Each value of (J X) is the CSR row of ShWavePP in slot order,
// ((cN xN + cS xS) + cW xW) + cE xE, without FMA (this file is built with
// contraction off), so it equals the reference's spmv entry
// (rom_compute_jacobians.hpp:90-101 computes J Phi with KokkosSparse::spmv).
//
// The product runs on v_mfma_f64_16x16x4f64. A workgroup (4 waves) owns a
// 64 x 64 output tile and one chunk of the rows; each stage stages 32 rows of
// w L and J X in LDS (coalesced along the rows, which are contiguous in
// both matrices) and every wave accumulates a 32 x 32 quarter of the tile
// from them. Chunks write partial tiles that a second kernel sums in chunk
// order, so the result is deterministic. The chunk count depends only on
// (op, n, Ka, Kb), so every caller of the same product gets the same bits.
//
// gemmNN (the POD's X W products) runs on the same MFMA.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rom_internal.hpp"
#include "shaw_common.hpp"

namespace shaw {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;       // output tile (rows a and columns b)
constexpr int kRows = 32;       // rows of L and J X staged per step
constexpr uint64_t kMaxChunks = 64;
constexpr uint64_t kPartBudget = (16ull << 20) / sizeof(double);  // partial tiles: 16 MiB of doubles

struct ProjArgs {
  const double* L;     // [n x Ka] column-major
  const double* w;     // [n] row weights or null (1)
  const double* X;     // [nX x Kb] column-major
  const double* coef;  // OP 0: coefVp [nVp][4]; OP 1: coefSp [nSp][2]
  double* part;        // [nch][Ka][Kb]
  uint64_t n, nX, chunk;
  uint32_t Ka, Kb, nr, nth;
};

__device__ __forceinline__ dbl4 mfma64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// stencil rows of J: neighbour indices into X and coefficients of row r
template <int OP>
__device__ __forceinline__ void stencilRow(const ProjArgs& p, uint64_t r, uint64_t (&nb)[4], double (&cf)[4]) {
  const uint64_t nr = p.nr, nth = p.nth, rs = 2 * nth - 1;
  if constexpr (OP == 0) {
    const uint64_t i = r % nth, j = r / nth;
    auto stp = [&](uint64_t ii, uint64_t jj) { return jj + 1 < nr ? jj * rs + 2 * ii + 1 : (nr - 1) * rs + ii; };
    nb[0] = (j < nr - 1 ? j : nr - 2) * rs + 2 * i;
    nb[1] = (j > 0 ? j - 1 : 0) * rs + 2 * i;
    nb[2] = stp(i > 0 ? i - 1 : 0, j);
    nb[3] = stp(i < nth - 2 ? i : nth - 2, j);
#pragma unroll
    for (int q = 0; q < 4; ++q) cf[q] = p.coef[4 * r + q];
  } else if constexpr (OP == 1) {
    const uint64_t last = (nr - 1) * rs;
    if (r < last) {
      const uint64_t j = r / rs, kk = r % rs;
      if ((kk & 1) == 0) {
        nb[0] = (j + 1) * nth + kk / 2;
        nb[1] = j * nth + kk / 2;
      } else {
        nb[0] = j * nth + (kk - 1) / 2;
        nb[1] = nb[0] + 1;
      }
    } else {
      nb[0] = (nr - 1) * nth + (r - last);
      nb[1] = nb[0] + 1;
    }
    cf[0] = p.coef[2 * r];
    cf[1] = p.coef[2 * r + 1];
    nb[2] = nb[3] = 0;
    cf[2] = cf[3] = 0.0;
  } else {
    nb[0] = r;
    nb[1] = nb[2] = nb[3] = 0;
    cf[0] = 1.0;
    cf[1] = cf[2] = cf[3] = 0.0;
  }
}

template <int OP>
__global__ void __launch_bounds__(256) k_proj_tn(ProjArgs p) {
  __shared__ double Ls[kRows][kTile + 1];
  __shared__ double Js[kRows][kTile + 1];
  const uint32_t a0 = blockIdx.x * kTile, b0 = blockIdx.y * kTile, ch = blockIdx.z;
  const uint64_t r0 = (uint64_t)ch * p.chunk;
  const uint64_t r1 = r0 + p.chunk < p.n ? r0 + p.chunk : p.n;
  const int t = threadIdx.x, rr = t % kRows, cg = t / kRows;  // staging: one row, 8 column groups
  const int lane = t % 64, wv = t / 64;
  const int wa = (wv & 1) * 32, wb = (wv >> 1) * 32;  // the wave's quarter of the tile
  dbl4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  for (uint64_t rb = r0; rb < r1; rb += kRows) {
    const uint64_t r = rb + rr;
    const bool live = r < r1;
    uint64_t nb[4] = {0, 0, 0, 0};
    double cf[4] = {0.0, 0.0, 0.0, 0.0};
    double wr = 1.0;
    if (live) {
      stencilRow<OP>(p, r, nb, cf);
      if (p.w) wr = p.w[r];
    }
#pragma unroll
    for (int u = 0; u < kTile / 8; ++u) {
      const uint32_t c = cg + 8 * u;
      double lv = 0.0, jv = 0.0;
      if (live && a0 + c < p.Ka) {
        const double x = p.L[r + p.n * (a0 + c)];
        lv = p.w ? wr * x : x;
      }
      if (live && b0 + c < p.Kb) {
        const double* x = p.X + p.nX * (b0 + c);
        if constexpr (OP == 0) {
          double s = cf[0] * x[nb[0]] + cf[1] * x[nb[1]];
          s = s + cf[2] * x[nb[2]];
          s = s + cf[3] * x[nb[3]];
          jv = s;
        } else if constexpr (OP == 1) {
          jv = cf[0] * x[nb[0]] + cf[1] * x[nb[1]];
        } else {
          jv = x[r];
        }
      }
      Ls[rr][c] = lv;
      Js[rr][c] = jv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kRows / 4; ++kk) {
      // A operand: lane l holds A[l % 16][l / 16] = L[row 4kk + l/16][col l % 16]
      // B operand: lane l holds B[l / 16][l % 16] = (J X)[row 4kk + l/16][col l % 16]
      const int k = 4 * kk + lane / 16, c = lane % 16;
      const double A0 = Ls[k][wa + c], A1 = Ls[k][wa + 16 + c];
      const double B0 = Js[k][wb + c], B1 = Js[k][wb + 16 + c];
      acc[0][0] = mfma64(A0, B0, acc[0][0]);
      acc[0][1] = mfma64(A0, B1, acc[0][1]);
      acc[1][0] = mfma64(A1, B0, acc[1][0]);
      acc[1][1] = mfma64(A1, B1, acc[1][1]);
    }
    __syncthreads();
  }
  // D register v of lane l holds D[l / 16 + 4 v][l % 16]
  double* P = p.part + (uint64_t)ch * p.Ka * p.Kb;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t a = a0 + wa + 16 * i + lane / 16 + 4 * v, b = b0 + wb + 16 * j + lane % 16;
        if (a < p.Ka && b < p.Kb) P[(uint64_t)a * p.Kb + b] = acc[i][j][v];
      }
}

// out[r + n c] = base + alpha * sum_k X[r + n k] W[k kout + c] on MFMA: the
// transposed product D = W^T X^T, so a result register holds 16 consecutive
// rows of one output column (coalesced stores). A workgroup owns 64 rows x
// 128 columns, each wave 32 rows x 64 columns; W is staged in LDS 16 k at a
// time, the wave's X values for the next stage are loaded during the
// current stage's products.
constexpr int nRows = 64, nCols = 128, nK = 16;
__global__ void __launch_bounds__(256) k_gemm_nn_mfma(const double* __restrict__ X, const double* __restrict__ W,
                                                      double* out, const double* base, double alpha, uint64_t n,
                                                      uint32_t kin, uint32_t kout) {
  __shared__ double Ws[nK][nCols + 2];
  const uint64_t r0 = (uint64_t)blockIdx.x * nRows;
  const uint32_t c0 = blockIdx.y * nCols;
  const int t = threadIdx.x, lane = t % 64, wv = t / 64;
  const int wr = (wv & 1) * 32, wc = (wv >> 1) * 64;
  dbl4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
  // B operand (4 k x 16 rows): lane l holds X[row l % 16][k l / 16]
  double xb[nK / 4][2];
  auto loadX = [&](uint32_t k0) {
#pragma unroll
    for (int kk = 0; kk < nK / 4; ++kk)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint64_t r = r0 + wr + 16 * j + lane % 16;
        const uint32_t k = k0 + 4 * kk + lane / 16;
        xb[kk][j] = r < n && k < kin ? X[r + n * k] : 0.0;
      }
  };
  loadX(0);
  for (uint32_t k0 = 0; k0 < kin; k0 += nK) {
    for (int e = t; e < nK * nCols; e += 256) {
      const uint32_t kk = k0 + e / nCols, cc = c0 + e % nCols;
      Ws[e / nCols][e % nCols] = kk < kin && cc < kout ? W[(uint64_t)kk * kout + cc] : 0.0;
    }
    double xc[nK / 4][2];
#pragma unroll
    for (int kk = 0; kk < nK / 4; ++kk)
#pragma unroll
      for (int j = 0; j < 2; ++j) xc[kk][j] = xb[kk][j];
    __syncthreads();
    if (k0 + nK < kin) loadX(k0 + nK);  // in flight during the products
#pragma unroll
    for (int kk = 0; kk < nK / 4; ++kk) {
      // A operand (16 columns x 4 k): lane l holds W[k l / 16][column l % 16]
      double A[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) A[i] = Ws[4 * kk + lane / 16][wc + 16 * i + lane % 16];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma64(A[i], xc[kk][j], acc[i][j]);
    }
    __syncthreads();
  }
  // D register v of lane l holds D[l / 16 + 4 v][l % 16]: column l / 16 + 4 v, row l % 16
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t c = c0 + wc + 16 * i + lane / 16 + 4 * v;
        const uint64_t r = r0 + wr + 16 * j + lane % 16;
        if (c < kout && r < n) {
          const uint64_t o = r + n * c;
          out[o] = (base ? base[o] : 0.0) + alpha * acc[i][j][v];
        }
      }
}

// out[e] = sum over chunks c = 0, 1, ... of part[c][e], in that order
__global__ void k_sum_chunks(const double* __restrict__ part, double* __restrict__ out, uint64_t ne, uint32_t nch) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = 0.0;
  for (uint32_t c = 0; c < nch; ++c) s += part[(uint64_t)c * ne + e];
  out[e] = s;
}

}  // namespace

uint32_t projChunks(uint64_t n, uint64_t Ka, uint64_t Kb) {
  // enough chunks that a product of 256-mode bases fills the GPU (16 tiles x
  // 32 chunks), as many as 16 MiB of partial tiles hold, at least 4096 rows each
  const uint64_t byBudget = kPartBudget / std::max<uint64_t>(1, Ka * Kb);
  const uint64_t byRows = (n + 4095) / 4096;
  return (uint32_t)std::max<uint64_t>(1, std::min({kMaxChunks, byBudget, byRows}));
}

uint64_t projPartDoubles(uint64_t n, uint64_t Ka, uint64_t Kb) { return (uint64_t)projChunks(n, Ka, Kb) * Ka * Kb; }

uint32_t gramChunks(uint64_t n, uint64_t Ka, uint64_t Kb) {
  // the POD's Gram products (op 2) of up to 4096 snapshots: up to 208 chunks
  // (at n = 8.4e6, Ka = Kb = 400: 160 ms against 211 ms with 13 chunks), as
  // many as 256 MiB of partial tiles hold, at least 4096 rows each
  const uint64_t byBudget = (256ull << 20) / sizeof(double) / std::max<uint64_t>(1, Ka * Kb);
  const uint64_t byRows = (n + 4095) / 4096;
  return (uint32_t)std::max<uint64_t>(1, std::min({(uint64_t)208, byBudget, byRows}));
}

uint64_t gramPartDoubles(uint64_t n, uint64_t Ka, uint64_t Kb) { return (uint64_t)gramChunks(n, Ka, Kb) * Ka * Kb; }

void projectTN(int op, uint32_t nr, uint32_t nth, uint64_t n, uint64_t nX, const double* coef, const double* L,
               uint32_t Ka, const double* w, const double* X, uint32_t Kb, double* part, double* out,
               hipStream_t stream) {
  SHAW_REQUIRE(op >= 0 && op <= 2 && n > 0 && Ka > 0 && Kb > 0, SHAW_ERR_ARG, "projection: bad shape");
  const uint32_t nch = op == 2 ? gramChunks(n, Ka, Kb) : projChunks(n, Ka, Kb);
  ProjArgs p{};
  p.L = L; p.w = w; p.X = X; p.coef = coef; p.part = part;
  p.n = n; p.nX = nX;
  p.chunk = (n + nch - 1) / nch;
  p.chunk = (p.chunk + kRows - 1) / kRows * kRows;
  p.Ka = Ka; p.Kb = Kb; p.nr = nr; p.nth = nth;
  const dim3 grid((Ka + kTile - 1) / kTile, (Kb + kTile - 1) / kTile, nch);
  if (op == 0) hipLaunchKernelGGL(k_proj_tn<0>, grid, dim3(256), 0, stream, p);
  else if (op == 1) hipLaunchKernelGGL(k_proj_tn<1>, grid, dim3(256), 0, stream, p);
  else hipLaunchKernelGGL(k_proj_tn<2>, grid, dim3(256), 0, stream, p);
  const uint64_t ne = (uint64_t)Ka * Kb;
  hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, stream, part, out, ne, nch);
  SHAW_HIP(hipGetLastError());
}

void gemmNN(const double* X, uint64_t n, uint32_t kin, const double* W, uint32_t kout, double* out,
            const double* base, double alpha, hipStream_t stream) {
  SHAW_REQUIRE(n > 0 && kin > 0 && kout > 0, SHAW_ERR_ARG, "gemm: bad shape");
  const dim3 grid((unsigned)((n + nRows - 1) / nRows), (kout + nCols - 1) / nCols);
  hipLaunchKernelGGL(k_gemm_nn_mfma, grid, dim3(256), 0, stream, X, W, out, base, alpha, n, kin, kout);
  SHAW_HIP(hipGetLastError());
}

}  // namespace shaw
