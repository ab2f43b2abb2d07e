// Galerkin ROM on gfx950.
//
// Projection (FP64): Ar_v = Phi_v^T (J_vp Phi_s), Ar_s = Phi_s^T (J_sp Phi_v)
// (rom_compute_jacobians.hpp:58-109), on FP64 MFMA with the FOM stencil
// applied to the basis columns inside the product (proj.hip); split-K over
// row chunks whose partial tiles are reduced in a fixed order.
//
// Reduced leapfrog (rom_run_rank_one.hpp:112-151 per column; batched as in
// torevise/run_rom_rank_two_forcing.hpp:149-188), for resident column slots:
//   q_v += dt * (Ar_v q_s) ; q_v += (f dt) b ; y = C q_v (sample steps) ;
//   q_s += dt * (Ar_s q_v)
// A persistent kernel: one workgroup (8 or 16 waves) owns 16 column slots and
// runs many steps, their q_v, q_s kept in FP64 for the whole launch (in LDS
// for FP64; for BF16 in the registers of the waves that update them, with
// their bf16 splits in LDS). Each wave owns up to 4 row tiles of 16 rows of
// each operator; operators are pre-packed in MFMA operand order as lane
// items and stream from L2 through a register ring per row tile, so the
// loads of later k-steps are in flight while the MFMAs of earlier ones run.
//   FP64:   v_mfma_f64_16x16x4f64, FP64 accumulation.
//   BF16x3: Ar = Ar_hi + Ar_lo and q = q_hi + q_lo in bf16; the product uses
//           Ar_hi q_hi + Ar_hi q_lo + Ar_lo q_hi on v_mfma_f32_16x16x32_bf16
//           with FP32 accumulation; the state update q += dt * acc is FP64.
//   BF16x4: as BF16x3 plus Ar_res q_hi, Ar_res = bf16(Ar - Ar_hi - Ar_lo):
//           the operator is represented to ~2^-24 instead of ~2^-17. Over
//           thousands of steps the operator's representation error acts as
//           a fixed frequency shift (a phase error that grows with time),
//           while the state's split error is new every step.
// Each column has its own step counter and horizon; a column past its
// horizon is frozen (serving with uneven horizons).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "rom_internal.hpp"
#include "shaw_app.hpp"
#include "shaw_common.hpp"

namespace shaw {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef short shrt4 __attribute__((ext_vector_type(4)));
typedef short shrt16 __attribute__((ext_vector_type(16)));

constexpr int kTileCols = 16;  // column slots per workgroup
constexpr int kWaves = 8;      // waves per ROM workgroup (16 at 9..16 row tiles)

// ------------------------------------------------------------ bf16 helpers
__host__ __device__ inline unsigned short f32ToBf16(float x) {
  // round to nearest even
  unsigned int u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf/nan
  const unsigned int lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf16ToF32(unsigned short h) {
  const unsigned int u = (unsigned int)h << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

// ------------------------------------------------------------ reconstruction
// out[r + n*c] = sum_j S[r + n*j] * W[j*k + c]   (S: n x ns col-major, W: ns x k row-major),
// each entry summed over j in order (the reconstruction tools' loop; the POD's
// products run on MFMA, gemmNN in proj.hip)
constexpr int NNC = 16;  // output columns per thread
// out = base + alpha * (S W); base may be null (treated as 0) or alias out
__global__ void __launch_bounds__(256) k_gemm_nn_rows(const double* __restrict__ S, const double* __restrict__ W,
                                                      double* out, const double* base, double alpha, uint64_t n,
                                                      uint32_t ns, uint32_t k) {
  __shared__ double Ws[64][NNC];
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c0 = blockIdx.y * NNC;
  double acc[NNC];
#pragma unroll
  for (int c = 0; c < NNC; ++c) acc[c] = 0.0;
  for (uint32_t j0 = 0; j0 < ns; j0 += 64) {
    for (int e = threadIdx.x; e < 64 * NNC; e += blockDim.x) {
      const uint32_t jj = j0 + e / NNC, cc = c0 + e % NNC;
      Ws[e / NNC][e % NNC] = (jj < ns && cc < k) ? W[(uint64_t)jj * k + cc] : 0.0;
    }
    __syncthreads();
    if (r < n) {
      const uint32_t jn = ns - j0 < 64 ? ns - j0 : 64;
      for (uint32_t j = 0; j < jn; ++j) {
        const double x = S[r + n * (j0 + j)];
#pragma unroll
        for (int c = 0; c < NNC; ++c) acc[c] += x * Ws[j][c];
      }
    }
    __syncthreads();
  }
  if (r < n)
    for (int c = 0; c < NNC; ++c)
      if (c0 + c < k) {
        const uint64_t o = r + n * (c0 + c);
        out[o] = (base ? base[o] : 0.0) + alpha * acc[c];
      }
}

// ------------------------------------------------------------ MFMA lane layouts
// v_mfma_f64_16x16x4f64 (one 16x16 output block, K = 4):
//   A: lane l holds A[l % 16][l / 16];  B: lane l holds B[l / 16][l % 16]
//   D: lane l, register r holds D[l / 16 + 4 * r][l % 16]
// v_mfma_f32_16x16x16bf16_1k:
//   A: lane l holds A[l % 16][4 * (l / 16) + e], e = 0..3
//   B: lane l holds B[4 * (l / 16) + e][l % 16]
//   D: lane l, register r holds D[4 * (l / 16) + r][l % 16]
// v_mfma_f32_16x16x32_bf16 (gfx950; twice the bf16 products per instruction
// of the 16x16x16 form):
//   A: lane l holds A[l % 16][8 * (l / 16) + e], e = 0..7
//   B: lane l holds B[8 * (l / 16) + e][l % 16]
//   D: as 16x16x16
// The ROM kernels rely only on A's value e of lane group l / 16 meeting B's
// value e of the same lane group, and on D's layout.
// These are checked on the device by shaw_debug_mfma (tests/test_gpu_rom.py).
typedef int int4v __attribute__((ext_vector_type(4)));
typedef short shrt8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ inline dbl4 mfmaF64(double a, double b, dbl4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline flt4 mfmaBf16(shrt4 a, shrt4 b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ inline flt4 mfmaBf16x32(int4v a, int4v b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
// result row of register r in lane l
__device__ inline int rowF64(int lane, int r) { return lane / 16 + 4 * r; }
__device__ inline int rowBf16(int lane, int r) { return 4 * (lane / 16) + r; }

__global__ void k_mfma_probe(int kind, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ d) {
  const int l = threadIdx.x;
  if (kind == 0) {
    dbl4 acc = {0, 0, 0, 0};
    acc = mfmaF64(a[l], b[l], acc);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
  } else if (kind == 1) {
    shrt4 x, y;
    for (int e = 0; e < 4; ++e) {
      x[e] = (short)f32ToBf16((float)a[4 * l + e]);
      y[e] = (short)f32ToBf16((float)b[4 * l + e]);
    }
    flt4 acc = {0, 0, 0, 0};
    acc = mfmaBf16(x, y, acc);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
  } else {
    shrt8 x, y;
    for (int e = 0; e < 8; ++e) {
      x[e] = (short)f32ToBf16((float)a[8 * l + e]);
      y[e] = (short)f32ToBf16((float)b[8 * l + e]);
    }
    flt4 acc = {0, 0, 0, 0};
    acc = mfmaBf16x32(__builtin_bit_cast(int4v, x), __builtin_bit_cast(int4v, y), acc);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
  }
}

// ------------------------------------------------------------ reduced leapfrog
// Operators are packed for the MFMA A operand in lane items, laid out
// [rowTile][group][lane] so that one wave reads contiguous memory per group:
//   FP64:   32-byte items; group g of row tile t covers k-steps 4g..4g+3 of
//           16 x 16 x 4; lane l holds A[16t + l%16][16g + 4s + l/16], s < 4.
//   BF16x3 / BF16x4: 2 / 3 parts of 16 bytes (hi, lo, residual bf16 parts);
//           group g covers k = 32g .. 32g + 31, one 16 x 16 x 32 product per
//           part; part p of lane l holds that part of
//           A[16t + l%16][32g + 16s + 4(l/16) + e] at position 4s + e,
//           s < 2, e < 4 (the state's split uses the same k order).
// Padded sizes are multiples of kRing groups (64 for FP64, 128 for BF16x3).
struct RomArgs {
  const void* av;       // Ar_v: rows Kvp, k-dim Ksp
  const void* as;       // Ar_s: rows Ksp, k-dim Kvp
  const void* c;        // C:    rows Rp,  k-dim Kvp
  double* qv;           // [slot][Kvp]
  double* qs;           // [slot][Ksp]
  const double* b;      // [slot][Kvp]
  const double* ftab;   // [slot][maxSteps]
  uint64_t* stepCol;    // [slot]
  const uint64_t* horizon;
  const uint64_t* every;
  double* seis;         // [slot][R][Tcap]
  uint32_t Kvp, Ksp, Rp, R;
  uint64_t maxSteps, Tcap, nsteps;
  double dt;
  // macro steps (BF16 kernels): the blocks of G = (F^m - I) / dt, F the
  // one-step matrix of [q_v; q_s]: G_vv [Kvp x Kvp], G_vs [Kvp x Ksp], G_sv,
  // G_ss, packed as the operators; m = 0 turns them off
  const void *gvv, *gvs, *gsv, *gss;
  const uint32_t* quiet;  // [slot] first step from which the column's forcing is negligible
  uint32_t macro;
  // time budget (macro steps on): 0, nsteps steps for every tile; w > 0, the
  // time of nsteps one-steps for every tile, a macro step costing w of them,
  // so tiles taking macro steps advance further (stepCol holds how far)
  uint32_t budget;
};

// A tile's macro plan for a budget of B steps: k one-steps first (until every
// active column is at a multiple of m and quiet), then k2 macro steps (every
// column still active then at least m k2 steps from its horizon, within B).
// k2 = 0 when the tile cannot take any: active columns at different
// residues mod m, or sampled at a period that is not a multiple of m. Lanes
// hold column lane % 16 (each column 4 times); every wave computes the same.
__device__ __forceinline__ void macroPlan(uint32_t m, uint32_t s, uint32_t hz, uint32_t ev, uint32_t quiet, uint64_t B,
                                          uint64_t& k, uint64_t& k2) {
  const bool act = s < hz;
  int bad = act && ((ev != 0 && ev % m != 0) || quiet == 0xffffffffu);
  const uint32_t r = s % m;
  uint32_t rmin = act ? r : 0xffffffffu, rmax = act ? r : 0u;
  uint32_t kq = act && quiet > s ? quiet - s : 0u;
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    const uint32_t a = (uint32_t)__shfl_xor((int)rmin, off), b = (uint32_t)__shfl_xor((int)rmax, off);
    const uint32_t c = (uint32_t)__shfl_xor((int)kq, off);
    rmin = a < rmin ? a : rmin;
    rmax = b > rmax ? b : rmax;
    kq = c > kq ? c : kq;
    bad |= __shfl_xor(bad, off);
  }
  k = B;
  k2 = 0;
  if (rmin > rmax || bad || rmin != rmax) return;  // no active column, or no common macro point
  const uint64_t k0 = (uint64_t)kq + (m - (rmin + kq % m) % m) % m;
  uint64_t lim = ~0ull;  // macro steps the columns still active at k0 allow
  if (act && (uint64_t)hz > (uint64_t)s + k0) lim = ((uint64_t)hz - s - k0) / m;
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    const uint64_t o = (uint64_t)__shfl_xor((long long)lim, off);
    lim = o < lim ? o : lim;
  }
  if (lim == ~0ull || k0 > B) return;
  const uint64_t n = (B - k0) / m < lim ? (B - k0) / m : lim;
  if (n == 0) return;
  k = k0;
  k2 = n;
}

constexpr int kRing = 4;  // operator groups kept in flight per row tile

// PARTS = 0: FP64 operators; 2: bf16 hi + lo (BF16x3); 3: bf16 hi + lo +
// residual (BF16x4). Ring items: a dbl4 (FP64) or PARTS 16-byte parts, part p
// holding that part's 8 bf16 values of group g (32 B or 48 B).
template <int PARTS>
struct BfItem {
  int4v p[PARTS];
};
template <int PARTS>
struct Prec {
  using Item = BfItem<PARTS>;
  using Acc = flt4;
};
template <>
struct Prec<0> {
  using Item = dbl4;
  using Acc = dbl4;
};

// B operand of k-step group g. FP64: X is the state [K][16] in LDS. BF16:
// X is the state's bf16 split (writeSplitRows): hi parts, then lo parts, each
// [K/32][4][16] items of 8 bf16 values in the operator's k order, so a
// lane's operand of a group is one 16-byte read.
template <int PARTS, int TPW>
__device__ __forceinline__ void groupProducts(const typename Prec<PARTS>::Item (&ag)[TPW], int g, const void* X,
                                              int loOff, int lane, typename Prec<PARTS>::Acc (&acc)[TPW]) {
  if constexpr (PARTS == 0) {
    const double* Xd = static_cast<const double*>(X);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double xb = Xd[((4 * g + s) * 4 + lane / 16) * kTileCols + lane % 16];
#pragma unroll
      for (int t = 0; t < TPW; ++t) acc[t] = mfmaF64(ag[t][s], xb, acc[t]);
    }
  } else {
    const int4v* Xh = static_cast<const int4v*>(X);
    const int item = (4 * g + lane / 16) * kTileCols + lane % 16;
    const int4v xh = Xh[item], xl = Xh[loOff + item];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      acc[t] = mfmaBf16x32(ag[t].p[0], xh, acc[t]);
      acc[t] = mfmaBf16x32(ag[t].p[0], xl, acc[t]);
      acc[t] = mfmaBf16x32(ag[t].p[1], xh, acc[t]);
      if constexpr (PARTS == 3) acc[t] = mfmaBf16x32(ag[t].p[2], xh, acc[t]);
    }
  }
}

// operator groups kept in flight per row tile: kRing, or 2 where the ring
// of all the wave's tiles would not fit the registers (register-resident
// states: beside one tile of 8 waves)
template <int PARTS, int TPW, int WAVES, bool RESIDENT>
constexpr int ringDepth() {
  return RESIDENT ? (TPW == 1 && WAVES <= 8 ? kRing : 2) : (PARTS * TPW > 9 ? 2 : kRing);
}

// issue the loads of groups 0 .. D-1 of the wave's row tiles
template <int PARTS, int TPW, int D>
__device__ __forceinline__ void preloadRing(const void* __restrict__ Araw, int kGroups, const int (&rt)[TPW], int lane,
                                            typename Prec<PARTS>::Item (&ring)[D][TPW]) {
  using Item = typename Prec<PARTS>::Item;
  const Item* A = reinterpret_cast<const Item*>(Araw);
#pragma unroll
  for (int d = 0; d < D; ++d) {
#pragma unroll
    for (int t = 0; t < TPW; ++t) ring[d][t] = A[((size_t)rt[t] * kGroups + d) * 64 + lane];
    __builtin_amdgcn_sched_barrier(0);  // issue order = consumption order
  }
}

// acc[t] = (row tile rt[t] of A) X over all kGroups groups; the ring holds
// groups 0 .. D-1 on entry (preloadRing) and is refilled D groups ahead
// (ZERO = false: add to acc instead, for a product of two operator blocks)
template <int PARTS, int TPW, int D, bool ZERO = true>
__device__ __forceinline__ void consumeRing(const void* __restrict__ Araw, int kGroups, const int (&rt)[TPW],
                                            const void* X, int loOff, int lane,
                                            typename Prec<PARTS>::Item (&ring)[D][TPW],
                                            typename Prec<PARTS>::Acc (&acc)[TPW]) {
  using Item = typename Prec<PARTS>::Item;
  const Item* A = reinterpret_cast<const Item*>(Araw);
  const Item* row[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    if constexpr (ZERO) acc[t] = typename Prec<PARTS>::Acc{0, 0, 0, 0};
    row[t] = A + (size_t)rt[t] * kGroups * 64 + lane;
  }
  int g0 = 0;
  for (; g0 + D < kGroups; g0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      groupProducts<PARTS, TPW>(ring[d], g0 + d, X, loOff, lane, acc);
#pragma unroll
      for (int t = 0; t < TPW; ++t) ring[d][t] = row[t][(g0 + d + D) * 64];
      // keep this group's refill here: the scheduler would otherwise sink all
      // refills of the round to its end, shortening the prefetch distance
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) groupProducts<PARTS, TPW>(ring[d], g0 + d, X, loOff, lane, acc);
}

template <int PARTS, int TPW, int D = ringDepth<PARTS, TPW, kWaves, false>()>
__device__ __forceinline__ void tileProducts(const void* __restrict__ Araw, int kGroups, const int (&rt)[TPW],
                                             const void* X, int loOff, int lane, typename Prec<PARTS>::Acc (&acc)[TPW]) {
  typename Prec<PARTS>::Item ring[D][TPW];
  preloadRing<PARTS, TPW, D>(Araw, kGroups, rt, lane, ring);
  consumeRing<PARTS, TPW, D>(Araw, kGroups, rt, X, loOff, lane, ring, acc);
}

__device__ __forceinline__ int accRow(int parts, int lane, int r) { return parts ? rowBf16(lane, r) : rowF64(lane, r); }

// Y[rows][cols] += dt * (A X) (+ (f dt) b for the velocity half), active
// columns only. Row tiles rt0 + t * WAVES, t < TPW, belong to this wave.
template <int PARTS, int TPW, int WAVES>
__device__ __forceinline__ void halfStep(const RomArgs& a, const void* A, int rowTiles, int kGroups, const void* Xs,
                                         int loOff, double* Ys, int wave, int lane, bool active, double fdt,
                                         const double* bcol, int slot0) {
  const int col = lane % 16;
  for (int rt0 = wave; rt0 < rowTiles; rt0 += WAVES * TPW) {
    int rt[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) rt[t] = rt0 + t * WAVES < rowTiles ? rt0 + t * WAVES : rt0;  // spare: recompute
    double bv[TPW][4];
    if (bcol)  // issued before the products, consumed after them
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) bv[t][r] = bcol[(uint64_t)(slot0 + col) * a.Kvp + rt[t] * 16 + accRow(PARTS, lane, r)];
    typename Prec<PARTS>::Acc acc[TPW];
    tileProducts<PARTS, TPW>(A, kGroups, rt, Xs, loOff, lane, acc);
#pragma unroll
    for (int t = 0; t < TPW; ++t)
      if (rt0 + t * WAVES < rowTiles && active)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rt[t] * 16 + accRow(PARTS, lane, r);
          double y = Ys[row * kTileCols + col] + a.dt * (double)acc[t][r];
          if (bcol) y = y + fdt * bv[t][r];
          Ys[row * kTileCols + col] = y;
        }
  }
}

template <int PARTS, int TPW, int WAVES, int D = ringDepth<PARTS, TPW, kWaves, false>()>
__device__ __forceinline__ void receivers(const RomArgs& a, const void* Xv, int loOff, int wave, int lane, bool sample,
                                          uint32_t sampleIdx, int slot0) {
  const int rowTiles = a.Rp / 16, kGroups = PARTS ? a.Kvp / 32 : a.Kvp / 16;
  for (int rt0 = wave; rt0 < rowTiles; rt0 += WAVES * TPW) {
    int rt[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) rt[t] = rt0 + t * WAVES < rowTiles ? rt0 + t * WAVES : rt0;
    typename Prec<PARTS>::Acc acc[TPW];
    tileProducts<PARTS, TPW, D>(a.c, kGroups, rt, Xv, loOff, lane, acc);
    const int col = lane % 16;
#pragma unroll
    for (int t = 0; t < TPW; ++t)
      if (rt0 + t * WAVES < rowTiles && sample)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rt[t] * 16 + accRow(PARTS, lane, r);
          // per slot LayoutLeft [R x Tcap]
          if (row < (int)a.R)
            a.seis[(uint64_t)(slot0 + col) * a.R * a.Tcap + (uint64_t)a.R * sampleIdx + row] = (double)acc[t][r];
        }
  }
}

// The FP64 reduced leapfrog. One workgroup (WAVES waves: 16 when the
// operators have 9 to 16 row tiles, one tile per wave, which doubles the
// loads in flight per CU for the same operator stream; else 8) owns
// kTileCols column slots and advances them a.nsteps steps, states resident
// in LDS; each wave owns up to TPW row tiles of each operator. Lane l works on column l % 16 in every product, so each
// lane keeps its column's step counter, horizon, sampling period and next
// forcing value in registers (the forcing of step s + 1 is loaded during
// step s); the block-uniform decisions (stop, sample) come from values every
// wave computes identically, so a step needs only the two barriers between
// the half-steps.
template <int PARTS, int TPW, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) k_rom_steps(RomArgs a) {
  static_assert(PARTS == 0, "the bf16 variants run k_rom_steps_bf");
  extern __shared__ double sm[];
  double* Qv = sm;                                   // [Kvp][16]
  double* Qs = Qv + (size_t)a.Kvp * kTileCols;       // [Ksp][16]
  const int slot0 = blockIdx.x * kTileCols;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64, col = lane % 16;
  for (int e = tid; e < (int)a.Kvp * kTileCols; e += blockDim.x) {
    const int row = e / kTileCols, c = e % kTileCols;
    Qv[e] = a.qv[(uint64_t)(slot0 + c) * a.Kvp + row];
  }
  for (int e = tid; e < (int)a.Ksp * kTileCols; e += blockDim.x) {
    const int row = e / kTileCols, c = e % kTileCols;
    Qs[e] = a.qs[(uint64_t)(slot0 + c) * a.Ksp + row];
  }
  // this lane's column (step counts fit 32 bits: the host checks horizons)
  const uint32_t s0 = (uint32_t)a.stepCol[slot0 + col];
  const uint32_t hz = (uint32_t)a.horizon[slot0 + col];
  const uint32_t ev = (uint32_t)a.every[slot0 + col];
  const double* fcol = a.ftab + (uint64_t)(slot0 + col) * a.maxSteps;
  // block-uniform: steps until every column is done (lanes l and l + 16k hold
  // the same column, so the wave's maximum is the maximum over the columns)
  uint32_t steps = hz > s0 ? hz - s0 : 0;
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    const uint32_t o = (uint32_t)__shfl_xor((int)steps, off);
    steps = o > steps ? o : steps;
  }
  if ((uint64_t)steps > a.nsteps) steps = (uint32_t)a.nsteps;
  __syncthreads();  // the states are in LDS
  const int kvTiles = a.Kvp / 16, ksTiles = a.Ksp / 16;
  const int kGroupsS = a.Ksp / 16, kGroupsV = a.Kvp / 16;
  double fcur = s0 < hz ? fcol[s0] : 0.0;
  uint32_t phase = ev ? (s0 + 1) % ev : 0;     // (s + 1) mod every
  uint32_t nextIdx = ev ? (s0 + 1) / ev : 0;   // index of the next sample
  for (uint32_t it = 0; it < steps; ++it) {
    const uint32_t s = s0 + it;  // steps already taken by this column
    const bool act = s < hz;
    const double fnext = act && s + 1 < hz ? fcol[s + 1] : 0.0;  // consumed next step
    const double fdt = act ? fcur * a.dt : 0.0;
    const bool smp = act && ev > 0 && phase == 0;
    // lanes 0..15 of every wave hold columns 0..15: the same answer in all waves
    const bool anySample = (__ballot(smp) & 0xffffull) != 0;
    // velocity: q_v += dt Ar_v q_s ; q_v += (f dt) b
    halfStep<PARTS, TPW, WAVES>(a, a.av, kvTiles, kGroupsS, Qs, 0, Qv, wave, lane, act, fdt, a.b, slot0);
    __syncthreads();
    if (anySample) receivers<PARTS, TPW, WAVES>(a, Qv, 0, wave, lane, smp, nextIdx - 1, slot0);
    // stress: q_s += dt Ar_s q_v
    halfStep<PARTS, TPW, WAVES>(a, a.as, ksTiles, kGroupsV, Qv, 0, Qs, wave, lane, act, 0.0, nullptr, slot0);
    __syncthreads();
    fcur = fnext;
    if (ev && act) {
      phase = phase + 1 == ev ? 0 : phase + 1;
      if (phase == 0) ++nextIdx;
    }
  }
  for (int e = tid; e < (int)a.Kvp * kTileCols; e += blockDim.x) {
    const int row = e / kTileCols, c = e % kTileCols;
    a.qv[(uint64_t)(slot0 + c) * a.Kvp + row] = Qv[e];
  }
  for (int e = tid; e < (int)a.Ksp * kTileCols; e += blockDim.x) {
    const int row = e / kTileCols, c = e % kTileCols;
    a.qs[(uint64_t)(slot0 + c) * a.Ksp + row] = Qs[e];
  }
  if (tid < kTileCols) {
    const uint64_t h = a.horizon[slot0 + tid];
    const uint64_t st = a.stepCol[slot0 + tid], s = st + a.nsteps;
    a.stepCol[slot0 + tid] = s < h ? s : (st < h ? h : st);
  }
}

// the bf16 split of the 4 state rows a lane holds after the products of row
// tile rt (rows 16 rt + 4 (lane / 16) + r, column lane % 16): x = float(q),
// hi = bf16(x), lo = bf16(x - hi), both rounded to nearest even. Item
// (4 g + q') * 16 + c of the split holds, at position 4 s + e, k = 32 g +
// 16 s + 4 q' + e of column c (the operator's k order, see the packing
// above); here g = rt / 2, s = rt % 2, q' = lane / 16, e = r
__device__ __forceinline__ void writeSplitRows(int4v* S, int loOff, int rt, int lane, const double (&q)[4]) {
  shrt4 h, l;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float x = (float)q[r];
    const unsigned short hb = f32ToBf16(x);
    h[r] = (short)hb;
    l[r] = (short)f32ToBf16(x - bf16ToF32(hb));
  }
  const int item = (4 * (rt / 2) + lane / 16) * kTileCols + lane % 16;
  *(reinterpret_cast<shrt4*>(S + item) + rt % 2) = h;
  *(reinterpret_cast<shrt4*>(S + loOff + item) + rt % 2) = l;
}

// The BF16x3 / BF16x4 reduced leapfrog with register-resident states. As
// k_rom_steps, one workgroup advances kTileCols column slots; here the FP64
// states live in the registers of the wave that computes them (lane l holds
// rows 16 rt + 4 (l / 16) + r of column l % 16 of each of its row tiles, the
// MFMA result layout), and LDS holds only the bf16 splits of q_v and q_s,
// each written by its owning waves right after their update. A step is then
// two barriers: velocity products (read the split of q_s) -> update q_v,
// write its split -> barrier -> stress products (read the split of q_v) ->
// update q_s, write its split -> barrier. The first operator groups of the
// next half-step are loaded before each barrier, so the barrier does not
// drain the operator stream. Every value equals k_rom_steps' (the same
// splits, products and FP64 updates in the same order).
template <int PARTS, int TPW, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) k_rom_steps_bf(RomArgs a) {
  // tiles per pass: at three or four tiles a wave the products run in passes
  // of one or two tiles, whose rings fit the registers beside the states
  // (K = 384 and 512, M = 4096: BF16x4 1.77e8 -> 2.06e8 and 7.6e7 -> 1.19e8,
  // BF16x3 3.02e8 -> 3.11e8 and 1.55e8 -> 1.72e8 forcing-steps/s)
  constexpr int TPP = TPW == 4 ? 2 : (TPW == 3 ? 1 : TPW), NP = TPW / TPP;
  // (BF16x4 at three tiles a wave: a 2-deep ring keeps it within 256 VGPRs
  // without spills; K = 384, M = 4096: 2.10e8 -> 2.30e8 forcing-steps/s)
  constexpr int D = (PARTS == 3 && TPW == 3) ? 2 : ringDepth<PARTS, TPP, WAVES, true>();
  using Item = typename Prec<PARTS>::Item;
  using Acc = typename Prec<PARTS>::Acc;
  extern __shared__ double sm[];
  // splits: hi items, then lo items, 2 K items of 16 bytes each per state
  int4v* Sv = reinterpret_cast<int4v*>(sm);
  int4v* Ss = Sv + 4 * (size_t)a.Kvp;
  const int loV = 2 * (int)a.Kvp, loS = 2 * (int)a.Ksp;
  const int slot0 = blockIdx.x * kTileCols;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64, col = lane % 16;
  const int kvTiles = a.Kvp / 16, ksTiles = a.Ksp / 16;
  const int kGroupsS = a.Ksp / 32, kGroupsV = a.Kvp / 32;
  // this wave's row tiles (the host launches WAVES * TPW >= both tile counts);
  // a spare slot repeats tile `wave` and neither stores nor writes a split
  int rtV[TPW], rtS[TPW];
  bool okV[TPW], okS[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    okV[t] = wave + t * WAVES < kvTiles;
    okS[t] = wave + t * WAVES < ksTiles;
    rtV[t] = okV[t] ? wave + t * WAVES : (wave < kvTiles ? wave : 0);
    rtS[t] = okS[t] ? wave + t * WAVES : (wave < ksTiles ? wave : 0);
  }
  int rtVp[NP][TPP], rtSp[NP][TPP];  // the tiles of each pass
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int j = 0; j < TPP; ++j) {
      rtVp[p][j] = rtV[p * TPP + j];
      rtSp[p][j] = rtS[p * TPP + j];
    }
  const uint64_t colV = (uint64_t)(slot0 + col) * a.Kvp, colS = (uint64_t)(slot0 + col) * a.Ksp;
  double qv[TPW][4], qs[TPW][4];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rowV = rtV[t] * 16 + rowBf16(lane, r), rowS = rtS[t] * 16 + rowBf16(lane, r);
      qv[t][r] = okV[t] ? a.qv[colV + rowV] : 0.0;
      qs[t][r] = okS[t] ? a.qs[colS + rowS] : 0.0;
    }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    if (okV[t]) writeSplitRows(Sv, loV, rtV[t], lane, qv[t]);
    if (okS[t]) writeSplitRows(Ss, loS, rtS[t], lane, qs[t]);
  }
  const uint32_t s0 = (uint32_t)a.stepCol[slot0 + col];
  const uint32_t hz = (uint32_t)a.horizon[slot0 + col];
  const uint32_t ev = (uint32_t)a.every[slot0 + col];
  const double* fcol = a.ftab + (uint64_t)(slot0 + col) * a.maxSteps;
  uint32_t steps = hz > s0 ? hz - s0 : 0;
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    const uint32_t o = (uint32_t)__shfl_xor((int)steps, off);
    steps = o > steps ? o : steps;
  }
  if ((uint64_t)steps > a.nsteps) steps = (uint32_t)a.nsteps;
  Item ring[D][TPP];
  Acc acc[TPP];
  if (steps) preloadRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[0], lane, ring);
  __syncthreads();  // both splits are in LDS
  double fcur = s0 < hz ? fcol[s0] : 0.0;
  uint32_t phase = ev ? (s0 + 1) % ev : 0;
  uint32_t nextIdx = ev ? (s0 + 1) / ev : 0;
  for (uint32_t it = 0; it < steps; ++it) {
    const uint32_t s = s0 + it;
    const bool act = s < hz;
    const double fnext = act && s + 1 < hz ? fcol[s + 1] : 0.0;
    const double fdt = act ? fcur * a.dt : 0.0;
    const bool smp = act && ev > 0 && phase == 0;
    const bool anySample = (__ballot(smp) & 0xffffull) != 0;
    // velocity: q_v += dt Ar_v q_s ; q_v += (f dt) b
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      consumeRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[p], Ss, loS, lane, ring, acc);
      double bv[TPP][4];  // b stays in L1/L2; loaded here to keep it out of the products' registers
#pragma unroll
      for (int j = 0; j < TPP; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) bv[j][r] = a.b[colV + rtVp[p][j] * 16 + rowBf16(lane, r)];
      if (p + 1 < NP) preloadRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[p + 1 < NP ? p + 1 : 0], lane, ring);
      else preloadRing<PARTS, TPP, D>(a.as, kGroupsV, rtSp[0], lane, ring);
#pragma unroll
      for (int j = 0; j < TPP; ++j) {
        const int t = p * TPP + j;
        if (okV[t]) {
          if (act)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              double y = qv[t][r] + a.dt * (double)acc[j][r];
              qv[t][r] = y + fdt * bv[j][r];
            }
          writeSplitRows(Sv, loV, rtV[t], lane, qv[t]);
        }
      }
    }
    __syncthreads();
    // stress: q_s += dt Ar_s q_v
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      consumeRing<PARTS, TPP, D>(a.as, kGroupsV, rtSp[p], Sv, loV, lane, ring, acc);
      if (p + 1 < NP) {
        preloadRing<PARTS, TPP, D>(a.as, kGroupsV, rtSp[p + 1 < NP ? p + 1 : 0], lane, ring);
      } else {
        // one receiver tile at a time with a short ring: sample steps are
        // rare, and the states stay in registers across them
        if (anySample) receivers<PARTS, 1, WAVES, 2>(a, Sv, loV, wave, lane, smp, nextIdx - 1, slot0);
        if (it + 1 < steps) preloadRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[0], lane, ring);
      }
#pragma unroll
      for (int j = 0; j < TPP; ++j) {
        const int t = p * TPP + j;
        if (okS[t]) {
          if (act)
#pragma unroll
            for (int r = 0; r < 4; ++r) qs[t][r] = qs[t][r] + a.dt * (double)acc[j][r];
          writeSplitRows(Ss, loS, rtS[t], lane, qs[t]);
        }
      }
    }
    __syncthreads();
    fcur = fnext;
    if (ev && act) {
      phase = phase + 1 == ev ? 0 : phase + 1;
      if (phase == 0) ++nextIdx;
    }
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (okV[t]) a.qv[colV + rtV[t] * 16 + rowBf16(lane, r)] = qv[t][r];
      if (okS[t]) a.qs[colS + rtS[t] * 16 + rowBf16(lane, r)] = qs[t][r];
    }
  if (tid < kTileCols) {
    const uint64_t h = a.horizon[slot0 + tid];
    const uint64_t st = a.stepCol[slot0 + tid], s = st + a.nsteps;
    a.stepCol[slot0 + tid] = s < h ? s : (st < h ? h : st);
  }
}

// k_rom_steps_bf with macro steps (the engine's m = a.macro > 0): each tile
// plans its launch (macroPlan) and runs k one-steps, k2 macro steps of m
// steps each, then one-steps to the end, all in this launch, so tiles at
// different stages do not wait for each other. A macro step is x += dt (G x)
// with x = [q_v; q_s] and G = (F^m - I) / dt, F the leapfrog's one-step
// matrix (no forcing: every active column is quiet); G's blocks G_vv, G_vs
// (rows of q_v) and G_sv, G_ss (rows of q_s) are two products a pass, read
// from the splits of the old state, and the new states stay in the registers
// until every wave is done with the old splits.
template <int PARTS, int TPW, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) k_rom_steps_bfm(RomArgs a) {
  // tiles per pass: at three or four tiles a wave the products run in passes
  // of one or two tiles, whose rings fit the registers beside the states
  // (K = 384 and 512, M = 4096: BF16x4 1.77e8 -> 2.06e8 and 7.6e7 -> 1.19e8,
  // BF16x3 3.02e8 -> 3.11e8 and 1.55e8 -> 1.72e8 forcing-steps/s)
  constexpr int TPP = TPW == 4 ? 2 : (TPW == 3 ? 1 : TPW), NP = TPW / TPP;
  // (BF16x4 at three tiles a wave: a 2-deep ring keeps it within 256 VGPRs
  // without spills; K = 384, M = 4096: 2.10e8 -> 2.30e8 forcing-steps/s)
  constexpr int D = (PARTS == 3 && TPW == 3) ? 2 : ringDepth<PARTS, TPP, WAVES, true>();
  using Item = typename Prec<PARTS>::Item;
  using Acc = typename Prec<PARTS>::Acc;
  extern __shared__ double sm[];
  // splits: hi items, then lo items, 2 K items of 16 bytes each per state
  int4v* Sv = reinterpret_cast<int4v*>(sm);
  int4v* Ss = Sv + 4 * (size_t)a.Kvp;
  const int loV = 2 * (int)a.Kvp, loS = 2 * (int)a.Ksp;
  const int slot0 = blockIdx.x * kTileCols;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64, col = lane % 16;
  const int kvTiles = a.Kvp / 16, ksTiles = a.Ksp / 16;
  const int kGroupsS = a.Ksp / 32, kGroupsV = a.Kvp / 32;
  // this wave's row tiles (the host launches WAVES * TPW >= both tile counts);
  // a spare slot repeats tile `wave` and neither stores nor writes a split
  int rtV[TPW], rtS[TPW];
  bool okV[TPW], okS[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    okV[t] = wave + t * WAVES < kvTiles;
    okS[t] = wave + t * WAVES < ksTiles;
    rtV[t] = okV[t] ? wave + t * WAVES : (wave < kvTiles ? wave : 0);
    rtS[t] = okS[t] ? wave + t * WAVES : (wave < ksTiles ? wave : 0);
  }
  int rtVp[NP][TPP], rtSp[NP][TPP];  // the tiles of each pass
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int j = 0; j < TPP; ++j) {
      rtVp[p][j] = rtV[p * TPP + j];
      rtSp[p][j] = rtS[p * TPP + j];
    }
  const uint64_t colV = (uint64_t)(slot0 + col) * a.Kvp, colS = (uint64_t)(slot0 + col) * a.Ksp;
  double qv[TPW][4], qs[TPW][4];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rowV = rtV[t] * 16 + rowBf16(lane, r), rowS = rtS[t] * 16 + rowBf16(lane, r);
      qv[t][r] = okV[t] ? a.qv[colV + rowV] : 0.0;
      qs[t][r] = okS[t] ? a.qs[colS + rowS] : 0.0;
    }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    if (okV[t]) writeSplitRows(Sv, loV, rtV[t], lane, qv[t]);
    if (okS[t]) writeSplitRows(Ss, loS, rtS[t], lane, qs[t]);
  }
  const uint32_t s0 = (uint32_t)a.stepCol[slot0 + col];
  const uint32_t hz = (uint32_t)a.horizon[slot0 + col];
  const uint32_t ev = (uint32_t)a.every[slot0 + col];
  const double* fcol = a.ftab + (uint64_t)(slot0 + col) * a.maxSteps;
  uint32_t span = hz > s0 ? hz - s0 : 0;  // steps to the tile's last horizon
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    const uint32_t o = (uint32_t)__shfl_xor((int)span, off);
    span = o > span ? o : span;
  }
  uint32_t steps = (uint64_t)span > a.nsteps ? (uint32_t)a.nsteps : span;
  uint64_t mk = 0, mk2 = 0;  // macro steps: mk one-steps, mk2 macro steps, then one-steps to the end
  if (a.macro && a.budget) {
    // the time of T one-steps: mk one-steps, mk2 macro steps of a.budget
    // one-steps' time each, one-steps for the rest (horizons below 2^31)
    const uint64_t T = a.nsteps < (1ull << 31) ? a.nsteps : (1ull << 31);
    macroPlan(a.macro, s0, hz, ev, a.quiet[slot0 + col], T * a.macro, mk, mk2);
    const uint64_t planned = mk2;  // what the horizons allow
    if (mk2) {
      const uint64_t cap = mk < T ? (T - mk) / a.budget : 0;
      mk2 = mk2 < cap ? mk2 : cap;
    }
    if (mk2) {
      // the time left goes to one-steps; unless a column nears its horizon,
      // whole multiples of m only, so the tile ends aligned for the next launch
      uint64_t rest = T - mk - a.budget * mk2;
      if (mk2 < planned) rest -= rest % a.macro;
      const uint64_t end = mk + (uint64_t)a.macro * mk2 + rest;
      steps = end < (uint64_t)span ? (uint32_t)end : span;
    }
  } else if (a.macro) {
    macroPlan(a.macro, s0, hz, ev, a.quiet[slot0 + col], a.nsteps, mk, mk2);
  }
  // the plan is the same in every lane: keep it in scalar registers
  steps = (uint32_t)__builtin_amdgcn_readfirstlane((int)steps);
  mk = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)mk);
  mk2 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)mk2);
  const uint32_t mSteps = a.macro;
  Item ring[D][TPP];
  Acc acc[TPP];
  for (int seg = 0; seg < 2; ++seg) {
    const uint32_t from = seg == 0 ? 0u : (uint32_t)(mk + (uint64_t)mSteps * mk2);
    const uint32_t to = seg == 0 ? (mk2 ? (uint32_t)mk : steps) : (mk2 ? steps : 0u);
    if (from < to) {
      preloadRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[0], lane, ring);
      __syncthreads();  // both splits are in LDS
      const uint32_t sf = s0 + from;
      double fcur = sf < hz ? fcol[sf] : 0.0;
      uint32_t phase = ev ? (sf + 1) % ev : 0;
      uint32_t nextIdx = ev ? (sf + 1) / ev : 0;
      for (uint32_t it = from; it < to; ++it) {
        const uint32_t s = s0 + it;
        const bool act = s < hz;
        const double fnext = act && s + 1 < hz ? fcol[s + 1] : 0.0;
        const double fdt = act ? fcur * a.dt : 0.0;
        const bool smp = act && ev > 0 && phase == 0;
        const bool anySample = (__ballot(smp) & 0xffffull) != 0;
        // velocity: q_v += dt Ar_v q_s ; q_v += (f dt) b
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          consumeRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[p], Ss, loS, lane, ring, acc);
          double bv[TPP][4];  // b stays in L1/L2; loaded here to keep it out of the products' registers
#pragma unroll
          for (int j = 0; j < TPP; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) bv[j][r] = a.b[colV + rtVp[p][j] * 16 + rowBf16(lane, r)];
          if (p + 1 < NP) preloadRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[p + 1 < NP ? p + 1 : 0], lane, ring);
          else preloadRing<PARTS, TPP, D>(a.as, kGroupsV, rtSp[0], lane, ring);
#pragma unroll
          for (int j = 0; j < TPP; ++j) {
            const int t = p * TPP + j;
            if (okV[t]) {
              if (act)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  double y = qv[t][r] + a.dt * (double)acc[j][r];
                  qv[t][r] = y + fdt * bv[j][r];
                }
              writeSplitRows(Sv, loV, rtV[t], lane, qv[t]);
            }
          }
        }
        __syncthreads();
        // stress: q_s += dt Ar_s q_v
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          consumeRing<PARTS, TPP, D>(a.as, kGroupsV, rtSp[p], Sv, loV, lane, ring, acc);
          if (p + 1 < NP) {
            preloadRing<PARTS, TPP, D>(a.as, kGroupsV, rtSp[p + 1 < NP ? p + 1 : 0], lane, ring);
          } else {
            // one receiver tile at a time with a short ring: sample steps are
            // rare, and the states stay in registers across them
            if (anySample) receivers<PARTS, 1, WAVES, 2>(a, Sv, loV, wave, lane, smp, nextIdx - 1, slot0);
            if (it + 1 < to) preloadRing<PARTS, TPP, D>(a.av, kGroupsS, rtVp[0], lane, ring);
          }
#pragma unroll
          for (int j = 0; j < TPP; ++j) {
            const int t = p * TPP + j;
            if (okS[t]) {
              if (act)
#pragma unroll
                for (int r = 0; r < 4; ++r) qs[t][r] = qs[t][r] + a.dt * (double)acc[j][r];
              writeSplitRows(Ss, loS, rtS[t], lane, qs[t]);
            }
          }
        }
        __syncthreads();
        fcur = fnext;
        if (ev && act) {
          phase = phase + 1 == ev ? 0 : phase + 1;
          if (phase == 0) ++nextIdx;
        }
        }
    }
    if (seg == 0 && mk2) {
      constexpr int DM = TPW == 3 ? 2 : D;
      const bool act = s0 + (uint32_t)mk < hz;
      Item mring[DM][TPP];
      preloadRing<PARTS, TPP, DM>(a.gvv, kGroupsV, rtVp[0], lane, mring);
      __syncthreads();  // both splits are in LDS
        for (uint64_t it = 0; it < mk2; ++it) {
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            consumeRing<PARTS, TPP, DM>(a.gvv, kGroupsV, rtVp[p], Sv, loV, lane, mring, acc);
            preloadRing<PARTS, TPP, DM>(a.gvs, kGroupsS, rtVp[p], lane, mring);
            consumeRing<PARTS, TPP, DM, false>(a.gvs, kGroupsS, rtVp[p], Ss, loS, lane, mring, acc);
            if (p + 1 < NP) preloadRing<PARTS, TPP, DM>(a.gvv, kGroupsV, rtVp[p + 1 < NP ? p + 1 : 0], lane, mring);
            else preloadRing<PARTS, TPP, DM>(a.gsv, kGroupsV, rtSp[0], lane, mring);
#pragma unroll
            for (int j = 0; j < TPP; ++j)
              if (act && okV[p * TPP + j])
#pragma unroll
                for (int r = 0; r < 4; ++r) qv[p * TPP + j][r] = qv[p * TPP + j][r] + a.dt * (double)acc[j][r];
          }
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            consumeRing<PARTS, TPP, DM>(a.gsv, kGroupsV, rtSp[p], Sv, loV, lane, mring, acc);
            preloadRing<PARTS, TPP, DM>(a.gss, kGroupsS, rtSp[p], lane, mring);
            consumeRing<PARTS, TPP, DM, false>(a.gss, kGroupsS, rtSp[p], Ss, loS, lane, mring, acc);
            if (p + 1 < NP) preloadRing<PARTS, TPP, DM>(a.gsv, kGroupsV, rtSp[p + 1 < NP ? p + 1 : 0], lane, mring);
            else if (it + 1 < mk2) preloadRing<PARTS, TPP, DM>(a.gvv, kGroupsV, rtVp[0], lane, mring);
#pragma unroll
            for (int j = 0; j < TPP; ++j)
              if (act && okS[p * TPP + j])
#pragma unroll
                for (int r = 0; r < 4; ++r) qs[p * TPP + j][r] = qs[p * TPP + j][r] + a.dt * (double)acc[j][r];
          }
          __syncthreads();  // every read of the old splits done
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            if (okV[t]) writeSplitRows(Sv, loV, rtV[t], lane, qv[t]);
            if (okS[t]) writeSplitRows(Ss, loS, rtS[t], lane, qs[t]);
          }
          __syncthreads();  // the new splits complete
          // the macro step's last step is sampled when it ends at a multiple of the period
          const uint64_t sn = (uint64_t)s0 + mk + (it + 1) * mSteps;
          const bool smp = act && ev > 0 && sn % ev == 0;
          if ((__ballot(smp) & 0xffffull) != 0)
            receivers<PARTS, 1, WAVES, 2>(a, Sv, loV, wave, lane, smp, smp ? (uint32_t)(sn / ev - 1) : 0u, slot0);
        }
      __syncthreads();  // the next one-steps write the split the receivers read
    }
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (okV[t]) a.qv[colV + rtV[t] * 16 + rowBf16(lane, r)] = qv[t][r];
      if (okS[t]) a.qs[colS + rtS[t] * 16 + rowBf16(lane, r)] = qs[t][r];
    }
  if (tid < kTileCols) {
    // the tile ran `steps` steps: a.nsteps, or fewer when every column
    // reached its horizon (the same counts), or more under a time budget
    const uint64_t h = a.horizon[slot0 + tid];
    const uint64_t st = a.stepCol[slot0 + tid], s = st + steps;
    a.stepCol[slot0 + tid] = s < h ? s : (st < h ? h : st);
  }
}

// admission: the first step from which the staged forcing series of
// column j ([n][nsteps]) stays at or below 2^-64 of its largest magnitude (0
// when it is all zero) goes to quiet[j]. A forcing value that small adds
// under 2^-64 of the column's response, far below the BF16 kernels'
// rounding, so a quiet column may take macro steps without its forcing.
__global__ void __launch_bounds__(256) k_quiet(const double* __restrict__ f, uint64_t nsteps,
                                               uint32_t* __restrict__ quiet) {
  __shared__ double peak[256];
  __shared__ uint64_t last[256];
  const uint64_t j = blockIdx.x;
  const unsigned t = threadIdx.x;
  const double* fj = f + j * nsteps;
  double mx = 0.0;
  for (uint64_t e = t; e < nsteps; e += 256) mx = fmax(mx, fabs(fj[e]));
  peak[t] = mx;
  __syncthreads();
  for (unsigned w = 128; w > 0; w /= 2) {
    if (t < w) peak[t] = fmax(peak[t], peak[t + w]);
    __syncthreads();
  }
  const double thr = peak[0] * 0x1p-64;
  uint64_t lastAbove = 0;  // 1 + the last index above the threshold
  for (uint64_t e = t; e < nsteps; e += 256)
    if (fabs(fj[e]) > thr) lastAbove = e + 1;
  last[t] = lastAbove;
  __syncthreads();
  for (unsigned w = 128; w > 0; w /= 2) {
    if (t < w) last[t] = last[t] > last[t + w] ? last[t] : last[t + w];
    __syncthreads();
  }
  if (t == 0) quiet[j] = last[0] < 0xffffffffull ? (uint32_t)last[0] : 0xffffffffu;
}

// admission: column j of the staged inputs (forcing column, series, quiet
// step) goes to slot slots[j]; its states start at zero
__global__ void k_admit_scatter(const uint64_t* __restrict__ slots, uint64_t n, const double* __restrict__ bStage,
                                const double* __restrict__ fStage, const uint32_t* __restrict__ quietStage,
                                uint32_t kvp, uint32_t ksp, uint64_t nsteps, uint64_t maxSteps,
                                double* __restrict__ b, double* __restrict__ ftab, double* __restrict__ qv,
                                double* __restrict__ qs, uint32_t* __restrict__ quiet) {
  const uint64_t j = blockIdx.y;
  if (j >= n) return;
  const uint64_t s = slots[j];
  if (blockIdx.x == 0 && threadIdx.x == 0) quiet[s] = quietStage[j];
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nsteps; e += (uint64_t)gridDim.x * blockDim.x)
    ftab[s * maxSteps + e] = fStage[j * nsteps + e];
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < kvp; e += (uint64_t)gridDim.x * blockDim.x) {
    b[s * kvp + e] = bStage[j * kvp + e];
    qv[s * kvp + e] = 0.0;
  }
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ksp; e += (uint64_t)gridDim.x * blockDim.x)
    qs[s * ksp + e] = 0.0;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t count) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = count;
    if (count) SHAW_HIP(hipMalloc((void**)&p, count * sizeof(T)));
  }
  void upload(const T* h, size_t count, size_t off = 0) {
    if (count) SHAW_HIP(hipMemcpy(p + off, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
};

uint32_t roundUp(uint64_t x, uint32_t m) { return (uint32_t)((x + m - 1) / m * m); }

// pack a row-major [rows x cols] operator (zero padded to rowsP x colsP,
// whole groups) into the 32-byte lane items described above; returns the
// packed bytes
std::vector<double> packF64(const double* A, uint64_t rows, uint64_t cols, uint32_t rowsP, uint32_t colsP) {
  const uint32_t rt = rowsP / 16, G = colsP / 16;
  std::vector<double> out((size_t)rt * G * 64 * 4, 0.0);
  for (uint32_t t = 0; t < rt; ++t)
    for (uint32_t g = 0; g < G; ++g)
      for (int l = 0; l < 64; ++l)
        for (int s = 0; s < 4; ++s) {
          const uint64_t r = t * 16 + l % 16, c = g * 16 + 4 * s + l / 16;
          out[(((size_t)t * G + g) * 64 + l) * 4 + s] = (r < rows && c < cols) ? A[r * cols + c] : 0.0;
        }
  return out;
}
std::vector<double> packBf16(const double* A, uint64_t rows, uint64_t cols, uint32_t rowsP, uint32_t colsP, int parts) {
  const uint32_t rt = rowsP / 16, G = colsP / 32;
  std::vector<double> out((size_t)rt * G * 64 * 2 * parts, 0.0);  // parts x 16 B per lane item
  unsigned short* o16 = reinterpret_cast<unsigned short*>(out.data());
  for (uint32_t t = 0; t < rt; ++t)
    for (uint32_t g = 0; g < G; ++g)
      for (int l = 0; l < 64; ++l)
        for (int s = 0; s < 2; ++s)
          for (int e = 0; e < 4; ++e) {
            const uint64_t r = t * 16 + l % 16, c = g * 32 + 16 * s + 4 * (l / 16) + e;
            const double v = (r < rows && c < cols) ? A[r * cols + c] : 0.0;
            const unsigned short h = f32ToBf16((float)v);
            const double rest = v - (double)bf16ToF32(h);
            const unsigned short lo = f32ToBf16((float)rest);
            const size_t o = (((size_t)t * G + g) * 64 + l) * (8 * parts) + 4 * s + e;
            o16[o] = h;
            o16[o + 8] = lo;
            if (parts == 3) o16[o + 16] = f32ToBf16((float)(rest - (double)bf16ToF32(lo)));
          }
  return out;
}

template <int PARTS, int TPW, int WAVES>
void launchRomKernel(dim3 grid, size_t lds, hipStream_t stream, const RomArgs& a) {
  auto go = [&](auto kernel) {
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kernel, grid, dim3(WAVES * 64), lds, stream, a);
  };
  if constexpr (PARTS == 0) go(k_rom_steps<0, TPW, WAVES>);
  else if (a.macro) go(k_rom_steps_bfm<PARTS, TPW, WAVES>);
  else go(k_rom_steps_bf<PARTS, TPW, WAVES>);
}

template <int PARTS>
void launchRomSteps(int tpw, dim3 grid, size_t lds, hipStream_t stream, const RomArgs& a) {
  switch (tpw) {
    case 0: launchRomKernel<PARTS, 1, 16>(grid, lds, stream, a); break;  // 9..16 row tiles: one per wave of 16
    case 1: launchRomKernel<PARTS, 1, kWaves>(grid, lds, stream, a); break;
    case 2: launchRomKernel<PARTS, 2, kWaves>(grid, lds, stream, a); break;
    case 3: launchRomKernel<PARTS, 3, kWaves>(grid, lds, stream, a); break;
    default: launchRomKernel<PARTS, 4, kWaves>(grid, lds, stream, a); break;
  }
}

}  // namespace
}  // namespace shaw

namespace shaw {
// POD of device snapshots dS [n x ns] column-major (overwritten by the
// deflation); phi (host, [n x k] column-major) and sigma (host, ns values).
void podDevice(double* dSp, uint64_t n, uint64_t ns, uint64_t k, double* phi, double* sigma, uint64_t* kOut) {
    SHAW_REQUIRE(dSp && phi && n > 0 && ns > 0 && k > 0, SHAW_ERR_ARG, "bad argument");
    SHAW_REQUIRE(k <= ns && k <= n, SHAW_ERR_ARG, "POD: number of modes exceeds the number of snapshots");
    DevBuf<double> d;
    d.alloc(n * k);
    uint64_t have = k;
    podDeviceTo(dSp, n, ns, k, d.p, sigma, kOut ? &have : nullptr);
    if (kOut) *kOut = have;
    SHAW_HIP(hipMemcpy(phi, d.p, n * have * 8, hipMemcpyDeviceToHost));
}

void podDeviceTo(double* dSp, uint64_t n, uint64_t ns, uint64_t k, double* dPhiOut, double* sigma, uint64_t* kOut) {
    SHAW_REQUIRE(dSp && dPhiOut && n > 0 && ns > 0 && k > 0, SHAW_ERR_ARG, "bad argument");
    SHAW_REQUIRE(k <= ns && k <= n, SHAW_ERR_ARG, "POD: number of modes exceeds the number of snapshots");
    SHAW_REQUIRE(ns <= 4096, SHAW_ERR_ARG, "POD: at most 4096 snapshots");
    DevBuf<double> part, dG, dW, dNew, dTmp;
    struct { double* p; } dS{dSp}, dPhi{dPhiOut};
    // host <- A^T B (A: n x ka, B: n x kb, column-major on the device), FP64 MFMA
    auto gramTN = [&](const double* A, uint32_t ka, const double* B, uint32_t kb, std::vector<double>& G) {
      const uint64_t need = gramPartDoubles(n, ka, kb);
      if (part.n < need) part.alloc(need);
      if (dG.n < (size_t)ka * kb) dG.alloc((size_t)ka * kb);
      projectTN(2, 0, 0, n, n, nullptr, A, ka, nullptr, B, kb, part.p, dG.p, 0);
      G.resize((size_t)ka * kb);
      SHAW_HIP(hipMemcpy(G.data(), dG.p, G.size() * 8, hipMemcpyDeviceToHost));
    };
    // out = base + alpha * X W (X: n x kin on the device, W: kin x kout row-major on the host)
    auto applyNN = [&](const double* X, uint32_t kin, const std::vector<double>& Wh, uint32_t kout, double* out,
                       const double* base, double alpha) {
      dW.alloc(Wh.size());
      dW.upload(Wh.data(), Wh.size());
      gemmNN(X, n, kin, dW.p, kout, out, base, alpha, 0);
    };
    // Phi_new <- Phi_new L^-T, Phi_new^T Phi_new = L L^T
    auto cholOrth = [&](double* P, uint32_t r) {
      std::vector<double> G2;
      gramTN(P, r, P, r, G2);
      std::vector<double> L((size_t)r * r, 0.0);
      for (uint32_t i = 0; i < r; ++i)
        for (uint32_t j = 0; j <= i; ++j) {
          double sum = G2[(size_t)i * r + j];
          for (uint32_t q = 0; q < j; ++q) sum -= L[(size_t)i * r + q] * L[(size_t)j * r + q];
          if (i == j) {
            SHAW_REQUIRE(sum > 0, SHAW_ERR_ARG, "POD: basis lost orthogonality (snapshots too ill-conditioned)");
            L[(size_t)i * r + i] = std::sqrt(sum);
          } else {
            L[(size_t)i * r + j] = sum / L[(size_t)j * r + j];
          }
        }
      std::vector<double> X((size_t)r * r, 0.0), W2((size_t)r * r);
      for (uint32_t c = 0; c < r; ++c)
        for (uint32_t i = 0; i < r; ++i) {
          double sum = i == c ? 1.0 : 0.0;
          for (uint32_t q = 0; q < i; ++q) sum -= L[(size_t)i * r + q] * X[(size_t)q * r + c];
          X[(size_t)i * r + c] = sum / L[(size_t)i * r + i];
        }
      for (uint32_t i = 0; i < r; ++i)
        for (uint32_t j = 0; j < r; ++j) W2[(size_t)i * r + j] = X[(size_t)j * r + i];
      dTmp.alloc((size_t)n * r);
      applyNN(P, r, W2, r, dTmp.p, nullptr, 1.0);
      SHAW_HIP(hipMemcpy(P, dTmp.p, (size_t)n * r * 8, hipMemcpyDeviceToDevice));
    };
    // Deflated method of snapshots: each pass takes the modes of the current
    // residual snapshots S_c whose singular values its Gram matrix resolves
    // to ~eps sigma_max(S_c) (sigma > 1e-3 sigma_max(S_c): the eigenvalue error
    // eps sigma_max^2 costs < 1e-13 sigma_max in sigma), orthogonalizes them and
    // removes them from S_c; smaller singular values are computed relative to
    // the residual, not to sigma_1 (the accuracy of a QR-based SVD).
    std::vector<double> sv;
    std::vector<double> lastTail;
    uint64_t have = 0;
    double sigma1 = 0.0;
    for (int pass = 0; pass < 24 && have < k; ++pass) {
      std::vector<double> G, w, V;
      gramTN(dS.p, (uint32_t)ns, dS.p, (uint32_t)ns, G);
      symmetricEigen(G, (int)ns, w, V);
      std::vector<int> order(ns);
      for (uint64_t i = 0; i < ns; ++i) order[i] = (int)i;
      std::sort(order.begin(), order.end(), [&](int x, int y) { return w[x] > w[y]; });
      std::vector<double> sc(ns);
      for (uint64_t i = 0; i < ns; ++i) sc[i] = std::sqrt(std::max(w[order[i]], 0.0));
      if (pass == 0) sigma1 = sc[0];
      lastTail.assign(sc.begin(), sc.end());
      uint64_t r = 0;
      while (r < k - have && sc[r] > 1e-3 * sc[0] && sc[r] > 1e-14 * sigma1) ++r;
      if (r == 0) break;
      std::vector<double> W(ns * r);
      for (uint64_t j = 0; j < ns; ++j)
        for (uint64_t c = 0; c < r; ++c) W[j * r + c] = V[(size_t)j * ns + order[c]] / sc[c];
      dNew.alloc(n * r);
      applyNN(dS.p, (uint32_t)ns, W, (uint32_t)r, dNew.p, nullptr, 1.0);
      if (have) {  // orthogonalize against the modes already taken
        std::vector<double> X;
        gramTN(dPhi.p, (uint32_t)have, dNew.p, (uint32_t)r, X);
        applyNN(dPhi.p, (uint32_t)have, X, (uint32_t)r, dNew.p, dNew.p, -1.0);
      }
      cholOrth(dNew.p, (uint32_t)r);
      SHAW_HIP(hipMemcpy(dPhi.p + n * have, dNew.p, n * r * 8, hipMemcpyDeviceToDevice));
      for (uint64_t c = 0; c < r; ++c) sv.push_back(sc[c]);
      have += r;
      if (have >= k) break;  // no further pass reads the residual
      // S_c <- S_c - Phi_new (Phi_new^T S_c)
      std::vector<double> Y;
      gramTN(dNew.p, (uint32_t)r, dS.p, (uint32_t)ns, Y);
      applyNN(dNew.p, (uint32_t)r, Y, (uint32_t)ns, dS.p, dS.p, -1.0);
    }
    // kOut: up to k modes (stop at the numerical rank); else exactly k
    SHAW_REQUIRE(kOut ? have > 0 : have == k, SHAW_ERR_ARG,
                 "POD: requested modes exceed the numerical rank of the snapshots (" + std::to_string(have) +
                     " resolved)");
    if (sigma) {
      for (uint64_t i = 0; i < ns; ++i) sigma[i] = i < have ? sv[i] : (i - have < lastTail.size() ? lastTail[i - have] : 0.0);
    }
    if (kOut) *kOut = have;
}
}  // namespace shaw

namespace shaw {
// host copy split over up to 8 threads (large retire copies are bound by a
// single core's memcpy rate otherwise)
void parallelCopy(void* dst, const void* src, size_t bytes) {
  const size_t nt = bytes < (16u << 20) ? 1 : std::min<size_t>(8, bytes / (8u << 20));
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  const size_t part = (bytes / nt + 4095) / 4096 * 4096;
  for (size_t t = 0; t < nt; ++t) {
    const size_t b0 = t * part;
    if (b0 >= bytes) break;
    const size_t len = std::min(part, bytes - b0);
    th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + b0, static_cast<const char*>(src) + b0, len); });
  }
  for (auto& x : th) x.join();
}
}  // namespace shaw

struct shaw_rom {
  uint64_t kv = 0, ks = 0, R = 0, maxCols = 0, maxSamples = 0, maxSteps = 0;
  uint32_t kvp = 0, ksp = 0, rp = 0;
  double dt = 0;
  int precision = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int tpw = 1;  // row tiles per wave
  shaw::DevBuf<double> av, as, c, qv, qs, b, ftab, seis;  // av/as/c: packed operator items
  // macro steps (BF16 engines): the FP64 operators they are built from, the
  // packed blocks of G = (F^m - I) / dt, m (0: off) and each slot's first
  // quiet step
  std::vector<double> hArV, hArS;
  shaw::DevBuf<double> gvv, gvs, gsv, gss;
  shaw::DevBuf<uint32_t> quiet, quietStage;
  uint32_t macro = 0, macroBuilt = 0;
  // time budget (shaw_rom_set_time_budget): 0 off, else a macro step's cost
  // in one-steps; launches then advance tiles by amounts the device records
  // in stepCol, read back into hStep once the stream is idle (settle)
  uint32_t timeBudget = 0;
  bool stepsOnDevice = false;
  shaw::DevBuf<uint64_t> stepCol, horizon, every;
  std::vector<uint64_t> hStep, hHorizon, hEvery;
  std::vector<uint8_t> used;
  shaw::DevBuf<double> stage;          // admission inputs, [n][Kvp] then [n][nsteps]
  shaw::DevBuf<uint64_t> stageSlots;
  void* pinned = nullptr;              // host staging of retired seismograms
  size_t pinnedBytes = 0;
  // pinned host staging of admissions (forcing columns, signal parameters,
  // slots) and of the slot tables, copied on the engine's stream without
  // waiting: nothing rewrites them before the next call on the engine,
  // which waits for the stream first (settle)
  void* admitPinned = nullptr;
  size_t admitPinnedBytes = 0;
  uint64_t* slotPinned = nullptr;      // [3][maxCols]: steps, horizons, sampling periods
  void* admitStage(size_t bytes) {
    if (bytes > admitPinnedBytes) {
      if (admitPinned) (void)hipHostFree(admitPinned);
      admitPinned = nullptr;
      admitPinnedBytes = 0;
      SHAW_HIP(hipHostMalloc(&admitPinned, bytes));
      admitPinnedBytes = bytes;
    }
    return admitPinned;
  }
  double* pinnedStage(size_t bytes) {
    if (bytes > pinnedBytes) {
      if (pinned) (void)hipHostFree(pinned);
      pinned = nullptr;
      pinnedBytes = 0;
      SHAW_HIP(hipHostMalloc(&pinned, bytes));
      pinnedBytes = bytes;
    }
    return static_cast<double*>(pinned);
  }
  void settle() {
    SHAW_HIP(hipStreamSynchronize(stream));
    if (stepsOnDevice) {
      // through the pinned slot-table buffer (its uploads are done: the
      // stream is idle), on the engine's stream
      if (!slotPinned) SHAW_HIP(hipHostMalloc((void**)&slotPinned, 3 * maxCols * sizeof(uint64_t)));
      SHAW_HIP(hipMemcpyAsync(slotPinned, stepCol.p, maxCols * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
      SHAW_HIP(hipStreamSynchronize(stream));
      std::memcpy(hStep.data(), slotPinned, maxCols * sizeof(uint64_t));
      stepsOnDevice = false;
    }
  }
  // stream-ordered upload of the slot tables (callers have settled the stream)
  void uploadSlotTables() {
    if (!slotPinned) SHAW_HIP(hipHostMalloc((void**)&slotPinned, 3 * maxCols * sizeof(uint64_t)));
    const size_t bytes = maxCols * sizeof(uint64_t);
    std::memcpy(slotPinned, hStep.data(), bytes);
    std::memcpy(slotPinned + maxCols, hHorizon.data(), bytes);
    std::memcpy(slotPinned + 2 * maxCols, hEvery.data(), bytes);
    SHAW_HIP(hipMemcpyAsync(stepCol.p, slotPinned, bytes, hipMemcpyHostToDevice, stream));
    SHAW_HIP(hipMemcpyAsync(horizon.p, slotPinned + maxCols, bytes, hipMemcpyHostToDevice, stream));
    SHAW_HIP(hipMemcpyAsync(every.p, slotPinned + 2 * maxCols, bytes, hipMemcpyHostToDevice, stream));
  }
  ~shaw_rom() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (pinned) (void)hipHostFree(pinned);
    if (admitPinned) (void)hipHostFree(admitPinned);
    if (slotPinned) (void)hipHostFree(slotPinned);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // LDS of one workgroup: FP64, the states (8 bytes an element); BF16, the
  // hi/lo splits of both states (4 bytes an element; the FP64 states are in
  // registers)
  size_t ldsBytes() const {
    return (size_t)(kvp + ksp) * shaw::kTileCols * (precision == SHAW_ROM_FP64 ? sizeof(double) : 4);
  }
  void launch(uint64_t nsteps) {
    using namespace shaw;
    RomArgs a{};
    a.av = reinterpret_cast<const dbl4*>(av.p);
    a.as = reinterpret_cast<const dbl4*>(as.p);
    a.c = reinterpret_cast<const dbl4*>(c.p);
    a.qv = qv.p; a.qs = qs.p; a.b = b.p; a.ftab = ftab.p;
    a.stepCol = stepCol.p; a.horizon = horizon.p; a.every = every.p; a.seis = seis.p;
    a.Kvp = kvp; a.Ksp = ksp; a.Rp = rp; a.R = (uint32_t)R;
    a.maxSteps = maxSteps; a.Tcap = maxSamples; a.nsteps = nsteps; a.dt = dt;
    a.gvv = gvv.p; a.gvs = gvs.p; a.gsv = gsv.p; a.gss = gss.p;
    a.quiet = quiet.p;
    a.macro = precision == SHAW_ROM_FP64 ? 0 : macro;
    a.budget = a.macro ? timeBudget : 0;
    const dim3 grid((unsigned)(maxCols / kTileCols));
    const size_t lds = ldsBytes();
    if (precision == SHAW_ROM_FP64) launchRomSteps<0>(tpw, grid, lds, stream, a);
    else if (precision == SHAW_ROM_BF16X3) launchRomSteps<2>(tpw, grid, lds, stream, a);
    else launchRomSteps<3>(tpw, grid, lds, stream, a);
    SHAW_HIP(hipGetLastError());
    if (a.budget) {
      stepsOnDevice = true;
      return;
    }
    for (uint64_t c = 0; c < maxCols; ++c) {
      const uint64_t s = hStep[c] + nsteps;
      hStep[c] = s < hHorizon[c] ? s : std::max(hStep[c], hHorizon[c]);
    }
  }
};

using shaw::guarded;

namespace shaw {
namespace {
// Ar_v = Phi_v^T (J_vp Phi_s), Ar_s = Phi_s^T (J_sp Phi_v) with the bases on
// the device (leading dimension nVp / nSp), results to the host (row-major)
void projectPair(shaw_app_t app, const double* pv, uint64_t kv, const double* ps, uint64_t ks, double* arV,
                 double* arS) {
  const Grid& g = app->ops.grid;
  const uint64_t nVp = g.nVp(), nSp = g.nSp();
  DevBuf<double> cv, cs, part, out;
  cv.alloc(app->ops.coefVp.size()); cv.upload(app->ops.coefVp.data(), cv.n);
  cs.alloc(app->ops.coefSp.size()); cs.upload(app->ops.coefSp.data(), cs.n);
  part.alloc(std::max(projPartDoubles(nVp, kv, ks), projPartDoubles(nSp, ks, kv)));
  out.alloc(kv * ks);
  projectTN(0, (uint32_t)g.nr, (uint32_t)g.nth, nVp, nSp, cv.p, pv, (uint32_t)kv, nullptr, ps, (uint32_t)ks,
            part.p, out.p, 0);
  SHAW_HIP(hipMemcpy(arV, out.p, kv * ks * sizeof(double), hipMemcpyDeviceToHost));
  projectTN(1, (uint32_t)g.nr, (uint32_t)g.nth, nSp, nVp, cs.p, ps, (uint32_t)ks, nullptr, pv, (uint32_t)kv,
            part.p, out.p, 0);
  SHAW_HIP(hipMemcpy(arS, out.p, kv * ks * sizeof(double), hipMemcpyDeviceToHost));
}

// the energy-consistent pair: Ar_v = (diag(Mv) Phi_v)^T (J_vp Phi_s); Ar_s = -Ar_v^T
void projectPairEnergy(shaw_app_t app, const double* pv, uint64_t kv, const double* ps, uint64_t ks,
                       const double* mv, double* arV, double* arS) {
  const Grid& g = app->ops.grid;
  const uint64_t nVp = g.nVp(), nSp = g.nSp();
  DevBuf<double> cv, w, part, out;
  cv.alloc(app->ops.coefVp.size()); cv.upload(app->ops.coefVp.data(), cv.n);
  w.alloc(nVp); w.upload(mv, nVp);
  part.alloc(projPartDoubles(nVp, kv, ks));
  out.alloc(kv * ks);
  projectTN(0, (uint32_t)g.nr, (uint32_t)g.nth, nVp, nSp, cv.p, pv, (uint32_t)kv, w.p, ps, (uint32_t)ks, part.p,
            out.p, 0);
  SHAW_HIP(hipMemcpy(arV, out.p, kv * ks * sizeof(double), hipMemcpyDeviceToHost));
  for (uint64_t r = 0; r < ks; ++r)
    for (uint64_t c = 0; c < kv; ++c) arS[r * kv + c] = -arV[c * ks + r];
}

// out[j][c] = phi[gid[j] + n c], c < k
__global__ void k_gather_rows(const double* __restrict__ phi, uint64_t n, const uint64_t* __restrict__ gid,
                              uint64_t m, uint64_t k, double* __restrict__ out) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * k) return;
  const uint64_t j = e / k, c = e % k;
  out[e] = phi[gid[j] + n * c];
}
}  // namespace
}  // namespace shaw

extern "C" {

int shaw_rom_project(shaw_app_t app, const double* phiV, uint64_t kv, const double* phiS, uint64_t ks, double* arV,
                     double* arS) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(app && phiV && phiS && arV && arS && kv > 0 && ks > 0, SHAW_ERR_ARG, "bad argument");
    const Grid& g = app->ops.grid;
    DevBuf<double> pv, ps;
    pv.alloc(g.nVp() * kv); pv.upload(phiV, pv.n);
    ps.alloc(g.nSp() * ks); ps.upload(phiS, ps.n);
    projectPair(app, pv.p, kv, ps.p, ks, arV, arS);
  });
}

int shaw_rom_project_energy(shaw_app_t app, const double* phiV, uint64_t kv, const double* phiS, uint64_t ks,
                            const double* mv, double* arV, double* arS) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(app && phiV && phiS && mv && arV && arS && kv > 0 && ks > 0, SHAW_ERR_ARG, "bad argument");
    const Grid& g = app->ops.grid;
    DevBuf<double> pv, ps;
    pv.alloc(g.nVp() * kv); pv.upload(phiV, pv.n);
    ps.alloc(g.nSp() * ks); ps.upload(phiS, ps.n);
    projectPairEnergy(app, pv.p, kv, ps.p, ks, mv, arV, arS);
  });
}

int shaw_basis_info(shaw_basis_t b, uint64_t* n_vp, uint64_t* n_sp, uint64_t* kv, uint64_t* ks, uint64_t* nsnap_vp,
                    uint64_t* nsnap_sp) {
  return guarded([&] {
    SHAW_REQUIRE(b, SHAW_ERR_ARG, "null basis");
    if (n_vp) *n_vp = b->n[0];
    if (n_sp) *n_sp = b->n[1];
    if (kv) *kv = b->k[0];
    if (ks) *ks = b->k[1];
    if (nsnap_vp) *nsnap_vp = b->ns[0];
    if (nsnap_sp) *nsnap_sp = b->ns[1];
  });
}

int shaw_basis_sigma(shaw_basis_t b, int dof, double* sigma) {
  return guarded([&] {
    SHAW_REQUIRE(b && sigma && (dof == 0 || dof == 1), SHAW_ERR_ARG, "bad argument");
    std::copy(b->sigma[dof].begin(), b->sigma[dof].end(), sigma);
  });
}

int shaw_basis_rows(shaw_basis_t b, int dof, const uint64_t* gids, uint64_t m, uint64_t k, double* out) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(b && (dof == 0 || dof == 1) && (m == 0 || (gids && out)), SHAW_ERR_ARG, "bad argument");
    SHAW_REQUIRE(k > 0 && k <= b->k[dof], SHAW_ERR_ARG, "more modes than the basis holds");
    for (uint64_t j = 0; j < m; ++j) SHAW_REQUIRE(gids[j] < b->n[dof], SHAW_ERR_ARG, "grid point out of range");
    if (m == 0) return;
    DevBuf<uint64_t> dg;
    DevBuf<double> dout;
    dg.alloc(m); dg.upload(gids, m);
    dout.alloc(m * k);
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((m * k + 255) / 256)), dim3(256), 0, 0, b->phi[dof], b->n[dof],
                       dg.p, m, k, dout.p);
    SHAW_HIP(hipGetLastError());
    SHAW_HIP(hipMemcpy(out, dout.p, m * k * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int shaw_basis_download(shaw_basis_t b, int dof, uint64_t k, double* phi) {
  return guarded([&] {
    SHAW_REQUIRE(b && phi && (dof == 0 || dof == 1), SHAW_ERR_ARG, "bad argument");
    SHAW_REQUIRE(k > 0 && k <= b->k[dof], SHAW_ERR_ARG, "more modes than the basis holds");
    SHAW_HIP(hipMemcpy(phi, b->phi[dof], b->n[dof] * k * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int shaw_basis_project(shaw_app_t app, shaw_basis_t b, uint64_t kv, uint64_t ks, const double* mv, double* ar_v,
                       double* ar_s) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(app && b && ar_v && ar_s && kv > 0 && ks > 0, SHAW_ERR_ARG, "bad argument");
    SHAW_REQUIRE(kv <= b->k[0] && ks <= b->k[1], SHAW_ERR_ARG, "more modes than the basis holds");
    SHAW_REQUIRE(b->n[0] == app->ops.grid.nVp() && b->n[1] == app->ops.grid.nSp(), SHAW_ERR_ARG,
                 "basis and app have different grids");
    if (mv) projectPairEnergy(app, b->phi[0], kv, b->phi[1], ks, mv, ar_v, ar_s);
    else projectPair(app, b->phi[0], kv, b->phi[1], ks, ar_v, ar_s);
  });
}

void shaw_basis_destroy(shaw_basis_t b) { delete b; }

int shaw_basis_orthonormalize(const double* phi, uint64_t n, uint64_t k, const double* w, double* phiOut,
                              double* rOut) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(phi && w && phiOut && n > 0 && k > 0 && k <= n && k < (1u << 16), SHAW_ERR_ARG, "bad argument");
    DevBuf<double> dPhi, dW, part, dG, dR, dOut;
    dPhi.alloc(n * k); dPhi.upload(phi, n * k);
    dW.alloc(n); dW.upload(w, n);
    part.alloc(gramPartDoubles(n, k, k));
    dG.alloc(k * k);
    // G = Phi^T diag(w) Phi (FP64 MFMA)
    projectTN(2, 0, 0, n, n, nullptr, dPhi.p, (uint32_t)k, dW.p, dPhi.p, (uint32_t)k, part.p, dG.p, 0);
    std::vector<double> G(k * k);
    SHAW_HIP(hipMemcpy(G.data(), dG.p, k * k * sizeof(double), hipMemcpyDeviceToHost));
    // G = R^T R, R upper triangular (Cholesky by rows of R)
    std::vector<double> R(k * k, 0.0);
    for (uint64_t i = 0; i < k; ++i) {
      double d = G[i * k + i];
      for (uint64_t q = 0; q < i; ++q) d -= R[q * k + i] * R[q * k + i];
      SHAW_REQUIRE(d > 0.0, SHAW_ERR_ARG, "basis is rank-deficient in the given weights");
      R[i * k + i] = std::sqrt(d);
      for (uint64_t j = i + 1; j < k; ++j) {
        double x = 0.5 * (G[i * k + j] + G[j * k + i]);
        for (uint64_t q = 0; q < i; ++q) x -= R[q * k + i] * R[q * k + j];
        R[i * k + j] = x / R[i * k + i];
      }
    }
    // Rinv (upper triangular), column by column: R Rinv = I
    std::vector<double> Ri(k * k, 0.0);
    for (uint64_t c = 0; c < k; ++c)
      for (uint64_t i = c + 1; i-- > 0;) {
        double x = i == c ? 1.0 : 0.0;
        for (uint64_t q = i + 1; q <= c; ++q) x -= R[i * k + q] * Ri[q * k + c];
        Ri[i * k + c] = x / R[i * k + i];
      }
    // Phi_out = Phi Rinv
    dR.alloc(k * k); dR.upload(Ri.data(), k * k);
    dOut.alloc(n * k);
    gemmNN(dPhi.p, n, (uint32_t)k, dR.p, (uint32_t)k, dOut.p, nullptr, 1.0, 0);
    SHAW_HIP(hipGetLastError());
    SHAW_HIP(hipMemcpy(phiOut, dOut.p, n * k * sizeof(double), hipMemcpyDeviceToHost));
    if (rOut) std::copy(R.begin(), R.end(), rOut);
  });
}

int shaw_rom_create(const shaw_rom_desc* d, shaw_rom_t* out) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(d && out && d->ar_v && d->ar_s, SHAW_ERR_ARG, "null argument");
    SHAW_REQUIRE(d->kv > 0 && d->ks > 0 && d->kv <= 512 && d->ks <= 512, SHAW_ERR_ARG, "ROM size must be in 1..512 modes per dof");
    SHAW_REQUIRE(d->precision == SHAW_ROM_FP64 || d->precision == SHAW_ROM_BF16X3 || d->precision == SHAW_ROM_BF16X4,
                 SHAW_ERR_ARG, "unknown ROM precision");
    SHAW_REQUIRE(d->max_cols > 0 && d->dt > 0, SHAW_ERR_ARG, "max_cols and dt must be > 0");
    SHAW_REQUIRE(d->r == 0 || d->c, SHAW_ERR_ARG, "receiver rows missing");
    int ndev = 0;
    SHAW_HIP(hipGetDeviceCount(&ndev));
    SHAW_REQUIRE(ndev > 0, SHAW_ERR_HIP, "no HIP device available");
    auto h = std::make_unique<shaw_rom>();
    h->kv = d->kv; h->ks = d->ks; h->R = d->r; h->dt = d->dt; h->precision = d->precision;
    // padded sizes: whole ring rounds of operator groups (16 k per FP64
    // group, 32 per BF16x3 group)
    const uint32_t kq = (d->precision == SHAW_ROM_FP64 ? 16 : 32) * kRing;
    h->kvp = roundUp(d->kv, kq);
    h->ksp = roundUp(d->ks, kq);
    h->rp = roundUp(std::max<uint64_t>(d->r, 1), 16);
    h->maxCols = roundUp(d->max_cols, kTileCols);
    h->maxSamples = d->max_samples;
    SHAW_REQUIRE(h->kvp <= 512 && h->ksp <= 512, SHAW_ERR_ARG, "ROM size must be at most 512 modes per dof");
    {
      const uint64_t tiles = std::max(h->kvp, h->ksp) / 16;
      h->tpw = tiles > (uint64_t)kWaves && tiles <= 16 ? 0 : (int)((tiles + kWaves - 1) / kWaves);
    }
    SHAW_REQUIRE(h->ldsBytes() <= 160 * 1024, SHAW_ERR_ARG, "ROM size too large for the LDS-resident state tile");
    SHAW_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    SHAW_HIP(hipEventCreate(&h->ev0));
    SHAW_HIP(hipEventCreate(&h->ev1));
    std::vector<double> cPad((size_t)h->rp * d->kv, 0.0);
    if (d->r) std::memcpy(cPad.data(), d->c, d->r * d->kv * sizeof(double));
    auto pack = [&](const double* A, uint64_t rows, uint64_t cols, uint32_t rowsP, uint32_t colsP,
                    shaw::DevBuf<double>& dst) {
      const std::vector<double> p = d->precision == SHAW_ROM_FP64
                                        ? packF64(A, rows, cols, rowsP, colsP)
                                        : packBf16(A, rows, cols, rowsP, colsP, d->precision == SHAW_ROM_BF16X4 ? 3 : 2);
      dst.alloc(p.size());
      dst.upload(p.data(), p.size());
    };
    pack(d->ar_v, d->kv, d->ks, h->kvp, h->ksp, h->av);
    pack(d->ar_s, d->ks, d->kv, h->ksp, h->kvp, h->as);
    if (d->precision != SHAW_ROM_FP64) {
      h->hArV.assign(d->ar_v, d->ar_v + d->kv * d->ks);
      h->hArS.assign(d->ar_s, d->ar_s + d->ks * d->kv);
    }
    pack(cPad.data(), h->rp, d->kv, h->rp, h->kvp, h->c);
    const uint64_t C = h->maxCols;
    h->qv.alloc(C * h->kvp); h->qs.alloc(C * h->ksp); h->b.alloc(C * h->kvp);
    SHAW_HIP(hipMemset(h->qv.p, 0, h->qv.n * 8)); SHAW_HIP(hipMemset(h->qs.p, 0, h->qs.n * 8));
    SHAW_HIP(hipMemset(h->b.p, 0, h->b.n * 8));
    h->stepCol.alloc(C); h->horizon.alloc(C); h->every.alloc(C);
    h->quiet.alloc(C);
    SHAW_HIP(hipMemset(h->quiet.p, 0xff, C * sizeof(uint32_t)));  // never quiet until admitted
    h->hStep.assign(C, 0); h->hHorizon.assign(C, 0); h->hEvery.assign(C, 0); h->used.assign(C, 0);
    h->stepCol.upload(h->hStep.data(), C); h->horizon.upload(h->hHorizon.data(), C); h->every.upload(h->hEvery.data(), C);
    h->seis.alloc(C * std::max<uint64_t>(h->R, 1) * std::max<uint64_t>(h->maxSamples, 1));
    *out = h.release();
  });
}

}  // extern "C"

namespace shaw {
namespace {
// when a signal (signalAt) falls for good to 2^-64 of its peak, in seconds,
// for ordering columns by when their forcing ends: a Ricker wavelet at
// |t - delay| = 2.23 periods (pi^2 f^2 tau^2 = 49), gaussDer at 6.87
// (f^2 tau^2 = 47.1); a sinusoid stays on
double signalEnd(int kind, double delay, double period) {
  if (kind == 0) return delay + 2.23 * period;
  if (kind == 2) return delay + 6.87 * period;
  return std::numeric_limits<double>::infinity();
}

// admission with the forcing series from a host table ftab [nsteps][n], or
// (ftab null) evaluated on the device from signal parameters
void admit(shaw_rom_t h, uint64_t n, const double* b, uint64_t nsteps, const double* ftab, const int* kinds,
           const double* delays, const double* periods, const uint64_t* horizons, uint64_t every, uint64_t* slots) {
    SHAW_REQUIRE(h && b && (ftab || (kinds && delays && periods)) && horizons && slots, SHAW_ERR_ARG,
                 "null argument");
    SHAW_REQUIRE(h->R == 0 || every > 0, SHAW_ERR_ARG, "sampling period must be > 0 when receivers are set");
    // the columns go to wholly free tiles first, then to the free slots of
    // the other tiles: a tile shared with columns at other stages takes
    // macro steps only when all of them allow it
    std::vector<uint64_t> free;
    std::vector<uint8_t> taken(h->used);
    const uint64_t tiles = h->maxCols / kTileCols;
    for (int pass = 0; pass < 2 && free.size() < n; ++pass)
      for (uint64_t t = 0; t < tiles && free.size() < n; ++t) {
        uint64_t nfree = 0;
        for (uint64_t c = t * kTileCols; c < (t + 1) * kTileCols; ++c) nfree += taken[c] ? 0 : 1;
        if (nfree == 0 || (pass == 0 && nfree < (uint64_t)kTileCols)) continue;
        for (uint64_t c = t * kTileCols; c < (t + 1) * kTileCols && free.size() < n; ++c)
          if (!taken[c]) {
            taken[c] = 1;
            free.push_back(c);
          }
      }
    SHAW_REQUIRE(free.size() == n, SHAW_ERR_MEMORY, "not enough free ROM column slots");
    for (uint64_t j = 0; j < n; ++j) {
      SHAW_REQUIRE(horizons[j] <= nsteps, SHAW_ERR_ARG, "horizon longer than the forcing table");
      SHAW_REQUIRE(horizons[j] < (1ull << 31), SHAW_ERR_ARG, "horizon must be below 2^31 steps");
      if (h->R) SHAW_REQUIRE(horizons[j] / every <= h->maxSamples, SHAW_ERR_ARG, "horizon exceeds the seismogram capacity");
    }
    if (n == 0) return;
    // a launch still in flight reads and writes back every slot of its
    // workgroups: admit only on an idle engine
    h->settle();
    const uint64_t need = std::max<uint64_t>(h->maxSteps, nsteps);
    if (need > h->maxSteps) {  // grow the per-slot forcing store, keeping admitted columns
      DevBuf<double> nb;
      nb.alloc(h->maxCols * need);
      SHAW_HIP(hipMemset(nb.p, 0, nb.n * 8));
      if (h->ftab.p && h->maxSteps)
        SHAW_HIP(hipMemcpy2D(nb.p, need * 8, h->ftab.p, h->maxSteps * 8, h->maxSteps * 8, h->maxCols, hipMemcpyDeviceToDevice));
      std::swap(nb.p, h->ftab.p);
      std::swap(nb.n, h->ftab.n);
      h->maxSteps = need;
    }
    // stage [n][Kvp] forcing columns and [n][nsteps] forcing series, one scatter.
    // With signals only b, the parameters [n][3] (after the series region)
    // and the slots cross PCIe, from the pinned admission buffer on the
    // engine's stream, and the series is evaluated on the device.
    const uint64_t stageNeed = n * (h->kvp + nsteps) + (ftab ? 0 : 3 * n);
    if (h->stage.n < stageNeed) h->stage.alloc(stageNeed);
    uint64_t* hSlots = nullptr;  // pinned slots of the signal path
    if (ftab) {
      std::vector<double> st(n * (h->kvp + nsteps), 0.0);
      for (uint64_t j = 0; j < n; ++j) {
        for (uint64_t k = 0; k < h->kv; ++k) st[j * h->kvp + k] = b[k * n + j];
        double* f = st.data() + n * h->kvp + j * nsteps;
        for (uint64_t t = 0; t < nsteps; ++t) f[t] = ftab[t * n + j];
      }
      h->stage.upload(st.data(), st.size());
    } else {
      const std::vector<double> par = signalParams(n, kinds, delays, periods);
      double* hb = static_cast<double*>(h->admitStage((n * h->kvp + 3 * n) * sizeof(double) + n * sizeof(uint64_t)));
      std::fill(hb, hb + n * h->kvp, 0.0);
      for (uint64_t j = 0; j < n; ++j)
        for (uint64_t k = 0; k < h->kv; ++k) hb[j * h->kvp + k] = b[k * n + j];
      std::copy(par.begin(), par.end(), hb + n * h->kvp);
      hSlots = reinterpret_cast<uint64_t*>(hb + n * h->kvp + 3 * n);
      double* dPar = h->stage.p + n * (h->kvp + nsteps);
      SHAW_HIP(hipMemcpyAsync(h->stage.p, hb, n * h->kvp * sizeof(double), hipMemcpyHostToDevice, h->stream));
      SHAW_HIP(hipMemcpyAsync(dPar, hb + n * h->kvp, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
      forcingTableFromParams(n, dPar, h->dt, nsteps, h->stage.p + n * h->kvp, nsteps, h->stream);
    }
    // each column's first quiet step, staged; the scatter puts it at the slot
    if (h->quietStage.n < n) h->quietStage.alloc(n);
    hipLaunchKernelGGL(k_quiet, dim3((unsigned)n), dim3(256), 0, h->stream, h->stage.p + n * h->kvp, nsteps,
                       h->quietStage.p);
    SHAW_HIP(hipGetLastError());
    std::vector<uint64_t> slotOf(free);  // column j -> slot
    if (h->macro) {
      // columns whose forcing ends at about the same step share tiles, so
      // that their tiles turn quiet together (a tile waits for its last
      // column): the free slots in tile order go to the columns by quiet
      // step; for signals by when the signal ends (no wait for the device)
      std::vector<double> q(n);
      if (ftab) {
        std::vector<uint32_t> qd(n);
        SHAW_HIP(hipMemcpyAsync(qd.data(), h->quietStage.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
        SHAW_HIP(hipStreamSynchronize(h->stream));
        for (uint64_t j = 0; j < n; ++j) q[j] = qd[j];
      } else {
        for (uint64_t j = 0; j < n; ++j) q[j] = signalEnd(kinds[j], delays[j], periods[j]);
      }
      std::vector<uint64_t> order(n);
      for (uint64_t j = 0; j < n; ++j) order[j] = j;
      std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return q[x] < q[y]; });
      for (uint64_t i = 0; i < n; ++i) slotOf[order[i]] = free[i];
    }
    if (h->stageSlots.n < n) h->stageSlots.alloc(n);
    if (hSlots) {
      std::copy(slotOf.begin(), slotOf.end(), hSlots);
      SHAW_HIP(hipMemcpyAsync(h->stageSlots.p, hSlots, n * sizeof(uint64_t), hipMemcpyHostToDevice, h->stream));
    } else {
      h->stageSlots.upload(slotOf.data(), n);
    }
    hipLaunchKernelGGL(k_admit_scatter, dim3(64, (unsigned)n), dim3(256), 0, h->stream, h->stageSlots.p, n, h->stage.p,
                       h->stage.p + n * h->kvp, h->quietStage.p, h->kvp, h->ksp, nsteps, h->maxSteps, h->b.p,
                       h->ftab.p, h->qv.p, h->qs.p, h->quiet.p);
    SHAW_HIP(hipGetLastError());
    for (uint64_t j = 0; j < n; ++j) {
      const uint64_t s = slotOf[j];
      h->hStep[s] = 0;
      h->hHorizon[s] = horizons[j];
      h->hEvery[s] = h->R ? every : 0;
      h->used[s] = 1;
      slots[j] = s;
    }
    h->uploadSlotTables();
}
}  // namespace
}  // namespace shaw

extern "C" {

int shaw_rom_admit(shaw_rom_t h, uint64_t n, const double* b, uint64_t nsteps, const double* ftab, const uint64_t* horizons,
                   uint64_t every, uint64_t* slots) {
  return guarded([&] {
    SHAW_REQUIRE(ftab, SHAW_ERR_ARG, "null argument");
    shaw::admit(h, n, b, nsteps, ftab, nullptr, nullptr, nullptr, horizons, every, slots);
  });
}

int shaw_rom_admit_signals(shaw_rom_t h, uint64_t n, const double* b, const int* kinds, const double* delays,
                           const double* periods, uint64_t nsteps, const uint64_t* horizons, uint64_t every,
                           uint64_t* slots) {
  return guarded([&] {
    SHAW_REQUIRE(kinds && delays && periods, SHAW_ERR_ARG, "null argument");
    shaw::admit(h, n, b, nsteps, nullptr, kinds, delays, periods, horizons, every, slots);
  });
}

int shaw_rom_set_macro_steps(shaw_rom_t h, uint32_t m) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    SHAW_REQUIRE(m == 0 || h->precision != SHAW_ROM_FP64, SHAW_ERR_ARG, "macro steps run in the BF16 kernels");
    SHAW_REQUIRE(m == 0 || (m >= 2 && m <= 64), SHAW_ERR_ARG, "macro steps: m must be 0 or 2..64");
    h->settle();
    if (m && m != h->macroBuilt) {
      std::vector<double> g[4];
      macroOperator(h->hArV.data(), h->hArS.data(), h->kv, h->ks, h->dt, m, g[0], g[1], g[2], g[3]);
      const int parts = h->precision == SHAW_ROM_BF16X4 ? 3 : 2;
      const uint64_t rows[4] = {h->kv, h->kv, h->ks, h->ks}, cols[4] = {h->kv, h->ks, h->kv, h->ks};
      const uint32_t rowsP[4] = {h->kvp, h->kvp, h->ksp, h->ksp}, colsP[4] = {h->kvp, h->ksp, h->kvp, h->ksp};
      DevBuf<double>* dst[4] = {&h->gvv, &h->gvs, &h->gsv, &h->gss};
      for (int q = 0; q < 4; ++q) {
        const std::vector<double> p = packBf16(g[q].data(), rows[q], cols[q], rowsP[q], colsP[q], parts);
        dst[q]->alloc(p.size());
        dst[q]->upload(p.data(), p.size());
      }
      h->macroBuilt = m;
    }
    h->macro = m;
  });
}

int shaw_rom_set_time_budget(shaw_rom_t h, uint32_t w) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    SHAW_REQUIRE(w <= 64, SHAW_ERR_ARG, "time budget: a macro step's cost must be 0..64 one-steps");
    h->settle();
    h->timeBudget = w;
  });
}

int shaw_rom_steps(shaw_rom_t h, uint64_t n, const uint64_t* slots, uint64_t* steps) {
  return guarded([&] {
    SHAW_REQUIRE(h && (n == 0 || (slots && steps)), SHAW_ERR_ARG, "null argument");
    for (uint64_t j = 0; j < n; ++j)
      SHAW_REQUIRE(slots[j] < h->maxCols && h->used[slots[j]], SHAW_ERR_ARG, "bad slot");
    h->settle();
    for (uint64_t j = 0; j < n; ++j) steps[j] = h->hStep[slots[j]];
  });
}

int shaw_rom_advance(shaw_rom_t h, uint64_t nsteps) {
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    if (nsteps) h->launch(nsteps);
  });
}

int shaw_rom_sync(shaw_rom_t h) {
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    h->settle();
  });
}

int shaw_rom_state(shaw_rom_t h, uint64_t slot, double* qv, double* qs, uint64_t* step) {
  return guarded([&] {
    SHAW_REQUIRE(h && slot < h->maxCols && h->used[slot], SHAW_ERR_ARG, "bad slot");
    h->settle();
    if (qv) SHAW_HIP(hipMemcpy(qv, h->qv.p + slot * h->kvp, h->kv * 8, hipMemcpyDeviceToHost));
    if (qs) SHAW_HIP(hipMemcpy(qs, h->qs.p + slot * h->ksp, h->ks * 8, hipMemcpyDeviceToHost));
    if (step) *step = h->hStep[slot];
  });
}

int shaw_rom_retire_batch(shaw_rom_t h, uint64_t n, const uint64_t* slots, uint64_t max_samples, double* seis,
                          uint64_t* samples) {
  return guarded([&] {
    SHAW_REQUIRE(h && (n == 0 || slots), SHAW_ERR_ARG, "null argument");
    for (uint64_t j = 0; j < n; ++j)
      SHAW_REQUIRE(slots[j] < h->maxCols && h->used[slots[j]], SHAW_ERR_ARG, "bad slot");
    h->settle();
    if (seis && h->R && max_samples) {
      // chunks of slots through a pinned buffer of at most 256 MB: async
      // copies of a chunk, one sync, then a threaded host copy
      const uint64_t per = h->R * max_samples;
      const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(n, (256ull << 20) / (per * sizeof(double))));
      double* st = h->pinnedStage(chunk * per * sizeof(double));
      for (uint64_t j0 = 0; j0 < n; j0 += chunk) {
        const uint64_t j1 = std::min(n, j0 + chunk);
        for (uint64_t j = j0; j < j1; ++j) {
          const uint64_t s = slots[j];
          const uint64_t T = std::min<uint64_t>(h->hEvery[s] ? h->hStep[s] / h->hEvery[s] : 0, max_samples);
          if (T)
            SHAW_HIP(hipMemcpyAsync(st + (j - j0) * per, h->seis.p + s * h->R * h->maxSamples, h->R * T * 8,
                                    hipMemcpyDeviceToHost, h->stream));
        }
        SHAW_HIP(hipStreamSynchronize(h->stream));
        shaw::parallelCopy(seis + j0 * per, st, (j1 - j0) * per * sizeof(double));
      }
    }
    for (uint64_t j = 0; j < n; ++j) {
      const uint64_t s = slots[j];
      if (samples) samples[j] = h->hEvery[s] ? std::min<uint64_t>(h->hStep[s] / h->hEvery[s], max_samples) : 0;
      h->used[s] = 0;
      h->hHorizon[s] = 0;
      h->hStep[s] = 0;
      h->hEvery[s] = 0;
    }
    h->uploadSlotTables();
  });
}

int shaw_rom_retire(shaw_rom_t h, uint64_t slot, double* seis, double* qv, double* qs, uint64_t* samples) {
  return guarded([&] {
    SHAW_REQUIRE(h && slot < h->maxCols && h->used[slot], SHAW_ERR_ARG, "bad slot");
    h->settle();
    const uint64_t T = h->hEvery[slot] ? h->hStep[slot] / h->hEvery[slot] : 0;
    if (seis && T && h->R)
      SHAW_HIP(hipMemcpy(seis, h->seis.p + slot * h->R * h->maxSamples, h->R * T * 8, hipMemcpyDeviceToHost));
    if (qv) SHAW_HIP(hipMemcpy(qv, h->qv.p + slot * h->kvp, h->kv * 8, hipMemcpyDeviceToHost));
    if (qs) SHAW_HIP(hipMemcpy(qs, h->qs.p + slot * h->ksp, h->ks * 8, hipMemcpyDeviceToHost));
    if (samples) *samples = T;
    h->used[slot] = 0;
    h->hHorizon[slot] = 0;
    h->hStep[slot] = 0;
    h->hEvery[slot] = 0;
    h->horizon.upload(&h->hHorizon[slot], 1, slot);
    h->stepCol.upload(&h->hStep[slot], 1, slot);
    h->every.upload(&h->hEvery[slot], 1, slot);
  });
}

int shaw_rom_info(shaw_rom_t h, uint64_t* kv, uint64_t* ks, uint64_t* r, uint64_t* maxCols, uint64_t* maxSamples) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    if (kv) *kv = h->kv;
    if (ks) *ks = h->ks;
    if (r) *r = h->R;
    if (maxCols) *maxCols = h->maxCols;
    if (maxSamples) *maxSamples = h->maxSamples;
  });
}

int shaw_rom_query(shaw_rom_t h, int* idle) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(h && idle, SHAW_ERR_ARG, "null argument");
    const hipError_t e = hipStreamQuery(h->stream);
    if (e == hipErrorNotReady) {
      *idle = 0;
      return;
    }
    SHAW_HIP(e);
    h->settle();  // idle: reads back the steps of a time-budget launch
    *idle = 1;
  });
}

int shaw_rom_free_slots(shaw_rom_t h, uint64_t* n) {
  return guarded([&] {
    SHAW_REQUIRE(h && n, SHAW_ERR_ARG, "null argument");
    *n = (uint64_t)std::count(h->used.begin(), h->used.end(), 0);
  });
}

int shaw_rom_time(shaw_rom_t h, uint64_t nsteps, double* total) {
  return guarded([&] {
    SHAW_REQUIRE(h && total, SHAW_ERR_ARG, "null argument");
    SHAW_HIP(hipEventRecord(h->ev0, h->stream));
    h->launch(nsteps);
    SHAW_HIP(hipEventRecord(h->ev1, h->stream));
    SHAW_HIP(hipEventSynchronize(h->ev1));
    float ms = 0;
    SHAW_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *total = ms;
  });
}

void shaw_rom_destroy(shaw_rom_t h) {
  if (h) guarded([&] { delete h; });
}


int shaw_pod_basis(const double* S, uint64_t n, uint64_t ns, uint64_t k, double* phi, double* sigma) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(S && phi && n > 0 && ns > 0 && k > 0, SHAW_ERR_ARG, "bad argument");
    int ndev = 0;
    SHAW_HIP(hipGetDeviceCount(&ndev));
    SHAW_REQUIRE(ndev > 0, SHAW_ERR_HIP, "no HIP device available");
    DevBuf<double> dS;
    dS.alloc(n * ns);
    dS.upload(S, n * ns);
    podDevice(dS.p, n, ns, k, phi, sigma, nullptr);
  });
}

int shaw_pod_basis_upto(const double* S, uint64_t n, uint64_t ns, uint64_t kmax, double* phi, double* sigma,
                        uint64_t* k) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(S && phi && k && n > 0 && ns > 0 && kmax > 0, SHAW_ERR_ARG, "bad argument");
    int ndev = 0;
    SHAW_HIP(hipGetDeviceCount(&ndev));
    SHAW_REQUIRE(ndev > 0, SHAW_ERR_HIP, "no HIP device available");
    DevBuf<double> dS;
    dS.alloc(n * ns);
    dS.upload(S, n * ns);
    podDevice(dS.p, n, ns, std::min(kmax, std::min(n, ns)), phi, sigma, k);
  });
}

int shaw_reconstruct(const double* phi, uint64_t n, uint64_t k, const double* q, uint64_t ncols, double* out) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(phi && q && out && n > 0 && k > 0 && ncols > 0, SHAW_ERR_ARG, "bad argument");
    SHAW_REQUIRE(k <= 0xffffffffull && ncols <= 0xffffffffull, SHAW_ERR_ARG, "too many modes or columns");
    SHAW_REQUIRE(ncols <= 65535ull * NNC, SHAW_ERR_ARG, "at most 1048560 columns per call");
    int ndev = 0;
    SHAW_HIP(hipGetDeviceCount(&ndev));
    SHAW_REQUIRE(ndev > 0, SHAW_ERR_HIP, "no HIP device available");
    // q column-major [k x ncols] -> row-major, the layout k_gemm_nn_rows reads
    std::vector<double> w((size_t)k * ncols);
    for (uint64_t c = 0; c < ncols; ++c)
      for (uint64_t j = 0; j < k; ++j) w[j * ncols + c] = q[j + k * c];
    DevBuf<double> dPhi, dW, dOut;
    dPhi.alloc(n * k);
    dPhi.upload(phi, n * k);
    dW.alloc(w.size());
    dW.upload(w.data(), w.size());
    dOut.alloc(n * ncols);
    hipLaunchKernelGGL(k_gemm_nn_rows, dim3((unsigned)((n + 255) / 256), (unsigned)((ncols + NNC - 1) / NNC)),
                       dim3(256), 0, 0, dPhi.p, dW.p, dOut.p, nullptr, 1.0, n, (uint32_t)k, (uint32_t)ncols);
    SHAW_HIP(hipGetLastError());
    SHAW_HIP(hipMemcpy(out, dOut.p, n * ncols * 8, hipMemcpyDeviceToHost));
  });
}

int shaw_debug_symeig(int n, const double* a, double* w, double* v) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(n > 0 && a && w && v, SHAW_ERR_ARG, "bad argument");
    std::vector<double> A(a, a + (size_t)n * n), ww, VV;
    symmetricEigen(A, n, ww, VV);
    std::copy(ww.begin(), ww.end(), w);
    std::copy(VV.begin(), VV.end(), v);
  });
}

int shaw_debug_mfma(int kind, const double* a, const double* b, double* d) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(a && b && d && kind >= 0 && kind <= 2, SHAW_ERR_ARG, "bad argument");
    const size_t na = kind == 0 ? 64 : kind == 1 ? 256 : 512;
    DevBuf<double> da, db, dd;
    da.alloc(na); db.alloc(na); dd.alloc(256);
    da.upload(a, na); db.upload(b, na);
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, kind, da.p, db.p, dd.p);
    SHAW_HIP(hipGetLastError());
    SHAW_HIP(hipMemcpy(d, dd.p, 256 * 8, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
