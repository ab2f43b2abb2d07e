// FOM leapfrog on gfx950: matrix-free stencil kernels for
//   x_v += dt J_vp x_s + dt rho^-1 f      (fom_update_kernels.hpp:60-90, 133-167)
//   x_s += dt J_sp x_v                     (fom_update_kernels.hpp:92-104, 169-181)
// for m forcing columns, plus the engine behind the shaw_fom_* C ABI.
//
// HBM layout. States are stored in column blocks of MC forcing columns:
//   x[b][gid][c]  (b = column / MC, c = column % MC), gids as in the mesh
// files, so MC*sizeof(T) contiguous bytes per grid point (16 B for the
// default FP64 MC = 2, 256 B for MC = 32). Coefficients: coefVp[gid][4] (slots N, S, W, E with merged
// entries in N / W), coefSp[gid][2]. Neighbours are addressed by the
// structured index formulas of the mesh (shaw_host.hpp).
//
// Arithmetic. Every per-point update evaluates the CSR row in the
// reference's order: s = ((cN*xN + cS*xS) + cW*xW) + cE*xE; x += dt*s, with
// FMA contraction disabled for this file, so FP64 results are bit-identical
// to a no-FMA CPU evaluation of the reference's CSR spmv.
//
// Kernels.
//  * k_velocity / k_stress: the reference's two half-steps, one thread per
//    (point, CPT columns), updating in place.
//  * k_fused: one launch per leapfrog step. A block owns a strip of P-1
//    theta points (plus one east halo point) x SEG radial rows x one column
//    block, marches up the rows keeping the previous row's new velocities,
//    the old srp row below and the old stp row in registers, exchanges the
//    theta neighbours through LDS, and writes new v and s to the other
//    ping-pong buffer. Each state value is read once and written once per
//    step (vs. 3 passes of x_s and x_v traffic for the split pair); the
//    halos cost one row per segment and one point per strip row.
//  * k_fused2 / k_fusedN: two, three or four steps per launch, the
//    intermediate levels kept in registers and LDS, so the states are read
//    and written once per launch (recomputed halo rows and points).
//  * k_fusedW: three to six steps per launch when a 64-point strip is one
//    wavefront (one thread per point, all columns of a 1- or 2-column block);
//    neighbours move between lanes, and the waves of a workgroup share only
//    a coefficient ring. The default: six steps per launch.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "rom_internal.hpp"
#include "shaw_app.hpp"
#include "shaw_common.hpp"

namespace shaw {
namespace {

template <class T, int N>
struct alignas(sizeof(T) * N) Vec {
  T v[N];
};

template <class T, int N>
__device__ __forceinline__ void load(T (&dst)[N], const T* p) {
  const Vec<T, N> x = *reinterpret_cast<const Vec<T, N>*>(p);
#pragma unroll
  for (int k = 0; k < N; ++k) dst[k] = x.v[k];
}
template <class T, int N>
__device__ __forceinline__ void store(T* p, const T (&src)[N]) {
  Vec<T, N> x;
#pragma unroll
  for (int k = 0; k < N; ++k) x.v[k] = src[k];
  *reinterpret_cast<Vec<T, N>*>(p) = x;
}

template <class T>
struct StepArgs {
  const T* vIn;
  T* vOut;
  const T* sIn;
  T* sOut;
  const T* cv;          // [nVp][4]
  const T* cs;          // [nSp][2]
  const int64_t* src;   // [nb*MC] source gid per column (-1: none)
  const T* srcRhoInv;   // [nb*MC]
  const double* frow;   // forcing values of this step, [M]
  const int32_t* recv;  // [nVp] receiver index or -1; nullptr: no sample this step
  int64_t recvRow;      // the one grid row holding every receiver, or -1
  // graph launches: the step is *stepDev + stepOff (1-based), and frow, recv
  // and sample follow from it (resolveStep); stepDev == nullptr: host-resolved
  const uint64_t* stepDev;
  uint64_t stepOff;
  const double* ftabBase;
  const int32_t* recvTable;
  uint64_t every;
  double* seis;         // LayoutLeft [R x Tcap x M]
  uint64_t sample, R, Tcap;
  uint32_t nr, nth, M, MC;
  T dt;
  int rank1;
  uint32_t P, SEG, nStrips, nSegs, nb;
};

// Input rows first..last through a row function whose row sets rotate (the
// row loads the next set while it computes on the current one). The loop
// runs whole rotations unconditionally, so every set stays in its registers
// across the back edge (a conditional row there makes the compiler copy the
// sets and wait for every load in flight); the rows left over follow it.
template <class I, class RS, class Row>
__device__ __forceinline__ void rotateRows(RS& A, RS& B, Row&& row, I first, I last) {
  I r = first;
  for (; r + 1 <= last; r += 2) {
    row(A, B, r);
    row(B, A, r + 1);
  }
  if (r <= last) row(A, B, r);
}
template <class I, class RS, class Row>
__device__ __forceinline__ void rotateRows(RS& A, RS& B, RS& C, Row&& row, I first, I last) {
  I r = first;
  for (; r + 2 <= last; r += 3) {
    row(A, C, r);
    row(B, A, r + 1);
    row(C, B, r + 2);
  }
  if (r <= last) row(A, C, r);
  if (r + 1 <= last) row(B, A, r + 1);
}

template <class T>
__device__ __forceinline__ void resolveStep(StepArgs<T>& a) {
  if (a.stepDev == nullptr) return;
  const uint64_t s = *a.stepDev + a.stepOff;
  a.frow = a.ftabBase + (s - 1) * a.M;
  const bool rec = a.R && a.every && s % a.every == 0;
  a.recv = rec ? a.recvTable : nullptr;
  a.sample = rec ? s / a.every - 1 : 0;
}

__device__ __forceinline__ uint64_t srpIdx(uint64_t i, uint64_t j, uint64_t rs) { return j * rs + 2 * i; }
__device__ __forceinline__ uint64_t stpIdx(uint64_t i, uint64_t j, uint64_t rs, uint64_t nr) {
  return j + 1 < nr ? j * rs + 2 * i + 1 : (nr - 1) * rs + i;
}

// Forcing of the velocity update for CPT columns. The per-column source gid
// and increment are loop invariants, loaded once per thread: the increment
// is (dt * rhoInv) * f for the rank-1 form `v = v + dt*rhoInv*f`
// (KokkosBlas::mult) and (rhoInv * f) * dt for the rank-2 form
// `v += rhoInv*f*dt` (AddForcingRank2), the same operations in the same
// order, so v + increment is bit-identical to either form.
template <class T, int CPT>
struct ColumnForcing {
  int64_t gid[CPT];
  T inc[CPT];
  __device__ __forceinline__ void load(const StepArgs<T>& a, uint32_t colBase) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const uint32_t col = colBase + k;
      gid[k] = col < a.M ? a.src[col] : -1;
      inc[k] = T(0);
      if (gid[k] >= 0) {
        const T f = (T)a.frow[col];
        inc[k] = a.rank1 ? a.dt * a.srcRhoInv[col] * f : a.srcRhoInv[col] * f * a.dt;
      }
    }
  }
  __device__ __forceinline__ void apply(T (&vn)[CPT], uint64_t p) const {
#pragma unroll
    for (int k = 0; k < CPT; ++k)
      if (gid[k] == (int64_t)p) vn[k] = vn[k] + inc[k];
  }
};

template <class T, int CPT>
__device__ __forceinline__ void recordReceivers(const StepArgs<T>& a, const T (&vn)[CPT], uint64_t p,
                                                uint32_t colBase) {
  if (a.recv == nullptr) return;
  const int32_t r = a.recv[p];
  if (r < 0) return;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const uint32_t col = colBase + k;
    if (col < a.M) a.seis[(uint64_t)r + a.R * (a.sample + a.Tcap * col)] = (double)vn[k];
  }
}

// ------------------------------------------------------------ split pair
template <class T, int CPT>
__global__ void __launch_bounds__(256) k_velocity(StepArgs<T> a) {
  const uint32_t tpp = a.MC / CPT;
  const uint64_t nVp = (uint64_t)a.nr * a.nth, nSp = (uint64_t)(a.nr - 1) * a.nth + (uint64_t)a.nr * (a.nth - 1);
  const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t p = gt / tpp;
  if (p >= nVp) return;
  const uint32_t cb = blockIdx.y, m0 = (uint32_t)(gt % tpp) * CPT;
  const uint64_t rs = 2ull * a.nth - 1, i = p % a.nth, j = p / a.nth;
  const uint64_t iN = srpIdx(i, j < a.nr - 1 ? j : a.nr - 2, rs);
  const uint64_t iS = srpIdx(i, j > 0 ? j - 1 : 0, rs);
  const uint64_t iW = stpIdx(i > 0 ? i - 1 : 0, j, rs, a.nr);
  const uint64_t iE = stpIdx(i < a.nth - 2 ? i : a.nth - 2, j, rs, a.nr);
  const T* s = a.sIn + (uint64_t)cb * nSp * a.MC + m0;
  T* v = a.vOut + ((uint64_t)cb * nVp + p) * a.MC + m0;
  T xN[CPT], xS[CPT], xW[CPT], xE[CPT], vo[CPT], vn[CPT];
  load(xN, s + iN * a.MC);
  load(xS, s + iS * a.MC);
  load(xW, s + iW * a.MC);
  load(xE, s + iE * a.MC);
  load(vo, v);
  T c[4];
  load(c, a.cv + 4 * p);
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    T sum = c[0] * xN[k] + c[1] * xS[k];
    sum = sum + c[2] * xW[k];
    sum = sum + c[3] * xE[k];
    vn[k] = vo[k] + a.dt * sum;
  }
  ColumnForcing<T, CPT> frc;
  frc.load(a, cb * a.MC + m0);
  frc.apply(vn, p);
  store(v, vn);
  recordReceivers<T, CPT>(a, vn, p, cb * a.MC + m0);
}

template <class T, int CPT>
__global__ void __launch_bounds__(256) k_stress(StepArgs<T> a) {
  const uint32_t tpp = a.MC / CPT;
  const uint64_t nVp = (uint64_t)a.nr * a.nth, nSp = (uint64_t)(a.nr - 1) * a.nth + (uint64_t)a.nr * (a.nth - 1);
  const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t q = gt / tpp;
  if (q >= nSp) return;
  const uint32_t cb = blockIdx.y, m0 = (uint32_t)(gt % tpp) * CPT;
  const uint64_t rs = 2ull * a.nth - 1, last = (uint64_t)(a.nr - 1) * rs;
  uint64_t n0, n1;
  if (q < last) {
    const uint64_t j = q / rs, k = q % rs;
    if ((k & 1) == 0) { n0 = (j + 1) * a.nth + k / 2; n1 = j * a.nth + k / 2; }   // srp: N, S
    else { n0 = j * a.nth + (k - 1) / 2; n1 = n0 + 1; }                           // stp: W, E
  } else {
    n0 = (uint64_t)(a.nr - 1) * a.nth + (q - last);
    n1 = n0 + 1;
  }
  const T* v = a.vIn + (uint64_t)cb * nVp * a.MC + m0;
  T* s = a.sOut + ((uint64_t)cb * nSp + q) * a.MC + m0;
  T x0[CPT], x1[CPT], so[CPT];
  load(x0, v + n0 * a.MC);
  load(x1, v + n1 * a.MC);
  load(so, s);
  T c[2];
  load(c, a.cs + 2 * q);
#pragma unroll
  for (int k = 0; k < CPT; ++k) so[k] = so[k] + a.dt * (c[0] * x0[k] + c[1] * x1[k]);
  store(s, so);
}

// ------------------------------------------------------------ fused step
// Old values of one grid row r for one thread, loaded one row ahead: the
// srp (north of r) and stp (east) stress, the velocity, the J_vp row, the
// west stp (used by the strip's first point only) and the J_sp rows of the
// stress points of row r - 1, which are updated while row r is processed.
template <class T, int CPT>
struct RowIn {
  T N[CPT], E[CPT], vo[CPT], W[CPT];
  T c[4], bS[2], bT[2];
};

template <class T, int CPT, int PF>
__global__ void __launch_bounds__(512) k_fused(StepArgs<T> a) {
  resolveStep(a);
  extern __shared__ unsigned char smem_raw[];
  T* stpRow = reinterpret_cast<T*>(smem_raw);    // [P+1][MC]: old stp(i0-1 .. i0+P-1) of row r
  T* vRow = stpRow + (a.P + 1) * a.MC;            // [P][MC]: new v of row jj

  const unsigned lb = xcdLogicalBlock(blockIdx.x, gridDim.x);
  const uint32_t strip = lb % a.nStrips;
  const uint32_t seg = (lb / a.nStrips) % a.nSegs;
  const uint32_t cb = lb / (a.nStrips * a.nSegs);

  const uint32_t tpp = a.MC / CPT;
  const uint32_t pl = threadIdx.x / tpp;
  const uint32_t m0 = (threadIdx.x % tpp) * CPT;
  const uint64_t nr = a.nr, nth = a.nth, rs = 2 * nth - 1;
  const uint64_t nVp = nr * nth, nSp = (nr - 1) * nth + nr * (nth - 1);
  const uint64_t i = (uint64_t)strip * (a.P - 1) + pl;
  const bool active = i < nth;
  const bool owner = active && pl + 1 < a.P;  // east halo point is recomputed, not owned
  const uint64_t j0 = (uint64_t)seg * a.SEG;
  const uint64_t j1 = j0 + a.SEG < nr ? j0 + a.SEG : nr;
  const uint32_t colBase = cb * a.MC + m0;

  const T* vIn = a.vIn + (uint64_t)cb * nVp * a.MC + m0;
  T* vOut = a.vOut + (uint64_t)cb * nVp * a.MC + m0;
  const T* sIn = a.sIn + (uint64_t)cb * nSp * a.MC + m0;
  T* sOut = a.sOut + (uint64_t)cb * nSp * a.MC + m0;

  // Loads are issued for every lane and every row with clamped indices (a
  // lane past the grid edge, a row past the segment or the grid): no branch
  // separates a load from its use, so the loads of row r + 1 stay in flight
  // while row r is computed.
  const uint64_t ic = active ? i : nth - 1;
  const uint64_t iE = ic < nth - 2 ? ic : nth - 2;        // east stp column
  const uint64_t iW = pl == 0 ? (ic > 0 ? ic - 1 : 0) : iE;  // west stp column (strip start only)
  auto fetch = [&](RowIn<T, CPT>& d, uint64_t r) {
    const uint64_t rc = r < nr ? r : nr - 1;
    const uint64_t p = rc * nth + ic;
    load(d.N, sIn + srpIdx(ic, rc < nr - 1 ? rc : nr - 2, rs) * a.MC);
    load(d.E, sIn + stpIdx(iE, rc, rs, nr) * a.MC);
    load(d.vo, vIn + p * a.MC);
    load(d.W, sIn + stpIdx(iW, rc, rs, nr) * a.MC);
    load(d.c, a.cv + 4 * p);
    // stress row r - 1, from the unclamped r (r = nr still has stress row nr - 1)
    const uint64_t rb = r == 0 ? 0 : (r - 1 < nr ? r - 1 : nr - 1);
    load(d.bS, a.cs + 2 * srpIdx(ic, rb < nr - 1 ? rb : nr - 2, rs));
    load(d.bT, a.cs + 2 * stpIdx(iE, rb, rs, nr));
  };

  ColumnForcing<T, CPT> frc;
  frc.load(a, colBase);
  T S[CPT], vj[CPT], Eold[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) { vj[k] = T(0); Eold[k] = T(0); }
  load(S, sIn + srpIdx(ic, j0 > 0 ? (j0 - 1 < nr - 1 ? j0 - 1 : nr - 2) : 0, rs) * a.MC);

  // one row: new v of row r = jj + 1 (all strip points incl. the halo), new
  // stress of row jj (owned points); `cur` holds row r, `nxt` receives row
  // r + PF (the set of row r - 1, consumed by the previous row)
  auto row = [&](const RowIn<T, CPT>& cur, RowIn<T, CPT>& nxt, int64_t jj) {
    const uint64_t r = (uint64_t)(jj + 1);
    const bool haveRow = r < nr;
    store(stpRow + (pl + 1) * a.MC + m0, cur.E);
    if (pl == 0) store(stpRow + m0, cur.W);
    fetch(nxt, r + PF);
    __syncthreads();
    T vn[CPT];
    if (active && haveRow) {
      const uint64_t p = r * nth + i;
      T W[CPT];
      load(W, stpRow + pl * a.MC + m0);
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        T sum = cur.c[0] * cur.N[k] + cur.c[1] * S[k];
        sum = sum + cur.c[2] * W[k];
        sum = sum + cur.c[3] * cur.E[k];
        vn[k] = cur.vo[k] + a.dt * sum;
      }
      frc.apply(vn, p);
      if (owner && r < j1) {
        store(vOut + p * a.MC, vn);
        if (a.recvRow < 0 || (int64_t)r == a.recvRow) recordReceivers<T, CPT>(a, vn, p, colBase);
      }
    }
    // stress rows jj (owned points): srp(i, jj) needs v_new rows jj and jj+1,
    // stp(i, jj) needs v_new(i, jj) and v_new(i+1, jj) (from LDS)
    if (owner && jj >= (int64_t)j0) {
      const uint64_t jr = (uint64_t)jj;
      if (jr + 1 < nr) {
        T so[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) so[k] = S[k] + a.dt * (cur.bS[0] * vn[k] + cur.bS[1] * vj[k]);
        store(sOut + srpIdx(i, jr, rs) * a.MC, so);
      }
      if (i + 1 < nth) {
        T vE[CPT], so[CPT];
        load(vE, vRow + (pl + 1) * a.MC + m0);
#pragma unroll
        for (int k = 0; k < CPT; ++k) so[k] = Eold[k] + a.dt * (cur.bT[0] * vj[k] + cur.bT[1] * vE[k]);
        store(sOut + stpIdx(i, jr, rs, nr) * a.MC, so);
      }
    }
    __syncthreads();
    if (active && haveRow) {
      store(vRow + pl * a.MC + m0, vn);
#pragma unroll
      for (int k = 0; k < CPT; ++k) { vj[k] = vn[k]; S[k] = cur.N[k]; Eold[k] = cur.E[k]; }
    }
  };

  // PF + 1 row sets rotate (no register copies)
  if constexpr (PF == 1) {
    RowIn<T, CPT> A, B;
    fetch(A, j0);
    rotateRows(A, B, row, (int64_t)j0 - 1, (int64_t)j1 - 1);
  } else {
    RowIn<T, CPT> A, B, C;
    fetch(A, j0);
    fetch(B, j0 + 1);
    rotateRows(A, B, C, row, (int64_t)j0 - 1, (int64_t)j1 - 1);
  }
}

// ------------------------------------------------------------ two fused steps
// One launch advances two leapfrog steps, n -> n+1 -> n+2. Level n+1 lives
// only in registers and LDS, so the states are read and written once per two
// steps. A block owns P-3 theta points of a strip (lanes: 1 west halo point,
// P-3 owned points, 2 east halo points) and SEG radial rows. Marching up
// input row r it computes, in order:
//   A  v1(r)       = v^n(r) + dt J_vp s^n + f(step n+1)
//   B  s1(row r-1) = s^n + dt J_sp v1          (srp: rows r-1, r of v1; stp: lanes i, i+1)
//   C  v2(r-1)     = v1(r-1) + dt J_vp s1 + f(step n+2)
//   D  s2(row r-2) = s1 + dt J_sp v2
// Every value is the expression the one-step kernel evaluates, in the same
// order, so two fused steps are bit-identical to two single steps. Rows
// j0-1 .. j1+1 of level n+1 and row j1 of level n+2 are recomputed halos.
template <class T>
struct StepArgs2 {
  StepArgs<T> s1, s2;  // the two steps' arguments (forcing rows, receiver sampling)
};

template <class T, int CPT>
__global__ void __launch_bounds__(512) k_fused2(StepArgs2<T> aa) {
  resolveStep(aa.s1);
  resolveStep(aa.s2);
  const StepArgs<T>& a = aa.s1;
  extern __shared__ unsigned char smem_raw[];
  T* stp0Row = reinterpret_cast<T*>(smem_raw);    // [P+1][MC]: s^n stp of row r (lane pl at pl+1, west load at 0)
  T* v1Row = stp0Row + (a.P + 1) * a.MC;          // [P][MC]: v1 of row r-1 (then r)
  T* stp1Row = v1Row + a.P * a.MC;                // [P+1][MC]: s1 stp of row r-1 (lane pl at pl+1)
  T* v2Row = stp1Row + (a.P + 1) * a.MC;          // [P][MC]: v2 of row r-2 (then r-1)

  const unsigned lb = xcdLogicalBlock(blockIdx.x, gridDim.x);
  const uint32_t strip = lb % a.nStrips;
  const uint32_t seg = (lb / a.nStrips) % a.nSegs;
  const uint32_t cb = lb / (a.nStrips * a.nSegs);

  const uint32_t tpp = a.MC / CPT;
  const uint32_t pl = threadIdx.x / tpp;
  const uint32_t m0 = (threadIdx.x % tpp) * CPT;
  const uint64_t nr = a.nr, nth = a.nth, rs = 2 * nth - 1;
  const uint64_t nVp = nr * nth, nSp = (nr - 1) * nth + nr * (nth - 1);
  const uint32_t Q = a.P - 3;  // owned points per strip
  const int64_t ii = (int64_t)strip * Q - 1 + pl;
  const bool active = ii >= 0 && ii < (int64_t)nth;
  const uint64_t i = active ? (uint64_t)ii : 0;
  const bool owner = active && pl >= 1 && pl <= Q;
  const bool doStp1 = active && pl + 1 < a.P && i + 1 < nth;  // s1 stp(i) needs v1(i+1)
  const bool doV2 = active && pl >= 1 && pl + 1 < a.P;        // needs s1 stp(i-1) of lane pl-1
  const uint64_t j0 = (uint64_t)seg * a.SEG;
  const uint64_t j1 = j0 + a.SEG < nr ? j0 + a.SEG : nr;
  const uint32_t colBase = cb * a.MC + m0;

  const T* vIn = a.vIn + (uint64_t)cb * nVp * a.MC + m0;
  T* vOut = a.vOut + (uint64_t)cb * nVp * a.MC + m0;
  const T* sIn = a.sIn + (uint64_t)cb * nSp * a.MC + m0;
  T* sOut = a.sOut + (uint64_t)cb * nSp * a.MC + m0;

  // clamped, branch-free loads as in the one-step kernel
  const uint64_t ic = active ? i : (ii < 0 ? 0 : nth - 1);
  const uint64_t iE = ic < nth - 2 ? ic : nth - 2;
  const uint64_t iW = pl == 0 ? (ic > 0 ? ic - 1 : 0) : iE;
  auto fetch = [&](RowIn<T, CPT>& d, uint64_t r) {
    const uint64_t rc = r < nr ? r : nr - 1;
    const uint64_t p = rc * nth + ic;
    load(d.N, sIn + srpIdx(ic, rc < nr - 1 ? rc : nr - 2, rs) * a.MC);
    load(d.E, sIn + stpIdx(iE, rc, rs, nr) * a.MC);
    load(d.vo, vIn + p * a.MC);
    load(d.W, sIn + stpIdx(iW, rc, rs, nr) * a.MC);
    load(d.c, a.cv + 4 * p);
    const uint64_t rb = r == 0 ? 0 : (r - 1 < nr ? r - 1 : nr - 1);
    load(d.bS, a.cs + 2 * srpIdx(ic, rb < nr - 1 ? rb : nr - 2, rs));
    load(d.bT, a.cs + 2 * stpIdx(iE, rb, rs, nr));
  };

  ColumnForcing<T, CPT> frc1, frc2;
  frc1.load(aa.s1, colBase);
  frc2.load(aa.s2, colBase);
  // carried values: level n at rows r-1 (S0: srp, E0: stp), v1 at row r-1,
  // level n+1 at row r-2 (S1: srp, E1: stp), v2 at row r-2, and the J_vp row
  // r-1 and J_sp rows r-2 the second step needs
  T S0[CPT], E0[CPT], vj1[CPT], S1[CPT], E1[CPT], vj2[CPT];
  T cPrev[4] = {T(0), T(0), T(0), T(0)}, bSPrev[2] = {T(0), T(0)}, bTPrev[2] = {T(0), T(0)};
#pragma unroll
  for (int k = 0; k < CPT; ++k) { E0[k] = vj1[k] = S1[k] = E1[k] = vj2[k] = T(0); }
  // the first input row is j0 - 1 (or 0): S0 is srp of the row below it
  const uint64_t rFirst = j0 > 0 ? j0 - 1 : 0;
  load(S0, sIn + srpIdx(ic, rFirst > 0 ? (rFirst - 1 < nr - 1 ? rFirst - 1 : nr - 2) : 0, rs) * a.MC);
  if (pl == 0) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) stp1Row[m0 + k] = T(0);
  }

  auto row = [&](const RowIn<T, CPT>& cur, RowIn<T, CPT>& nxt, uint64_t r) {
    const bool haveRow = r < nr;
    store(stp0Row + (pl + 1) * a.MC + m0, cur.E);
    if (pl == 0) store(stp0Row + m0, cur.W);
    fetch(nxt, r + 1);
    __syncthreads();  // (1) s^n stp of row r visible
    // A: v1(r), every lane (halo points included)
    T v1[CPT];
    if (haveRow) {
      T W[CPT];
      load(W, stp0Row + pl * a.MC + m0);
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        T sum = cur.c[0] * cur.N[k] + cur.c[1] * S0[k];
        sum = sum + cur.c[2] * W[k];
        sum = sum + cur.c[3] * cur.E[k];
        v1[k] = cur.vo[k] + a.dt * sum;
      }
      frc1.apply(v1, r * nth + i);
      if (owner && r >= j0 && r < j1 && (a.recvRow < 0 || (int64_t)r == a.recvRow))
        recordReceivers<T, CPT>(aa.s1, v1, r * nth + i, colBase);
    } else {
#pragma unroll
      for (int k = 0; k < CPT; ++k) v1[k] = T(0);
    }
    // B: s1 of row r-1 (srp needs v1 rows r-1 and r; stp needs v1(i+1, r-1) from LDS)
    T srp1[CPT], stp1[CPT];
    const bool rowB = r >= 1 && r - 1 >= rFirst && r - 1 < nr;
    const uint64_t rB = r - 1;
#pragma unroll
    for (int k = 0; k < CPT; ++k) { srp1[k] = S1[k]; stp1[k] = T(0); }
    if (rowB) {
      if (rB + 1 < nr) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) srp1[k] = S0[k] + a.dt * (cur.bS[0] * v1[k] + cur.bS[1] * vj1[k]);
      }
      if (doStp1) {
        T vE[CPT];
        load(vE, v1Row + (pl + 1) * a.MC + m0);
#pragma unroll
        for (int k = 0; k < CPT; ++k) stp1[k] = E0[k] + a.dt * (cur.bT[0] * vj1[k] + cur.bT[1] * vE[k]);
      }
    }
    store(stp1Row + (pl + 1) * a.MC + m0, stp1);
    __syncthreads();  // (2) s1 stp of row r-1 visible
    // C: v2(r-1), lanes with both stp neighbours
    T v2[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) v2[k] = T(0);
    const bool rowC = r >= 1 && r - 1 >= j0 && r - 1 <= j1 && r - 1 < nr;
    if (rowC && doV2) {
      const uint64_t rc = r - 1;
      T W[CPT];
      load(W, stp1Row + pl * a.MC + m0);
      // north srp of row nr-1 is srp(nr-2) (the merged surface slot), which is S1
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const T Nk = rc + 1 < nr ? srp1[k] : S1[k];
        T sum = cPrev[0] * Nk + cPrev[1] * S1[k];
        sum = sum + cPrev[2] * W[k];
        sum = sum + cPrev[3] * stp1[k];
        v2[k] = vj1[k] + a.dt * sum;
      }
      frc2.apply(v2, rc * nth + i);
      if (owner && rc < j1) {
        store(vOut + (rc * nth + i) * a.MC, v2);
        if (a.recvRow < 0 || (int64_t)rc == a.recvRow) recordReceivers<T, CPT>(aa.s2, v2, rc * nth + i, colBase);
      }
    }
    // D: s2 of row r-2, owned points (srp needs v2 rows r-2, r-1; stp needs v2(i+1, r-2) from LDS)
    if (owner && r >= 2 && r - 2 >= j0 && r - 2 < j1) {
      const uint64_t rd = r - 2;
      if (rd + 1 < nr) {
        T so[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) so[k] = S1[k] + a.dt * (bSPrev[0] * v2[k] + bSPrev[1] * vj2[k]);
        store(sOut + srpIdx(i, rd, rs) * a.MC, so);
      }
      if (i + 1 < nth) {
        T vE[CPT], so[CPT];
        load(vE, v2Row + (pl + 1) * a.MC + m0);
#pragma unroll
        for (int k = 0; k < CPT; ++k) so[k] = E1[k] + a.dt * (bTPrev[0] * vj2[k] + bTPrev[1] * vE[k]);
        store(sOut + stpIdx(i, rd, rs, nr) * a.MC, so);
      }
    }
    __syncthreads();  // (3) readers of v1Row / v2Row done
    store(v1Row + pl * a.MC + m0, v1);
    store(v2Row + pl * a.MC + m0, v2);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      vj1[k] = v1[k]; S0[k] = cur.N[k]; E0[k] = cur.E[k];
      vj2[k] = v2[k]; S1[k] = srp1[k]; E1[k] = stp1[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) cPrev[k] = cur.c[k];
    bSPrev[0] = cur.bS[0]; bSPrev[1] = cur.bS[1];
    bTPrev[0] = cur.bT[0]; bTPrev[1] = cur.bT[1];
  };

  // input rows rFirst .. j1 + 1 (D of row j1 - 1 happens at r = j1 + 1)
  const uint64_t rLast = j1 + 1;
  RowIn<T, CPT> A, B;
  fetch(A, rFirst);
  rotateRows(A, B, row, rFirst, rLast);
}

// ------------------------------------------------------------ NS fused steps
// One launch advances NS leapfrog steps (NS = 3 or 4), n -> n+NS; the
// intermediate levels live only in registers and LDS, so the states are read
// and written once per NS steps. A block owns Q = P - 2 NS + 1 theta points
// of a strip (lanes: NS-1 west halo points, Q owned, NS east halo points)
// and SEG radial rows. Marching up input row r, level l = 1..NS computes
//   v_l(r-l+1) = v_{l-1}(r-l+1) + dt J_vp s_{l-1} + f(step n+l)
//   s_l(r-l)   = s_{l-1}(r-l)   + dt J_sp v_l
// each as the one-step kernel evaluates it, in the same order, so NS fused
// steps are bit-identical to NS single steps. Level l is valid on lanes
// l-1 .. P-l (velocity, srp) and l-1 .. P-l-1 (stp), and from row
// j0-(NS-l) (recomputed halo rows below the segment). The coefficient rows
// of the last NS input rows sit in an LDS ring (loaded once per grid point
// rather than once per column thread: they cost no registers), and the stp
// rows that consecutive levels exchange alternate between two LDS buffers.
template <class T, int NS>
struct StepArgsN {
  StepArgs<T> s[NS];  // the steps' arguments (forcing rows, receiver sampling)
};

// ColumnForcing for NS steps. The source of a column is hit when the lane's
// grid column is the source's and the row matches: a 32-bit row compare per
// level instead of a 64-bit gid compare. One increment per step, each the
// one-step kernel's expression.
template <class T, int CPT, int NS>
struct ColumnForcingN {
  uint32_t row[CPT];  // the source's grid row if it lies in this lane's grid column, else ~0u
  T inc[NS][CPT];
  bool any;           // a lane of the wave holds a source (wave-uniform)
  __device__ __forceinline__ void load(const StepArgsN<T, NS>& aa, uint32_t colBase, uint32_t i, bool active) {
    const StepArgs<T>& a = aa.s[0];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const uint32_t col = colBase + k;
      const int64_t g = col < a.M ? a.src[col] : -1;
      row[k] = ~0u;
#pragma unroll
      for (int l = 0; l < NS; ++l) inc[l][k] = T(0);
      if (g >= 0) {
        if (active && (uint32_t)(g % a.nth) == i) row[k] = (uint32_t)(g / a.nth);
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const T f = (T)aa.s[l].frow[col];
          inc[l][k] = a.rank1 ? a.dt * a.srcRhoInv[col] * f : a.srcRhoInv[col] * f * a.dt;
        }
      }
    }
    bool hit = false;
#pragma unroll
    for (int k = 0; k < CPT; ++k) hit = hit || row[k] != ~0u;
    any = __ballot(hit) != 0;
  }
  __device__ __forceinline__ void apply(int l, T (&vn)[CPT], uint32_t r) const {
    if (!any) return;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const T added = vn[k] + inc[l][k];
      vn[k] = row[k] == r ? added : vn[k];
    }
  }
};

// old state of one input row for one thread (coefficients go to the LDS ring)
template <class T, int CPT>
struct RowState {
  T N[CPT], E[CPT], vo[CPT], W[CPT];
};

// Addressing: every row index is block-uniform and every offset within a
// column block fits 32 bits (the host checks nSp*MC), so a load address is a
// scalar row offset plus a per-lane offset fixed for the whole launch.
template <class T, int CPT, int NS>
__global__ void __launch_bounds__(512) k_fusedN(StepArgsN<T, NS> aa) {
#pragma unroll
  for (int l = 0; l < NS; ++l) resolveStep(aa.s[l]);
  const StepArgs<T>& a = aa.s[0];
  const uint32_t P = a.P, MC = a.MC;
  extern __shared__ unsigned char smem_raw[];
  // the ring first: its 8-value entries stay aligned for any MC
  T* ring = reinterpret_cast<T*>(smem_raw);  // [NS][P][8]: J_vp row rho, J_sp rows rho-1 of input row rho
  T* stpEven = ring + NS * P * 8;            // [P+1][MC]: stp of levels 0, 2 (lane pl at pl+1; level 0's west load at 0)
  T* stpOdd = stpEven + (P + 1) * MC;        // [P+1][MC]: stp of levels 1, 3
  T* vRow = stpOdd + (P + 1) * MC;           // [NS][P][MC]: v_l of the previous input row

  const unsigned lb = xcdLogicalBlock(blockIdx.x, gridDim.x);
  const uint32_t strip = lb % a.nStrips;
  const uint32_t seg = (lb / a.nStrips) % a.nSegs;
  const uint32_t cb = lb / (a.nStrips * a.nSegs);

  const uint32_t tpp = MC / CPT;
  const uint32_t pl = threadIdx.x / tpp;
  const uint32_t m0 = (threadIdx.x % tpp) * CPT;
  const uint32_t nr = a.nr, nth = a.nth, rs = 2 * nth - 1;
  const uint32_t Q = P - (2 * NS - 1);  // owned points per strip
  const int32_t ii = (int32_t)(strip * Q) - (NS - 1) + (int32_t)pl;
  const bool active = ii >= 0 && ii < (int32_t)nth;
  const uint32_t i = active ? (uint32_t)ii : 0;
  const bool owner = active && pl >= NS - 1 && pl < NS - 1 + Q;
  const uint32_t j0 = seg * a.SEG;
  const uint32_t j1 = j0 + a.SEG < nr ? j0 + a.SEG : nr;

  const uint64_t nVp = (uint64_t)nr * nth, nSp = (uint64_t)(nr - 1) * nth + (uint64_t)nr * (nth - 1);
  const T* vIn = a.vIn + (uint64_t)cb * nVp * MC;
  T* vOut = a.vOut + (uint64_t)cb * nVp * MC;
  const T* sIn = a.sIn + (uint64_t)cb * nSp * MC;
  T* sOut = a.sOut + (uint64_t)cb * nSp * MC;

  // per-lane offsets (clamped, so every lane loads in bounds and no branch
  // separates a load from its use); a v row is nth*MC values, a stress row
  // rs*MC, and the surface stress row nr-1 holds stp only
  const uint32_t ic = active ? i : (ii < 0 ? 0 : nth - 1);
  const uint32_t iE = ic < nth - 2 ? ic : nth - 2;        // east stp column
  const uint32_t iW = pl == 0 ? (ic > 0 ? ic - 1 : 0) : iE;  // west stp column (strip start only)
  const uint32_t vStride = nth * MC, sStride = rs * MC, topS = (nr - 1) * sStride;
  const uint32_t oV = ic * MC + m0, oSrp = 2 * ic * MC + m0;
  const uint32_t oStpE = (2 * iE + 1) * MC + m0, oStpETop = iE * MC + m0;
  const uint32_t oStpW = (2 * iW + 1) * MC + m0, oStpWTop = iW * MC + m0;
  const uint32_t cvStride = 4 * nth, csStride = 2 * rs, csTop = (nr - 1) * csStride;
  const uint32_t cV = 4 * ic, cSrp = 4 * ic, cStp = 2 * (2 * iE + 1), cStpTop = 2 * iE;

  T cf[8];  // coefficients of the next input row, on their way to the ring
  auto fetch = [&](RowState<T, CPT>& d, uint32_t r) {
    const uint32_t rc = r < nr ? r : nr - 1;
    const uint32_t rn = rc < nr - 1 ? rc : nr - 2;  // north srp row (the surface's is row nr-2)
    const bool top = rc + 1 >= nr;
    load(d.N, sIn + (rn * sStride + oSrp));
    load(d.E, sIn + (top ? topS + oStpETop : rc * sStride + oStpE));
    load(d.vo, vIn + (rc * vStride + oV));
    load(d.W, sIn + (top ? topS + oStpWTop : rc * sStride + oStpW));
    T c4[4], bS[2], bT[2];
    load(c4, a.cv + (rc * cvStride + cV));
    // stress row r - 1, from the unclamped r (r = nr still has stress row nr - 1)
    const uint32_t rb = r == 0 ? 0 : (r - 1 < nr ? r - 1 : nr - 1);
    const uint32_t rbn = rb < nr - 1 ? rb : nr - 2;
    load(bS, a.cs + (rbn * csStride + cSrp));
    load(bT, a.cs + (rb + 1 >= nr ? csTop + cStpTop : rb * csStride + cStp));
    cf[0] = c4[0]; cf[1] = c4[1]; cf[2] = c4[2]; cf[3] = c4[3];
    cf[4] = bS[0]; cf[5] = bS[1]; cf[6] = bT[0]; cf[7] = bT[1];
  };

  ColumnForcingN<T, CPT, NS> frc;
  frc.load(aa, cb * MC + m0, i, active);
  // carries from the previous input row r-1: S[l], E[l] = srp, stp of level
  // l at row r-1-l (l = 0..NS-1); vj[l-1] = v_l at row r-l (l = 1..NS)
  T S[NS][CPT], E[NS][CPT], vj[NS][CPT];
#pragma unroll
  for (int l = 0; l < NS; ++l)
#pragma unroll
    for (int k = 0; k < CPT; ++k) S[l][k] = E[l][k] = vj[l][k] = T(0);
  // first row of level l: j0 - (NS - l), at least 0; the first input row is lo[1]
  uint32_t lo[NS + 1];
#pragma unroll
  for (int l = 0; l <= NS; ++l) lo[l] = j0 >= (uint32_t)(NS - l) ? j0 - (NS - l) : 0;
  const uint32_t rFirst = lo[1];
  {  // S[0] is srp of the row below the first input row
    const uint32_t rb = rFirst > 0 ? (rFirst - 1 < nr - 1 ? rFirst - 1 : nr - 2) : 0;
    load(S[0], sIn + (rb * sStride + oSrp));
  }
  if (pl == 0) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) stpOdd[m0 + k] = T(0);
  }
  // receivers, compacted so the row loop keeps few scalars live: every
  // sampling level reads the same receiver table; a mask bit per level
  const int64_t recvRow = a.recvRow;
  const uint32_t colBase = cb * MC + m0;
  const int32_t* recvTab = nullptr;
  uint32_t recMask = 0, sample[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    if (aa.s[l].recv) { recvTab = aa.s[l].recv; recMask |= 1u << l; }
    sample[l] = (uint32_t)aa.s[l].sample;
  }
  double* const seis = a.seis;
  const uint32_t R = (uint32_t)a.R, Tcap = (uint32_t)a.Tcap, M = a.M;
  auto record = [&](int l, const T (&vn)[CPT], uint32_t p) {
    const int32_t q = recvTab[p];
    if (q < 0) return;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const uint32_t col = colBase + k;
      if (col < M) seis[(uint64_t)q + (uint64_t)R * (sample[l] + (uint64_t)Tcap * col)] = (double)vn[k];
    }
  };

  auto row = [&](const RowState<T, CPT>& cur, RowState<T, CPT>& nxt, uint32_t r) {
    store(stpEven + (pl + 1) * MC + m0, cur.E);
    if (pl == 0) store(stpEven + m0, cur.W);
    const uint32_t slot = r % NS;
    if (m0 == 0) store(ring + (slot * P + pl) * 8, cf);
    fetch(nxt, r + 1);
    __syncthreads();  // level-0 stp and the coefficient row of r visible
    T vNew[NS][CPT], srpN[NS][CPT], stpN[NS][CPT];
#pragma unroll
    for (int l = 1; l <= NS; ++l) {
      const T* stpIn = ((l - 1) & 1) ? stpOdd : stpEven;
      T* stpOut = (l & 1) ? stpOdd : stpEven;
      // velocity of level l at row rv (input row rv's coefficients: ring slot (r+1-l) mod NS)
      const bool past = r + 1 >= (uint32_t)l;
      const uint32_t rv = r + 1 - l;
      const bool rowV = past && rv >= lo[l] && rv < nr;
      const uint32_t slotL = (slot + NS - (l - 1)) % NS;
      T cl[8];
      load(cl, ring + (slotL * P + pl) * 8);
      // Row conditions are block-uniform branches. Lanes outside a level's
      // valid range (velocity and srp: lanes l-1 .. P-l; stp: l-1 .. P-l-1
      // and east of the last grid point) compute too: their values are finite (every
      // load is clamped into the grid), only lanes outside the next level's
      // narrower range read them, and nothing of theirs is stored. An LDS
      // read of such a lane stays inside its array.
      T v[CPT];
#pragma unroll
      for (int k = 0; k < CPT; ++k) v[k] = T(0);
      if (rowV) {
        T W[CPT];
        load(W, stpIn + pl * MC + m0);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
          // level 1 reads the loaded row; level l > 1 the level below, this
          // row (north srp, own stp) and the previous row (south srp, v)
          const T Nk = l == 1 ? cur.N[k] : srpN[l - 1][k];
          const T Sk = S[l - 1][k];
          const T Ek = l == 1 ? cur.E[k] : stpN[l - 1][k];
          const T vo = l == 1 ? cur.vo[k] : vj[l >= 2 ? l - 2 : 0][k];
          T sum = cl[0] * Nk + cl[1] * Sk;
          sum = sum + cl[2] * W[k];
          sum = sum + cl[3] * Ek;
          v[k] = vo + a.dt * sum;
        }
        frc.apply(l - 1, v, rv);
        if (rv >= j0 && rv < j1) {
          if (((recMask >> (l - 1)) & 1u) && (recvRow < 0 || (int64_t)rv == recvRow) && owner)
            record(l - 1, v, rv * nth + i);
          if (l == NS && owner) store(vOut + (rv * vStride + oV), v);
        }
      }
      // stress of level l at row rsr = r - l (srp needs v_l rows rsr, rsr+1;
      // stp needs v_l(i+1, rsr) of the previous input row, from LDS)
      const uint32_t rsr = r - l;
      const bool rowS = r >= (uint32_t)l && rsr >= lo[l] && rsr < nr;
      if (l < NS) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) { srpN[l][k] = S[l][k]; stpN[l][k] = T(0); }
        if (rowS) {
          if (rsr + 1 < nr) {
#pragma unroll
            for (int k = 0; k < CPT; ++k) srpN[l][k] = S[l - 1][k] + a.dt * (cl[4] * v[k] + cl[5] * vj[l - 1][k]);
          }
          T vE[CPT];  // lane P-1 reads the next level's first entry: unused, in bounds (l < NS)
          load(vE, vRow + ((l - 1) * P + pl + 1) * MC + m0);
#pragma unroll
          for (int k = 0; k < CPT; ++k) stpN[l][k] = E[l - 1][k] + a.dt * (cl[6] * vj[l - 1][k] + cl[7] * vE[k]);
        }
        store(stpOut + (pl + 1) * MC + m0, stpN[l]);
        __syncthreads();  // stp of level l visible to level l + 1
      } else if (owner && rowS && rsr >= j0 && rsr < j1) {
        if (rsr + 1 < nr) {
          T so[CPT];
#pragma unroll
          for (int k = 0; k < CPT; ++k) so[k] = S[l - 1][k] + a.dt * (cl[4] * v[k] + cl[5] * vj[l - 1][k]);
          store(sOut + (rsr * sStride + oSrp), so);
        }
        if (i + 1 < nth) {  // owner and i <= nth-2: iE == i
          T vE[CPT], so[CPT];
          load(vE, vRow + ((l - 1) * P + pl + 1) * MC + m0);
#pragma unroll
          for (int k = 0; k < CPT; ++k) so[k] = E[l - 1][k] + a.dt * (cl[6] * vj[l - 1][k] + cl[7] * vE[k]);
          store(sOut + (rsr + 1 < nr ? rsr * sStride + oStpE : topS + oStpETop), so);
        }
      }
#pragma unroll
      for (int k = 0; k < CPT; ++k) vNew[l - 1][k] = v[k];
    }
    __syncthreads();  // readers of vRow and of the ring slot of row r+1-NS done
#pragma unroll
    for (int l = 0; l < NS; ++l) store(vRow + (l * P + pl) * MC + m0, vNew[l]);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      S[0][k] = cur.N[k];
      E[0][k] = cur.E[k];
    }
#pragma unroll
    for (int l = 1; l < NS; ++l)
#pragma unroll
      for (int k = 0; k < CPT; ++k) { S[l][k] = srpN[l][k]; E[l][k] = stpN[l][k]; }
#pragma unroll
    for (int l = 0; l < NS; ++l)
#pragma unroll
      for (int k = 0; k < CPT; ++k) vj[l][k] = vNew[l][k];
  };

  // input rows rFirst .. j1 + NS - 1 (level NS's stress of row j1 - 1 happens at r = j1 - 1 + NS)
  const uint32_t rLast = j1 + NS - 1;
  RowState<T, CPT> A, B;
  fetch(A, rFirst);
  rotateRows(A, B, row, rFirst, rLast);
}

// ------------------------------------------------------------ wave strips
// k_fusedW: the NS-step kernel when a strip is one wavefront (one thread per
// grid point carrying all CPT = MC columns of its column block, 64-point
// strips). Lane p of a wave is point p of the strip (NS-1 west halo points,
// Q = 65 - 2 NS owned, NS east halo points); the G waves of a workgroup
// (G = 1, 2 or 4) are consecutive column blocks of the same strip and
// segment. Theta neighbours pass between lanes with ds_bpermute, so the stp
// rows that consecutive levels exchange and the east v of the previous row
// never touch an LDS array or a barrier. The waves share only the coefficient
// ring, each loading 4/G of the four 16-byte coefficient pairs of every input
// row; with at least NS+1 ring slots one barrier per input row orders both
// the ring writes after the last reads of the slot and the reads after the
// writes.
// Every value is the one-step kernel's expression in its order, as in
// k_fusedN.

// Neighbour exchange along a one-wavefront strip without LDS: DPP whole-wave
// rotates by one lane (gfx9 family), a VALU move instead of a round trip
// through the LDS crossbar. fromWest: lane p receives lane p - 1 (lane 0
// receives lane 63, wave_ror:1); fromEast: lane p receives lane p + 1 (lane
// 63 receives lane 0, wave_rol:1).
constexpr int kDppWaveRol1 = 0x134, kDppWaveRor1 = 0x13C;
template <int CTRL>
__device__ __forceinline__ int dppMove(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dppMove(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const int lo = dppMove<CTRL>((int)(uint32_t)u);
  const int hi = dppMove<CTRL>((int)(uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}
template <int CTRL>
__device__ __forceinline__ float dppMove(float x) {
  return __int_as_float(dppMove<CTRL>(__float_as_int(x)));
}
template <class T>
__device__ __forceinline__ T fromWest(T x) { return dppMove<kDppWaveRor1>(x); }
template <class T>
__device__ __forceinline__ T fromEast(T x) { return dppMove<kDppWaveRol1>(x); }

// ColumnForcing for k_fusedW: the increments of the NS steps sit in a small
// LDS table of the wave ([NS][CPT], written once by the wave's first CPT
// lanes and read only by the same wave, whose LDS operations are ordered),
// so they hold no registers; a wave reads them only when one of its lanes
// holds a source (a wave-uniform test). Each increment is the one-step
// kernel's expression.
template <class T, int CPT, int NS>
struct ColumnForcingW {
  uint32_t row[CPT];  // the source's grid row if it lies in this lane's grid column, else ~0u
  bool any;           // a lane of the wave holds a source (wave-uniform)
  const T* inc;       // the wave's table
  __device__ __forceinline__ void load(const StepArgsN<T, NS>& aa, uint32_t colBase, uint32_t i, bool active,
                                       uint32_t lane, T* table) {
    const StepArgs<T>& a = aa.s[0];
    bool hit = false;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const uint32_t col = colBase + k;
      const int64_t g = col < a.M ? a.src[col] : -1;
      row[k] = ~0u;
      if (g >= 0 && active && (uint32_t)(g % a.nth) == i) row[k] = (uint32_t)(g / a.nth);
      hit = hit || row[k] != ~0u;
    }
    any = __ballot(hit) != 0;
    if (lane < CPT) {
      const uint32_t col = colBase + lane;
      const int64_t g = col < a.M ? a.src[col] : -1;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        T v = T(0);
        if (g >= 0) {
          const T f = (T)aa.s[l].frow[col];
          v = a.rank1 ? a.dt * a.srcRhoInv[col] * f : a.srcRhoInv[col] * f * a.dt;
        }
        table[l * CPT + lane] = v;
      }
    }
    inc = table;
  }
  __device__ __forceinline__ void apply(int l, T (&vn)[CPT], uint32_t r) const {
    if (!any) return;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const T added = vn[k] + inc[l * CPT + k];
      vn[k] = row[k] == r ? added : vn[k];
    }
  }
};

// old state of one input row for one wave-strip thread, with the NP
// coefficient pairs of that row its wave loads
template <class T, int CPT, int NP>
struct RowStateW {
  T N[CPT], E[CPT], vo[CPT], W[CPT], c[NP][2];
};

// PF: input rows loaded ahead (1 or 2); PF+1 row states rotate. G: waves
// (column blocks) per workgroup.
template <class T, int CPT, int NS, int PF, int G>
__global__ void __launch_bounds__(256) k_fusedW(StepArgsN<T, NS> aa) {
#pragma unroll
  for (int l = 0; l < NS; ++l) resolveStep(aa.s[l]);
  const StepArgs<T>& a = aa.s[0];
  // RS >= NS+1 coefficient ring slots, a power of two (slot = row & (RS-1))
  constexpr uint32_t P = 64, RS = NS + 1 <= 4 ? 4 : 8, MC = CPT;
  constexpr int NP = 4 / G;  // coefficient pairs a wave loads per row
  extern __shared__ unsigned char smem_raw[];
  // [RS][4][64] pairs: pair q of point p of input row rho (slot rho mod RS)
  // is (J_vp row rho)[2q, 2q+1] for q < 2, the srp J_sp pair of stress row
  // rho-1 for q = 2, its stp pair for q = 3
  T* ring = reinterpret_cast<T*>(smem_raw);

  const unsigned lb = xcdLogicalBlock(blockIdx.x, gridDim.x);
  const uint32_t strip = lb % a.nStrips;
  const uint32_t seg = (lb / a.nStrips) % a.nSegs;
  const uint32_t grp = lb / (a.nStrips * a.nSegs);
  // the wave's index, a wave-uniform (scalar) value
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / P)), pl = threadIdx.x % P;
  const uint32_t cbRaw = grp * G + w;
  const bool live = cbRaw < a.nb;  // waves past the last column block load and compute, store nothing
  const uint32_t cb = live ? cbRaw : a.nb - 1;

  const uint32_t nr = a.nr, nth = a.nth, rs = 2 * nth - 1;
  const uint32_t Q = P - (2 * NS - 1);
  const int32_t ii = (int32_t)(strip * Q) - (NS - 1) + (int32_t)pl;
  const bool active = ii >= 0 && ii < (int32_t)nth;
  const uint32_t i = active ? (uint32_t)ii : 0;
  const bool owner = live && active && pl >= NS - 1 && pl < NS - 1 + Q;
  const uint32_t j0 = seg * a.SEG;
  const uint32_t j1 = j0 + a.SEG < nr ? j0 + a.SEG : nr;

  const uint64_t nVp = (uint64_t)nr * nth, nSp = (uint64_t)(nr - 1) * nth + (uint64_t)nr * (nth - 1);
  const T* vIn = a.vIn + (uint64_t)cb * nVp * MC;
  T* vOut = a.vOut + (uint64_t)cb * nVp * MC;
  const T* sIn = a.sIn + (uint64_t)cb * nSp * MC;
  T* sOut = a.sOut + (uint64_t)cb * nSp * MC;

  // per-lane offsets, clamped into the grid (as in k_fusedN, m0 = 0)
  const uint32_t ic = active ? i : (ii < 0 ? 0 : nth - 1);
  const uint32_t iE = ic < nth - 2 ? ic : nth - 2;
  const uint32_t iW = pl == 0 ? (ic > 0 ? ic - 1 : 0) : iE;
  const uint32_t vStride = nth * MC, sStride = rs * MC, topS = (nr - 1) * sStride;
  const uint32_t oV = ic * MC, oSrp = 2 * ic * MC;
  const uint32_t oStpE = (2 * iE + 1) * MC, oStpETop = iE * MC;
  const uint32_t oStpW = (2 * iW + 1) * MC, oStpWTop = iW * MC;
  // the coefficient pairs q = w + jG this wave loads (J_vp halves q = 0, 1,
  // the srp J_sp pair q = 2, the stp J_sp pair q = 3): base, row stride and
  // lane offset chosen once
  uint32_t cq[NP], cStr[NP], cOff[NP];
  const T* cBase[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint32_t q = w + j * G;
    cq[j] = q;
    cBase[j] = q < 2 ? a.cv : a.cs;
    cStr[j] = q < 2 ? 4 * nth : 2 * rs;
    cOff[j] = q < 2 ? 4 * ic + 2 * q : (q == 2 ? 4 * ic : 2 * (2 * iE + 1));
  }
  const uint32_t cTop = (nr - 1) * 2 * rs + 2 * iE;  // the surface stp row's pair (q = 3)

  auto fetch = [&](RowStateW<T, CPT, NP>& d, uint32_t r) {
    const uint32_t rc = r < nr ? r : nr - 1;
    const uint32_t rn = rc < nr - 1 ? rc : nr - 2;
    const bool top = rc + 1 >= nr;
    load(d.N, sIn + (rn * sStride + oSrp));
    load(d.E, sIn + (top ? topS + oStpETop : rc * sStride + oStpE));
    load(d.vo, vIn + (rc * vStride + oV));
    load(d.W, sIn + (top ? topS + oStpWTop : rc * sStride + oStpW));
    const uint32_t rb = r == 0 ? 0 : (r - 1 < nr ? r - 1 : nr - 1);
    const uint32_t rbn = rb < nr - 1 ? rb : nr - 2;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const uint32_t rowC = cq[j] < 2 ? rc : (cq[j] == 2 ? rbn : rb);
      const bool topC = cq[j] == 3 && rb + 1 >= nr;
      load(d.c[j], cBase[j] + (topC ? cTop : rowC * cStr[j] + cOff[j]));
    }
  };

  ColumnForcingW<T, CPT, NS> frc;  // increment table after the ring
  frc.load(aa, cb * MC, i, active, pl, ring + RS * 4 * P * 2 + w * NS * CPT);
  T S[NS][CPT], E[NS][CPT], vj[NS][CPT];
#pragma unroll
  for (int l = 0; l < NS; ++l)
#pragma unroll
    for (int k = 0; k < CPT; ++k) S[l][k] = E[l][k] = vj[l][k] = T(0);
  uint32_t lo[NS + 1];
#pragma unroll
  for (int l = 0; l <= NS; ++l) lo[l] = j0 >= (uint32_t)(NS - l) ? j0 - (NS - l) : 0;
  const uint32_t rFirst = lo[1];
  {
    const uint32_t rb = rFirst > 0 ? (rFirst - 1 < nr - 1 ? rFirst - 1 : nr - 2) : 0;
    load(S[0], sIn + (rb * sStride + oSrp));
  }
  const int64_t recvRow = a.recvRow;
  const uint32_t colBase = cb * MC;
  const int32_t* recvTab = nullptr;
  uint32_t recMask = 0, sample[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    if (aa.s[l].recv) { recvTab = aa.s[l].recv; recMask |= 1u << l; }
    sample[l] = (uint32_t)aa.s[l].sample;
  }
  double* const seis = a.seis;
  const uint32_t R = (uint32_t)a.R, Tcap = (uint32_t)a.Tcap, M = a.M;
  auto record = [&](int l, const T (&vn)[CPT], uint32_t p) {
    const int32_t q = recvTab[p];
    if (q < 0) return;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const uint32_t col = colBase + k;
      if (col < M) seis[(uint64_t)q + (uint64_t)R * (sample[l] + (uint64_t)Tcap * col)] = (double)vn[k];
    }
  };

  // ST: a steady row (j0 + NS <= r < j1): every level's velocity and stress
  // rows are valid, inside the segment and below the surface, so none of the
  // row conditions is evaluated
  auto row = [&](auto steady, const RowStateW<T, CPT, NP>& cur, RowStateW<T, CPT, NP>& nxt, uint32_t r) {
    constexpr bool ST = decltype(steady)::value;
    const uint32_t slot = r & (RS - 1);
#pragma unroll
    for (int j = 0; j < NP; ++j) store(ring + ((slot * 4 + cq[j]) * P + pl) * 2, cur.c[j]);
    // the next row's loads are issued at raised wave priority, so the
    // arbiter lets them out ahead of the other wave's arithmetic (3-4 %
    // faster at 1024 x 4096)
    __builtin_amdgcn_s_setprio(2);
    fetch(nxt, r + PF);
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();  // the coefficient row of r visible; every read of the slot it replaced done
    T srpN[NS][CPT], stpN[NS][CPT], vNew[NS][CPT];
#pragma unroll
    for (int l = 1; l <= NS; ++l) {
      const bool past = r + 1 >= (uint32_t)l;
      const uint32_t rv = r + 1 - l;
      const bool rowV = ST || (past && rv >= lo[l] && rv < nr);
      const uint32_t slotL = rv & (RS - 1);  // (r + 1 - l) mod RS, also when it wraps
      T cl[8];
      {
        const T* e = ring + (slotL * 4 * P + pl) * 2;
        T c0[2], c1[2], c2[2], c3[2];
        load(c0, e);
        load(c1, e + 2 * P);
        load(c2, e + 4 * P);
        load(c3, e + 6 * P);
        cl[0] = c0[0]; cl[1] = c0[1]; cl[2] = c1[0]; cl[3] = c1[1];
        cl[4] = c2[0]; cl[5] = c2[1]; cl[6] = c3[0]; cl[7] = c3[1];
      }
      // Row conditions are block-uniform, so every lane takes part in the
      // lane reads. Lanes outside a level's valid range compute finite values
      // that only lanes outside the next level's range read.
      T v[CPT];
#pragma unroll
      for (int k = 0; k < CPT; ++k) v[k] = T(0);
      if (rowV) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
          // west stp of level l-1: the west lane's (level 1: lane 0 loaded its own)
          const T Ek = l == 1 ? cur.E[k] : stpN[l - 1][k];
          const T fromW = fromWest(Ek);
          const T Wk = (l == 1 && pl == 0) ? cur.W[k] : fromW;
          const T Nk = l == 1 ? cur.N[k] : srpN[l - 1][k];
          const T Sk = S[l - 1][k];
          const T vo = l == 1 ? cur.vo[k] : vj[l >= 2 ? l - 2 : 0][k];
          T sum = cl[0] * Nk + cl[1] * Sk;
          sum = sum + cl[2] * Wk;
          sum = sum + cl[3] * Ek;
          v[k] = vo + a.dt * sum;
        }
        frc.apply(l - 1, v, rv);
        if (ST || (rv >= j0 && rv < j1)) {
          if (((recMask >> (l - 1)) & 1u) && (recvRow < 0 || (int64_t)rv == recvRow) && owner)
            record(l - 1, v, rv * nth + i);
          if (l == NS && owner) store(vOut + (rv * vStride + oV), v);
        }
      }
      const uint32_t rsr = r - l;
      const bool rowS = ST || (r >= (uint32_t)l && rsr >= lo[l] && rsr < nr);
      if (l < NS) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) { srpN[l][k] = S[l][k]; stpN[l][k] = T(0); }
        if (rowS) {
          if (ST || rsr + 1 < nr) {
#pragma unroll
            for (int k = 0; k < CPT; ++k) srpN[l][k] = S[l - 1][k] + a.dt * (cl[4] * v[k] + cl[5] * vj[l - 1][k]);
          }
#pragma unroll
          for (int k = 0; k < CPT; ++k) {
            const T vE = fromEast(vj[l - 1][k]);  // v_l of the east point, previous input row
            stpN[l][k] = E[l - 1][k] + a.dt * (cl[6] * vj[l - 1][k] + cl[7] * vE);
          }
        }
      } else if (ST || (rowS && rsr >= j0 && rsr < j1)) {
        T vE[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) vE[k] = fromEast(vj[l - 1][k]);
        if (owner) {
          if (ST || rsr + 1 < nr) {
            T so[CPT];
#pragma unroll
            for (int k = 0; k < CPT; ++k) so[k] = S[l - 1][k] + a.dt * (cl[4] * v[k] + cl[5] * vj[l - 1][k]);
            store(sOut + (rsr * sStride + oSrp), so);
          }
          if (i + 1 < nth) {
            T so[CPT];
#pragma unroll
            for (int k = 0; k < CPT; ++k) so[k] = E[l - 1][k] + a.dt * (cl[6] * vj[l - 1][k] + cl[7] * vE[k]);
            store(sOut + ((ST || rsr + 1 < nr) ? rsr * sStride + oStpE : topS + oStpETop), so);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < CPT; ++k) vNew[l - 1][k] = v[k];
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      S[0][k] = cur.N[k];
      E[0][k] = cur.E[k];
    }
#pragma unroll
    for (int l = 1; l < NS; ++l)
#pragma unroll
      for (int k = 0; k < CPT; ++k) { S[l][k] = srpN[l][k]; E[l][k] = stpN[l][k]; }
#pragma unroll
    for (int l = 0; l < NS; ++l)
#pragma unroll
      for (int k = 0; k < CPT; ++k) vj[l][k] = vNew[l][k];
  };

  auto anyRow = [&](const RowStateW<T, CPT, NP>& cur, RowStateW<T, CPT, NP>& nxt, uint32_t r) {
    if (r >= j0 + NS && r < j1) row(std::true_type{}, cur, nxt, r);
    else row(std::false_type{}, cur, nxt, r);
  };
  const uint32_t rLast = j1 + NS - 1;
  if constexpr (PF == 1) {
    RowStateW<T, CPT, NP> A, B;
    fetch(A, rFirst);
    rotateRows(A, B, anyRow, rFirst, rLast);
  } else {
    RowStateW<T, CPT, NP> A, B, C;
    fetch(A, rFirst);
    fetch(B, rFirst + 1);
    rotateRows(A, B, C, anyRow, rFirst, rLast);
  }
}

// ------------------------------------------------------------ gathers
template <class T>
__global__ void k_snapshot(const T* __restrict__ x, uint64_t n, uint32_t MC, const uint32_t* __restrict__ cols,
                           uint32_t ncols, double* __restrict__ out, uint64_t sample, uint64_t Tcap) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = blockIdx.y;
  if (g >= n || c >= ncols) return;
  const uint32_t col = cols[c];
  out[g + n * (sample + Tcap * c)] = (double)x[((uint64_t)(col / MC) * n + g) * MC + col % MC];
}

// the device step counter of graph launches (one lane, a vector store)
__global__ void k_step_set(uint64_t* __restrict__ step, uint64_t value) {
  if (threadIdx.x == 0) step[threadIdx.x] = value;
}
__global__ void k_step_add(uint64_t* __restrict__ step, uint64_t n) {
  if (threadIdx.x == 0) step[threadIdx.x] = step[threadIdx.x] + n;
}

// x[r + n*c] *= w[r] for every column c (grid-stride over columns)
__global__ void k_scale_rows(double* __restrict__ x, const double* __restrict__ w, uint64_t n, uint64_t cols) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double f = w[r];
  for (uint64_t c = blockIdx.y; c < cols; c += gridDim.y) x[r + n * c] *= f;
}

__global__ void k_to_float(const double* __restrict__ x, float* __restrict__ y, uint64_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = (float)x[e];
}

// x[r + n*c] = x[r + n*c] / w[r] for every column c (grid-stride over columns)
__global__ void k_div_rows(double* __restrict__ x, const double* __restrict__ w, uint64_t n, uint64_t cols) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double f = w[r];
  for (uint64_t c = blockIdx.y; c < cols; c += gridDim.y) x[r + n * c] = x[r + n * c] / f;
}

// plain streaming copy (16 B per lane per iteration), for bandwidth and
// counter calibration
__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ src, double2* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ------------------------------------------------------------ engine
struct FomBase {
  virtual ~FomBase() = default;
};

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
};

}  // namespace
}  // namespace shaw

struct shaw_fom {
  virtual ~shaw_fom() = default;
  virtual void setSources(const uint64_t* gid) = 0;
  virtual void setForcing(uint64_t nsteps, const double* tab) = 0;
  // forcing table generated on the device from signal parameters
  virtual void setSignals(const int* kinds, const double* delays, const double* periods, uint64_t nsteps) = 0;
  // operators of another material on the same mesh, assembled on the device
  virtual void setMaterial(shaw_app& app, const shaw::Material& mat) = 0;
  virtual void setReceivers(uint64_t R, const uint64_t* gid, uint64_t every, uint64_t maxSamples) = 0;
  virtual void setSnapshots(int dof, uint64_t every, uint64_t ncols, const uint64_t* cols, uint64_t maxSamples) = 0;
  virtual void reset() = 0;
  virtual void advance(uint64_t nsteps) = 0;
  virtual void sync() = 0;
  virtual uint64_t step() const = 0;
  virtual uint64_t getSeismogram(double* out) = 0;
  virtual uint64_t getSnapshots(int dof, double* out) = 0;
  virtual void getState(double* vp, double* sp) = 0;
  virtual void setState(const double* vp, const double* sp) = 0;
  virtual void time(uint64_t nsteps, double* total, double* perStep) = 0;
  virtual void timeChunks(uint64_t nsteps, uint64_t chunk, double* ms) = 0;
  virtual void bytes(double* kernel, double* model) const = 0;
  virtual void config(int* kernel, int* cpt, int* strip, int* seg, int* pf, int* spl, int* colBlock) const = 0;
  virtual void pod(int dof, uint64_t k, const double* w, double* phi, double* sigma) = 0;
  // POD into a device buffer dPhi [n x k]; podSize: snapshots of the dof (n: rows)
  virtual void podTo(int dof, uint64_t k, const double* w, double* dPhi, double* sigma) = 0;
  virtual uint64_t podSize(int dof, uint64_t* n) const = 0;
  virtual void autotune(uint64_t trialSteps, double* bestMs) = 0;
};

namespace shaw {
namespace {

template <class T>
class FomEngine final : public shaw_fom {
 public:
  FomEngine(const shaw_app& app, uint64_t M, double dt, const shaw_fom_options& o) {
    const Grid& g = app.ops.grid;
    nr_ = g.nr;
    nth_ = g.nth;
    nVp_ = g.nVp();
    nSp_ = g.nSp();
    M_ = M;
    dt_ = dt;
    rank1_ = o.rank1_forcing;
    SHAW_REQUIRE(M >= 1 && M <= (1u << 20), SHAW_ERR_ARG, "forcing column count out of range");
    SHAW_REQUIRE(nr_ >= 2 && nth_ >= 3 && nVp_ < (1ull << 40), SHAW_ERR_ARG, "bad grid");
    // column block: MC*sizeof(T) <= 256 B, power of two >= M when M is small
    // (or the requested power of two)
    const uint32_t cap = 256 / sizeof(T);
    SHAW_REQUIRE(o.col_block >= 0 && (uint32_t)o.col_block <= cap && (o.col_block & (o.col_block - 1)) == 0,
                 SHAW_ERR_ARG, "col_block must be a power of two of at most 256 bytes of state");
    // default: 2-column blocks, one thread per grid point at 2 columns per
    // thread, so a 64-point strip is one wavefront (k_fusedW)
    const uint32_t want = o.col_block > 0 ? (uint32_t)o.col_block : 2;
    MC_ = 1;
    while (MC_ < M && MC_ < want) MC_ *= 2;
    nb_ = (uint32_t)((M + MC_ - 1) / MC_);
    kernel_ = o.kernel == SHAW_KERNEL_SPLIT ? SHAW_KERNEL_SPLIT : SHAW_KERNEL_FUSED;
    // defaults from sweeps of the fused kernels on MI355X: at 1024 x 4096,
    // M = 32, 2-column blocks, 2 columns per thread, 64-point strips, 128-row
    // segments and six steps per launch; at 200 x 1000, M = 1, four steps per
    // launch two rows ahead, with segments of about 1700 wavefronts a launch
    cpt_ = o.cols_per_thread > 0 ? o.cols_per_thread : (int)std::min<uint32_t>(MC_, sizeof(T) == 8 ? 2 : 4);
    SHAW_REQUIRE(cpt_ == 1 || cpt_ == 2 || cpt_ == 4 || cpt_ == 8, SHAW_ERR_ARG, "cols_per_thread must be 1, 2, 4 or 8");
    SHAW_REQUIRE(MC_ % cpt_ == 0, SHAW_ERR_ARG, "cols_per_thread must divide the column block width");
    const uint32_t tpp = MC_ / cpt_;
    if (kernel_ == SHAW_KERNEL_FUSED) {
      P_ = o.strip > 0 ? (uint32_t)o.strip : std::max<uint32_t>(2, std::min<uint32_t>(64, 512 / tpp));
      SHAW_REQUIRE(P_ >= 2 && P_ * tpp <= 512, SHAW_ERR_ARG, "strip too wide for a 512-thread block");
      pf_ = o.prefetch_rows > 0 ? o.prefetch_rows : 1;
      SHAW_REQUIRE(pf_ == 1 || pf_ == 2, SHAW_ERR_ARG, "prefetch_rows must be 1 or 2");
      // one-wavefront strips: six steps per launch where a launch of 64-row
      // segments already fills the chip, else four steps loaded two rows
      // ahead (small grids, where a launch is latency-bound); other strips:
      // four steps per launch where the strip and the 32-bit addressing allow
      const bool wave = waveStrip(cpt_, P_) && fits32();
      if (o.steps_per_launch > 0) {
        spl_ = o.steps_per_launch;
      } else if (wave && o.prefetch_rows == 0 && launchWaves(cpt_, P_, 64, 6) >= kMinWaves) {
        spl_ = 6;
      } else if (wave) {
        spl_ = 4;
        if (o.prefetch_rows == 0) pf_ = 2;
      } else {
        spl_ = pf_ != 1 ? 1 : (P_ >= 8 && fits32() ? 4 : (P_ >= 4 ? 2 : 1));
      }
      SHAW_REQUIRE(spl_ >= 1 && spl_ <= kMaxSpl, SHAW_ERR_ARG, "steps_per_launch must be 1 to 7");
      SHAW_REQUIRE(spl_ == 1 || P_ >= 2u * spl_, SHAW_ERR_ARG,
                   "k steps per launch need a strip of at least 2k points");
      SHAW_REQUIRE(spl_ <= 4 || (waveStrip(cpt_, P_) && pf_ == 1), SHAW_ERR_ARG,
                   "5 to 7 steps per launch need one thread per point (cols_per_thread = col_block <= 2), "
                   "64-point strips and prefetch_rows 1");
      graphStepsOpt_ = o.graph_steps;
      graphSteps_ = graphStepsFor(spl_);
      SHAW_REQUIRE(graphSteps_ % (2 * spl_) == 0, SHAW_ERR_ARG,
                   "graph_steps must be a multiple of 2 x steps per launch (an even number of launches)");
      // segment: 64 rows (several steps per launch) or 32, halved while a
      // launch has fewer than kMinWaves wavefronts (small grids and few
      // columns: 200 x 1000 with M = 1 is fastest with 2-row segments)
      // (a segment of a k-step launch recomputes 2k-1 halo rows, so it is not
      // halved below k rows)
      if (o.segment > 0) {
        SEG_ = (uint32_t)o.segment;
      } else {
        // (six steps: 128 rows for FP64, 64 for FP32, the faster in sweeps at 1024 x 4096)
        SEG_ = spl_ >= 5 ? (sizeof(T) == 8 ? 128 : 64) : (spl_ >= 2 ? 64 : 32);
        while (SEG_ > std::max(2, spl_) && launchWaves(cpt_, P_, SEG_, spl_) < kMinWaves) SEG_ /= 2;
      }
      SHAW_REQUIRE(ldsBytes(cpt_, P_, spl_) <= kMaxLds, SHAW_ERR_ARG, "tile needs more LDS than a workgroup has");
      SHAW_REQUIRE(spl_ <= 2 || fits32(), SHAW_ERR_ARG,
                   "3 or 4 steps per launch need a column block of fewer than 2^32 stress values");
      setStrips();
    }
    SHAW_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SHAW_HIP(hipEventCreate(&ev0_));
    SHAW_HIP(hipEventCreate(&ev1_));
    const size_t nv = (size_t)nb_ * nVp_ * MC_, ns = (size_t)nb_ * nSp_ * MC_;
    const int nbuf = kernel_ == SHAW_KERNEL_FUSED ? 2 : 1;
    for (int b = 0; b < nbuf; ++b) { v_[b].alloc(nv); s_[b].alloc(ns); }
    // coefficients in the state precision
    std::vector<T> cv(app.ops.coefVp.begin(), app.ops.coefVp.end());
    std::vector<T> cs(app.ops.coefSp.begin(), app.ops.coefSp.end());
    cv_.alloc(cv.size());
    cs_.alloc(cs.size());
    SHAW_HIP(hipMemcpy(cv_.p, cv.data(), cv.size() * sizeof(T), hipMemcpyHostToDevice));
    SHAW_HIP(hipMemcpy(cs_.p, cs.data(), cs.size() * sizeof(T), hipMemcpyHostToDevice));
    rhoInv_ = app.ops.rhoInv;
    src_.alloc((size_t)nb_ * MC_);
    srcRho_.alloc((size_t)nb_ * MC_);
    std::vector<int64_t> none((size_t)nb_ * MC_, -1);
    SHAW_HIP(hipMemcpy(src_.p, none.data(), none.size() * 8, hipMemcpyHostToDevice));
    SHAW_HIP(hipMemset(srcRho_.p, 0, srcRho_.n * sizeof(T)));
    reset();
  }
  ~FomEngine() override {
    if (stream_) (void)hipStreamSynchronize(stream_);
    dropGraph();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  void setSources(const uint64_t* gid) override {
    srcGid_.assign(gid, gid + M_);
    std::vector<int64_t> s((size_t)nb_ * MC_, -1);
    std::vector<T> r((size_t)nb_ * MC_, T(0));
    for (uint64_t m = 0; m < M_; ++m) {
      SHAW_REQUIRE(gid[m] < nVp_, SHAW_ERR_ARG, "source gid out of range");
      s[m] = (int64_t)gid[m];
      r[m] = (T)rhoInv_[gid[m]];
    }
    SHAW_HIP(hipMemcpyAsync(src_.p, s.data(), s.size() * 8, hipMemcpyHostToDevice, stream_));
    SHAW_HIP(hipMemcpyAsync(srcRho_.p, r.data(), r.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
    SHAW_HIP(hipStreamSynchronize(stream_));
    haveSources_ = true;
  }
  void setForcing(uint64_t nsteps, const double* tab) override {
    sync();
    dropGraph();  // the graph holds the table's address
    ftab_.alloc(nsteps * M_);
    SHAW_HIP(hipMemcpy(ftab_.p, tab, nsteps * M_ * sizeof(double), hipMemcpyHostToDevice));
    ftabSteps_ = nsteps;
  }
  void setSignals(const int* kinds, const double* delays, const double* periods, uint64_t nsteps) override {
    sync();
    dropGraph();  // the graph holds the table's address
    ftab_.alloc(nsteps * M_);
    forcingTableDevice(M_, kinds, delays, periods, dt_, nsteps, ftab_.p, 0, stream_);
    ftabSteps_ = nsteps;
  }
  void setMaterial(shaw_app& app, const Material& mat) override {
    SHAW_REQUIRE(app.mesh.grid.nVp() == nVp_ && app.mesh.grid.nSp() == nSp_, SHAW_ERR_ARG,
                 "the app's mesh is not this FOM's mesh");
    sync();
    dropGraph();
    DevBuf<double> cv, cs, ri;
    cv.alloc(nVp_ * 4);
    cs.alloc(nSp_ * 2);
    ri.alloc(nVp_);
    assembleOperatorsDevice(app, mat, cv.p, cs.p, ri.p, stream_);
    if constexpr (sizeof(T) == 8) {
      SHAW_HIP(hipMemcpyAsync(cv_.p, cv.p, nVp_ * 4 * 8, hipMemcpyDeviceToDevice, stream_));
      SHAW_HIP(hipMemcpyAsync(cs_.p, cs.p, nSp_ * 2 * 8, hipMemcpyDeviceToDevice, stream_));
    } else {  // the state precision, as the host cast at creation
      hipLaunchKernelGGL(k_to_float, dim3((unsigned)((nVp_ * 4 + 255) / 256)), dim3(256), 0, stream_, cv.p,
                         reinterpret_cast<float*>(cv_.p), nVp_ * 4);
      hipLaunchKernelGGL(k_to_float, dim3((unsigned)((nSp_ * 2 + 255) / 256)), dim3(256), 0, stream_, cs.p,
                         reinterpret_cast<float*>(cs_.p), nSp_ * 2);
      SHAW_HIP(hipGetLastError());
    }
    SHAW_HIP(hipMemcpyAsync(rhoInv_.data(), ri.p, nVp_ * 8, hipMemcpyDeviceToHost, stream_));
    SHAW_HIP(hipStreamSynchronize(stream_));
    if (haveSources_) setSources(srcGid_.data());  // 1/rho at the source points of the new material
  }
  void setReceivers(uint64_t R, const uint64_t* gid, uint64_t every, uint64_t maxSamples) override {
    SHAW_REQUIRE(R == 0 || every > 0, SHAW_ERR_ARG, "receiver sampling period must be > 0");
    sync();
    dropGraph();  // the graph holds the receiver table and seismogram addresses
    std::vector<int32_t> idx(nVp_, -1);
    dupRecv_.clear();
    for (uint64_t r = 0; r < R; ++r) {
      SHAW_REQUIRE(gid[r] < nVp_, SHAW_ERR_ARG, "receiver gid out of range");
      // several receivers on one grid point: keep each, record via a duplicate list
      if (idx[gid[r]] >= 0) dupRecv_.push_back({(uint32_t)r, (uint32_t)idx[gid[r]]});
      else idx[gid[r]] = (int32_t)r;
    }
    recv_.alloc(R ? nVp_ : 0);
    if (R) SHAW_HIP(hipMemcpy(recv_.p, idx.data(), nVp_ * 4, hipMemcpyHostToDevice));
    // receivers sit on the surface row in every reference input; the fused
    // kernel then looks the table up on that row only
    recvRow_ = R ? (int64_t)(gid[0] / nth_) : -1;
    for (uint64_t r = 0; r < R; ++r)
      if ((int64_t)(gid[r] / nth_) != recvRow_) recvRow_ = -1;
    R_ = R;
    every_ = every;
    Tcap_ = maxSamples;
    seis_.alloc(R * maxSamples * M_);
    if (seis_.n) SHAW_HIP(hipMemset(seis_.p, 0, seis_.n * sizeof(double)));
  }
  void setSnapshots(int dof, uint64_t every, uint64_t ncols, const uint64_t* cols, uint64_t maxSamples) override {
    SHAW_REQUIRE(dof == 0 || dof == 1, SHAW_ERR_ARG, "dof must be 0 (vp) or 1 (sp)");
    SHAW_REQUIRE(ncols == 0 || every > 0, SHAW_ERR_ARG, "snapshot period must be > 0");
    Snap& s = snap_[dof];
    std::vector<uint32_t> c(ncols);
    for (uint64_t k = 0; k < ncols; ++k) {
      SHAW_REQUIRE(cols[k] < M_, SHAW_ERR_ARG, "snapshot column out of range");
      c[k] = (uint32_t)cols[k];
    }
    s.cols.alloc(ncols);
    if (ncols) SHAW_HIP(hipMemcpy(s.cols.p, c.data(), ncols * 4, hipMemcpyHostToDevice));
    s.ncols = ncols;
    s.every = every;
    s.Tcap = maxSamples;
    const uint64_t n = dof == 0 ? nVp_ : nSp_;
    s.buf.alloc(n * maxSamples * ncols);
  }
  void reset() override {
    for (int b = 0; b < 2; ++b) {
      if (v_[b].p) SHAW_HIP(hipMemsetAsync(v_[b].p, 0, v_[b].n * sizeof(T), stream_));
      if (s_[b].p) SHAW_HIP(hipMemsetAsync(s_[b].p, 0, s_[b].n * sizeof(T), stream_));
    }
    SHAW_HIP(hipStreamSynchronize(stream_));
    step_ = 0;
    cur_ = 0;
  }

  void advance(uint64_t nsteps) override {
    SHAW_REQUIRE(haveSources_, SHAW_ERR_STATE, "shaw_fom_set_sources must be called before advancing");
    SHAW_REQUIRE(ftab_.p && step_ + nsteps <= ftabSteps_, SHAW_ERR_STATE,
                 "forcing table shorter than the requested steps");
    const uint64_t last = step_ + nsteps;
    if (R_) SHAW_REQUIRE(last / every_ <= Tcap_, SHAW_ERR_STATE, "seismogram buffer too small for the requested steps");
    for (int d = 0; d < 2; ++d)
      if (snap_[d].ncols) SHAW_REQUIRE(last / snap_[d].every <= snap_[d].Tcap, SHAW_ERR_STATE, "snapshot buffer too small");
    uint64_t k = 0;
    // whole chunks of graphStep_ steps replay one captured graph (fused
    // kernels, no snapshots: a snapshot kernel would depend on the step)
    const bool graphs = graphSteps_ > 0 && kernel_ == SHAW_KERNEL_FUSED && !snap_[0].ncols && !snap_[1].ncols &&
                        nsteps >= (uint64_t)graphSteps_;
    if (graphs) {
      ensureGraph();
      hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, stream_, stepDev_.p, step_);
      // leave the last chunk to the eager packs when the remainder alone
      // would be a pack of one or two steps
      const uint64_t rem = nsteps % graphSteps_;
      const uint64_t whole = nsteps / graphSteps_ - (rem > 0 && rem < 3 && nsteps / graphSteps_ > 0 ? 1 : 0);
      for (; k + graphSteps_ <= whole * graphSteps_; k += graphSteps_) {
        SHAW_HIP(hipGraphLaunch(gexec_, stream_));
        step_ += graphSteps_;
      }
    }
    // eager launches of spl_ steps; a launch of the multi-step kernels costs
    // about the same whatever its steps (its time follows the bytes it
    // moves: 1.50-1.57 ms at 1024 x 4096 for 3 to 6 steps), so the last
    // steps (at most two of the longest packs) go in as few launches as
    // possible, split evenly: 20 steps at six per launch run 6 + 7 + 7, 8
    // run 4 + 4. No pack has a snapshot due at an intermediate step.
    for (; k < nsteps;) {
      const uint64_t left = nsteps - k;
      int ns = spl_;
      if (spl_ >= 3) {
        const uint64_t top = longestPack();
        if (left <= top) ns = (int)left;                           // one launch
        else if (left <= 2 * top) ns = (int)((left + 1) / 2);      // two, balanced
      }
      while (ns > 1 && (k + ns > nsteps || packBlocked(ns))) --ns;
      switch (ns) {
        case 1: launchStep(); break;
        case 2: launchPair(); break;
        case 3: launchPackN<3>(); break;
        case 4: launchPackN<4>(); break;
        case 5: launchPackN<5>(); break;
        case 6: launchPackN<6>(); break;
        default: launchPackN<7>(); break;
      }
      k += ns;
    }
    SHAW_HIP(hipGetLastError());
  }
  // the most steps one launch of the current shape takes: at six steps per
  // launch the wave kernel (one-wavefront strips, one row loaded ahead) also
  // runs seven-step packs for the last steps of a run
  uint64_t longestPack() const {
    return (spl_ >= 6 && waveStrip(cpt_, P_) && pf_ == 1) ? (uint64_t)kMaxSpl : (uint64_t)spl_;
  }
  // a snapshot due at one of the intermediate steps of a pack of ns steps
  bool packBlocked(int ns) const {
    for (int j = 1; j < ns; ++j)
      if (snapDue(step_ + j)) return true;
    return false;
  }
  // A graph of graphSteps_ steps of the current shape, captured from the
  // current buffer parity; every launch reads its step from stepDev_ (set
  // before the replays, advanced by the graph's last node), so one graph
  // serves every chunk. It holds an even number of launches, so the
  // ping-pong parity is the same after a replay.
  void ensureGraph() {
    const int key[7] = {cpt_, (int)P_, (int)SEG_, pf_, spl_, cur_, graphSteps_};
    if (gexec_ && std::equal(key, key + 7, graphKey_)) return;
    dropGraph();
    if (!stepDev_.p) stepDev_.alloc(1);
    const int cur0 = cur_;
    SHAW_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < graphSteps_; k += spl_) {
      if (spl_ == 2) {
        StepArgs2<T> aa;
        aa.s1 = deviceStepArgs(k + 1);
        aa.s2 = deviceStepArgs(k + 2);
        dispatchPair(aa, false);
      } else if (spl_ >= 3) {
        switch (spl_) {
          case 3: captureN<3>(k); break;
          case 4: captureN<4>(k); break;
          case 5: captureN<5>(k); break;
          case 6: captureN<6>(k); break;
          default: captureN<7>(k); break;
        }
      } else {
        StepArgs<T> a = deviceStepArgs(k + 1);
        dispatchStep(a, false);
      }
    }
    hipLaunchKernelGGL(k_step_add, dim3(1), dim3(64), 0, stream_, stepDev_.p, (uint64_t)graphSteps_);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream_, &g);
    cur_ = cur0;  // capture launched nothing
    SHAW_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SHAW_HIP(ei);
    std::copy(key, key + 7, graphKey_);
  }
  void dropGraph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    gexec_ = nullptr;
  }
  StepArgs<T> deviceStepArgs(uint64_t off) {
    StepArgs<T> a = args(1);  // shape fields; the step-dependent ones come from stepDev_
    a.stepDev = stepDev_.p;
    a.stepOff = off;
    a.ftabBase = ftab_.p;
    a.recvTable = R_ ? recv_.p : nullptr;
    a.every = R_ ? every_ : 0;
    return a;
  }
  void dispatchStep(StepArgs<T>& a, bool observe) {
    switch (cpt_) {
      case 1: launchFused<1>(a, observe); break;
      case 2: launchFused<2>(a, observe); break;
      case 4: launchFused<4>(a, observe); break;
      default: launchFused<8>(a, observe); break;
    }
  }
  template <int NS>
  void dispatchN(StepArgsN<T, NS>& aa, bool observe) {
    switch (cpt_) {
      case 1: launchFusedN<1, NS>(aa, observe); break;
      case 2: launchFusedN<2, NS>(aa, observe); break;
      case 4: launchFusedN<4, NS>(aa, observe); break;
      default: launchFusedN<8, NS>(aa, observe); break;
    }
  }
  void dispatchPair(StepArgs2<T>& aa, bool observe) {
    switch (cpt_) {
      case 1: launchFused2<1>(aa, observe); break;
      case 2: launchFused2<2>(aa, observe); break;
      case 4: launchFused2<4>(aa, observe); break;
      default: launchFused2<8>(aa, observe); break;
    }
  }
  // a snapshot at the intermediate step of a pair needs that level in HBM
  bool snapDue(uint64_t stepNo) const {
    for (int d = 0; d < 2; ++d)
      if (snap_[d].ncols && stepNo % snap_[d].every == 0) return true;
    return false;
  }
  void sync() override { SHAW_HIP(hipStreamSynchronize(stream_)); }
  uint64_t step() const override { return step_; }

  uint64_t getSeismogram(double* out) override {
    sync();
    const uint64_t T_ = R_ ? step_ / every_ : 0;
    // device [R x Tcap x M] -> host [R x T x M]
    for (uint64_t m = 0; m < M_ && T_; ++m)
      SHAW_HIP(hipMemcpy(out + R_ * T_ * m, seis_.p + R_ * Tcap_ * m, R_ * T_ * sizeof(double), hipMemcpyDeviceToHost));
    // receivers sharing a grid point with an earlier receiver
    for (const auto& d : dupRecv_)
      for (uint64_t m = 0; m < M_; ++m)
        for (uint64_t t = 0; t < T_; ++t) out[d.first + R_ * (t + T_ * m)] = out[d.second + R_ * (t + T_ * m)];
    return T_;
  }
  uint64_t getSnapshots(int dof, double* out) override {
    sync();
    const Snap& s = snap_[dof];
    const uint64_t T_ = s.ncols ? step_ / s.every : 0, n = dof == 0 ? nVp_ : nSp_;
    for (uint64_t c = 0; c < s.ncols && T_; ++c)
      SHAW_HIP(hipMemcpy(out + n * T_ * c, s.buf.p + n * s.Tcap * c, n * T_ * sizeof(double), hipMemcpyDeviceToHost));
    return T_;
  }
  void getState(double* vp, double* sp) override {
    sync();
    copyState(vp, v_[cur_].p, nVp_, true);
    copyState(sp, s_[cur_].p, nSp_, true);
  }
  void setState(const double* vp, const double* sp) override {
    sync();
    copyState(const_cast<double*>(vp), v_[cur_].p, nVp_, false);
    copyState(const_cast<double*>(sp), s_[cur_].p, nSp_, false);
  }
  void time(uint64_t nsteps, double* total, double* perStep) override {
    SHAW_HIP(hipEventRecord(ev0_, stream_));
    advance(nsteps);
    SHAW_HIP(hipEventRecord(ev1_, stream_));
    SHAW_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    SHAW_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    *total = ms;
    *perStep = nsteps ? ms / (double)nsteps : 0.0;
  }
  // the run split into chunks of `chunk` steps with an event between
  // chunks, all enqueued before one wait: ms[i] = duration of chunk i
  void timeChunks(uint64_t nsteps, uint64_t chunk, double* ms) override {
    const uint64_t n = (nsteps + chunk - 1) / chunk;
    std::vector<hipEvent_t> ev(n + 1, nullptr);
    try {
      for (auto& e : ev) SHAW_HIP(hipEventCreate(&e));
      SHAW_HIP(hipEventRecord(ev[0], stream_));
      for (uint64_t i = 0; i < n; ++i) {
        advance(std::min(chunk, nsteps - i * chunk));
        SHAW_HIP(hipEventRecord(ev[i + 1], stream_));
      }
      SHAW_HIP(hipEventSynchronize(ev[n]));
      for (uint64_t i = 0; i < n; ++i) {
        float t = 0;
        SHAW_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
        ms[i] = t;
      }
    } catch (...) {
      (void)hipStreamSynchronize(stream_);
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
      throw;
    }
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  void bytes(double* kernel, double* model) const override {
    // algorithmic minimum: the M state columns and the coefficients once
    // (a launch over several column blocks re-reads the coefficients per
    // block, and a strip re-reads its halo points: that is traffic above it)
    const double s = sizeof(T), cols = (double)M_;
    const double coef = (4.0 * nVp_ + 2.0 * nSp_) * s;
    if (kernel) {
      // fused: read + write each state once; split: v (r+w) + s (r) / v (r) + s (r+w)
      // (per step: a launch of the k-step kernel moves these bytes for k steps)
      const double passes = kernel_ == SHAW_KERNEL_FUSED ? 2.0 : 3.0;
      *kernel = (s * cols * passes * (double)(nVp_ + nSp_) + coef) / (kernel_ == SHAW_KERNEL_FUSED ? spl_ : 1);
    }
    if (model) *model = s * (double)M_ * 3.0 * (double)(nVp_ + nSp_) + 8.0 * (4.0 * nVp_ + 2.0 * nSp_);
  }
  uint64_t podSize(int dof, uint64_t* n) const override {
    const Snap& s = snap_[dof];
    if (n) *n = dof == 0 ? nVp_ : nSp_;
    return s.every ? (step_ / s.every) * s.ncols : 0;
  }
  void pod(int dof, uint64_t k, const double* w, double* phi, double* sigma) override {
    uint64_t n = 0;
    podSize(dof, &n);
    DevBuf<double> d;
    d.alloc(n * k);
    podTo(dof, k, w, d.p, sigma);
    SHAW_HIP(hipMemcpy(phi, d.p, n * k * sizeof(double), hipMemcpyDeviceToHost));
  }
  void podTo(int dof, uint64_t k, const double* w, double* dPhi, double* sigma) override {
    SHAW_REQUIRE(sizeof(T) == 8, SHAW_ERR_ARG, "POD of snapshots needs the FP64 FOM");
    Snap& s = snap_[dof];
    SHAW_REQUIRE(s.ncols > 0, SHAW_ERR_STATE, "no snapshots were configured for this dof");
    const uint64_t T_ = step_ / s.every, n = dof == 0 ? nVp_ : nSp_;
    SHAW_REQUIRE(T_ == s.Tcap, SHAW_ERR_STATE, "POD needs a full snapshot buffer (advance to max_samples)");
    sync();
    DevBuf<double> dw;
    const uint64_t cols = T_ * s.ncols;
    if (w) {  // POD in the inner product diag(w): modes of diag(sqrt w) S, then unscaled
      std::vector<double> sw(n);
      for (uint64_t i = 0; i < n; ++i) {
        SHAW_REQUIRE(w[i] > 0 && std::isfinite(w[i]), SHAW_ERR_ARG, "POD weights must be positive and finite");
        sw[i] = std::sqrt(w[i]);
      }
      dw.alloc(n);
      SHAW_HIP(hipMemcpy(dw.p, sw.data(), n * sizeof(double), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((n + 255) / 256), (unsigned)std::min<uint64_t>(cols, 65535)),
                         dim3(256), 0, 0, s.buf.p, dw.p, n, cols);
      SHAW_HIP(hipGetLastError());
      SHAW_HIP(hipDeviceSynchronize());
    }
    podDeviceTo(s.buf.p, n, cols, k, dPhi, sigma, nullptr);
    if (w) {
      hipLaunchKernelGGL(k_div_rows, dim3((unsigned)((n + 255) / 256), (unsigned)std::min<uint64_t>(k, 65535)),
                         dim3(256), 0, 0, dPhi, dw.p, n, k);
      SHAW_HIP(hipGetLastError());
      SHAW_HIP(hipDeviceSynchronize());
    }
    s.ncols = 0;  // the buffer was consumed
  }
  void config(int* kernel, int* cpt, int* strip, int* seg, int* pf, int* spl, int* colBlock) const override {
    if (colBlock) *colBlock = (int)MC_;
    if (kernel) *kernel = kernel_;
    if (cpt) *cpt = cpt_;
    if (strip) *strip = (int)P_;
    if (seg) *seg = (int)SEG_;
    if (pf) *pf = pf_;
    if (spl) *spl = kernel_ == SHAW_KERNEL_FUSED ? spl_ : 1;
  }

 private:
  struct Snap {
    DevBuf<uint32_t> cols;
    DevBuf<double> buf;
    uint64_t ncols = 0, every = 0, Tcap = 0;
  };

  StepArgs<T> args(uint64_t stepNo) {
    StepArgs<T> a{};
    a.cv = cv_.p;
    a.cs = cs_.p;
    a.src = src_.p;
    a.srcRhoInv = srcRho_.p;
    a.frow = ftab_.p + (stepNo - 1) * M_;
    const bool rec = R_ && stepNo % every_ == 0;
    a.recv = rec ? recv_.p : nullptr;
    a.recvRow = recvRow_;
    a.seis = seis_.p;
    a.sample = rec ? stepNo / every_ - 1 : 0;
    a.R = R_;
    a.Tcap = Tcap_;
    a.nr = (uint32_t)nr_;
    a.nth = (uint32_t)nth_;
    a.M = (uint32_t)M_;
    a.MC = MC_;
    a.dt = (T)dt_;
    a.rank1 = rank1_;
    a.P = P_;
    a.SEG = SEG_;
    a.nStrips = nStrips_;
    a.nSegs = nSegs_;
    a.nb = nb_;
    return a;
  }

  template <int CPT>
  void launchSplit(StepArgs<T>& a) {
    const uint64_t tpp = MC_ / CPT;
    dim3 gv((unsigned)((nVp_ * tpp + 255) / 256), nb_), gs((unsigned)((nSp_ * tpp + 255) / 256), nb_);
    a.vIn = a.vOut = v_[0].p;
    a.sIn = a.sOut = s_[0].p;
    hipLaunchKernelGGL((k_velocity<T, CPT>), gv, dim3(256), 0, stream_, a);
    snapshot(0, a.vOut);
    hipLaunchKernelGGL((k_stress<T, CPT>), gs, dim3(256), 0, stream_, a);
    snapshot(1, a.sOut);
  }
  // Tile autotuner of the fused kernel: every valid (columns per thread,
  // strip, segment, rows prefetched) is timed for `trialSteps` steps on the
  // current state (after one untimed step), the fastest is kept, and the
  // state is restored: the trials do not advance the run, record receivers
  // or take snapshots.
  void autotune(uint64_t trialSteps, double* bestMs) override {
    SHAW_REQUIRE(kernel_ == SHAW_KERNEL_FUSED, SHAW_ERR_STATE, "the autotuner tunes the fused kernel");
    SHAW_REQUIRE(trialSteps >= 1 && trialSteps <= 1000, SHAW_ERR_ARG, "trial steps must be in 1..1000");
    SHAW_REQUIRE(haveSources_ && ftab_.p && step_ + 2 <= ftabSteps_, SHAW_ERR_STATE,
                 "set sources and a forcing table covering the next two steps before tuning");
    sync();
    const int cur0 = cur_;
    DevBuf<T> sv, ss;
    sv.alloc(v_[cur0].n);
    ss.alloc(s_[cur0].n);
    SHAW_HIP(hipMemcpy(sv.p, v_[cur0].p, sv.n * sizeof(T), hipMemcpyDeviceToDevice));
    SHAW_HIP(hipMemcpy(ss.p, s_[cur0].p, ss.n * sizeof(T), hipMemcpyDeviceToDevice));
    struct Cfg { int cpt; uint32_t P, SEG; int pf, spl; };
    Cfg best{cpt_, P_, SEG_, pf_, spl_};
    double bestT = 1e300;
    auto apply = [&](const Cfg& c) {
      cpt_ = c.cpt; P_ = c.P; SEG_ = c.SEG; pf_ = c.pf; spl_ = c.spl;
      graphSteps_ = graphStepsFor(spl_);
      setStrips();
    };
    for (int cpt : {1, 2, 4, 8}) {
      if (MC_ % cpt) continue;
      const uint32_t tpp = MC_ / cpt;
      for (uint32_t P : {8u, 16u, 32u, 64u, 128u, 256u}) {
        if (P * tpp > 512 || P * tpp < 64) continue;
        for (uint32_t SEG : {2u, 4u, 8u, 16u, 32u, 64u, 128u, 256u})
          for (int pf : {1, 2})
            for (int spl = 1; spl <= kMaxSpl; ++spl) {
              // the multi-step kernels load one row ahead, k_fusedW one or two
              if (spl >= 2 && pf == 2 && !(spl >= 3 && spl <= 4 && waveStrip(cpt, P))) continue;
              if (spl >= 5 && !waveStrip(cpt, P)) continue;  // 5 to 7 steps: one-wavefront strips only
              if (spl >= 2 && P < 2u * spl) continue;
              if (step_ + spl > ftabSteps_) continue;  // a trial reads the forcing of its steps
              if (graphStepsFor(spl) % (2 * spl)) continue;  // a requested graph length must fit
              if (ldsBytes(cpt, P, spl) > kMaxLds) continue;
              if (spl >= 3 && !fits32()) continue;
              // short segments only where the grid is too small to fill the chip
              if (SEG < 16 && launchWaves(cpt, P, SEG, spl) > 16 * kMinWaves) continue;
              const Cfg c{cpt, P, SEG, pf, spl};
              apply(c);
              trialLaunch();
              SHAW_HIP(hipEventRecord(ev0_, stream_));
              for (uint64_t k = 0; k < trialSteps; ++k) trialLaunch();
              SHAW_HIP(hipEventRecord(ev1_, stream_));
              SHAW_HIP(hipEventSynchronize(ev1_));
              float ms = 0;
              SHAW_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
              const double perStep = ms / (double)(trialSteps * spl);
              if (perStep < bestT) { bestT = perStep; best = c; }
            }
      }
    }
    apply(best);
    SHAW_HIP(hipMemcpy(v_[cur0].p, sv.p, sv.n * sizeof(T), hipMemcpyDeviceToDevice));
    SHAW_HIP(hipMemcpy(s_[cur0].p, ss.p, ss.n * sizeof(T), hipMemcpyDeviceToDevice));
    cur_ = cur0;
    if (bestMs) *bestMs = bestT;
  }
  // graph chunk for `spl` steps per launch: the requested length, or 64
  // rounded up to an even number of launches; 0 = no graphs
  int graphStepsFor(int spl) const {
    if (graphStepsOpt_ < 0) return 0;
    if (graphStepsOpt_ > 0) return graphStepsOpt_;
    return (64 + 2 * spl - 1) / (2 * spl) * (2 * spl);
  }
  // strips of every kernel for the current strip width and segment
  void setStrips() {
    nStrips_ = (uint32_t)((nth_ + (P_ - 1) - 1) / (P_ - 1));
    for (int ns = 2; ns <= kMaxSpl; ++ns) {
      const uint32_t q = P_ >= 2u * ns ? P_ - (2 * ns - 1) : 0;
      nStripsK_[ns] = q ? (uint32_t)((nth_ + q - 1) / q) : 0;
    }
    nStrips2_ = nStripsK_[2];
    nSegs_ = (uint32_t)((nr_ + SEG_ - 1) / SEG_);
  }
  // dynamic LDS of one fused launch
  size_t ldsBytes(int cpt, uint32_t P, int spl) const {
    if (spl == 1) return (size_t)(2 * P + 1) * MC_ * sizeof(T);
    if (spl == 2) return (size_t)(4 * P + 2) * MC_ * sizeof(T);
    return waveStrip(cpt, P) ? ldsBytesW(spl) : ldsBytesN(P, spl);
  }
  size_t ldsBytesN(uint32_t P, int ns) const {
    return ((size_t)ns * P * 8 + 2 * (size_t)(P + 1) * MC_ + (size_t)ns * P * MC_) * sizeof(T);
  }
  // k_fusedW: a ring of 4 (3 steps) or 8 (4 to 7 steps) coefficient rows of a 64-point strip
  // and the forcing increments ([4 waves][steps][2 columns] at most)
  size_t ldsBytesW(int ns) const { return ((size_t)(ns + 1 <= 4 ? 4 : 8) * 64 * 8 + 4 * ns * 2) * sizeof(T); }
  // 3 to 7 steps per launch run k_fusedW when a strip is one wavefront: one
  // thread per point (all columns of the block) and 64-point strips
  bool waveStrip(int cpt, uint32_t P) const { return (uint32_t)cpt == MC_ && P == 64 && cpt <= 2; }
  // column blocks (waves) per k_fusedW workgroup
  uint32_t waveGroup() const { return nb_ >= 4 ? 4 : (nb_ >= 2 ? 2 : 1); }
  // k_fusedN addresses a column block with 32-bit offsets
  bool fits32() const {
    const uint64_t lim = 1ull << 32;
    return (uint64_t)nSp_ * MC_ + 2ull * nth_ * MC_ < lim && 4ull * nVp_ < lim && 2ull * nSp_ + 4ull * nth_ < lim;
  }
  // wavefronts of one fused launch of the given shape
  uint64_t launchWaves(int cpt, uint32_t P, uint32_t SEG, int spl) const {
    const uint32_t halo = spl == 1 ? 1 : 2 * spl - 1;
    const uint64_t strips = (nth_ + (P - halo) - 1) / (P - halo), segs = (nr_ + SEG - 1) / SEG;
    if (spl >= 3 && waveStrip(cpt, P))  // whole groups of column blocks, one wave each
      return strips * segs * ((nb_ + waveGroup() - 1) / waveGroup()) * waveGroup();
    const uint64_t wavesPerBlock = ((uint64_t)P * (MC_ / cpt) + 63) / 64;
    return strips * segs * nb_ * wavesPerBlock;
  }
  // one launch of the current shape (one step or a pair) that records
  // nothing and does not advance the run
  void trialLaunch() {
    if (spl_ >= 3) {
      switch (spl_) {
        case 3: trialPackN<3>(); break;
        case 4: trialPackN<4>(); break;
        case 5: trialPackN<5>(); break;
        case 6: trialPackN<6>(); break;
        default: trialPackN<7>(); break;
      }
    } else if (spl_ == 2) {
      StepArgs2<T> aa;
      pairArgs(aa);
      aa.s1.recv = aa.s2.recv = nullptr;
      switch (cpt_) {
        case 1: launchFused2<1>(aa, false); break;
        case 2: launchFused2<2>(aa, false); break;
        case 4: launchFused2<4>(aa, false); break;
        default: launchFused2<8>(aa, false); break;
      }
      SHAW_HIP(hipGetLastError());
    } else {
      trialStep();
    }
  }
  template <int NS>
  void trialPackN() {
    StepArgsN<T, NS> aa;
    packArgs(aa);
    for (int l = 0; l < NS; ++l) aa.s[l].recv = nullptr;
    dispatchN<NS>(aa, false);
    SHAW_HIP(hipGetLastError());
  }
  void trialStep() {
    StepArgs<T> a = args(step_ + 1);
    a.recv = nullptr;
    switch (cpt_) {
      case 1: launchFused<1>(a, false); break;
      case 2: launchFused<2>(a, false); break;
      case 4: launchFused<4>(a, false); break;
      default: launchFused<8>(a, false); break;
    }
    SHAW_HIP(hipGetLastError());
  }
  template <int CPT>
  void launchFused(StepArgs<T>& a, bool observe = true) {
    const uint32_t tpp = MC_ / CPT;
    a.vIn = v_[cur_].p;
    a.sIn = s_[cur_].p;
    a.vOut = v_[cur_ ^ 1].p;
    a.sOut = s_[cur_ ^ 1].p;
    const unsigned blocks = nStrips_ * nSegs_ * nb_;
    const size_t lds = (size_t)(2 * P_ + 1) * MC_ * sizeof(T);
    if (pf_ == 2) hipLaunchKernelGGL((k_fused<T, CPT, 2>), dim3(blocks), dim3(P_ * tpp), lds, stream_, a);
    else hipLaunchKernelGGL((k_fused<T, CPT, 1>), dim3(blocks), dim3(P_ * tpp), lds, stream_, a);
    cur_ ^= 1;
    if (observe) {
      snapshot(0, v_[cur_].p);
      snapshot(1, s_[cur_].p);
    }
  }
  template <int CPT>
  void launchFused2(StepArgs2<T>& aa, bool observe = true) {
    const uint32_t tpp = MC_ / CPT;
    for (StepArgs<T>* a : {&aa.s1, &aa.s2}) {
      a->vIn = v_[cur_].p;
      a->sIn = s_[cur_].p;
      a->vOut = v_[cur_ ^ 1].p;
      a->sOut = s_[cur_ ^ 1].p;
      a->nStrips = nStrips2_;
    }
    const unsigned blocks = nStrips2_ * nSegs_ * nb_;
    const size_t lds = (size_t)(4 * P_ + 2) * MC_ * sizeof(T);
    hipLaunchKernelGGL((k_fused2<T, CPT>), dim3(blocks), dim3(P_ * tpp), lds, stream_, aa);
    cur_ ^= 1;
    if (observe) {
      snapshot(0, v_[cur_].p);
      snapshot(1, s_[cur_].p);
    }
  }
  template <int CPT, int NS>
  void launchFusedN(StepArgsN<T, NS>& aa, bool observe = true) {
    const uint32_t tpp = MC_ / CPT;
    for (int l = 0; l < NS; ++l) {
      StepArgs<T>& a = aa.s[l];
      a.vIn = v_[cur_].p;
      a.sIn = s_[cur_].p;
      a.vOut = v_[cur_ ^ 1].p;
      a.sOut = s_[cur_ ^ 1].p;
      a.nStrips = nStripsK_[NS];
    }
    if constexpr (CPT <= 2) {
      if (waveStrip(CPT, P_)) {
        switch (waveGroup()) {
          case 1: launchWave<CPT, NS, 1>(aa); break;
          case 2: launchWave<CPT, NS, 2>(aa); break;
          default: launchWave<CPT, NS, 4>(aa); break;
        }
        cur_ ^= 1;
        if (observe) {
          snapshot(0, v_[cur_].p);
          snapshot(1, s_[cur_].p);
        }
        return;
      }
    }
    if constexpr (NS <= 4) {
      const unsigned blocks = nStripsK_[NS] * nSegs_ * nb_;
      hipLaunchKernelGGL((k_fusedN<T, CPT, NS>), dim3(blocks), dim3(P_ * tpp), ldsBytesN(P_, NS), stream_, aa);
    } else {
      (void)tpp;
      throw Error(SHAW_ERR_STATE, "5 to 7 steps per launch run only on one-wavefront strips");
    }
    cur_ ^= 1;
    if (observe) {
      snapshot(0, v_[cur_].p);
      snapshot(1, s_[cur_].p);
    }
  }
  template <int CPT, int NS, int G>
  void launchWave(StepArgsN<T, NS>& aa) {
    const unsigned blocks = nStripsK_[NS] * nSegs_ * ((nb_ + G - 1) / G);
    if constexpr (NS <= 4) {
      if (pf_ == 2) {
        hipLaunchKernelGGL((k_fusedW<T, CPT, NS, 2, G>), dim3(blocks), dim3(G * 64), ldsBytesW(NS), stream_, aa);
        return;
      }
    }
    hipLaunchKernelGGL((k_fusedW<T, CPT, NS, 1, G>), dim3(blocks), dim3(G * 64), ldsBytesW(NS), stream_, aa);
  }
  template <int NS>
  void captureN(int k) {
    StepArgsN<T, NS> aa;
    for (int l = 0; l < NS; ++l) aa.s[l] = deviceStepArgs(k + 1 + l);
    dispatchN<NS>(aa, false);
  }
  template <int NS>
  void packArgs(StepArgsN<T, NS>& aa) {
    for (int l = 0; l < NS; ++l) aa.s[l] = args(step_ + 1 + l);
  }
  template <int NS>
  void launchPackN() {
    StepArgsN<T, NS> aa;
    packArgs(aa);
    stepForSnap_ = step_ + NS;
    dispatchN<NS>(aa, true);
    step_ += NS;
  }
  void pairArgs(StepArgs2<T>& aa) {
    aa.s1 = args(step_ + 1);
    aa.s2 = args(step_ + 2);
  }
  void launchPair() {
    StepArgs2<T> aa;
    pairArgs(aa);
    stepForSnap_ = step_ + 2;
    switch (cpt_) {
      case 1: launchFused2<1>(aa); break;
      case 2: launchFused2<2>(aa); break;
      case 4: launchFused2<4>(aa); break;
      default: launchFused2<8>(aa); break;
    }
    step_ += 2;
  }
  void launchStep() {
    const uint64_t stepNo = step_ + 1;
    stepForSnap_ = stepNo;
    StepArgs<T> a = args(stepNo);
    const bool fused = kernel_ == SHAW_KERNEL_FUSED;
    switch (cpt_) {
      case 1: fused ? launchFused<1>(a) : launchSplit<1>(a); break;
      case 2: fused ? launchFused<2>(a) : launchSplit<2>(a); break;
      case 4: fused ? launchFused<4>(a) : launchSplit<4>(a); break;
      default: fused ? launchFused<8>(a) : launchSplit<8>(a); break;
    }
    step_ = stepNo;
  }
  void snapshot(int dof, const T* x) {
    Snap& s = snap_[dof];
    if (!s.ncols || stepForSnap_ % s.every != 0) return;
    const uint64_t n = dof == 0 ? nVp_ : nSp_;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)s.ncols);
    hipLaunchKernelGGL(k_snapshot<T>, grid, dim3(256), 0, stream_, x, n, MC_, s.cols.p, (uint32_t)s.ncols, s.buf.p,
                       stepForSnap_ / s.every - 1, s.Tcap);
  }
  // host [n][M] double <-> device [nb][n][MC] T
  void copyState(double* host, T* dev, uint64_t n, bool toHost) {
    if (!host) return;
    std::vector<T> tmp((size_t)nb_ * n * MC_, T(0));
    if (toHost) {
      SHAW_HIP(hipMemcpy(tmp.data(), dev, tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
      for (uint64_t g = 0; g < n; ++g)
        for (uint64_t m = 0; m < M_; ++m) host[g * M_ + m] = (double)tmp[((m / MC_) * n + g) * MC_ + m % MC_];
    } else {
      for (uint64_t g = 0; g < n; ++g)
        for (uint64_t m = 0; m < M_; ++m) tmp[((m / MC_) * n + g) * MC_ + m % MC_] = (T)host[g * M_ + m];
      SHAW_HIP(hipMemcpy(dev, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    }
  }

  uint64_t nr_ = 0, nth_ = 0, nVp_ = 0, nSp_ = 0, M_ = 0;
  uint32_t MC_ = 1, nb_ = 1, P_ = 0, SEG_ = 0, nStrips_ = 0, nStrips2_ = 0, nSegs_ = 0;
  int kernel_ = SHAW_KERNEL_FUSED, cpt_ = 1, rank1_ = 0, pf_ = 1, spl_ = 1;
  double dt_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  DevBuf<T> v_[2], s_[2], cv_, cs_, srcRho_;
  DevBuf<int64_t> src_;
  DevBuf<double> ftab_, seis_;
  DevBuf<int32_t> recv_;
  int64_t recvRow_ = -1;
  DevBuf<uint64_t> stepDev_;
  hipGraphExec_t gexec_ = nullptr;
  int graphKey_[7] = {0, 0, 0, 0, 0, -1, 0};
  int graphSteps_ = 64, graphStepsOpt_ = 0;
  uint32_t nStripsK_[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // strips of the k-step kernels (k = 2..7)
  static constexpr int kMaxSpl = 7;
  static constexpr size_t kMaxLds = 160 * 1024;
  static constexpr uint64_t kMinWaves = 1536;  // 6 per CU of the 256
  std::vector<double> rhoInv_;
  std::vector<uint64_t> srcGid_;
  std::vector<std::pair<uint32_t, uint32_t>> dupRecv_;
  uint64_t ftabSteps_ = 0, R_ = 0, every_ = 1, Tcap_ = 0, step_ = 0, stepForSnap_ = 0;
  int cur_ = 0;
  bool haveSources_ = false;
  Snap snap_[2];
};

}  // namespace
}  // namespace shaw

// ------------------------------------------------------------------ C ABI
using shaw::guarded;

extern "C" {

int shaw_fom_create(shaw_app_t app, uint64_t m, double dt, const shaw_fom_options* opts, shaw_fom_t* out) {
  return guarded([&] {
    SHAW_REQUIRE(app && out, SHAW_ERR_ARG, "null argument");
    SHAW_REQUIRE(dt > 0, SHAW_ERR_ARG, "dt must be > 0");
    // the options are validated before the engine's first HIP call, so an
    // invalid shape is refused with its reason also where no device is present
    shaw_fom_options o{};
    if (opts) o = *opts;
    SHAW_REQUIRE(o.precision == SHAW_FP64 || o.precision == SHAW_FP32, SHAW_ERR_ARG, "unknown precision");
    if (o.precision == SHAW_FP64) *out = new shaw::FomEngine<double>(*app, m, dt, o);
    else *out = new shaw::FomEngine<float>(*app, m, dt, o);
  });
}
int shaw_fom_set_sources(shaw_fom_t h, const uint64_t* gid) {
  return guarded([&] { SHAW_REQUIRE(h && gid, SHAW_ERR_ARG, "null argument"); h->setSources(gid); });
}
int shaw_fom_set_forcing(shaw_fom_t h, uint64_t nsteps, const double* tab) {
  return guarded([&] { SHAW_REQUIRE(h && tab, SHAW_ERR_ARG, "null argument"); h->setForcing(nsteps, tab); });
}
int shaw_fom_set_receivers(shaw_fom_t h, uint64_t r, const uint64_t* gid, uint64_t every, uint64_t maxSamples) {
  return guarded([&] {
    SHAW_REQUIRE(h && (r == 0 || gid), SHAW_ERR_ARG, "null argument");
    h->setReceivers(r, gid, every, maxSamples);
  });
}
int shaw_fom_set_snapshots(shaw_fom_t h, int dof, uint64_t every, uint64_t ncols, const uint64_t* cols,
                           uint64_t maxSamples) {
  return guarded([&] {
    SHAW_REQUIRE(h && (ncols == 0 || cols), SHAW_ERR_ARG, "null argument");
    h->setSnapshots(dof, every, ncols, cols, maxSamples);
  });
}
int shaw_fom_reset(shaw_fom_t h) {
  return guarded([&] { SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle"); h->reset(); });
}
int shaw_fom_advance(shaw_fom_t h, uint64_t nsteps) {
  return guarded([&] { SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle"); h->advance(nsteps); });
}
int shaw_fom_sync(shaw_fom_t h) {
  return guarded([&] { SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle"); h->sync(); });
}
int shaw_fom_step(shaw_fom_t h, uint64_t* step) {
  return guarded([&] { SHAW_REQUIRE(h && step, SHAW_ERR_ARG, "null argument"); *step = h->step(); });
}
int shaw_fom_get_seismogram(shaw_fom_t h, double* out, uint64_t* samples) {
  return guarded([&] {
    SHAW_REQUIRE(h && out, SHAW_ERR_ARG, "null argument");
    const uint64_t t = h->getSeismogram(out);
    if (samples) *samples = t;
  });
}
int shaw_fom_get_snapshots(shaw_fom_t h, int dof, double* out, uint64_t* samples) {
  return guarded([&] {
    SHAW_REQUIRE(h && out && (dof == 0 || dof == 1), SHAW_ERR_ARG, "bad argument");
    const uint64_t t = h->getSnapshots(dof, out);
    if (samples) *samples = t;
  });
}
int shaw_fom_get_state(shaw_fom_t h, double* vp, double* sp) {
  return guarded([&] { SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle"); h->getState(vp, sp); });
}
int shaw_fom_set_state(shaw_fom_t h, const double* vp, const double* sp) {
  return guarded([&] { SHAW_REQUIRE(h && vp && sp, SHAW_ERR_ARG, "null argument"); h->setState(vp, sp); });
}
int shaw_fom_time(shaw_fom_t h, uint64_t nsteps, double* total, double* perStep) {
  return guarded([&] {
    SHAW_REQUIRE(h && total && perStep, SHAW_ERR_ARG, "null argument");
    h->time(nsteps, total, perStep);
  });
}
int shaw_fom_time_chunks(shaw_fom_t h, uint64_t nsteps, uint64_t chunk, double* chunk_ms) {
  return guarded([&] {
    SHAW_REQUIRE(h && chunk_ms && chunk > 0, SHAW_ERR_ARG, "bad argument");
    h->timeChunks(nsteps, chunk, chunk_ms);
  });
}
int shaw_fom_bytes_per_step(shaw_fom_t h, double* kernel, double* model) {
  return guarded([&] { SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle"); h->bytes(kernel, model); });
}
int shaw_fom_config(shaw_fom_t h, int* kernel, int* cpt, int* strip, int* seg, int* pf, int* spl, int* col_block) {
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    h->config(kernel, cpt, strip, seg, pf, spl, col_block);
  });
}
int shaw_fom_autotune(shaw_fom_t h, uint64_t trial_steps, double* best_ms_per_step) {
  return guarded([&] {
    SHAW_REQUIRE(h, SHAW_ERR_ARG, "null handle");
    h->autotune(trial_steps, best_ms_per_step);
  });
}
int shaw_fom_pod(shaw_fom_t h, int dof, uint64_t k, double* phi, double* sigma) {
  return shaw_fom_pod_weighted(h, dof, k, nullptr, phi, sigma);
}
int shaw_fom_pod_weighted(shaw_fom_t h, int dof, uint64_t k, const double* w, double* phi, double* sigma) {
  return guarded([&] {
    SHAW_REQUIRE(h && phi && (dof == 0 || dof == 1), SHAW_ERR_ARG, "bad argument");
    h->pod(dof, k, w, phi, sigma);
  });
}
int shaw_fom_set_signals(shaw_fom_t h, const int* kinds, const double* delays, const double* periods,
                         uint64_t nsteps) {
  return guarded([&] {
    SHAW_REQUIRE(h && kinds && delays && periods, SHAW_ERR_ARG, "null argument");
    h->setSignals(kinds, delays, periods, nsteps);
  });
}
int shaw_fom_set_material(shaw_fom_t h, shaw_app_t app, const shaw_material* mat) {
  return guarded([&] {
    SHAW_REQUIRE(h && app && mat, SHAW_ERR_ARG, "null argument");
    shaw::Material m = shaw::materialFromC(mat);
    m.rSurfM = app->mesh.info.bounds[3];
    m.validate(app->mesh.info);
    h->setMaterial(*app, m);
  });
}
int shaw_fom_pod_basis(shaw_fom_t h, uint64_t kv, uint64_t ks, const double* wv, const double* ws,
                       shaw_basis_t* out) {
  return guarded([&] {
    SHAW_REQUIRE(h && out && kv > 0 && ks > 0 && (!wv) == (!ws), SHAW_ERR_ARG, "bad argument");
    auto b = std::make_unique<shaw_basis>();
    const uint64_t kk[2] = {kv, ks};
    const double* ww[2] = {wv, ws};
    for (int d = 0; d < 2; ++d) {
      uint64_t n = 0;
      const uint64_t ns = h->podSize(d, &n);
      SHAW_REQUIRE(ns > 0, SHAW_ERR_STATE, "no snapshots were configured for this dof");
      SHAW_REQUIRE(kk[d] <= ns && kk[d] <= n, SHAW_ERR_ARG, "POD: number of modes exceeds the number of snapshots");
      b->n[d] = n;
      b->k[d] = kk[d];
      b->ns[d] = ns;
      SHAW_HIP(hipMalloc((void**)&b->phi[d], n * kk[d] * sizeof(double)));
      b->sigma[d].assign(ns, 0.0);
      h->podTo(d, kk[d], ww[d], b->phi[d], b->sigma[d].data());
    }
    *out = b.release();
  });
}
int shaw_debug_copy(uint64_t bytes, int reps, double* ms_per_copy) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(ms_per_copy && reps > 0 && bytes >= 16, SHAW_ERR_ARG, "bad argument");
    const uint64_t n = bytes / 16;
    DevBuf<double> a, b;
    a.alloc(2 * n);
    b.alloc(2 * n);
    SHAW_HIP(hipMemset(a.p, 0, n * 16));
    hipEvent_t e0, e1;
    SHAW_HIP(hipEventCreate(&e0));
    SHAW_HIP(hipEventCreate(&e1));
    // one 16-byte element per thread (the kernel's grid-stride loop then runs
    // once): the fastest plain copy shape measured on MI355X
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 0x7fffffffu);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, (const double2*)a.p, (double2*)b.p, n);  // warm
    SHAW_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, (const double2*)a.p, (double2*)b.p, n);
    SHAW_HIP(hipEventRecord(e1, 0));
    SHAW_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SHAW_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

void shaw_fom_destroy(shaw_fom_t h) {
  if (h) guarded([&] { delete h; });
}

}  // extern "C"
