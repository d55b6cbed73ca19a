// Shared by the HEX8, TET4, QUAD4 and HEX20 element kernels: the per-node gathers of
// element matrices into the node's three CSR rows (deterministic: the
// node's elements in ascending order, no atomics), the row finish with the
// Dirichlet elimination, and the Dirichlet rows of the residual.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "common.h"

namespace fx {
namespace {

constexpr int kNodesPerBlock = 4;  // one wave per node, 4 waves per block
constexpr int kMaxDeg = 64;        // LDS accumulator: 9 * kMaxDeg values per wave

// The node's three CSR rows from its LDS accumulator (acc: row i, block kb,
// column j at i * len + 3 kb + j): Jacobi dinv, the difference form, and the
// store with the Dirichlet elimination of feax/DCboundary.py:147-199 (the
// difference form is stored block by block: its 3x3 blocks contiguous).  The
// constrained columns are found once per neighbour block (one lane per
// block reads the block's column node, a ballot per component collects the
// Dirichlet flags), not once per value.
// First column of neighbour block kb = 64 w + lane of the node's rows
// (V m for column node m), -1 past the node's degree: loaded before the
// element loop so that the load overlaps it.
template <int V, int MAXDEG>
__device__ __forceinline__ void block_first_cols(int lane, int len, int row0, const int* __restrict__ indices,
                                                 int (&col)[(MAXDEG + 63) / 64]) {
#pragma unroll
  for (int w = 0; w < (MAXDEG + 63) / 64; ++w) {
    const int kb = 64 * w + lane;
    col[w] = kb < len / V ? indices[row0 + V * kb] : -V;
  }
}

// Dirichlet flags of the node's V rows and of the V columns of each of its
// neighbour blocks (block kb = 64 w + lane), as stored in bc_mask.  All the
// loads are issued together, unconditionally (a lane past the node's degree
// reads the node's own flags and is ignored), so that a caller can issue
// them early and let their latency overlap other loads.
template <int V, int KW>
struct DirichletFlags {
  uint8_t row[V];
  uint8_t col[V][KW];
};

template <int V, int MAXDEG>
__device__ __forceinline__ DirichletFlags<V, (MAXDEG + 63) / 64> load_dirichlet_flags(
    int nd, const int (&col)[(MAXDEG + 63) / 64], const uint8_t* __restrict__ bc_mask) {
  DirichletFlags<V, (MAXDEG + 63) / 64> f;
  if (bc_mask) {
#pragma unroll
    for (int i = 0; i < V; ++i) f.row[i] = bc_mask[V * nd + i];
#pragma unroll
    for (int w = 0; w < (MAXDEG + 63) / 64; ++w) {
      const int m = col[w] >= 0 ? col[w] / V : nd;
#pragma unroll
      for (int d = 0; d < V; ++d) f.col[d][w] = bc_mask[V * m + d];
    }
  } else {
#pragma unroll
    for (int i = 0; i < V; ++i) f.row[i] = 0;
#pragma unroll
    for (int w = 0; w < (MAXDEG + 63) / 64; ++w)
#pragma unroll
      for (int d = 0; d < V; ++d) f.col[d][w] = 0;
  }
  return f;
}

template <typename T, typename C, int NPE = 8, int V = 3, int SHIFT = 3, int MAXDEG = kMaxDeg>
__device__ __forceinline__ void finish_rows(int nd, int lane, C* acc, int len, int row0, int nval,
                                            const int* __restrict__ inc_ptr, const int* __restrict__ inc,
                                            const uint8_t* __restrict__ rank, const int* __restrict__ indices,
                                            const uint8_t* __restrict__ bc_mask, int symmetric, int diff_form,
                                            double jacobi_shift, double* __restrict__ dinv, T* __restrict__ data,
                                            const int* col0 = nullptr, const int* __restrict__ up_ptr = nullptr,
                                            const DirichletFlags<V, (MAXDEG + 63) / 64>* pre = nullptr) {
  constexpr int kW = (MAXDEG + 63) / 64;
  __builtin_amdgcn_wave_barrier();
  const int deg = len / V;
  int col[kW];
  if (col0) {
#pragma unroll
    for (int w = 0; w < kW; ++w) col[w] = col0[w];
  } else {
    block_first_cols<V, MAXDEG>(lane, len, row0, indices, col);
  }
  // Per neighbour block kb (one lane each): its column node m.  kself = the
  // block of node nd itself (its diagonal block); cb[d][w] bit t: column DOF
  // (m, d) of block kb = 64 w + t is constrained.
  int kself = 0;
  unsigned long long cb[V][kW];
  bool rbc[V];
  const DirichletFlags<V, kW> f = pre ? *pre : load_dirichlet_flags<V, MAXDEG>(nd, col, bc_mask);
#pragma unroll
  for (int i = 0; i < V; ++i) rbc[i] = f.row[i] != 0;
#pragma unroll
  for (int w = 0; w < kW; ++w) {
    const int kb = 64 * w + lane;
    const bool in = kb < deg;
    const int m = col[w] / V;
    const unsigned long long self = __ballot(in && m == nd);
    if (self) kself = 64 * w + __builtin_ctzll(self);
#pragma unroll
    for (int d = 0; d < V; ++d) cb[d][w] = bc_mask ? __ballot(in && f.col[d][w] != 0) : 0ull;
  }
  if (dinv && lane < V) {
    // Jacobi: 1 / (K_bc[r, r] + shift) (feax/solvers/common.py:86-94); BC rows have diagonal 1.
    const int r = V * nd + lane;
    const double d = (bc_mask && bc_mask[r]) ? 1.0 : static_cast<double>(acc[lane * len + V * kself + lane]);
    dinv[r] = 1.0 / (d + jacobi_shift);
  }
  if (diff_form) {
    // Difference form: the diagonal block slot holds S = block row sums of
    // the BC-eliminated matrix.  Unconstrained element rows sum to zero
    // (sum_b grad N_b = 0), so for a free row S[c][d] = -(sum of its entries
    // in eliminated columns (m, d)), added in ascending block order; a BC
    // row has S[c][d] = delta_cd.
    C S = C(0);
    const int c = lane / V, d = lane - V * (lane / V);
    if (lane < V * V) {
      bool rowbc = false;
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (i == c) rowbc = rbc[i];
      if (rowbc) {
        S = (c == d) ? C(1) : C(0);
      } else {
#pragma unroll
        for (int w = 0; w < kW; ++w) {
          unsigned long long bits = 0;
#pragma unroll
          for (int dd = 0; dd < V; ++dd)
            if (dd == d) bits = cb[dd][w];
          while (bits) {
            const int t = __builtin_ctzll(bits);
            bits &= bits - 1;
            S -= acc[c * len + V * (64 * w + t) + d];
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < V * V) acc[c * len + V * kself + d] = S;
    __builtin_amdgcn_wave_barrier();
  }
  // Output position o of the node's V rows: CSR order (row i, block kb,
  // column d at i len + V kb + d), or, for the difference form, block
  // order (the V x V block kb contiguous, row-major, at V^2 kb + V i + d).
  // up_ptr (difference form of a symmetric matrix): only the blocks from
  // the node's own block on (columns >= nd: the lower blocks are the
  // transposes of blocks other nodes store) go to the upper-block storage,
  // from block up_ptr[nd] on.
  const bool up = diff_form && up_ptr;
  const int o_first = up ? V * V * kself : 0;
  const long out0 = up ? static_cast<long>(V * V) * up_ptr[nd] : static_cast<long>(row0);
  // Dirichlet elimination of value (row i, block kb, column d).
  bool any_bc = false;
#pragma unroll
  for (int i = 0; i < V; ++i) any_bc = any_bc || rbc[i];
#pragma unroll
  for (int w = 0; w < kW; ++w)
#pragma unroll
    for (int d = 0; d < V; ++d) any_bc = any_bc || cb[d][w] != 0ull;
  auto eliminate = [&](C v, int i, int kb, int d) {
    const bool own = kb == kself;
    if (!(diff_form && own)) {
      bool zr = false, zc = false;
#pragma unroll
      for (int ii = 0; ii < V; ++ii)
        if (ii == i) zr = rbc[ii];
#pragma unroll
      for (int w = 0; w < kW; ++w)
#pragma unroll
        for (int dd = 0; dd < V; ++dd)
          if (w == (kb >> 6) && dd == d) zc = (cb[dd][w] >> (kb & 63)) & 1ull;
      if (zr || (symmetric && zc)) v = C(0);
      if (zr && own && i == d) v = C(1);
    }
    return v;
  };
  if (diff_form) {
    for (int o = o_first + lane; o < nval; o += 64) {
      const int kb = o / (V * V);
      const int rem = o - V * V * kb;
      const int i = rem / V, d = rem - V * i;
      C v = acc[i * len + V * kb + d];
      if (any_bc) v = eliminate(v, i, kb, d);
      data[out0 + (o - o_first)] = static_cast<T>(v);
    }
  } else {
    // row by row: no division by the (run-time) row length
#pragma unroll
    for (int i = 0; i < V; ++i) {
      for (int p = lane; p < len; p += 64) {
        C v = acc[i * len + p];
        if (any_bc) {
          const int kb = p / V, d = p - V * kb;
          v = eliminate(v, i, kb, d);
        }
        data[out0 + i * len + p] = static_cast<T>(v);
      }
    }
  }
}

// Position of block (a, b), a <= b, in the packed upper block triangle of
// an NPE-node element matrix (row a of blocks starts after the
// NPE + (NPE - 1) + ... + (NPE - a + 1) blocks of the rows above).
template <int NPE>
__host__ __device__ __forceinline__ int packed_block(int a, int b) {
  return a * (2 * NPE - a + 1) / 2 + (b - a);
}

// NPE nodes per element (8: HEX8, 4: TET4 / QUAD4, 20: HEX20), V DOFs per
// node (3, or 2 in the plane): element matrices of V^2 NPE^2 values, blocks
// (a, b) at V^2 (NPE a + b), ranks at NPE (NPE e + a) + b.  PACK: the
// element matrices are exactly symmetric and only their upper block
// triangle is stored (V^2 NPE (NPE + 1) / 2 values per element, block
// (a, b), a <= b, at V^2 packed_block(a, b)); block (a, b) with a > b is
// read as the transpose of block (b, a), which equals it bit for bit.  An
// incidence entry is (e << SHIFT) | a; nodes have at most MAXDEG
// neighbours.
//
// order (nullable): the nodes in a space-filling-curve order; block b then
// runs the nodes order[4 b' ..] of its XCD-contiguous index b', so each XCD
// gathers one compact region of the mesh and an element's workspace, read
// by its NPE nodes, is an L2 hit after its first reader (HEX8 at 1M DOF,
// gather_wide_rows_kernel: FETCH 2.52 -> 1.16 GB with the same time, about
// 0.48 ms: the gather waits on its chain of dependent loads, not on HBM, so
// the assembly passes no order by default).  up_ptr (nullable):
// upper-block store of a symmetric difference form (finish_rows).
// HEX20 runs this kernel; elements of 4 or 8 nodes run
// gather_wide_rows_kernel below.
template <typename T, typename C, int NPE = 8, int V = 3, int SHIFT = 3, int MAXDEG = kMaxDeg, bool PACK = false>
__global__ __launch_bounds__(64 * kNodesPerBlock) void gather_rows_kernel(
    int num_nodes, const int* __restrict__ order, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
    const uint8_t* __restrict__ rank, const int* __restrict__ indptr, const int* __restrict__ indices,
    const C* __restrict__ ke, const uint8_t* __restrict__ bc_mask, int symmetric, int diff_form,
    double jacobi_shift, double* __restrict__ dinv, T* __restrict__ data, const int* __restrict__ up_ptr) {
  __shared__ C acc_all[kNodesPerBlock][9 * MAXDEG];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = (order ? xcd_block(blockIdx.x, gridDim.x) : static_cast<int>(blockIdx.x)) * kNodesPerBlock + wid;
  if (slot >= num_nodes) return;
  const int nd = order ? order[slot] : slot;
  C* acc = acc_all[wid];
  const int row0 = indptr[V * nd];
  const int len = indptr[V * nd + 1] - row0;  // V * deg
  const int nval = V * len;
  if (len > V * MAXDEG) return;  // host guarantees deg <= MAXDEG
  int col[(MAXDEG + 63) / 64];
  block_first_cols<V, MAXDEG>(lane, len, row0, indices, col);
  for (int i = lane; i < nval; i += 64) acc[i] = C(0);
  __builtin_amdgcn_wave_barrier();
  // V^2 NPE values per element: lane idx -> block b = idx / V^2, entry
  // t = idx % V^2; distinct lanes touch distinct accumulator entries.  The
  // values and ranks of kPf elements are loaded before any is added (their
  // loads overlap), then added in ascending element order.
  constexpr int kVal = V * V * NPE, kPer = (kVal + 63) / 64, kPf = 4;  // 8: 35 % slower (occupancy 8 -> 5)
  constexpr long kElem = PACK ? V * V * NPE * (NPE + 1) / 2 : V * V * NPE * NPE;  // workspace values per element
  const int kend = inc_ptr[nd + 1];
  for (int k0 = inc_ptr[nd]; k0 < kend; k0 += kPf) {
    C val[kPf][kPer];
    int pos[kPf][kPer];
#pragma unroll
    for (int u = 0; u < kPf; ++u) {
      const int k = k0 + u;
      const int pk = k < kend ? inc[k] : 0;
      const int e = pk >> SHIFT, a = pk & ((1 << SHIFT) - 1);
      const C* src = ke + kElem * e;
      const uint8_t* rk = rank + NPE * NPE * static_cast<long>(e) + NPE * a;
#pragma unroll
      for (int p = 0; p < kPer; ++p) {
        const int idx = lane + 64 * p;
        const bool live = k < kend && idx < kVal;
        const int b = idx / (V * V), t = idx - V * V * b;
        const int i = t / V, j = t - V * i;
        int off;
        if constexpr (PACK)
          off = a <= b ? V * V * packed_block<NPE>(a, b) + t : V * V * packed_block<NPE>(b, a) + V * j + i;
        else
          off = kVal * a + idx;  // blocks (a, b), b < NPE
        val[u][p] = live ? src[off] : C(0);
        pos[u][p] = live ? i * len + V * rk[b] + j : -1;
      }
    }
#pragma unroll
    for (int u = 0; u < kPf; ++u) {
#pragma unroll
      for (int p = 0; p < kPer; ++p)
        if (pos[u][p] >= 0) acc[pos[u][p]] += val[u][p];
      __builtin_amdgcn_wave_barrier();
    }
  }
  finish_rows<T, C, NPE, V, SHIFT, MAXDEG>(nd, lane, acc, len, row0, nval, inc_ptr, inc, rank, indices, bc_mask,
                                          symmetric, diff_form, jacobi_shift, dinv, data, col, up_ptr);
}

// The same gather for elements of NPE = 4 or 8 nodes (64 / NPE elements per
// pass): lane (k, b) of a pass holds block (a, b) of the node's element k
// (a = the node's local index in it), so one round of loads brings every
// value of up to 64 / NPE elements, and each lane forms its addresses once
// per block instead of once per value.  The blocks are then added into the
// LDS rows element by element (the lanes of element k add their NPE blocks
// while the others wait; the NPE blocks of one element go to distinct
// neighbour slots), in ascending element order: the sums are those of
// gather_rows_kernel bit for bit.  The node's Dirichlet flags are loaded
// with the first pass, so the wave waits on three rounds of dependent loads
// (row range and incidences; columns, element blocks and flags; stores)
// instead of one per element pass and one per flag.  HEX8 at 1M DOF: about
// 250 VALU instructions per node instead of 690, 0.48 ms instead of 0.78.
template <typename T, typename C, int NPE, int V = 3, int SHIFT = 3, int MAXDEG = kMaxDeg, bool PACK = false>
__global__ __launch_bounds__(64 * kNodesPerBlock) void gather_wide_rows_kernel(
    int num_nodes, const int* __restrict__ order, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
    const uint8_t* __restrict__ rank, const int* __restrict__ indptr, const int* __restrict__ indices,
    const C* __restrict__ ke, const uint8_t* __restrict__ bc_mask, int symmetric, int diff_form,
    double jacobi_shift, double* __restrict__ dinv, T* __restrict__ data, const int* __restrict__ up_ptr) {
  static_assert(NPE == 4 || NPE == 8, "lanes (k, b): NPE must divide the wave");
  constexpr int kEpp = 64 / NPE;  // elements per pass
  constexpr long kElem = PACK ? V * V * NPE * (NPE + 1) / 2 : V * V * NPE * NPE;  // workspace values per element
  __shared__ C acc_all[kNodesPerBlock][9 * MAXDEG];
  // the wave index is uniform: the node, its row range and its element
  // range come through the scalar unit
  const int wid = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int slot = (order ? xcd_block(blockIdx.x, gridDim.x) : static_cast<int>(blockIdx.x)) * kNodesPerBlock + wid;
  if (slot >= num_nodes) return;
  const int nd = order ? order[slot] : slot;
  C* acc = acc_all[wid];
  const int row0 = indptr[V * nd];
  const int len = indptr[V * nd + 1] - row0;  // V * deg
  const int nval = V * len;
  if (len > V * MAXDEG) return;  // host guarantees deg <= MAXDEG
  const int kbeg = inc_ptr[nd], kend = inc_ptr[nd + 1];
  const int kl = lane / NPE, b = lane - NPE * kl;
  // lanes past the node's last element load that element again and add nothing
  auto incidence = [&](int k0) { return k0 < kend ? inc[k0 + kl < kend ? k0 + kl : kend - 1] : 0; };
  int col[(MAXDEG + 63) / 64];
  block_first_cols<V, MAXDEG>(lane, len, row0, indices, col);
  int pk = incidence(kbeg);
  for (int i = lane; i < nval; i += 64) acc[i] = C(0);
  // The Dirichlet flags are loaded with the first element pass's values (or
  // alone, for a node without elements): finish_rows reads them.
  DirichletFlags<V, (MAXDEG + 63) / 64> flags;
  if (kbeg >= kend) flags = load_dirichlet_flags<V, MAXDEG>(nd, col, bc_mask);
  for (int k0 = kbeg; k0 < kend; k0 += kEpp) {
    const int e = pk >> SHIFT, a = pk & ((1 << SHIFT) - 1);
    const int q = rank[NPE * NPE * static_cast<long>(e) + NPE * a + b];
    const bool tr = PACK && a > b;  // block (a, b) read as the transpose of block (b, a)
    const C* src = ke + kElem * e + V * V * (PACK ? packed_block<NPE>(tr ? b : a, tr ? a : b) : NPE * a + b);
    C v[V * V];
#pragma unroll
    for (int t = 0; t < V * V; ++t) v[t] = src[t];
    if (k0 == kbeg) flags = load_dirichlet_flags<V, MAXDEG>(nd, col, bc_mask);
    if constexpr (PACK) {
#pragma unroll
      for (int i = 0; i < V; ++i)
#pragma unroll
        for (int j = i + 1; j < V; ++j) {
          const C lo = v[V * i + j], hi = v[V * j + i];
          v[V * i + j] = tr ? hi : lo;
          v[V * j + i] = tr ? lo : hi;
        }
    }
    C* dst = acc + V * q;
    const int nk = kend - k0 < kEpp ? kend - k0 : kEpp;
    pk = incidence(k0 + kEpp);
    __builtin_amdgcn_wave_barrier();
    for (int r = 0; r < nk; ++r) {
      if (kl == r) {
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int j = 0; j < V; ++j) dst[i * len + j] += v[V * i + j];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  finish_rows<T, C, NPE, V, SHIFT, MAXDEG>(nd, lane, acc, len, row0, nval, inc_ptr, inc, rank, indices, bc_mask,
                                          symmetric, diff_form, jacobi_shift, dinv, data, col, up_ptr, &flags);
}

__global__ void bc_res_kernel(int nbc, const int* __restrict__ rows, const double* __restrict__ vals,
                              const double* __restrict__ u, int negate, double* __restrict__ res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbc) return;
  const int r = rows[i];
  const double v = (u ? u[r] : 0.0) - vals[i];
  res[r] = negate ? -v : v;
}

}  // namespace
}  // namespace fx
