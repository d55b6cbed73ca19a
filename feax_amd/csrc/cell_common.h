// Shared by the HEX8, TET4, QUAD4 and HEX20 element kernels: the per-node gather of
// element matrices into the node's three CSR rows (deterministic: the
// node's elements in ascending order, no atomics), the row finish with the
// Dirichlet elimination, and the Dirichlet rows of the residual.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace fx {
namespace {

constexpr int kNodesPerBlock = 4;  // one wave per node, 4 waves per block
constexpr int kMaxDeg = 64;        // LDS accumulator: 9 * kMaxDeg values per wave

// The node's three CSR rows from its LDS accumulator (acc: row i, block kb,
// column j at i * len + 3 kb + j): Jacobi dinv, the difference form, and the
// store with the Dirichlet elimination of feax/DCboundary.py:147-199.
template <typename T, typename C, int NPE = 8, int V = 3, int SHIFT = 3>
__device__ __forceinline__ void finish_rows(int nd, int lane, C* acc, int len, int row0, int nval,
                                            const int* __restrict__ inc_ptr, const int* __restrict__ inc,
                                            const uint8_t* __restrict__ rank, const int* __restrict__ indices,
                                            const uint8_t* __restrict__ bc_mask, int symmetric, int diff_form,
                                            double jacobi_shift, double* __restrict__ dinv, T* __restrict__ data) {
  __builtin_amdgcn_wave_barrier();
  // Position of node nd in its own neighbour list (its diagonal 3x3 block).
  int kself = 0;
  {
    const int pk = inc[inc_ptr[nd]];
    const int a0 = pk & ((1 << SHIFT) - 1);
    kself = rank[NPE * NPE * static_cast<long>(pk >> SHIFT) + (NPE + 1) * a0];
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
    // in eliminated columns (m, d)); a BC row has S[c][d] = delta_cd.
    C S = C(0);
    const int c = lane / V, d = lane - V * (lane / V);
    if (lane < V * V) {
      if (bc_mask && bc_mask[V * nd + c]) {
        S = (c == d) ? C(1) : C(0);
      } else if (bc_mask) {
        for (int kb = 0; kb < len / V; ++kb) {
          const int m = indices[row0 + V * kb] / V;
          if (bc_mask[V * m + d]) S -= acc[c * len + V * kb + d];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < V * V) acc[c * len + V * kself + d] = S;
    __builtin_amdgcn_wave_barrier();
  }
  for (int idx = lane; idx < nval; idx += 64) {
    C v = acc[idx];
    const int slot = row0 + idx;
    const bool own = ((idx % len) / V) == kself;
    if (bc_mask && !(diff_form && own)) {
      const int r = V * nd + idx / len;
      const int c = indices[slot];
      const bool zr = bc_mask[r] != 0;
      if (zr || (symmetric && bc_mask[c] != 0)) v = C(0);
      if (zr && r == c) v = C(1);
    }
    data[slot] = static_cast<T>(v);
  }
}

// NPE nodes per element (8: HEX8, 4: TET4 / QUAD4, 20: HEX20), V DOFs per
// node (3, or 2 in the plane): element matrices of V^2 NPE^2 values, blocks
// (a, b) at V^2 (NPE a + b), ranks at NPE (NPE e + a) + b.  An incidence
// entry is (e << SHIFT) | a; nodes have at most MAXDEG neighbours.
template <typename T, typename C, int NPE = 8, int V = 3, int SHIFT = 3, int MAXDEG = kMaxDeg>
__global__ __launch_bounds__(64 * kNodesPerBlock) void gather_rows_kernel(
    int num_nodes, const int* __restrict__ inc_ptr, const int* __restrict__ inc, const uint8_t* __restrict__ rank,
    const int* __restrict__ indptr, const int* __restrict__ indices, const C* __restrict__ ke,
    const uint8_t* __restrict__ bc_mask, int symmetric, int diff_form, double jacobi_shift,
    double* __restrict__ dinv, T* __restrict__ data) {
  __shared__ C acc_all[kNodesPerBlock][9 * MAXDEG];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = blockIdx.x * kNodesPerBlock + wid;
  if (nd >= num_nodes) return;
  C* acc = acc_all[wid];
  const int row0 = indptr[V * nd];
  const int len = indptr[V * nd + 1] - row0;  // V * deg
  const int nval = V * len;
  if (len > V * MAXDEG) return;  // host guarantees deg <= MAXDEG
  for (int i = lane; i < nval; i += 64) acc[i] = C(0);
  __builtin_amdgcn_wave_barrier();
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    const int e = pk >> SHIFT, a = pk & ((1 << SHIFT) - 1);
    const C* src = ke + V * V * NPE * NPE * static_cast<long>(e) + V * V * NPE * a;  // blocks (a, b), b < NPE
    const uint8_t* rk = rank + NPE * NPE * static_cast<long>(e) + NPE * a;
    // V^2 NPE values: lane idx -> block b = idx / V^2, entry t = idx % V^2;
    // distinct lanes touch distinct accumulator entries.
    for (int idx = lane; idx < V * V * NPE; idx += 64) {
      const int b = idx / (V * V), t = idx - V * V * b;
      const int i = t / V, j = t - V * i;
      acc[i * len + V * rk[b] + j] += src[idx];
    }
    __builtin_amdgcn_wave_barrier();
  }
  finish_rows<T, C, NPE, V, SHIFT>(nd, lane, acc, len, row0, nval, inc_ptr, inc, rank, indices, bc_mask,
                                  symmetric, diff_form, jacobi_shift, dinv, data);
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
