// SpMV row kernels shared by the PCG drivers (krylov.hip, batch.hip): CSR
// rows in feax's layout (feax/csr.py:81-90) and the block-3 form of HEX8
// patterns with vec = 3.
#pragma once

#include <hip/hip_runtime.h>

namespace fx {

template <typename T, int L>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
  return v;
}

// One row, L lanes, value type V, arithmetic type T.
template <typename V, typename T, int L>
__device__ __forceinline__ T csr_row(const int* __restrict__ indptr, const int* __restrict__ indices,
                                     const V* __restrict__ data, const T* __restrict__ x, int row, int gl) {
  const int s = indptr[row], e = indptr[row + 1];
  T acc = T(0);
  for (int k = s + gl; k < e; k += L) acc += static_cast<T>(data[k]) * x[indices[k]];
  return group_sum<T, L>(acc);
}

// Rows 3m..3m+2 of a block-3 pattern; one lane per 3x3 block (L lanes per
// node): one column index, three consecutive x entries and three runs of
// three consecutive values.  DIFF: difference form -- the node's own block
// slot holds S (block row sums), the others act on x_m - x_n:
//   y = sum_{m != n} K_nm (x_m - x_n) + S x_n.
template <typename V, typename T, int L, bool DIFF>
__device__ __forceinline__ void b3_rows(const int* __restrict__ indptr, const int* __restrict__ bcol,
                                        const V* __restrict__ data, const T* __restrict__ x, int m, int gl,
                                        T& y0, T& y1, T& y2) {
  const int r0 = indptr[3 * m];
  const int len = indptr[3 * m + 1] - r0;  // 3 * deg
  const int nb = len / 3;
  const int* bc = bcol + r0 / 9;
  const V* d0 = data + r0;
  const V* d1 = d0 + len;
  const V* d2 = d1 + len;
  T a0 = T(0), a1 = T(0), a2 = T(0);
  T xs0 = T(0), xs1 = T(0), xs2 = T(0);
  if constexpr (DIFF) {
    xs0 = x[3 * m];
    xs1 = x[3 * m + 1];
    xs2 = x[3 * m + 2];
  }
  for (int kb = gl; kb < nb; kb += L) {
    const int c = bc[kb];
    T x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
    if constexpr (DIFF) {
      if (c != m) {
        x0 -= xs0;
        x1 -= xs1;
        x2 -= xs2;
      }
    }
    const int o = 3 * kb;
    a0 += static_cast<T>(d0[o]) * x0 + static_cast<T>(d0[o + 1]) * x1 + static_cast<T>(d0[o + 2]) * x2;
    a1 += static_cast<T>(d1[o]) * x0 + static_cast<T>(d1[o + 1]) * x1 + static_cast<T>(d1[o + 2]) * x2;
    a2 += static_cast<T>(d2[o]) * x0 + static_cast<T>(d2[o + 1]) * x1 + static_cast<T>(d2[o + 2]) * x2;
  }
  y0 = group_sum<T, L>(a0);
  y1 = group_sum<T, L>(a1);
  y2 = group_sum<T, L>(a2);
}

}  // namespace fx
