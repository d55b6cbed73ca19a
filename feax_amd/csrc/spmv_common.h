// SpMV row kernels shared by the PCG drivers (krylov.hip, batch.hip): CSR
// rows in feax's layout (feax/csr.py:81-90) and the block-3 form of HEX8
// patterns with vec = 3.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

// Values of 3x3 block kb of the node whose rows start at r0 (len = 3 deg
// values per row): rows of the CSR layout (three runs of three values), or
// the block layout of the difference form (nine consecutive values,
// row-major).  v[3 i + j] = entry (i, j).
template <typename V, bool BLK>
__device__ __forceinline__ void b3_block(const V* __restrict__ data, int r0, int len, int kb, V (&v)[9]) {
  if constexpr (BLK) {
    const V* b = data + r0 + 9 * kb;
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = b[t];
  } else {
    const V* b = data + r0 + 3 * kb;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) v[3 * i + j] = b[i * len + j];
  }
}

// Rows 3m..3m+2 of a block-3 pattern; one lane per 3x3 block (L lanes per
// node): one column index, three consecutive x entries and the block's
// nine values.  DIFF: difference form (values in the block layout) -- the
// node's own block slot holds S (block row sums), the others act on
// x_m - x_n:
//   y = sum_{m != n} K_nm (x_m - x_n) + S x_n.
template <typename V, typename T, int L, bool DIFF>
__device__ __forceinline__ void b3_rows(const int* __restrict__ indptr, const int* __restrict__ bcol,
                                        const V* __restrict__ data, const T* __restrict__ x, int m, int gl,
                                        T& y0, T& y1, T& y2) {
  const int r0 = indptr[3 * m];
  const int len = indptr[3 * m + 1] - r0;  // 3 * deg
  const int nb = len / 3;
  const int* bc = bcol + r0 / 9;
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
    V v[9];
    b3_block<V, DIFF>(data, r0, len, kb, v);
    a0 += static_cast<T>(v[0]) * x0 + static_cast<T>(v[1]) * x1 + static_cast<T>(v[2]) * x2;
    a1 += static_cast<T>(v[3]) * x0 + static_cast<T>(v[4]) * x1 + static_cast<T>(v[5]) * x2;
    a2 += static_cast<T>(v[6]) * x0 + static_cast<T>(v[7]) * x1 + static_cast<T>(v[8]) * x2;
  }
  y0 = group_sum<T, L>(a0);
  y1 = group_sum<T, L>(a1);
  y2 = group_sum<T, L>(a2);
}

// One workgroup of the block-3 SpMV through a block-column map: the rows of
// nodes first .. end - 1 (end - first <= BS / L), q = A p.  The workgroup's
// nu unique column nodes ucol[0 .. nu) (sorted, so the copy reads nearby
// addresses) are copied from p into LDS (xs: 3 nu values) once, and every
// 3x3 block reads its three x values from xs at lidx (the position of its
// column node in that list).  own[i] = the list position of node first + i
// (its diagonal block; the difference form subtracts x_n from the other
// blocks' x).  Products and sums in the order of b3_rows.  Returns this
// thread's share of p . q (nonzero in the first lane of a node only).
// Contains a __syncthreads: every thread of the workgroup calls it.
template <typename V, typename T, int L, bool DIFF, int BS>
__device__ __forceinline__ double b3map_rows(int first, int end, const int* __restrict__ indptr,
                                             const uint16_t* __restrict__ lidx, const int* __restrict__ ucol,
                                             int nu, const V* __restrict__ data, const T* __restrict__ p,
                                             T* __restrict__ q, T* xs, int* own) {
  for (int u = threadIdx.x; u < nu; u += BS) {
    const int c = ucol[u];
    xs[3 * u] = p[3 * c];
    xs[3 * u + 1] = p[3 * c + 1];
    xs[3 * u + 2] = p[3 * c + 2];
    if (c >= first && c < end) own[c - first] = u;
  }
  __syncthreads();
  const int gl = threadIdx.x % L;
  const int g = first + threadIdx.x / L;
  double dot = 0.0;
  if (g < end) {
    const int r0 = indptr[3 * g];
    const int len = indptr[3 * g + 1] - r0;  // 3 * deg
    const int nb = len / 3;
    const uint16_t* li = lidx + r0 / 9;
    const int lo = own[g - first];
    const T xs0 = xs[3 * lo], xs1 = xs[3 * lo + 1], xs2 = xs[3 * lo + 2];
    T a0 = T(0), a1 = T(0), a2 = T(0);
    for (int kb = gl; kb < nb; kb += L) {
      const int l = li[kb];
      T x0 = xs[3 * l], x1 = xs[3 * l + 1], x2 = xs[3 * l + 2];
      if constexpr (DIFF) {
        if (l != lo) {
          x0 -= xs0;
          x1 -= xs1;
          x2 -= xs2;
        }
      }
      V v[9];
      b3_block<V, DIFF>(data, r0, len, kb, v);
      a0 += static_cast<T>(v[0]) * x0 + static_cast<T>(v[1]) * x1 + static_cast<T>(v[2]) * x2;
      a1 += static_cast<T>(v[3]) * x0 + static_cast<T>(v[4]) * x1 + static_cast<T>(v[5]) * x2;
      a2 += static_cast<T>(v[6]) * x0 + static_cast<T>(v[7]) * x1 + static_cast<T>(v[8]) * x2;
    }
    const T y0 = group_sum<T, L>(a0), y1 = group_sum<T, L>(a1), y2 = group_sum<T, L>(a2);
    if (gl == 0) {
      q[3 * g] = y0;
      q[3 * g + 1] = y1;
      q[3 * g + 2] = y2;
      dot = static_cast<double>(xs0) * static_cast<double>(y0) + static_cast<double>(xs1) * static_cast<double>(y1) +
            static_cast<double>(xs2) * static_cast<double>(y2);
    }
  }
  return dot;
}

// Upper-block storage of a bitwise-symmetric block-3 difference-form
// matrix (fx_b3sym): node n's blocks with column node c >= n (its own S
// block first, then its upper neighbours in column order) lie contiguously,
// nine FP32 values each (row-major), from block up_ptr[n]; a lower block
// (n, c), c < n, is read as the transpose of the stored block (c, n).
// code[kSymSlots n + kb] describes node n's block kb (kSymNoBlock past its
// degree): (l << 6) | (lower << 5) | r, with l the position of the column
// node in the workgroup's unique column list and r the position of the
// stored block in its owner's upper list (owner = c for a lower block,
// else n); usup[u] = up_ptr[ucol[u]], the first stored block of the list's
// column node u.
//
// One workgroup of BS threads runs G = BS / L consecutive nodes (workgroup
// index b; the caller deals b to XCDs in contiguous ranges, so that the
// upper blocks a lower block reads were fetched into the same L2 by the
// workgroups just before).  Lane gl of a node takes blocks gl + j L,
// j < U (degree <= U L = kSymSlots, checked by the entry point).  Load
// order (the kernel is latency-bound, so every independent chain starts
// first): the column list entries and their first stored blocks, the
// node's first stored block and its codes (none depends on another load);
// the x values of the first BS columns and the upper blocks; then the LDS
// writes and the barrier; then the lower blocks (their owner's first block
// comes from LDS).  Returns this thread's share of p . q (first lane of a
// node).  Contains a __syncthreads.
// V = float with DIFF: the FP32 difference form (the diagonal slot holds
// the block row sums S, the other blocks act on x_m - x_n); V = double
// without DIFF: the plain block-3 matrix y = sum_m K_nm x_m.
constexpr int kSymSlots = 32;         // code entries per node: the degree bound
constexpr int kSymNoBlock = 0xFFFF;  // code past a node's degree (l <= 1022 keeps real codes below it)
template <typename V, bool DIFF, int L, int BS, int U>
__device__ __forceinline__ double b3sym_rows(int b, int nnodes, const uint16_t* __restrict__ code,
                                             const int* __restrict__ uptr, const int* __restrict__ ucol,
                                             const int* __restrict__ usup, const int* __restrict__ up_ptr,
                                             const V* __restrict__ vals, const double* __restrict__ p,
                                             double* __restrict__ q, double* xs, int* sup, int* own) {
  static_assert(L * U == kSymSlots, "one code slot per lane and block");
  constexpr int G = BS / L;
  const int first = b * G;
  const int end = min(first + G, nnodes);
  const int u0 = uptr[b];
  const int nu = uptr[b + 1] - u0;
  const bool st0 = static_cast<int>(threadIdx.x) < nu;
  int cu = 0, su = 0;
  if (st0) {
    cu = ucol[u0 + threadIdx.x];
    su = usup[u0 + threadIdx.x];
  }
  const int gl = threadIdx.x % L;
  const int g = first + threadIdx.x / L;
  const bool act = g < end;
  const int base = act ? up_ptr[g] : 0;
  int w[U];
#pragma unroll
  for (int j = 0; j < U; ++j) w[j] = act ? static_cast<int>(code[kSymSlots * g + gl + j * L]) : kSymNoBlock;
  double sx0 = 0.0, sx1 = 0.0, sx2 = 0.0;
  if (st0) {
    sx0 = p[3 * cu];
    sx1 = p[3 * cu + 1];
    sx2 = p[3 * cu + 2];
  }
  V v[U][9];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    if (w[j] != kSymNoBlock && !(w[j] & 32)) {
      const V* bp = vals + 9 * static_cast<size_t>(base + (w[j] & 31));
#pragma unroll
      for (int t = 0; t < 9; ++t) v[j][t] = bp[t];
    }
  }
  if (st0) {
    xs[3 * threadIdx.x] = sx0;
    xs[3 * threadIdx.x + 1] = sx1;
    xs[3 * threadIdx.x + 2] = sx2;
    sup[threadIdx.x] = su;
    if (cu >= first && cu < end) own[cu - first] = threadIdx.x;
  }
  for (int u = threadIdx.x + BS; u < nu; u += BS) {
    const int c = ucol[u0 + u];
    xs[3 * u] = p[3 * c];
    xs[3 * u + 1] = p[3 * c + 1];
    xs[3 * u + 2] = p[3 * c + 2];
    sup[u] = usup[u0 + u];
    if (c >= first && c < end) own[c - first] = u;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < U; ++j) {
    if (w[j] != kSymNoBlock && (w[j] & 32)) {
      const V* bp = vals + 9 * static_cast<size_t>(sup[w[j] >> 6] + (w[j] & 31));
#pragma unroll
      for (int t = 0; t < 9; ++t) v[j][t] = bp[t];
    }
  }
  double dot = 0.0;
  if (act) {
    const int lo = own[g - first];
    const double xs0 = xs[3 * lo], xs1 = xs[3 * lo + 1], xs2 = xs[3 * lo + 2];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (w[j] != kSymNoBlock) {
        const int l = w[j] >> 6;
        double x0 = xs[3 * l], x1 = xs[3 * l + 1], x2 = xs[3 * l + 2];
        if constexpr (DIFF) {
          if (l != lo) {
            x0 -= xs0;
            x1 -= xs1;
            x2 -= xs2;
          }
        }
        const bool low = (w[j] & 32) != 0;
        const V* e = v[j];
        const V e01 = low ? e[3] : e[1], e10 = low ? e[1] : e[3];
        const V e02 = low ? e[6] : e[2], e20 = low ? e[2] : e[6];
        const V e12 = low ? e[7] : e[5], e21 = low ? e[5] : e[7];
        a0 += static_cast<double>(e[0]) * x0 + static_cast<double>(e01) * x1 + static_cast<double>(e02) * x2;
        a1 += static_cast<double>(e10) * x0 + static_cast<double>(e[4]) * x1 + static_cast<double>(e12) * x2;
        a2 += static_cast<double>(e20) * x0 + static_cast<double>(e21) * x1 + static_cast<double>(e[8]) * x2;
      }
    }
    const double y0 = group_sum<double, L>(a0), y1 = group_sum<double, L>(a1), y2 = group_sum<double, L>(a2);
    if (gl == 0) {
      q[3 * g] = y0;
      q[3 * g + 1] = y1;
      q[3 * g + 2] = y2;
      dot = xs0 * y0 + xs1 * y1 + xs2 * y2;
    }
  }
  return dot;
}

}  // namespace fx
