// CSR SpMV, Jacobi diagonal and Jacobi-preconditioned CG for gfx950.
//
// SpMV replaces CSRMatrix.__matmul__ (feax/csr.py:81-90): y = A x with the
// CSR triple (data, indptr, indices) in feax's exact layout (int32 indptr /
// indices, row-major, sorted columns).
//
// Two kernels:
//  * scalar CSR: a group of L lanes of a 64-wide wavefront owns one row and
//    walks it with coalesced loads of data/indices.
//  * block-3 (HEX8, vec = 3): rows 3m..3m+2 of a node share one column list
//    made of whole 3-blocks.  A group of L lanes owns a node; lane t reads
//    value position k of all three rows (coalesced), one column per 3x3
//    block from `bcol` (nse/9 int32 built once by fx_csr_block3_index), and
//    each gathered x entry is used three times.  Bytes per nonzero:
//    s_v + 4/9 instead of s_v + 4.
// One row / node per group and as many blocks as groups need, with plain
// loads: on MI355X this measured faster than a persistent grid, an
// XCD-contiguous block remap and non-temporal loads (the FP64 block-3 kernel
// reads at the measured streaming rate).
//
// PCG follows jax.scipy.sparse.linalg.cg (feax/solvers/common.py:229-239)
// with scalars kept on the device.  Per iteration on a CSR matrix, three
// launches: the SpMV fused with the p.q block partials; the vector update
// (each block sums the p.q partials itself for alpha) with the r.z / r.r
// partials; the direction update (each block sums those for beta; block 0
// advances the iteration and the stop test).  With feax's matrix-free
// J_bc v (hex8.hip) as the operator: operator, 1-block finalize (alpha),
// update, 1-block finalize (beta, stop test), direction.  FP64 partial
// sums, fixed order: bitwise repeatable.
#include <algorithm>
#include <cstdint>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.h"
#include "pcg_common.h"
#include "reduce.h"
#include "spmv_common.h"

namespace fx {

// Streaming read of data used once (the recorded Krylov directions).
template <typename T>
__device__ __forceinline__ T ldnt(const T* p) {
  return __builtin_nontemporal_load(p);
}

struct Range {
  int begin, end;
};

// Contiguous range of items owned by this (XCD-remapped) block (vector kernels).
__device__ __forceinline__ Range block_range(int ngroups, int& lblock) {
  lblock = xcd_block(blockIdx.x, gridDim.x);
  const int chunk = (ngroups + gridDim.x - 1) / gridDim.x;
  Range r;
  r.begin = min(ngroups, lblock * chunk);
  r.end = min(ngroups, r.begin + chunk);
  return r;
}

// SpMV body: one group of L lanes per row (scalar) or node (block-3), one
// row / node per group; kBlock / L groups per block and as many blocks as
// needed (a grid-stride or persistent loop measured slower here).  When
// DOT, the block partial of sum_i x_i (A x)_i goes to partials[blockIdx.x].
template <typename V, typename T, int L, bool B3, bool DIFF, bool DOT>
__device__ __forceinline__ void spmv_body(int n, const int* __restrict__ indptr, const int* __restrict__ idx,
                                          const V* __restrict__ data, const T* __restrict__ x, T* __restrict__ y,
                                          double* partials) {
  const int ngroups = B3 ? n / 3 : n;
  const int gl = threadIdx.x % L;
  const int g = blockIdx.x * (kBlock / L) + threadIdx.x / L;
  DVec<1> acc;
  acc.v[0] = 0.0;
  if (g < ngroups) {
    if constexpr (B3) {
      T y0, y1, y2;
      b3_rows<V, T, L, DIFF>(indptr, idx, data, x, g, gl, y0, y1, y2);
      if (gl == 0) {
        y[3 * g] = y0;
        y[3 * g + 1] = y1;
        y[3 * g + 2] = y2;
        if constexpr (DOT)
          acc.v[0] = static_cast<double>(x[3 * g]) * static_cast<double>(y0) +
                     static_cast<double>(x[3 * g + 1]) * static_cast<double>(y1) +
                     static_cast<double>(x[3 * g + 2]) * static_cast<double>(y2);
      }
    } else {
      const T v = csr_row<V, T, L>(indptr, idx, data, x, g, gl);
      if (gl == 0) {
        y[g] = v;
        if constexpr (DOT) acc.v[0] = static_cast<double>(x[g]) * static_cast<double>(v);
      }
    }
  }
  if constexpr (DOT) store_partials<1>(block_sum<1, kBlock>(acc), blockIdx.x, partials);
}

template <typename V, typename T, int L, bool B3, bool DIFF>
__global__ __launch_bounds__(kBlock) void spmv_kernel(int n, const int* __restrict__ indptr,
                                                      const int* __restrict__ idx, const V* __restrict__ data,
                                                      const T* __restrict__ x, T* __restrict__ y) {
  spmv_body<V, T, L, B3, DIFF, false>(n, indptr, idx, data, x, y, nullptr);
}

// Lanes per group.  Scalar CSR: from the mean row length.  Block-3: from the
// mean number of 3x3 blocks per node (27 for interior HEX8 nodes): 32 lanes
// (one pass) for FP64 values, 16 lanes (two passes) for FP32 values -- the
// faster mappings measured on MI355X for the 1M-DOF cantilever.
inline int pick_lanes(double mean_row, bool b3, bool v64) {
  if (b3) {
    const double nb = mean_row / 3.0;
    if (v64) return nb > 16 ? 32 : (nb > 8 ? 16 : 8);
    return nb > 8 ? 16 : 8;
  }
  if (mean_row > 96) return 64;
  if (mean_row > 40) return 32;
  if (mean_row > 14) return 16;
  if (mean_row > 6) return 8;
  return 4;
}

inline int lanes_for(int n, long long nse, bool b3, bool v64) {
  // Tuning override (FEAXHIP_SPMV_LANES / FEAXHIP_SPMV_B3_LANES), read once.
  static const int ov_csr = getenv("FEAXHIP_SPMV_LANES") ? atoi(getenv("FEAXHIP_SPMV_LANES")) : 0;
  static const int ov_b3 = getenv("FEAXHIP_SPMV_B3_LANES") ? atoi(getenv("FEAXHIP_SPMV_B3_LANES")) : 0;
  const int ov = b3 ? ov_b3 : ov_csr;
  if (ov == 4 || ov == 8 || ov == 16 || ov == 32 || ov == 64) return ov;
  const double mean = n ? static_cast<double>(nse) / n : 0.0;
  return pick_lanes(mean, b3, v64);
}

// Blocks of an SpMV launch: one group per row / node.
inline int grid_for(int ngroups, int lanes) {
  const int gpb = kBlock / lanes;
  return std::max(1, (ngroups + gpb - 1) / gpb);
}

// Upper bound of grid_for over all lane counts (>= 4): SpMV partial slots.
inline size_t spmv_partial_cap(int n) { return static_cast<size_t>(n) / 4 + 64; }

inline int vec_grid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, kMaxGridBlocks)); }

#define FX_LANE_SWITCH(lanes, ...)                                            \
  switch (lanes) {                                                            \
    case 4: { constexpr int LL = 4; __VA_ARGS__; break; }                     \
    case 8: { constexpr int LL = 8; __VA_ARGS__; break; }                     \
    case 16: { constexpr int LL = 16; __VA_ARGS__; break; }                   \
    case 32: { constexpr int LL = 32; __VA_ARGS__; break; }                   \
    case 64: { constexpr int LL = 64; __VA_ARGS__; break; }                   \
    default: return fail(FX_EINVAL, "unsupported lane count %d", lanes);     \
  }

template <typename V, typename T, bool B3, bool DIFF>
int launch_spmv(hipStream_t st, int n, const int* indptr, const int* idx, const V* data, const T* x, T* y,
                int lanes) {
  const int ngroups = B3 ? n / 3 : n;
  if (ngroups == 0) return FX_OK;
  const int grid = grid_for(ngroups, lanes);
  FX_LANE_SWITCH(lanes, hipLaunchKernelGGL((spmv_kernel<V, T, LL, B3, DIFF>), dim3(grid), dim3(kBlock), 0, st, n,
                                            indptr, idx, data, x, y));
  FX_LAUNCHED("spmv_kernel");
  return FX_OK;
}

// bcol[indptr[3m]/9 + k] = column node of the k-th 3x3 block of node m.
__global__ __launch_bounds__(kBlock) void block3_index_kernel(int nnodes, const int* __restrict__ indptr,
                                                              const int* __restrict__ indices,
                                                              int* __restrict__ bcol) {
  const int m = blockIdx.x * (kBlock / 32) + threadIdx.x / 32;
  if (m >= nnodes) return;
  const int r0 = indptr[3 * m];
  const int deg = (indptr[3 * m + 1] - r0) / 3;
  for (int k = threadIdx.x % 32; k < deg; k += 32) bcol[r0 / 9 + k] = indices[r0 + 3 * k] / 3;
}

// ---------------------------------------------------------------- Jacobi
// 16 lanes per row scan the row's slots for the (i, i) entries; the row
// holds one (the CSR is deduplicated), so the lane sum is that value.
constexpr int kDiagLanes = 16;
inline int diag_grid(int n) { return (n + kBlock / kDiagLanes - 1) / (kBlock / kDiagLanes); }

template <typename T>
__global__ __launch_bounds__(kBlock) void jacobi_kernel(int n, const int* __restrict__ indptr,
                                                        const int* __restrict__ indices, const T* __restrict__ data,
                                                        double shift, T* __restrict__ dinv) {
  const int i = blockIdx.x * (kBlock / kDiagLanes) + threadIdx.x / kDiagLanes;
  const int gl = threadIdx.x % kDiagLanes;
  if (i >= n) return;
  // feax CSRMatrix.diagonal (csr.py:115-121): sum of the row's (i, i) slots.
  double d = 0.0;
  for (int k = indptr[i] + gl; k < indptr[i + 1]; k += kDiagLanes)
    if (indices[k] == i) d += static_cast<double>(data[k]);
  d = group_sum<double, kDiagLanes>(d);
  if (gl == 0) dinv[i] = static_cast<T>(1.0 / (d + shift));
}

template <typename T>
__global__ __launch_bounds__(kBlock) void diag_kernel(int n, const int* __restrict__ indptr,
                                                      const int* __restrict__ indices, const T* __restrict__ data,
                                                      T* __restrict__ diag) {
  const int i = blockIdx.x * (kBlock / kDiagLanes) + threadIdx.x / kDiagLanes;
  const int gl = threadIdx.x % kDiagLanes;
  if (i >= n) return;
  T d = T(0);
  for (int k = indptr[i] + gl; k < indptr[i + 1]; k += kDiagLanes)
    if (indices[k] == i) d += data[k];
  d = group_sum<T, kDiagLanes>(d);
  if (gl == 0) diag[i] = d;
}

// ---------------------------------------------------------------- PCG

constexpr size_t kStateBytes = 256;
// Vector-kernel partial slots (x kMaxGridBlocks doubles): r.z r.r [0, 2), init sums [2, 5).
constexpr size_t kPartialBytes = 5 * kMaxGridBlocks * sizeof(double);

inline size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Workspace: state | vector partials | SpMV partials (one per SpMV block) | r z p q.
inline size_t spmv_partial_bytes(int n) { return align256(spmv_partial_cap(n) * sizeof(double)); }
inline size_t pcg_ws_bytes(int n, size_t elem) {
  return kStateBytes + kPartialBytes + spmv_partial_bytes(n) + 4 * align256(static_cast<size_t>(n) * elem);
}

template <typename V, typename T, int L, bool B3, bool DIFF>
__global__ __launch_bounds__(kBlock) void pcg_spmv_kernel(int n, const int* __restrict__ indptr,
                                                          const int* __restrict__ idx, const V* __restrict__ data,
                                                          const T* __restrict__ p, T* __restrict__ q,
                                                          const PcgState* __restrict__ st, double* partials) {
  if (st->done) return;
  spmv_body<V, T, L, B3, DIFF, true>(n, indptr, idx, data, p, q, partials);
}

__global__ __launch_bounds__(kFinBlock) void pcg_fin_alpha(PcgState* st, const double* __restrict__ partials,
                                                           int nparts, double* __restrict__ rec_pq, int rec_max) {
  if (st->done) return;
  DVec<1> t = sum_partials<1>(partials, nparts);
  if (threadIdx.x == 0) {
    st->alpha = st->gamma / t.v[0];
    if (rec_pq && st->k < rec_max) rec_pq[st->k] = t.v[0];
  }
}

// x += alpha p; r -= alpha q; z = M r; partials of r.z and r.r.
template <typename T>
__global__ __launch_bounds__(kBlock) void pcg_update_kernel(int n, T* __restrict__ x, T* __restrict__ r,
                                                            T* __restrict__ z, const T* __restrict__ p,
                                                            const T* __restrict__ q, const T* __restrict__ dinv,
                                                            const PcgState* __restrict__ st, double* partials,
                                                            float* __restrict__ rec_p, int rec_max) {
  if (st->done) return;
  int lb;
  const Range rg = block_range(n, lb);
  const T alpha = static_cast<T>(st->alpha);
  // Recording: direction p_k of this iteration, FP32, slot k.
  float* rec = (rec_p && st->k < rec_max) ? rec_p + static_cast<size_t>(st->k) * n : nullptr;
  DVec<2> acc;
  acc.v[0] = 0.0;
  acc.v[1] = 0.0;
  for (int i = rg.begin + threadIdx.x; i < rg.end; i += kBlock) {
    const T pi = p[i];
    if (rec) rec[i] = static_cast<float>(pi);
    x[i] += alpha * pi;
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    acc.v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    acc.v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  store_partials<2>(block_sum<2, kBlock>(acc), lb, partials);
}

// The PCG SpMV of a block-3 matrix with its block-column map (fx_b3map):
// workgroup b runs the rows of nodes b G .. b G + G - 1 (G = BS / L) with
// x of its unique column nodes ucol[uptr[b] .. uptr[b + 1]) staged in LDS
// (b3map_rows).  (Dealing the workgroups to the XCDs in contiguous ranges,
// so that neighbouring blocks share an L2, measured 5 % slower; loading the
// node's 3x3 blocks into registers ahead of the x copy, 4 % slower.)  PCG: skip when st->done and write the block partials of
// p . q; else a plain q = A p.
template <typename V, typename T, int L, bool DIFF, bool PCG, int BS>
__global__ __launch_bounds__(BS) void pcg_spmv_map_kernel(int n, const int* __restrict__ indptr,
                                                              const uint16_t* __restrict__ lidx,
                                                              const int* __restrict__ uptr,
                                                              const int* __restrict__ ucol,
                                                              const V* __restrict__ data, const T* __restrict__ p,
                                                              T* __restrict__ q, const PcgState* __restrict__ st,
                                                              double* partials) {
  if constexpr (PCG) {
    if (st->done) return;
  }
  constexpr int G = BS / L;
  extern __shared__ unsigned char smem[];
  __shared__ int own[G];
  const int first = blockIdx.x * G;
  const int u0 = uptr[blockIdx.x];
  DVec<1> acc;
  acc.v[0] = b3map_rows<V, T, L, DIFF, BS>(first, min(first + G, n / 3), indptr, lidx, ucol + u0,
                                           uptr[blockIdx.x + 1] - u0, data, p, q, reinterpret_cast<T*>(smem), own);
  if constexpr (PCG) store_partials<1>(block_sum<1, BS>(acc), blockIdx.x, partials);
}

// Upper-block (symmetric) storage (fx_b3sym): FP32 difference form with
// 8 lanes per node, 64 nodes per 512-thread workgroup, four blocks per lane;
// FP64 plain block-3 values with 16 lanes per node, 32 nodes per workgroup,
// two blocks per lane (the nine FP64 values of four blocks would take the
// registers of two waves).  Degree <= 32 either way.  Workgroups are dealt
// to the XCDs in contiguous node ranges (xcd_block), so the stored block a
// lower block reads was brought into this XCD's L2 by a workgroup shortly
// before: HBM sees the upper half of the matrix once.
template <typename V>
struct SymCfg;
template <>
struct SymCfg<float> {
  static constexpr int L = 8, U = 4, BS = 512;
  static constexpr bool DIFF = true;
};
template <>
struct SymCfg<double> {
  static constexpr int L = 16, U = 2, BS = 512;
  static constexpr bool DIFF = false;
};
template <typename V>
constexpr int sym_nodes() {
  return SymCfg<V>::BS / SymCfg<V>::L;
}
constexpr int kSymThreads = 512, kSymMaxDegree = kSymSlots;
static_assert(SymCfg<float>::L * SymCfg<float>::U == kSymMaxDegree && SymCfg<double>::L * SymCfg<double>::U ==
              kSymMaxDegree, "degree bound");

// Seven or more waves per SIMD (at most 72 VGPRs): the kernel is latency-bound.
template <typename V, bool PCG>
__global__ __launch_bounds__(kSymThreads) __attribute__((amdgpu_waves_per_eu(7, 8))) void spmv_sym_kernel(
    int n, int max_unique, const int* __restrict__ indptr, fx_b3sym sym, const V* __restrict__ vals,
    const double* __restrict__ p, double* __restrict__ q, const PcgState* __restrict__ st, double* partials) {
  using C = SymCfg<V>;
  if constexpr (PCG) {
    if (st->done) return;
  }
  extern __shared__ unsigned char smem[];
  __shared__ int own[sym_nodes<V>()];
  double* xs = reinterpret_cast<double*>(smem);
  int* sup = reinterpret_cast<int*>(xs + 3 * static_cast<size_t>(max_unique));
  DVec<1> acc;
  acc.v[0] = b3sym_rows<V, C::DIFF, C::L, C::BS, C::U>(xcd_block(blockIdx.x, gridDim.x), n / 3, sym.code, sym.uptr,
                                                         sym.ucol, sym.usup, sym.up_ptr, vals, p, q, xs, sup, own);
  if constexpr (PCG) store_partials<1>(block_sum<1, kSymThreads>(acc), blockIdx.x, partials);
}

// Upper blocks of a full block-3 matrix into the compact storage, and a
// bitwise symmetry check: bad[0] counts lower blocks (n, c) that differ from
// the transpose of (c, n), bad[1] blocks whose code does not match the
// pattern (counters zeroed by the caller).  V = float: the difference form
// (block layout); V = double: feax's CSR rows.  8 lanes per node.
constexpr int kCompactLanes = 8, kCompactNodes = kSymThreads / kCompactLanes;
template <typename V>
__global__ __launch_bounds__(kSymThreads) void sym_compact_kernel(int nnodes, const int* __restrict__ indptr,
                                                                  const int* __restrict__ bcol, fx_b3sym sym,
                                                                  const V* __restrict__ data, V* __restrict__ upper,
                                                                  int* __restrict__ bad) {
  constexpr bool BLK = SymCfg<V>::DIFF;
  using U = typename std::conditional<sizeof(V) == 8, uint64_t, uint32_t>::type;
  const int g = blockIdx.x * kCompactNodes + threadIdx.x / kCompactLanes;
  const int gl = threadIdx.x % kCompactLanes;
  if (g >= nnodes) return;
  const int r0 = indptr[3 * g];
  const int len = indptr[3 * g + 1] - r0;
  const int k0 = r0 / 9;
  const int nb = len / 3;
  const int base = sym.up_ptr[g];
  const int nlow = nb - (sym.up_ptr[g + 1] - base);
  int nbad = 0, nsym = 0;
  for (int kb = gl; kb < nb; kb += kCompactLanes) {
    const int c = bcol[k0 + kb];
    const int w = sym.code[kSymSlots * g + kb];
    const int r = w & 31;
    const bool low = (w & 32) != 0;
    V blk[9];
    b3_block<V, BLK>(data, r0, len, kb, blk);
    if (low != (c < g) || low != (kb < nlow)) {
      ++nbad;
    } else if (!low) {
      if (r != kb - nlow) {
        ++nbad;
      } else {
        V* dst = upper + 9 * static_cast<size_t>(base + r);
#pragma unroll
        for (int t = 0; t < 9; ++t) dst[t] = blk[t];
      }
    } else {
      const int c0 = indptr[3 * c];
      const int lenc = indptr[3 * c + 1] - c0;
      const int nlow_c = lenc / 3 - (sym.up_ptr[c + 1] - sym.up_ptr[c]);
      const int km = nlow_c + r;
      if (bcol[c0 / 9 + km] != g) {
        ++nbad;
      } else {
        V mb[9];
        b3_block<V, BLK>(data, c0, lenc, km, mb);
        bool eq = true;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            U a, m;
            __builtin_memcpy(&a, &blk[3 * i + j], sizeof(V));
            __builtin_memcpy(&m, &mb[3 * j + i], sizeof(V));
            eq = eq && a == m;
          }
        nsym += eq ? 0 : 1;
      }
    }
  }
  if (nsym) atomicAdd(bad, nsym);
  if (nbad) atomicAdd(bad + 1, nbad);
}

// CSR iterations: three launches, the two 1-block finalizes folded into
// the vector kernels.  Each of these kernels runs upd_grid(n) blocks of
// kFinBlock threads, and every block sums all partials of the previous
// kernel itself (same fixed order in every block, so all blocks hold the
// same scalar).  gamma alternates between st->g[0] and st->g[1] by the
// parity of the iteration: the direction kernel of iteration k writes
// g[(k + 1) % 2] while its other blocks still read g[k % 2].
constexpr int kUpdBlocks = 256;  // one block per CU (384 / 512 / 1024 measured 1.4-2x slower)
constexpr int kPre = 4;  // rows per thread loaded ahead of the partial sums (1M rows: 3.9 per thread)
inline int upd_grid(int n) {
  // Tuning override (FEAXHIP_UPD_BLOCKS, 64..2048), read once.
  static const int ov = getenv("FEAXHIP_UPD_BLOCKS") ? atoi(getenv("FEAXHIP_UPD_BLOCKS")) : 0;
  const int nb = (ov >= 64 && ov <= kMaxGridBlocks) ? ov : kUpdBlocks;
  return std::max(1, std::min(nb, (n + kFinBlock - 1) / kFinBlock));
}

__device__ __forceinline__ Range fin_block_range(int n, int& lblock) {
  lblock = xcd_block(blockIdx.x, gridDim.x);
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  Range r;
  r.begin = min(n, lblock * chunk);
  r.end = min(n, r.begin + chunk);
  return r;
}

// Rows above which the update kernel reads and writes x, r, q and D^-1
// non-temporally: their five vectors then exceed the L2 of the eight XCDs
// (8 x 4 MB), and keeping them out of the caches leaves the SpMV's matrix
// there (c2, 1M rows: 74.1 -> 72.5 ms); smaller systems keep them in L2.
constexpr int kNtRows = 800000;
template <typename T, bool NT>
__device__ __forceinline__ T vload(const T* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <typename T, bool NT>
__device__ __forceinline__ void vstore(T v, T* p) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// alpha = g[par] / (p.q) from the operator partials; x += alpha p;
// r -= alpha q; z = M r; partials of r.z and r.r.
template <typename T, bool NT>
__global__ __launch_bounds__(kFinBlock) void pcg_update_alpha_kernel(
    int n, int par, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z, const T* __restrict__ p,
    const T* __restrict__ q, const T* __restrict__ dinv, PcgState* __restrict__ st,
    const double* __restrict__ op_partials, int op_nparts, double* partials, float* __restrict__ rec_p,
    double* __restrict__ rec_pq, int rec_max) {
  if (st->done) return;
  int lb;
  const Range rg = fin_block_range(n, lb);
  // The thread's first kPre rows are loaded before the partial sums, so
  // their HBM latency overlaps the (latency-bound) sum of the SpMV partials.
  T pv[kPre] = {}, qv[kPre] = {}, xv[kPre] = {}, rv[kPre] = {}, dv[kPre] = {};
#pragma unroll
  for (int j = 0; j < kPre; ++j) {
    const int i = rg.begin + threadIdx.x + j * kFinBlock;
    if (i < rg.end) {
      pv[j] = p[i];
      qv[j] = vload<T, NT>(q + i);
      xv[j] = vload<T, NT>(x + i);
      rv[j] = vload<T, NT>(r + i);
      if (dinv) dv[j] = vload<T, NT>(dinv + i);
    }
  }
  __shared__ double s_alpha;
  {
    const DVec<1> t = sum_partials<1>(op_partials, op_nparts);
    if (threadIdx.x == 0) {
      s_alpha = st->g[par] / t.v[0];
      if (blockIdx.x == 0) {
        st->alpha = s_alpha;
        if (rec_pq && st->k < rec_max) rec_pq[st->k] = t.v[0];
      }
    }
  }
  __syncthreads();
  const T alpha = static_cast<T>(s_alpha);
  float* rec = (rec_p && st->k < rec_max) ? rec_p + static_cast<size_t>(st->k) * n : nullptr;
  DVec<2> acc;
  acc.v[0] = 0.0;
  acc.v[1] = 0.0;
#pragma unroll
  for (int j = 0; j < kPre; ++j) {
    const int i = rg.begin + threadIdx.x + j * kFinBlock;
    if (i < rg.end) {
      if (rec) rec[i] = static_cast<float>(pv[j]);
      vstore<T, NT>(xv[j] + alpha * pv[j], x + i);
      const T ri = rv[j] - alpha * qv[j];
      vstore<T, NT>(ri, r + i);
      const T zi = dinv ? dv[j] * ri : ri;
      z[i] = zi;
      acc.v[0] += static_cast<double>(ri) * static_cast<double>(zi);
      acc.v[1] += static_cast<double>(ri) * static_cast<double>(ri);
    }
  }
  for (int i = rg.begin + threadIdx.x + kPre * kFinBlock; i < rg.end; i += kFinBlock) {
    const T pi = p[i];
    if (rec) rec[i] = static_cast<float>(pi);
    x[i] += alpha * pi;
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    acc.v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    acc.v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  store_partials<2>(block_sum<2, kFinBlock>(acc), lb, partials);
}

// gamma' = r.z, beta = gamma' / g[par], p = z + beta p; block 0 advances
// the iteration (g[1 - par], k, rs, stop flag).  A block that starts after
// block 0 raised the stop flag skips the (then unused) direction update.
template <typename T>
__global__ __launch_bounds__(kFinBlock) void pcg_dir_beta_kernel(int n, int par, T* __restrict__ p,
                                                                 const T* __restrict__ z, PcgState* __restrict__ st,
                                                                 const double* __restrict__ partials, int nparts) {
  if (st->done) return;
  int lb;
  const Range rg = fin_block_range(n, lb);
  T zv[kPre] = {}, pv[kPre] = {};
#pragma unroll
  for (int j = 0; j < kPre; ++j) {
    const int i = rg.begin + threadIdx.x + j * kFinBlock;
    if (i < rg.end) {
      zv[j] = z[i];
      pv[j] = p[i];
    }
  }
  __shared__ double s_beta;
  {
    const DVec<2> t = sum_partials<2>(partials, nparts);
    if (threadIdx.x == 0) {
      const double g_new = t.v[0];
      s_beta = g_new / st->g[par];
      if (blockIdx.x == 0) {
        const int k = st->k + 1;
        st->g[1 - par] = g_new;
        st->gamma = g_new;
        st->beta = s_beta;
        st->rs = t.v[1];
        st->k = k;
        st->done = !(t.v[1] > st->atol2 && k < st->maxiter);
      }
    }
  }
  __syncthreads();
  const T beta = static_cast<T>(s_beta);
#pragma unroll
  for (int j = 0; j < kPre; ++j) {
    const int i = rg.begin + threadIdx.x + j * kFinBlock;
    if (i < rg.end) p[i] = zv[j] + beta * pv[j];
  }
  for (int i = rg.begin + threadIdx.x + kPre * kFinBlock; i < rg.end; i += kFinBlock) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(kFinBlock) void pcg_fin_beta(PcgState* st, const double* __restrict__ partials,
                                                          int nparts) {
  if (st->done) return;
  DVec<2> t = sum_partials<2>(partials, nparts);
  if (threadIdx.x == 0) {
    const double g_new = t.v[0];
    st->beta = g_new / st->gamma;
    st->gamma = g_new;
    st->rs = t.v[1];
    st->k += 1;
    st->done = !(st->rs > st->atol2 && st->k < st->maxiter);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void pcg_dir_kernel(int n, T* __restrict__ p, const T* __restrict__ z,
                                                         const PcgState* __restrict__ st) {
  if (st->done) return;
  const T beta = static_cast<T>(st->beta);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = z[i] + beta * p[i];
}

// r = b - q (q = A x0), z = M r, p = z; partials of r.z, r.r, b.b.
template <typename T>
__global__ __launch_bounds__(kBlock) void pcg_init_kernel(int n, const T* __restrict__ b, const T* __restrict__ q,
                                                          T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                                                          const T* __restrict__ dinv, double* partials) {
  int lb;
  const Range rg = block_range(n, lb);
  DVec<3> acc;
  acc.v[0] = acc.v[1] = acc.v[2] = 0.0;
  for (int i = rg.begin + threadIdx.x; i < rg.end; i += kBlock) {
    const T bi = b[i];
    const T ri = bi - q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    p[i] = zi;
    acc.v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    acc.v[1] += static_cast<double>(ri) * static_cast<double>(ri);
    acc.v[2] += static_cast<double>(bi) * static_cast<double>(bi);
  }
  store_partials<3>(block_sum<3, kBlock>(acc), lb, partials);
}

__global__ __launch_bounds__(kFinBlock) void pcg_fin_init(PcgState* st, const double* __restrict__ partials,
                                                          int nparts, double tol, double atol, int maxiter) {
  DVec<3> t = sum_partials<3>(partials, nparts);
  if (threadIdx.x == 0) {
    st->alpha = st->beta = 0.0;
    st->gamma = t.v[0];
    st->g[0] = t.v[0];
    st->g[1] = 0.0;
    st->rs = t.v[1];
    st->bb = t.v[2];
    st->atol2 = fmax(tol * tol * t.v[2], atol * atol);
    st->k = 0;
    st->maxiter = maxiter;
    st->done = !(st->rs > st->atol2 && 0 < maxiter);
  }
}

struct PcgArgs {
  int n;
  long long nse;
  int flags;
  const int* indptr;
  const int* idx;  // indices (scalar kernel) or bcol (block-3 kernels)
  float* rec_p;    // optional: directions p_k (rec_max x n FP32)
  double* rec_pq;  // optional: p_k . A p_k
  int rec_max;
};

// PCG operators: apply(x, y) computes y = A x; iter(p, q, st, partials)
// computes q = A p unless st->done and writes grid() block partials of p.q.
template <typename V, typename T, int L, bool B3, bool DIFF>
struct CsrOp {
  int n;
  const int* indptr;
  const int* idx;
  const V* data;
  int grid() const { return grid_for(B3 ? n / 3 : n, L); }
  void apply(hipStream_t s, const T* xv, T* yv) const {
    hipLaunchKernelGGL((spmv_kernel<V, T, L, B3, DIFF>), dim3(grid()), dim3(kBlock), 0, s, n, indptr, idx, data, xv,
                       yv);
  }
  void iter(hipStream_t s, const T* p, T* q, const PcgState* st, double* partials) const {
    hipLaunchKernelGGL((pcg_spmv_kernel<V, T, L, B3, DIFF>), dim3(grid()), dim3(kBlock), 0, s, n, indptr, idx, data,
                       p, q, st, partials);
  }
  static constexpr bool kThreeLaunch = true;
};

// The block-3 PCG operator with a block-column map (LDS-staged x); BS
// threads per workgroup = the map's nodes per block x L.
template <typename V, typename T, int L, bool DIFF, int BS>
struct CsrMapOp {
  int n;
  const int* indptr;
  fx_b3map map;
  const V* data;
  int grid() const { return std::max(1, (n / 3 + BS / L - 1) / (BS / L)); }
  size_t shm() const { return 3 * sizeof(T) * static_cast<size_t>(map.max_unique); }
  void apply(hipStream_t s, const T* xv, T* yv) const {
    hipLaunchKernelGGL((pcg_spmv_map_kernel<V, T, L, DIFF, false, BS>), dim3(grid()), dim3(BS), shm(), s, n,
                       indptr, map.lidx, map.uptr, map.ucol, data, xv, yv, nullptr, nullptr);
  }
  void iter(hipStream_t s, const T* p, T* q, const PcgState* st, double* partials) const {
    hipLaunchKernelGGL((pcg_spmv_map_kernel<V, T, L, DIFF, true, BS>), dim3(grid()), dim3(BS), shm(), s, n,
                       indptr, map.lidx, map.uptr, map.ucol, data, p, q, st, partials);
  }
  static constexpr bool kThreeLaunch = true;
};

// The PCG operator on upper-block storage (fx_b3sym): the FP32 difference
// form (V = float) or the FP64 block-3 matrix (V = double).
template <typename V>
struct CsrSymOp {
  int n;
  const int* indptr;
  fx_b3sym sym;
  const V* vals;
  int grid() const { return std::max(1, (n / 3 + sym_nodes<V>() - 1) / sym_nodes<V>()); }
  size_t shm() const { return (3 * sizeof(double) + sizeof(int)) * static_cast<size_t>(sym.max_unique); }
  void apply(hipStream_t s, const double* xv, double* yv) const {
    hipLaunchKernelGGL((spmv_sym_kernel<V, false>), dim3(grid()), dim3(kSymThreads), shm(), s, n, sym.max_unique,
                       indptr, sym, vals, xv, yv, nullptr, nullptr);
  }
  void iter(hipStream_t s, const double* p, double* q, const PcgState* st, double* partials) const {
    hipLaunchKernelGGL((spmv_sym_kernel<V, true>), dim3(grid()), dim3(kSymThreads), shm(), s, n, sym.max_unique,
                       indptr, sym, vals, p, q, st, partials);
  }
  static constexpr bool kThreeLaunch = true;
};

// Matrix-free J_bc v of the HEX8 mesh (feax's Krylov operator).
struct MatfreeOp {
  static constexpr bool kThreeLaunch = false;
  Hex8Op op;
  int grid() const { return hex8_op_blocks(op.num_nodes); }
  void apply(hipStream_t s, const double* xv, double* yv) const { launch_hex8_op(s, op, xv, yv, nullptr, nullptr); }
  void iter(hipStream_t s, const double* p, double* q, const PcgState* st, double* partials) const {
    launch_hex8_op(s, op, p, q, st, partials);
  }
};

// One iteration (it = its index from 0, for the gamma parity).  CSR
// operators, three launches: q = A p with the p.q partials; alpha and the
// x, r, z updates; beta and the direction.  Matrix-free operator, five:
// q = A p; finalize alpha; updates; finalize beta; direction.
template <typename T, typename Op>
void launch_pcg_iter(hipStream_t s, const PcgArgs& a, const Op& op, int it, int grid_v, const T* dinv, T* x, T* r,
                     T* z, T* p, T* q, PcgState* st, double* partials, double* op_partials, hipEvent_t ev0,
                     hipEvent_t ev1) {
  if (ev0) (void)hipEventRecord(ev0, s);
  op.iter(s, p, q, st, op_partials);
  if (ev1) (void)hipEventRecord(ev1, s);
  if constexpr (Op::kThreeLaunch) {
    const int gu = upd_grid(a.n), par = it & 1;
    if (a.n > kNtRows)
      hipLaunchKernelGGL((pcg_update_alpha_kernel<T, true>), dim3(gu), dim3(kFinBlock), 0, s, a.n, par, x, r, z, p, q,
                         dinv, st, op_partials, op.grid(), partials, a.rec_p, a.rec_pq, a.rec_max);
    else
      hipLaunchKernelGGL((pcg_update_alpha_kernel<T, false>), dim3(gu), dim3(kFinBlock), 0, s, a.n, par, x, r, z, p, q,
                         dinv, st, op_partials, op.grid(), partials, a.rec_p, a.rec_pq, a.rec_max);
    hipLaunchKernelGGL((pcg_dir_beta_kernel<T>), dim3(gu), dim3(kFinBlock), 0, s, a.n, par, p, z, st, partials, gu);
  } else {
    hipLaunchKernelGGL(pcg_fin_alpha, dim3(1), dim3(kFinBlock), 0, s, st, op_partials, op.grid(), a.rec_pq,
                       a.rec_max);
    hipLaunchKernelGGL((pcg_update_kernel<T>), dim3(grid_v), dim3(kBlock), 0, s, a.n, x, r, z, p, q, dinv, st,
                       partials, a.rec_p, a.rec_max);
    hipLaunchKernelGGL(pcg_fin_beta, dim3(1), dim3(kFinBlock), 0, s, st, partials, grid_v);
    hipLaunchKernelGGL((pcg_dir_kernel<T>), dim3(grid_v), dim3(kBlock), 0, s, a.n, p, z, st);
  }
}

template <typename T, typename Op>
int pcg_loop(hipStream_t s, const PcgArgs& a, const Op& op, const T* dinv, const T* b, const T* x0, double tol,
             double atol, int maxiter, int check_every, T* x, int* iters, double* rs, void* workspace) {
  const int n = a.n;
  if (static_cast<size_t>(op.grid()) > spmv_partial_cap(n))
    return fail(FX_ELIMIT, "pcg: operator grid %d exceeds the partial slots", op.grid());
  char* w = static_cast<char*>(workspace);
  PcgState* st = reinterpret_cast<PcgState*>(w);
  double* partials = reinterpret_cast<double*>(w + kStateBytes);
  double* op_partials = reinterpret_cast<double*>(w + kStateBytes + kPartialBytes);
  const size_t vb = align256(static_cast<size_t>(n) * sizeof(T));
  T* r = reinterpret_cast<T*>(w + kStateBytes + kPartialBytes + spmv_partial_bytes(n));
  T* z = reinterpret_cast<T*>(reinterpret_cast<char*>(r) + vb);
  T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(z) + vb);
  T* q = reinterpret_cast<T*>(reinterpret_cast<char*>(p) + vb);
  const int grid_v = vec_grid(n);
  if (x0) {
    if (x0 != x) FX_HIP(hipMemcpyAsync(x, x0, static_cast<size_t>(n) * sizeof(T), hipMemcpyDeviceToDevice, s));
  } else {
    FX_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * sizeof(T), s));
  }
  op.apply(s, x, q);
  FX_LAUNCHED("pcg_operator");
  hipLaunchKernelGGL((pcg_init_kernel<T>), dim3(grid_v), dim3(kBlock), 0, s, n, b, q, r, z, p, dinv,
                     partials + 2 * kMaxGridBlocks);
  hipLaunchKernelGGL(pcg_fin_init, dim3(1), dim3(kFinBlock), 0, s, st, partials + 2 * kMaxGridBlocks, grid_v, tol,
                     atol, maxiter);
  FX_LAUNCHED("pcg_init");
  if (check_every <= 0) check_every = 32;
  // A captured chunk starts at an even iteration (the gamma parity of the
  // three-launch CSR iteration repeats with every replay).
  if (Op::kThreeLaunch && (check_every & 1)) ++check_every;
  struct {
    int k, done;
  } h = {0, 0};
  FX_HIP(hipMemcpyAsync(&h, &st->k, sizeof(h), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  // FX_PCG_TIME: HIP events around every operator launch of the executed
  // iterations; average in rs[1] (microseconds), count in rs[2].
  const bool timing = (a.flags & FX_PCG_TIME) != 0;
  EventGuard evg;
  std::vector<hipEvent_t>& ev = evg.ev;
  if (timing) {
    ev.assign(2 * static_cast<size_t>(check_every), nullptr);
    for (auto& e : ev) FX_HIP(hipEventCreate(&e));
  }
  double t_sum_ms = 0.0;
  long t_count = 0;
  int launched = 0;
  int rc_loop = FX_OK;
  // Without timing, one chunk of check_every iterations is captured once as
  // a HIP graph (on a private stream of default priority) and replayed on s:
  // the same kernels in the same order, fewer launch gaps.  Iterations past
  // convergence or maxiter exit at their first instruction.
  GraphGuard gg;
  if (!timing && !(a.flags & FX_PCG_NO_GRAPH) && !h.done && maxiter >= check_every) {
    hipError_t e = hipStreamCreateWithFlags(&gg.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(gg.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int c = 0; c < check_every; ++c)
        launch_pcg_iter<T, Op>(gg.stream, a, op, c, grid_v, dinv, x, r, z, p, q, st, partials, op_partials, nullptr,
                               nullptr);
      e = hipStreamEndCapture(gg.stream, &gg.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc_loop = hip_fail(e, "pcg graph capture");
  }
  const hipGraphExec_t gexec = gg.exec;
  while (rc_loop == FX_OK && !h.done && launched < maxiter) {
    const int chunk = std::min(check_every, maxiter - launched);
    const int k_before = h.k;
    if (gexec && chunk == check_every) {
      hipError_t e = hipGraphLaunch(gexec, s);
      if (e != hipSuccess) {
        rc_loop = hip_fail(e, "hipGraphLaunch");
        break;
      }
    } else {
      for (int c = 0; c < chunk; ++c) {
        hipEvent_t e0 = timing ? ev[2 * c] : nullptr, e1 = timing ? ev[2 * c + 1] : nullptr;
        launch_pcg_iter<T, Op>(s, a, op, launched + c, grid_v, dinv, x, r, z, p, q, st, partials, op_partials, e0,
                               e1);
      }
    }
    launched += chunk;
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
      rc_loop = hip_fail(le, "pcg_iter");
      break;
    }
    FX_HIP(hipMemcpyAsync(&h, &st->k, sizeof(h), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    if (timing) {
      for (int c = 0; c < chunk && k_before + c < h.k; ++c) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[2 * c], ev[2 * c + 1]) == hipSuccess) {
          t_sum_ms += ms;
          ++t_count;
        }
      }
    }
  }
  if (rc_loop) return rc_loop;
  double hrs = 0.0;
  FX_HIP(hipMemcpyAsync(&hrs, &st->rs, sizeof(double), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  *iters = h.k;
  rs[0] = hrs;
  if (timing) {
    rs[1] = t_count ? 1e3 * t_sum_ms / t_count : 0.0;
    rs[2] = static_cast<double>(t_count);
  }
  return FX_OK;
}

template <typename V, typename T>
int run_pcg(void* stream, int n, long long nse, int flags, const int* indptr, const int* indices, const int* bcol,
            const V* data, const T* dinv, const T* b, const T* x0, double tol, double atol, int maxiter,
            int check_every, T* x, int* iters, double* rs, float* rec_p, double* rec_pq, int rec_max,
            void* workspace, size_t ws_bytes) {
  hipStream_t s = as_stream(stream);
  if ((rec_p == nullptr) != (rec_pq == nullptr) || (rec_p && rec_max < 0))
    return fail(FX_EINVAL, "pcg: recording needs rec_p, rec_pq and rec_max >= 0");
  if (n < 0 || !indptr || !data || !b || !x || !iters || !rs) return fail(FX_EINVAL, "pcg: bad arguments");
  if (maxiter < 0) return fail(FX_EINVAL, "pcg: maxiter < 0");
  const bool b3 = (flags & (FX_PCG_BLOCK3 | FX_PCG_DIFF3)) != 0;
  const bool diff = (flags & FX_PCG_DIFF3) != 0;
  if (b3 && (n % 3 || !bcol)) return fail(FX_EINVAL, "pcg: block-3 kernels need n %% 3 == 0 and bcol");
  if (!b3 && !indices) return fail(FX_EINVAL, "pcg: indices missing");
  const size_t need = pcg_ws_bytes(n, sizeof(T));
  if (!workspace || ws_bytes < need) return fail(FX_EWORKSPACE, "pcg: workspace %zu < %zu bytes", ws_bytes, need);
  if (n == 0) {
    *iters = 0;
    rs[0] = 0.0;
    if (flags & FX_PCG_TIME) rs[1] = rs[2] = 0.0;
    return FX_OK;
  }
  const PcgArgs a{n, nse, flags, indptr, b3 ? bcol : indices, rec_p, rec_pq, rec_p ? rec_max : 0};
  const int lanes = lanes_for(n, nse, b3, sizeof(V) == 8);
  if (diff) {
    FX_LANE_SWITCH(lanes, return pcg_loop<T>(s, a, CsrOp<V, T, LL, true, true>{n, indptr, a.idx, data}, dinv, b, x0,
                                             tol, atol, maxiter, check_every, x, iters, rs, workspace));
  } else if (b3) {
    FX_LANE_SWITCH(lanes, return pcg_loop<T>(s, a, CsrOp<V, T, LL, true, false>{n, indptr, a.idx, data}, dinv, b,
                                             x0, tol, atol, maxiter, check_every, x, iters, rs, workspace));
  } else {
    FX_LANE_SWITCH(lanes, return pcg_loop<T>(s, a, CsrOp<V, T, LL, false, false>{n, indptr, a.idx, data}, dinv, b,
                                             x0, tol, atol, maxiter, check_every, x, iters, rs, workspace));
  }
  return FX_OK;
}

// run_pcg on a block-3 matrix through its block-column map.
template <typename V, typename T>
int run_pcg_map(void* stream, int n, long long nse, int flags, const int* indptr, const fx_b3map* map,
                const V* data, const T* dinv, const T* b, const T* x0, double tol, double atol, int maxiter,
                int check_every, T* x, int* iters, double* rs, float* rec_p, double* rec_pq, int rec_max,
                void* workspace, size_t ws_bytes) {
  hipStream_t s = as_stream(stream);
  if ((rec_p == nullptr) != (rec_pq == nullptr) || (rec_p && rec_max < 0))
    return fail(FX_EINVAL, "pcg: recording needs rec_p, rec_pq and rec_max >= 0");
  if (n < 0 || !indptr || !map || !data || !b || !x || !iters || !rs) return fail(FX_EINVAL, "pcg: bad arguments");
  if (maxiter < 0) return fail(FX_EINVAL, "pcg: maxiter < 0");
  if (!(flags & (FX_PCG_BLOCK3 | FX_PCG_DIFF3)) || n % 3)
    return fail(FX_EINVAL, "pcg_b3map: a block-3 matrix (FX_PCG_BLOCK3 or FX_PCG_DIFF3, n %% 3 == 0)");
  if (!map->lidx || !map->uptr || !map->ucol || map->max_unique < 0)
    return fail(FX_EINVAL, "pcg_b3map: incomplete map");
  const size_t need = pcg_ws_bytes(n, sizeof(T));
  if (!workspace || ws_bytes < need) return fail(FX_EWORKSPACE, "pcg: workspace %zu < %zu bytes", ws_bytes, need);
  const int lanes = lanes_for(n, nse, true, sizeof(V) == 8);
  const int bs = map->nodes_per_block * lanes;
  if (bs != 256 && bs != 512)
    return fail(FX_EINVAL, "pcg_b3map: %d nodes per block x %d lanes is not a 256- or 512-thread workgroup",
                map->nodes_per_block, lanes);
  if (3 * sizeof(T) * static_cast<size_t>(map->max_unique) > 64 * 1024)
    return fail(FX_ELIMIT, "pcg_b3map: %d unique columns per block exceed the LDS staging", map->max_unique);
  if (n == 0) {
    *iters = 0;
    rs[0] = 0.0;
    if (flags & FX_PCG_TIME) rs[1] = rs[2] = 0.0;
    return FX_OK;
  }
  const PcgArgs a{n, nse, flags, indptr, nullptr, rec_p, rec_pq, rec_p ? rec_max : 0};
#define FX_MAP_LOOP(D, B)                                                                                     \
  FX_LANE_SWITCH(lanes, return pcg_loop<T>(s, a, CsrMapOp<V, T, LL, D, B>{n, indptr, *map, data}, dinv, b, x0, tol, \
                                           atol, maxiter, check_every, x, iters, rs, workspace))
  const bool diff = (flags & FX_PCG_DIFF3) != 0;
  if (diff && bs == 512) {
    FX_MAP_LOOP(true, 512);
  } else if (diff) {
    FX_MAP_LOOP(true, 256);
  } else if (bs == 512) {
    FX_MAP_LOOP(false, 512);
  } else {
    FX_MAP_LOOP(false, 256);
  }
#undef FX_MAP_LOOP
  return FX_OK;
}

// Host-side checks of an fx_b3sym against the kernels' fixed shape.
template <typename V>
int check_b3sym(const char* who, int n, const fx_b3sym* sym) {
  if (!sym || !sym->code || !sym->uptr || !sym->ucol || !sym->usup || !sym->up_ptr)
    return fail(FX_EINVAL, "%s: incomplete map", who);
  if (n % 3) return fail(FX_EINVAL, "%s: n %% 3 != 0", who);
  if (sym->nodes_per_block != sym_nodes<V>())
    return fail(FX_EINVAL, "%s: the map has %d nodes per block, the kernels run %d", who, sym->nodes_per_block,
                sym_nodes<V>());
  if (sym->max_degree > kSymMaxDegree || sym->max_degree < 0)
    return fail(FX_ELIMIT, "%s: node degree %d exceeds %d blocks", who, sym->max_degree, kSymMaxDegree);
  if (sym->max_unique < 0 || sym->max_unique > 1023 ||
      (3 * sizeof(double) + sizeof(int)) * static_cast<size_t>(sym->max_unique) > 64 * 1024)
    return fail(FX_ELIMIT, "%s: %d unique columns per block exceed the code or the LDS staging", who,
                sym->max_unique);
  return FX_OK;
}

template <typename V>
int run_pcg_sym(void* stream, int n, long long nse, int flags, const int* indptr, const fx_b3sym* sym,
                const V* vals, const double* dinv, const double* b, const double* x0, double tol, double atol,
                int maxiter, int check_every, double* x, int* iters, double* rs, float* rec_p, double* rec_pq,
                int rec_max, void* workspace, size_t ws_bytes) {
  hipStream_t s = as_stream(stream);
  if ((rec_p == nullptr) != (rec_pq == nullptr) || (rec_p && rec_max < 0))
    return fail(FX_EINVAL, "pcg: recording needs rec_p, rec_pq and rec_max >= 0");
  if (n < 0 || !indptr || !vals || !b || !x || !iters || !rs) return fail(FX_EINVAL, "pcg: bad arguments");
  if (maxiter < 0) return fail(FX_EINVAL, "pcg: maxiter < 0");
  const int want = SymCfg<V>::DIFF ? FX_PCG_DIFF3 : FX_PCG_BLOCK3;
  if (!(flags & want) || (flags & (FX_PCG_DIFF3 | FX_PCG_BLOCK3)) != want)
    return fail(FX_EINVAL, "pcg_b3sym: %s", SymCfg<V>::DIFF ? "FP32 upper-block storage is of the difference form "
                                                              "(FX_PCG_DIFF3)"
                                                            : "FP64 upper-block storage is of the block-3 matrix "
                                                              "(FX_PCG_BLOCK3)");
  if (int rc = check_b3sym<V>("pcg_b3sym", n, sym)) return rc;
  const size_t need = pcg_ws_bytes(n, sizeof(double));
  if (!workspace || ws_bytes < need) return fail(FX_EWORKSPACE, "pcg: workspace %zu < %zu bytes", ws_bytes, need);
  if (n == 0) {
    *iters = 0;
    rs[0] = 0.0;
    if (flags & FX_PCG_TIME) rs[1] = rs[2] = 0.0;
    return FX_OK;
  }
  const PcgArgs a{n, nse, flags, indptr, nullptr, rec_p, rec_pq, rec_p ? rec_max : 0};
  return pcg_loop<double>(s, a, CsrSymOp<V>{n, indptr, *sym, vals}, dinv, b, x0, tol, atol, maxiter, check_every,
                          x, iters, rs, workspace);
}

template <typename V>
int b3sym_compact(void* stream, int32_t n, const int32_t* indptr, const int32_t* bcol, const fx_b3sym* sym,
                  const V* data, V* upper, int32_t* bad) {
  if (int rc = check_b3sym<V>("fx_b3sym_compact", n, sym)) return rc;
  if (!indptr || !bcol || !data || !upper || !bad) return fail(FX_EINVAL, "fx_b3sym_compact: bad arguments");
  hipStream_t s = as_stream(stream);
  FX_HIP(hipMemsetAsync(bad, 0, 2 * sizeof(int32_t), s));
  const int nn = n / 3;
  if (nn == 0) return FX_OK;
  hipLaunchKernelGGL((sym_compact_kernel<V>), dim3((nn + kCompactNodes - 1) / kCompactNodes), dim3(kSymThreads), 0, s,
                     nn, indptr, bcol, *sym, data, upper, bad);
  FX_LAUNCHED("sym_compact_kernel");
  return FX_OK;
}

template <typename V>
int spmv_b3sym(void* stream, int32_t n, const int32_t* indptr, const fx_b3sym* sym, const V* upper, const double* x,
               double* y) {
  if (int rc = check_b3sym<V>("fx_csr_spmv_b3sym", n, sym)) return rc;
  if (!indptr || !upper || !x || !y) return fail(FX_EINVAL, "fx_csr_spmv_b3sym: bad arguments");
  if (n == 0) return FX_OK;
  CsrSymOp<V>{n, indptr, *sym, upper}.apply(as_stream(stream), x, y);
  FX_LAUNCHED("spmv_sym_kernel");
  return FX_OK;
}

// ---------------------------------------------------------------- vectors
__global__ __launch_bounds__(kBlock) void dot_kernel(int n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* partials) {
  int lb;
  const Range rg = block_range(n, lb);
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = rg.begin + threadIdx.x; i < rg.end; i += kBlock) acc.v[0] += a[i] * b[i];
  store_partials<1>(block_sum<1, kBlock>(acc), lb, partials);
}

__global__ __launch_bounds__(kFinBlock) void dot_fin(const double* __restrict__ partials, int nparts, double* out) {
  DVec<1> t = sum_partials<1>(partials, nparts);
  if (threadIdx.x == 0) *out = t.v[0];
}

__global__ __launch_bounds__(kBlock) void axpy_f32_f64_kernel(int n, double scale, const float* __restrict__ x,
                                                              double* __restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] += scale * static_cast<double>(x[i]);
}

__global__ __launch_bounds__(kBlock) void convert_kernel(int n, double scale, const double* __restrict__ x,
                                                         float* __restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] = static_cast<float>(scale * x[i]);
}

__global__ __launch_bounds__(kBlock) void axpby_kernel(int n, double a, const double* __restrict__ x, double b,
                                                       const double* __restrict__ y, double* __restrict__ out) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) out[i] = a * x[i] + b * y[i];
}

// ---------------------------------------------------------------- recycling
// Galerkin projection onto recorded A-conjugate directions:
//   x0 = sum_k p_k (p_k . r) / (p_k . A p_k).
constexpr int kProjSplit = 64;  // blocks per direction for the dot products

__global__ __launch_bounds__(kBlock) void proj_dot_kernel(int n, const float* __restrict__ P,
                                                          const double* __restrict__ r, double* partials) {
  const int k = blockIdx.x, b = blockIdx.y;
  const float* pk = P + static_cast<size_t>(k) * n;
  const int chunk = (n + kProjSplit - 1) / kProjSplit;
  const int beg = b * chunk, end = min(n, beg + chunk);
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = beg + threadIdx.x; i < end; i += kBlock) acc.v[0] += static_cast<double>(pk[i]) * r[i];
  DVec<1> t = block_sum<1, kBlock>(acc);
  if (threadIdx.x == 0) partials[static_cast<size_t>(k) * kProjSplit + b] = t.v[0];
}

__global__ __launch_bounds__(kBlock) void proj_coef_kernel(int nvec, const double* __restrict__ partials,
                                                           const double* __restrict__ pq, double* __restrict__ c) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nvec) return;
  double s = 0.0;
  for (int b = 0; b < kProjSplit; ++b) s += partials[static_cast<size_t>(k) * kProjSplit + b];
  c[k] = pq[k] != 0.0 ? s / pq[k] : 0.0;
}

__global__ __launch_bounds__(kBlock) void proj_combine_kernel(int n, int nvec, const float* __restrict__ P,
                                                              const double* __restrict__ c, double* __restrict__ x0) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  // four independent partial sums (more loads in flight), added in a fixed order
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k = 0;
  for (; k + 3 < nvec; k += 4) {
    s0 += c[k] * static_cast<double>(ldnt(P + static_cast<size_t>(k) * n + i));
    s1 += c[k + 1] * static_cast<double>(ldnt(P + static_cast<size_t>(k + 1) * n + i));
    s2 += c[k + 2] * static_cast<double>(ldnt(P + static_cast<size_t>(k + 2) * n + i));
    s3 += c[k + 3] * static_cast<double>(ldnt(P + static_cast<size_t>(k + 3) * n + i));
  }
  for (; k < nvec; ++k) s0 += c[k] * static_cast<double>(ldnt(P + static_cast<size_t>(k) * n + i));
  x0[i] = (s0 + s1) + (s2 + s3);
}

// x0 = 0 except x0[bc_rows] = bc_vals - u0[bc_rows] (feax/solvers/common.py:23-53).
__global__ __launch_bounds__(kBlock) void bc_lift_kernel(int nbc, const int* __restrict__ rows,
                                                         const double* __restrict__ vals,
                                                         const double* __restrict__ u0, double* __restrict__ x0) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nbc) x0[rows[i]] = vals[i] - (u0 ? u0[rows[i]] : 0.0);
}

}  // namespace fx

namespace fx {

constexpr size_t kNoAllocScratchBytes = size_t(40) << 20;

// Device scratch of fx_pcg_f64_noalloc: the PCG state, reduction partials
// and the r, z, p, q vectors live here instead of in a caller workspace.
// Each GPU holds its own copy for the life of the process.  In host code
// the name is only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch()).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (looked up on the first
// call on each device, with that device current).
int noalloc_scratch(void** out) {
  int dev = 0;
  FX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FX_ELIMIT, "noalloc: device %d >= %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FX_EINVAL, "noalloc: hipGetSymbolAddress returned a null address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

int fx_csr_spmv_f64(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* indices,
                    const double* data, const double* x, double* y) {
  return launch_spmv<double, double, false, false>(as_stream(stream), n, indptr, indices, data, x, y,
                                                   lanes_for(n, nse, false, true));
}

int fx_csr_spmv_f32(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* indices,
                    const float* data, const float* x, float* y) {
  return launch_spmv<float, float, false, false>(as_stream(stream), n, indptr, indices, data, x, y,
                                                 lanes_for(n, nse, false, false));
}

int fx_csr_block3_index(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* bcol) {
  if (n % 3) return fail(FX_EINVAL, "fx_csr_block3_index: n %% 3 != 0");
  const int nn = n / 3;
  if (nn == 0) return FX_OK;
  hipLaunchKernelGGL(block3_index_kernel, dim3((nn + kBlock / 32 - 1) / (kBlock / 32)), dim3(kBlock), 0,
                     as_stream(stream), nn, indptr, indices, bcol);
  FX_LAUNCHED("block3_index_kernel");
  return FX_OK;
}

int fx_csr_spmv_b3_f64(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* bcol,
                       const double* data, const double* x, double* y) {
  if (n % 3) return fail(FX_EINVAL, "spmv_b3: n %% 3 != 0");
  return launch_spmv<double, double, true, false>(as_stream(stream), n, indptr, bcol, data, x, y,
                                                  lanes_for(n, nse, true, true));
}

int fx_csr_spmv_b3_f32(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* bcol,
                       const float* data, const float* x, float* y) {
  if (n % 3) return fail(FX_EINVAL, "spmv_b3: n %% 3 != 0");
  return launch_spmv<float, float, true, false>(as_stream(stream), n, indptr, bcol, data, x, y,
                                                lanes_for(n, nse, true, false));
}

int fx_csr_spmv_diff3_f32f64(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* bcol,
                             const float* data, const double* x, double* y) {
  if (n % 3) return fail(FX_EINVAL, "spmv_diff3: n %% 3 != 0");
  return launch_spmv<float, double, true, true>(as_stream(stream), n, indptr, bcol, data, x, y,
                                                lanes_for(n, nse, true, false));
}

int fx_csr_jacobi_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                      double shift, double* dinv) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL((jacobi_kernel<double>), dim3(diag_grid(n)), dim3(kBlock), 0, as_stream(stream),
                     n, indptr, indices, data, shift, dinv);
  FX_LAUNCHED("jacobi_kernel");
  return FX_OK;
}

int fx_csr_jacobi_f32(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const float* data,
                      double shift, float* dinv) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL((jacobi_kernel<float>), dim3(diag_grid(n)), dim3(kBlock), 0, as_stream(stream),
                     n, indptr, indices, data, shift, dinv);
  FX_LAUNCHED("jacobi_kernel");
  return FX_OK;
}

int fx_csr_diagonal_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                        double* diag) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL((diag_kernel<double>), dim3(diag_grid(n)), dim3(kBlock), 0, as_stream(stream), n,
                     indptr, indices, data, diag);
  FX_LAUNCHED("diag_kernel");
  return FX_OK;
}

int fx_csr_diagonal_f32(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const float* data,
                        float* diag) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL((diag_kernel<float>), dim3(diag_grid(n)), dim3(kBlock), 0, as_stream(stream), n,
                     indptr, indices, data, diag);
  FX_LAUNCHED("diag_kernel");
  return FX_OK;
}

size_t fx_pcg_workspace_bytes(int32_t n, int32_t elem_bytes) {
  return pcg_ws_bytes(std::max(n, 0), static_cast<size_t>(elem_bytes));
}

int fx_pcg_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const int32_t* indices,
               const int32_t* bcol, const double* data, const double* dinv, const double* b, const double* x0,
               double tol, double atol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
               float* rec_p, double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3) return fail(FX_EINVAL, "fx_pcg_f64: FX_PCG_DIFF3 is for FP32 matrices");
  return run_pcg<double, double>(stream, n, nse, flags, indptr, indices, bcol, data, dinv, b, x0, tol, atol, maxiter,
                                 check_every, x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_pcg_f64_noalloc(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                       const int32_t* indices, const int32_t* bcol, const double* data, const double* dinv,
                       const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                       int32_t check_every, double* x, int32_t* iters, double* rs) {
  if (flags & FX_PCG_DIFF3) return fail(FX_EINVAL, "fx_pcg_f64_noalloc: FX_PCG_DIFF3 is for FP32 matrices");
  if (flags & FX_PCG_TIME) return fail(FX_EINVAL, "fx_pcg_f64_noalloc: FX_PCG_TIME creates events");
  const size_t need = pcg_ws_bytes(std::max(n, 0), sizeof(double));
  if (need > kNoAllocScratchBytes)
    return fail(FX_ELIMIT, "fx_pcg_f64_noalloc: n = %d needs %zu bytes of scratch, more than its %zu", n, need,
                kNoAllocScratchBytes);
  void* ws = nullptr;
  const int rc = noalloc_scratch(&ws);
  if (rc) return rc;
  // No graph capture (a graph and its capture stream are allocations): the
  // chunks of check_every iterations are launched directly on the stream.
  return run_pcg<double, double>(stream, n, nse, flags | FX_PCG_NO_GRAPH, indptr, indices, bcol, data, dinv, b, x0,
                                 tol, atol, maxiter, check_every, x, iters, rs, nullptr, nullptr, 0, ws,
                                 kNoAllocScratchBytes);
}

int fx_pcg_f32(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const int32_t* indices,
               const int32_t* bcol, const float* data, const float* dinv, const float* b, const float* x0, double tol,
               double atol, int32_t maxiter, int32_t check_every, float* x, int32_t* iters, double* rs,
               float* rec_p, double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3) return fail(FX_EINVAL, "fx_pcg_f32: use fx_pcg_f32f64 for the difference form");
  return run_pcg<float, float>(stream, n, nse, flags, indptr, indices, bcol, data, dinv, b, x0, tol, atol, maxiter,
                               check_every, x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_pcg_f32f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const int32_t* indices,
                  const int32_t* bcol, const float* data, const double* dinv, const double* b, const double* x0,
                  double tol, double atol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters,
                  double* rs, float* rec_p, double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes) {
  return run_pcg<float, double>(stream, n, nse, flags, indptr, indices, bcol, data, dinv, b, x0, tol, atol, maxiter,
                                check_every, x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_csr_block3_lanes(int32_t n, int64_t nse, int32_t value_bytes) {
  return lanes_for(n, nse, true, value_bytes == 8);
}

int fx_pcg_b3map_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const fx_b3map* map,
                     const double* data, const double* dinv, const double* b, const double* x0, double tol,
                     double atol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
                     float* rec_p, double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3) return fail(FX_EINVAL, "fx_pcg_b3map_f64: FX_PCG_DIFF3 is for FP32 matrices");
  return run_pcg_map<double, double>(stream, n, nse, flags, indptr, map, data, dinv, b, x0, tol, atol, maxiter,
                                     check_every, x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_pcg_b3map_f32f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                        const fx_b3map* map, const float* data, const double* dinv, const double* b, const double* x0,
                        double tol, double atol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters,
                        double* rs, float* rec_p, double* rec_pq, int32_t rec_max, void* workspace,
                        size_t ws_bytes) {
  return run_pcg_map<float, double>(stream, n, nse, flags, indptr, map, data, dinv, b, x0, tol, atol, maxiter,
                                    check_every, x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_b3sym_compact_f32(void* stream, int32_t n, const int32_t* indptr, const int32_t* bcol, const fx_b3sym* sym,
                         const float* data, float* upper, int32_t* bad) {
  return b3sym_compact<float>(stream, n, indptr, bcol, sym, data, upper, bad);
}

int fx_b3sym_compact_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* bcol, const fx_b3sym* sym,
                         const double* data, double* upper, int32_t* bad) {
  return b3sym_compact<double>(stream, n, indptr, bcol, sym, data, upper, bad);
}

int fx_csr_spmv_b3sym_f32f64(void* stream, int32_t n, const int32_t* indptr, const fx_b3sym* sym, const float* upper,
                             const double* x, double* y) {
  return spmv_b3sym<float>(stream, n, indptr, sym, upper, x, y);
}

int fx_csr_spmv_b3sym_f64(void* stream, int32_t n, const int32_t* indptr, const fx_b3sym* sym, const double* upper,
                          const double* x, double* y) {
  return spmv_b3sym<double>(stream, n, indptr, sym, upper, x, y);
}

int fx_pcg_b3sym_f32f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                        const fx_b3sym* sym, const float* upper, const double* dinv, const double* b,
                        const double* x0, double tol, double atol, int32_t maxiter, int32_t check_every, double* x,
                        int32_t* iters, double* rs, float* rec_p, double* rec_pq, int32_t rec_max, void* workspace,
                        size_t ws_bytes) {
  return run_pcg_sym<float>(stream, n, nse, flags, indptr, sym, upper, dinv, b, x0, tol, atol, maxiter, check_every,
                            x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_pcg_b3sym_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const fx_b3sym* sym,
                     const double* upper, const double* dinv, const double* b, const double* x0, double tol,
                     double atol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
                     float* rec_p, double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes) {
  return run_pcg_sym<double>(stream, n, nse, flags, indptr, sym, upper, dinv, b, x0, tol, atol, maxiter, check_every,
                             x, iters, rs, rec_p, rec_pq, rec_max, workspace, ws_bytes);
}

int fx_pcg_matfree_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                       const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                       const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                       int32_t flags, const double* dinv, const double* b, const double* x0, double tol, double atol,
                       int32_t maxiter, int32_t check_every, double* xout, int32_t* iters, double* rs,
                       void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !b ||
      !xout || !iters || !rs)
    return fail(FX_EINVAL, "fx_pcg_matfree_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_pcg_matfree_f64: unknown material %d", mat->kind);
  const bool nh = mat->kind == FX_MAT_NEO_HOOKEAN;
  if (nh && !u) return fail(FX_EINVAL, "fx_pcg_matfree_f64: the Neo-Hookean tangent needs u");
  if (maxiter < 0) return fail(FX_EINVAL, "fx_pcg_matfree_f64: maxiter < 0");
  const int n = 3 * num_nodes;
  const size_t need = pcg_ws_bytes(n, sizeof(double));
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_pcg_matfree_f64: workspace %zu < %zu bytes", ws_bytes, need);
  MatParams mp;
  mp.kind = mat->kind;
  mp.loc = mat->param_loc;
  mp.nu = mat->nu;
  mp.E0 = mat->E0;
  mp.Emin = mat->Emin;
  mp.penal = mat->penal;
  mp.kappa = mat->kappa;
  const MatfreeOp op{Hex8Op{num_nodes, x, y, z, cells, inc_ptr, inc, mp, param, nh ? u : nullptr, bc_mask}};
  const PcgArgs a{n, 0, flags & FX_PCG_TIME, nullptr, nullptr, nullptr, nullptr, 0};
  return pcg_loop<double>(as_stream(stream), a, op, dinv, b, x0, tol, atol, maxiter, check_every, xout, iters, rs,
                          workspace);
}

size_t fx_krylov_project_workspace_bytes(int32_t nvec) {
  return (static_cast<size_t>(std::max(nvec, 0)) * (kProjSplit + 1) + 32) * sizeof(double);
}

int fx_krylov_project_f32(void* stream, int32_t n, int32_t nvec, const float* P, const double* pq, const double* r,
                          double* x0, void* workspace, size_t ws_bytes) {
  if (n < 0 || nvec < 0 || (nvec > 0 && (!P || !pq)) || !r || !x0)
    return fail(FX_EINVAL, "fx_krylov_project_f32: bad arguments");
  if (!workspace || ws_bytes < fx_krylov_project_workspace_bytes(nvec))
    return fail(FX_EWORKSPACE, "fx_krylov_project_f32: workspace too small");
  hipStream_t s = as_stream(stream);
  if (n == 0) return FX_OK;
  if (nvec == 0) {
    FX_HIP(hipMemsetAsync(x0, 0, static_cast<size_t>(n) * sizeof(double), s));
    return FX_OK;
  }
  double* partials = static_cast<double*>(workspace);
  double* c = partials + static_cast<size_t>(nvec) * kProjSplit;
  hipLaunchKernelGGL(proj_dot_kernel, dim3(nvec, kProjSplit), dim3(kBlock), 0, s, n, P, r, partials);
  hipLaunchKernelGGL(proj_coef_kernel, dim3((nvec + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nvec, partials, pq, c);
  hipLaunchKernelGGL(proj_combine_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, nvec, P, c, x0);
  FX_LAUNCHED("proj kernels");
  return FX_OK;
}

int fx_dot_f64(void* stream, int32_t n, const double* a, const double* b, double* out, void* workspace,
               size_t ws_bytes) {
  if (!workspace || ws_bytes < sizeof(double) * kMaxGridBlocks) return fail(FX_EWORKSPACE, "fx_dot_f64: workspace");
  hipStream_t s = as_stream(stream);
  const int grid = vec_grid(n);
  double* partials = static_cast<double*>(workspace);
  hipLaunchKernelGGL(dot_kernel, dim3(grid), dim3(kBlock), 0, s, n, a, b, partials);
  hipLaunchKernelGGL(dot_fin, dim3(1), dim3(kFinBlock), 0, s, partials, grid, out);
  FX_LAUNCHED("dot_kernel");
  return FX_OK;
}

int fx_axpy_f32_into_f64(void* stream, int32_t n, double scale, const float* x, double* y) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL(axpy_f32_f64_kernel, dim3(vec_grid(n)), dim3(kBlock), 0, as_stream(stream), n, scale, x, y);
  FX_LAUNCHED("axpy_f32_f64_kernel");
  return FX_OK;
}

int fx_convert_f64_to_f32(void* stream, int32_t n, double scale, const double* x, float* y) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL(convert_kernel, dim3(vec_grid(n)), dim3(kBlock), 0, as_stream(stream), n, scale, x, y);
  FX_LAUNCHED("convert_kernel");
  return FX_OK;
}

int fx_axpby_f64(void* stream, int32_t n, double a, const double* x, double b, const double* y, double* out) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(vec_grid(n)), dim3(kBlock), 0, as_stream(stream), n, a, x, b, y, out);
  FX_LAUNCHED("axpby_kernel");
  return FX_OK;
}

int fx_bc_lift_f64(void* stream, int32_t n, int32_t nbc, const int32_t* bc_rows, const double* bc_vals,
                   const double* u0, double* x0) {
  if (n <= 0) return FX_OK;
  hipStream_t s = as_stream(stream);
  FX_HIP(hipMemsetAsync(x0, 0, static_cast<size_t>(n) * sizeof(double), s));
  if (nbc > 0) {
    if (!bc_rows || !bc_vals) return fail(FX_EINVAL, "fx_bc_lift_f64: bc arrays missing");
    hipLaunchKernelGGL(bc_lift_kernel, dim3((nbc + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nbc, bc_rows, bc_vals,
                       u0, x0);
    FX_LAUNCHED("bc_lift_kernel");
  }
  return FX_OK;
}

}  // extern "C"
