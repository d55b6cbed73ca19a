// Batched uneven instances for gfx950: one Jacobi-PCG over many independent
// HEX8 systems at once.
//
// feax batches with jax.vmap over the parameters of one mesh, where a
// batched while_loop runs until every lane has converged (test_vmap.py of
// feax, feax/solvers/common.py:229-239 per lane).  Here the instances --
// different meshes allowed -- are concatenated into one block-diagonal
// system (node ranges seg_node_ptr[s] .. seg_node_ptr[s+1], CSR in feax's
// layout, block-3 form) and every CG scalar is per instance:
//
//  * every block of every kernel belongs to exactly one instance (block
//    tables built on the host), so its partial sums belong to that instance;
//  * an iteration is three launches: q = A p with the p.q partials; alpha
//    and the x, r, z updates; beta and the new direction.  Every block of
//    the two vector kernels sums its own instance's partials of the
//    previous kernel (same fixed order in every block, so all blocks of an
//    instance hold the same scalar); the instance's first vector block
//    advances its state.  gamma alternates between g[0] and g[1] by the
//    parity of the iteration, as in the single-system driver;
//  * each instance has its own alpha, beta, gamma, stopping test (jax cg
//    semantics) and done flag; converged instances' blocks exit at their
//    first instruction while the others iterate;
//  * a chunk of check_every (even) iterations is captured once as a HIP
//    graph and replayed on the caller's stream until all instances are done
//    (one host read of the per-instance flags per chunk);
//  * the SpMV runs L lanes per node and BS threads per workgroup (BS / L
//    consecutive nodes of one instance per workgroup), reading x either
//    through bcol or, with a block-column map built on the device for that
//    partition, from an LDS copy of the workgroup's unique column nodes;
//    FX_PCG_AUTOTUNE times lanes 8 / 16 / 32 x workgroups of 256 / 512
//    threads x map on / off on the instance set and keeps the fastest per
//    problem shape;
//  * optionally each instance's search directions p_k (FP32) and p_k.A p_k
//    are recorded at its own iteration count k, and fx_seg_project_f32
//    forms the per-instance Galerkin projection onto them (Krylov recycling
//    for the later inner solves of a mixed-precision defect correction).
// FP64 partial sums in a fixed order: bitwise repeatable, with or without
// the graph, and independent of the SpMV configuration except through the
// order of the p.q partials (one per SpMV workgroup).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "common.h"
#include "pcg_common.h"
#include "reduce.h"
#include "spmv_common.h"

namespace fx {
namespace {

constexpr int kSegBlock = 256;  // threads per block: init, projection and map-build kernels
constexpr int kSegVec = 512;    // threads per block of the iteration's vector kernels (256 / 1024 slower)
constexpr int kVecRows = 4096;  // rows per vector-kernel block
constexpr int kMapWords = 2048; // column bitmap of the map build: 65536 nodes
constexpr int kMapLds = 62 * 1024;  // LDS staging limit of the map SpMV (bytes; 64 KiB with the static LDS)

struct BlockRange {
  int seg, begin, end;  // instance, [begin, end) of nodes (SpMV) or rows (vector kernels)
};

// Fixed-order sums of cnt partials starting at p[k] (one block of BS
// threads, one block reduction for all NV quantities); valid in thread 0.
template <int NV, int BS>
__device__ __forceinline__ DVec<NV> seg_sums(const double* const (&p)[NV], int cnt) {
  DVec<NV> acc;
#pragma unroll
  for (int k = 0; k < NV; ++k) acc.v[k] = 0.0;
  for (int i = threadIdx.x; i < cnt; i += BS) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc.v[k] += p[k][i];
  }
  return block_sum<NV, BS>(acc);
}

// q = A p over the nodes of one table block (BS / L nodes, L lanes each).
// DOT: skip a converged instance and write the block partial of p . q.
template <typename V, typename T, int L, bool DIFF, bool DOT, int BS>
__global__ __launch_bounds__(BS) void seg_spmv_kernel(const BlockRange* __restrict__ tab,
                                                      const int* __restrict__ indptr, const int* __restrict__ bcol,
                                                      const V* __restrict__ data, const T* __restrict__ x,
                                                      T* __restrict__ y, const PcgState* __restrict__ st,
                                                      double* partials) {
  const int lb = blockIdx.x;
  const BlockRange br = tab[lb];
  if constexpr (DOT) {
    if (st[br.seg].done) return;  // uniform over the block
  }
  const int gl = threadIdx.x % L;
  const int g = br.begin + threadIdx.x / L;
  DVec<1> acc;
  acc.v[0] = 0.0;
  if (g < br.end) {
    T y0, y1, y2;
    b3_rows<V, T, L, DIFF>(indptr, bcol, data, x, g, gl, y0, y1, y2);
    if (gl == 0) {
      y[3 * g] = y0;
      y[3 * g + 1] = y1;
      y[3 * g + 2] = y2;
      if constexpr (DOT)
        acc.v[0] = static_cast<double>(x[3 * g]) * static_cast<double>(y0) +
                   static_cast<double>(x[3 * g + 1]) * static_cast<double>(y1) +
                   static_cast<double>(x[3 * g + 2]) * static_cast<double>(y2);
    }
  }
  if constexpr (DOT) {
    const DVec<1> t = block_sum<1, BS>(acc);
    if (threadIdx.x == 0) partials[lb] = t.v[0];
  }
}

// The same through the block-column map of the table: block b's unique
// column nodes are ucol[e0 .. e0 + ucnt[b]) with e0 = its first 3x3 block
// (indptr[3 begin] / 9; a block has at least as many 3x3 blocks as unique
// columns, so the lists of the blocks never overlap).
template <typename V, typename T, int L, bool DIFF, bool DOT, int BS>
__global__ __launch_bounds__(BS) void seg_spmv_map_kernel(const BlockRange* __restrict__ tab,
                                                          const int* __restrict__ indptr,
                                                          const uint16_t* __restrict__ lidx,
                                                          const int* __restrict__ ucnt, const int* __restrict__ ucol,
                                                          const V* __restrict__ data, const T* __restrict__ x,
                                                          T* __restrict__ y, const PcgState* __restrict__ st,
                                                          double* partials) {
  const int lb = blockIdx.x;
  const BlockRange br = tab[lb];
  if constexpr (DOT) {
    if (st[br.seg].done) return;
  }
  extern __shared__ unsigned char smem[];
  __shared__ int own[BS / L];
  const int e0 = indptr[3 * br.begin] / 9;
  DVec<1> acc;
  acc.v[0] = b3map_rows<V, T, L, DIFF, BS>(br.begin, br.end, indptr, lidx, ucol + e0, ucnt[lb], data, x, y,
                                           reinterpret_cast<T*>(smem), own);
  if constexpr (DOT) {
    const DVec<1> t = block_sum<1, BS>(acc);
    if (threadIdx.x == 0) partials[lb] = t.v[0];
  }
}

// Exclusive prefix sum over the kSegBlock threads of a block; *total = sum.
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[kSegBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kSegBlock / 64; ++w) {
    base += w < wid ? wsum[w] : 0;
    tot += wsum[w];
  }
  *total = tot;
  return base + x - v;
}

// Block-column map of an SpMV table, one workgroup per table block: the
// block's column nodes are marked in an LDS bitmap over [cmin, cmax]; the
// rank of a column among the set bits is its position in the sorted unique
// list.  Writes lidx for every 3x3 block of the block's nodes, the list to
// ucol[e0 ..] and its length to ucnt[b]; info[0] = max list length,
// info[1] = 1 when some block's column span exceeds the bitmap (the map is
// then unusable).
__global__ __launch_bounds__(kSegBlock) void seg_map_build_kernel(const BlockRange* __restrict__ tab,
                                                                  const int* __restrict__ indptr,
                                                                  const int* __restrict__ bcol,
                                                                  uint16_t* __restrict__ lidx,
                                                                  int* __restrict__ ucnt, int* __restrict__ ucol,
                                                                  int* info) {
  __shared__ unsigned bits[kMapWords];
  __shared__ int wpre[kMapWords];
  __shared__ int rmin[kSegBlock / 64], rmax[kSegBlock / 64];
  const BlockRange br = tab[blockIdx.x];
  const int e0 = indptr[3 * br.begin] / 9, e1 = indptr[3 * br.end] / 9;
  int cmin = INT_MAX, cmax = INT_MIN;
  for (int e = e0 + threadIdx.x; e < e1; e += kSegBlock) {
    const int c = bcol[e];
    cmin = min(cmin, c);
    cmax = max(cmax, c);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cmin = min(cmin, __shfl_xor(cmin, off, 64));
    cmax = max(cmax, __shfl_xor(cmax, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    rmin[threadIdx.x >> 6] = cmin;
    rmax[threadIdx.x >> 6] = cmax;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kSegBlock / 64; ++w) {
    cmin = min(cmin, rmin[w]);
    cmax = max(cmax, rmax[w]);
  }
  if (e1 <= e0) {
    if (threadIdx.x == 0) ucnt[blockIdx.x] = 0;
    return;
  }
  const long span = static_cast<long>(cmax) - cmin + 1;
  if (span > 32L * kMapWords) {
    if (threadIdx.x == 0) {
      ucnt[blockIdx.x] = 0;
      atomicOr(info + 1, 1);
    }
    return;
  }
  const int nw = static_cast<int>((span + 31) / 32);
  for (int w = threadIdx.x; w < nw; w += kSegBlock) bits[w] = 0u;
  __syncthreads();
  for (int e = e0 + threadIdx.x; e < e1; e += kSegBlock) {
    const int d = bcol[e] - cmin;
    atomicOr(&bits[d >> 5], 1u << (d & 31));
  }
  __syncthreads();
  // each thread owns a run of consecutive words
  const int per = (nw + kSegBlock - 1) / kSegBlock;
  const int w0 = min(nw, threadIdx.x * per), w1 = min(nw, w0 + per);
  int mine = 0;
  for (int w = w0; w < w1; ++w) mine += __popc(bits[w]);
  int total;
  int base = block_excl_scan(mine, &total);
  for (int w = w0; w < w1; ++w) {
    wpre[w] = base;
    const unsigned word = bits[w];
    unsigned rest = word;
    int pos = base;
    while (rest) {
      const int bit = __ffs(rest) - 1;
      ucol[e0 + pos++] = cmin + 32 * w + bit;
      rest &= rest - 1u;
    }
    base += __popc(word);
  }
  __syncthreads();
  for (int e = e0 + threadIdx.x; e < e1; e += kSegBlock) {
    const int d = bcol[e] - cmin;
    const int w = d >> 5;
    const unsigned below = (1u << (d & 31)) - 1u;
    lidx[e] = static_cast<uint16_t>(wpre[w] + __popc(bits[w] & below));
  }
  if (threadIdx.x == 0) {
    ucnt[blockIdx.x] = total;
    atomicMax(info, total);
  }
}

// Recording of search directions: P[k n + i] = p_k[i] (FP32) for the rows
// of an instance at its iteration k < rec_max, pq[k nseg + s] = p_k.A p_k.
struct SegRecord {
  float* P;
  double* pq;
  int rec_max, n, nseg;
};

// alpha = g[par] / (p.q) of the block's instance (its SpMV partials
// mv_ptr[s] .. mv_ptr[s + 1]); x += alpha p; r -= alpha q; z = M r;
// partials of r.z and r.r.
template <typename T>
__global__ __launch_bounds__(kSegVec) void seg_update_alpha_kernel(
    const BlockRange* __restrict__ vtab, int par, const int* __restrict__ mv_ptr, const int* __restrict__ v_ptr,
    const double* __restrict__ pmv, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z, const T* __restrict__ p,
    const T* __restrict__ q, const T* __restrict__ dinv, PcgState* __restrict__ st, double* prz, double* prr,
    SegRecord rec) {
  const BlockRange br = vtab[blockIdx.x];
  const int s = br.seg;
  if (st[s].done) return;
  __shared__ double s_alpha;
  {
    const double* src[1] = {pmv + mv_ptr[s]};
    const DVec<1> pq = seg_sums<1, kSegVec>(src, mv_ptr[s + 1] - mv_ptr[s]);
    if (threadIdx.x == 0) {
      s_alpha = st[s].g[par] / pq.v[0];
      if (static_cast<int>(blockIdx.x) == v_ptr[s]) {
        st[s].alpha = s_alpha;
        if (rec.pq && st[s].k < rec.rec_max) rec.pq[static_cast<size_t>(st[s].k) * rec.nseg + s] = pq.v[0];
      }
    }
  }
  __syncthreads();
  const T alpha = static_cast<T>(s_alpha);
  const int k = st[s].k;
  float* rp = (rec.P && k < rec.rec_max) ? rec.P + static_cast<size_t>(k) * rec.n : nullptr;
  DVec<2> acc;
  acc.v[0] = acc.v[1] = 0.0;
  for (int i = br.begin + threadIdx.x; i < br.end; i += kSegVec) {
    const T pi = p[i];
    if (rp) rp[i] = static_cast<float>(pi);
    x[i] += alpha * pi;
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    acc.v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    acc.v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const DVec<2> t = block_sum<2, kSegVec>(acc);
  if (threadIdx.x == 0) {
    prz[blockIdx.x] = t.v[0];
    prr[blockIdx.x] = t.v[1];
  }
}

// gamma' = r.z of the block's instance, beta = gamma' / g[par], p = z +
// beta p; the instance's first block advances its state (g[1 - par], k,
// rs, stop flag).  A block that starts after that one raised the stop flag
// skips the (then unused) direction update.
template <typename T>
__global__ __launch_bounds__(kSegVec) void seg_dir_beta_kernel(const BlockRange* __restrict__ vtab, int par,
                                                               const int* __restrict__ v_ptr,
                                                               const double* __restrict__ prz,
                                                               const double* __restrict__ prr, T* __restrict__ p,
                                                               const T* __restrict__ z, PcgState* __restrict__ st) {
  const BlockRange br = vtab[blockIdx.x];
  const int s = br.seg;
  if (st[s].done) return;
  __shared__ double s_beta;
  {
    const int b0 = v_ptr[s], cnt = v_ptr[s + 1] - v_ptr[s];
    const double* src[2] = {prz + b0, prr + b0};
    const DVec<2> t = seg_sums<2, kSegVec>(src, cnt);
    if (threadIdx.x == 0) {
      const double g_new = t.v[0];
      s_beta = g_new / st[s].g[par];
      if (static_cast<int>(blockIdx.x) == b0) {
        PcgState& S = st[s];
        const int k = S.k + 1;
        S.g[1 - par] = g_new;
        S.gamma = g_new;
        S.beta = s_beta;
        S.rs = t.v[1];
        S.k = k;
        S.done = !(t.v[1] > S.atol2 && k < S.maxiter);
      }
    }
  }
  __syncthreads();
  const T beta = static_cast<T>(s_beta);
  for (int i = br.begin + threadIdx.x; i < br.end; i += kSegVec) p[i] = z[i] + beta * p[i];
}

// r = b - q (q = A x0), z = M r, p = z; partials of r.z, r.r, b.b.
template <typename T>
__global__ __launch_bounds__(kSegBlock) void seg_init_kernel(const BlockRange* __restrict__ vtab,
                                                             const T* __restrict__ b, const T* __restrict__ q,
                                                             T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                                                             const T* __restrict__ dinv, double* prz, double* prr,
                                                             double* pbb) {
  const BlockRange br = vtab[blockIdx.x];
  DVec<3> acc;
  acc.v[0] = acc.v[1] = acc.v[2] = 0.0;
  for (int i = br.begin + threadIdx.x; i < br.end; i += kSegBlock) {
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
  const DVec<3> t = block_sum<3, kSegBlock>(acc);
  if (threadIdx.x == 0) {
    prz[blockIdx.x] = t.v[0];
    prr[blockIdx.x] = t.v[1];
    pbb[blockIdx.x] = t.v[2];
  }
}

__global__ __launch_bounds__(kFinBlock) void seg_fin_init(PcgState* st, const int* __restrict__ v_ptr,
                                                          const double* __restrict__ prz,
                                                          const double* __restrict__ prr,
                                                          const double* __restrict__ pbb, double tol, double atol,
                                                          int maxiter) {
  const int s = blockIdx.x;
  const int b0 = v_ptr[s], cnt = v_ptr[s + 1] - v_ptr[s];
  const double* src[3] = {prz + b0, prr + b0, pbb + b0};
  const DVec<3> t = seg_sums<3, kFinBlock>(src, cnt);
  if (threadIdx.x == 0) {
    PcgState& S = st[s];
    S.alpha = S.beta = 0.0;
    S.gamma = t.v[0];
    S.g[0] = t.v[0];
    S.g[1] = 0.0;
    S.rs = t.v[1];
    S.bb = t.v[2];
    S.atol2 = fmax(tol * tol * S.bb, atol * atol);
    S.k = 0;
    S.maxiter = maxiter;
    S.nparts = cnt;
    S.done = !(S.rs > S.atol2 && 0 < maxiter);
  }
}

// Per-instance dot products out[s] = sum_{i in rows of s} a_i b_i.
__global__ __launch_bounds__(kFinBlock) void seg_dot_kernel(const int* __restrict__ row_ptr,
                                                            const double* __restrict__ a,
                                                            const double* __restrict__ b, double* __restrict__ out) {
  const int s = blockIdx.x;
  const int r0 = row_ptr[s], r1 = row_ptr[s + 1];
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += kFinBlock) acc.v[0] += a[i] * b[i];
  const DVec<1> t = block_sum<1, kFinBlock>(acc);
  if (threadIdx.x == 0) out[s] = t.v[0];
}

// Per-instance Galerkin projection onto recorded directions:
//   x0[i] = sum_{k < cnt[s]} c[k, s] P[k, i],  c[k, s] = (p_k . r)_s / pq[k, s]
// for the rows i of instance s.  Dots per (vector block, k), then a fixed
// order sum over the instance's blocks.
__global__ __launch_bounds__(kSegBlock) void seg_proj_dot_kernel(const BlockRange* __restrict__ vtab, int n, int nv,
                                                                 const int* __restrict__ cnt,
                                                                 const float* __restrict__ P,
                                                                 const double* __restrict__ r, double* partials) {
  const BlockRange br = vtab[blockIdx.x];
  const int k = blockIdx.y;
  DVec<1> acc;
  acc.v[0] = 0.0;
  if (k < cnt[br.seg]) {
    const float* pk = P + static_cast<size_t>(k) * n;
    for (int i = br.begin + threadIdx.x; i < br.end; i += kSegBlock)
      acc.v[0] += static_cast<double>(pk[i]) * r[i];
  }
  const DVec<1> t = block_sum<1, kSegBlock>(acc);
  if (threadIdx.x == 0) partials[static_cast<size_t>(k) * nv + blockIdx.x] = t.v[0];
}

__global__ __launch_bounds__(kSegBlock) void seg_proj_coef_kernel(int nseg, int kmax, int nv,
                                                                  const int* __restrict__ vptr,
                                                                  const int* __restrict__ cnt,
                                                                  const double* __restrict__ partials,
                                                                  const double* __restrict__ pq,
                                                                  double* __restrict__ c) {
  const long idx = static_cast<long>(blockIdx.x) * kSegBlock + threadIdx.x;
  if (idx >= static_cast<long>(kmax) * nseg) return;
  const int k = static_cast<int>(idx / nseg), s = static_cast<int>(idx % nseg);
  double v = 0.0;
  if (k < cnt[s]) {
    const double* src = partials + static_cast<size_t>(k) * nv;
    double d = 0.0;
    for (int b = vptr[s]; b < vptr[s + 1]; ++b) d += src[b];
    const double den = pq[static_cast<size_t>(k) * nseg + s];
    v = den != 0.0 ? d / den : 0.0;
  }
  c[idx] = v;
}

__global__ __launch_bounds__(kSegBlock) void seg_proj_combine_kernel(const BlockRange* __restrict__ vtab, int n,
                                                                     int nseg, const int* __restrict__ cnt,
                                                                     const float* __restrict__ P,
                                                                     const double* __restrict__ c,
                                                                     double* __restrict__ x0) {
  const BlockRange br = vtab[blockIdx.x];
  const int K = cnt[br.seg];
  for (int i = br.begin + threadIdx.x; i < br.end; i += kSegBlock) {
    // four independent partial sums (more loads in flight), added in a fixed order
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int k = 0;
    for (; k + 3 < K; k += 4) {
      v0 += c[static_cast<size_t>(k) * nseg + br.seg] * static_cast<double>(P[static_cast<size_t>(k) * n + i]);
      v1 += c[static_cast<size_t>(k + 1) * nseg + br.seg] * static_cast<double>(P[static_cast<size_t>(k + 1) * n + i]);
      v2 += c[static_cast<size_t>(k + 2) * nseg + br.seg] * static_cast<double>(P[static_cast<size_t>(k + 2) * n + i]);
      v3 += c[static_cast<size_t>(k + 3) * nseg + br.seg] * static_cast<double>(P[static_cast<size_t>(k + 3) * n + i]);
    }
    for (; k < K; ++k)
      v0 += c[static_cast<size_t>(k) * nseg + br.seg] * static_cast<double>(P[static_cast<size_t>(k) * n + i]);
    x0[i] = (v0 + v1) + (v2 + v3);
  }
}

// ----------------------------------------------------------------- host side
inline size_t al256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

struct SegLayout {
  int nseg = 0, n = 0, nmv = 0, nv = 0;  // instances, rows, SpMV blocks (worst case), vector blocks
  long long nb3 = 0;                     // 3x3 blocks of the matrix (nse / 9)
  size_t off_st = 0, off_mvtab = 0, off_mvptr = 0, off_vtab = 0, off_vptr = 0, off_rowptr = 0, off_pmv = 0,
         off_pv = 0, off_lidx = 0, off_ucol = 0, off_ucnt = 0, off_info = 0, off_vec = 0, vec_bytes = 0, total = 0;
};

// Worst-case SpMV block count over the configurations used (at least 8
// nodes per block: 32 lanes x 256 threads).
inline int mv_blocks_cap(const int* seg_node_ptr, int nseg) {
  long total = 0;
  for (int s = 0; s < nseg; ++s) total += (seg_node_ptr[s + 1] - seg_node_ptr[s] + 7) / 8;
  return static_cast<int>(std::max<long>(total, 1));
}

inline int v_blocks(const int* seg_node_ptr, int nseg) {
  long total = 0;
  for (int s = 0; s < nseg; ++s) total += (3L * (seg_node_ptr[s + 1] - seg_node_ptr[s]) + kVecRows - 1) / kVecRows;
  return static_cast<int>(std::max<long>(total, 1));
}

inline SegLayout seg_layout(int nseg, const int* seg_node_ptr, long long nse, size_t elem) {
  SegLayout L;
  L.nseg = nseg;
  L.n = 3 * seg_node_ptr[nseg];
  L.nmv = mv_blocks_cap(seg_node_ptr, nseg);
  L.nv = v_blocks(seg_node_ptr, nseg);
  L.nb3 = std::max<long long>(nse / 9, 1);
  size_t o = 0;
  L.off_st = o;
  o += al256(static_cast<size_t>(nseg) * sizeof(PcgState));
  L.off_mvtab = o;
  o += al256(static_cast<size_t>(L.nmv) * sizeof(BlockRange));
  L.off_mvptr = o;
  o += al256(static_cast<size_t>(nseg + 1) * sizeof(int));
  L.off_vtab = o;
  o += al256(static_cast<size_t>(L.nv) * sizeof(BlockRange));
  L.off_vptr = o;
  o += al256(static_cast<size_t>(nseg + 1) * sizeof(int));
  L.off_rowptr = o;
  o += al256(static_cast<size_t>(nseg + 1) * sizeof(int));
  L.off_pmv = o;
  o += al256(static_cast<size_t>(L.nmv) * sizeof(double));
  L.off_pv = o;
  o += al256(3 * static_cast<size_t>(L.nv) * sizeof(double));
  L.off_lidx = o;
  o += al256(static_cast<size_t>(L.nb3) * sizeof(uint16_t));
  L.off_ucol = o;
  o += al256(static_cast<size_t>(L.nb3) * sizeof(int));
  L.off_ucnt = o;
  o += al256(static_cast<size_t>(L.nmv) * sizeof(int));
  L.off_info = o;
  o += al256(2 * sizeof(int));
  L.off_vec = o;
  L.vec_bytes = al256(static_cast<size_t>(L.n) * elem);
  o += 4 * L.vec_bytes;
  L.total = o;
  return L;
}

// SpMV block table: gpb consecutive nodes of one instance per block.
inline void mv_table(int nseg, const int* seg_node_ptr, int gpb, std::vector<BlockRange>& tab, std::vector<int>& ptr) {
  tab.clear();
  ptr.assign(nseg + 1, 0);
  for (int s = 0; s < nseg; ++s) {
    ptr[s] = static_cast<int>(tab.size());
    for (int a = seg_node_ptr[s]; a < seg_node_ptr[s + 1]; a += gpb)
      tab.push_back(BlockRange{s, a, std::min(a + gpb, seg_node_ptr[s + 1])});
  }
  ptr[nseg] = static_cast<int>(tab.size());
}

inline void v_table(int nseg, const int* seg_node_ptr, std::vector<BlockRange>& tab, std::vector<int>& ptr,
                    std::vector<int>& rowptr) {
  tab.clear();
  ptr.assign(nseg + 1, 0);
  rowptr.assign(nseg + 1, 0);
  for (int s = 0; s < nseg; ++s) {
    ptr[s] = static_cast<int>(tab.size());
    const int r0 = 3 * seg_node_ptr[s], r1 = 3 * seg_node_ptr[s + 1];
    rowptr[s] = r0;
    for (int a = r0; a < r1; a += kVecRows) tab.push_back(BlockRange{s, a, std::min(a + kVecRows, r1)});
  }
  ptr[nseg] = static_cast<int>(tab.size());
  rowptr[nseg] = 3 * seg_node_ptr[nseg];
}

// SpMV configuration: lanes per node, threads per workgroup, map on / off.
struct SegCfg {
  int lanes, bs;
  bool map;
};

// Block-column map storage (in the workspace) and the built map.
struct SegMap {
  uint16_t* lidx;
  int *ucol, *ucnt, *info;
  int max_unique;  // longest unique-column list; -1: unusable
};

// Builds the map of the SpMV table dtab (nb blocks, already on the device)
// and reads back whether it is usable: every block's column span within
// the bitmap and its LDS copy (3 x max_unique values of T) within kMapLds.
template <typename T>
int build_seg_map(hipStream_t s, const BlockRange* dtab, int nb, const int* indptr, const int* bcol, SegMap& m) {
  FX_HIP(hipMemsetAsync(m.info, 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(seg_map_build_kernel, dim3(nb), dim3(kSegBlock), 0, s, dtab, indptr, bcol, m.lidx, m.ucnt,
                     m.ucol, m.info);
  FX_LAUNCHED("seg_map_build_kernel");
  int h[2] = {0, 0};
  FX_HIP(hipMemcpyAsync(h, m.info, sizeof h, hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  m.max_unique = (h[1] || 3 * sizeof(T) * static_cast<size_t>(h[0]) > static_cast<size_t>(kMapLds)) ? -1 : h[0];
  return FX_OK;
}

template <typename V, typename T, int L, bool DIFF, int BS, bool MAP>
struct SegOp {
  const BlockRange* tab;
  int nblocks;
  const int *indptr, *bcol;
  const V* data;
  SegMap map;
  size_t shm() const { return MAP ? 3 * sizeof(T) * static_cast<size_t>(map.max_unique) : 0; }
  template <bool DOT>
  void launch(hipStream_t s, const T* x, T* y, const PcgState* st, double* partials) const {
    if constexpr (MAP)
      hipLaunchKernelGGL((seg_spmv_map_kernel<V, T, L, DIFF, DOT, BS>), dim3(nblocks), dim3(BS), shm(), s, tab,
                         indptr, map.lidx, map.ucnt, map.ucol, data, x, y, st, partials);
    else
      hipLaunchKernelGGL((seg_spmv_kernel<V, T, L, DIFF, DOT, BS>), dim3(nblocks), dim3(BS), 0, s, tab, indptr, bcol,
                         data, x, y, st, partials);
  }
  void apply(hipStream_t s, const T* x, T* y) const { launch<false>(s, x, y, nullptr, nullptr); }
  void iter(hipStream_t s, const T* p, T* q, const PcgState* st, double* partials) const {
    launch<true>(s, p, q, st, partials);
  }
};

template <int K>
using IntC = std::integral_constant<int, K>;
template <bool K>
using BoolC = std::integral_constant<bool, K>;

// f(IntC<L>, IntC<BS>, BoolC<MAP>) for the configuration c.
template <typename F>
int with_cfg(const SegCfg& c, F&& f) {
  auto by_bs = [&](auto lc) -> int {
    if (c.bs == 512) return c.map ? f(lc, IntC<512>{}, BoolC<true>{}) : f(lc, IntC<512>{}, BoolC<false>{});
    if (c.bs == 256) return c.map ? f(lc, IntC<256>{}, BoolC<true>{}) : f(lc, IntC<256>{}, BoolC<false>{});
    return fail(FX_EINVAL, "pcg_batched: %d threads per workgroup", c.bs);
  };
  switch (c.lanes) {
    case 8:
      return by_bs(IntC<8>{});
    case 16:
      return by_bs(IntC<16>{});
    case 32:
      return by_bs(IntC<32>{});
    default:
      return fail(FX_EINVAL, "pcg_batched: lane count %d", c.lanes);
  }
}

struct SegBuffers {
  PcgState* st;
  BlockRange *mvtab, *vtab;
  int *mvptr, *vptr;
  double *pmv, *prz, *prr, *pbb;
  int nseg, nv;
  SegRecord rec;
};

// Iteration `it` (from 0; its parity selects the gamma slot).
template <typename T, typename Op>
void seg_iteration(hipStream_t s, const Op& op, const SegBuffers& B, int it, const T* dinv, T* x, T* r, T* z, T* p,
                   T* q) {
  const int par = it & 1;
  op.iter(s, p, q, B.st, B.pmv);
  hipLaunchKernelGGL((seg_update_alpha_kernel<T>), dim3(B.nv), dim3(kSegVec), 0, s, B.vtab, par, B.mvptr, B.vptr,
                     B.pmv, x, r, z, p, q, dinv, B.st, B.prz, B.prr, B.rec);
  hipLaunchKernelGGL((seg_dir_beta_kernel<T>), dim3(B.nv), dim3(kSegVec), 0, s, B.vtab, par, B.vptr, B.prz, B.prr,
                     p, z, B.st);
}

template <typename T, typename Op>
int seg_pcg_loop(hipStream_t s, const Op& op, const SegBuffers& B, const T* dinv, const T* b, const T* x0, T* x,
                 T* r, T* z, T* p, T* q, int n, double tol, double atol, int maxiter, int check_every, bool graph,
                 int* iters, double* rs) {
  if (x0) {
    if (x0 != x) FX_HIP(hipMemcpyAsync(x, x0, static_cast<size_t>(n) * sizeof(T), hipMemcpyDeviceToDevice, s));
  } else {
    FX_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * sizeof(T), s));
  }
  op.apply(s, x, q);
  hipLaunchKernelGGL((seg_init_kernel<T>), dim3(B.nv), dim3(kSegBlock), 0, s, B.vtab, b, q, r, z, p, dinv, B.prz,
                     B.prr, B.pbb);
  hipLaunchKernelGGL(seg_fin_init, dim3(B.nseg), dim3(kFinBlock), 0, s, B.st, B.vptr, B.prz, B.prr, B.pbb, tol, atol,
                     maxiter);
  FX_LAUNCHED("seg_pcg_init");
  if (check_every <= 0) check_every = 32;
  if (check_every & 1) ++check_every;  // every replay starts at an even iteration
  std::vector<PcgState> h(B.nseg);
  auto all_done = [&]() {
    for (const auto& S : h)
      if (!S.done) return false;
    return true;
  };
  FX_HIP(hipMemcpyAsync(h.data(), B.st, B.nseg * sizeof(PcgState), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  GraphGuard gg;
  int rc = FX_OK;
  if (graph && !all_done()) {
    // Capture one chunk of iterations on a private stream (default
    // priority); it is replayed on the caller's stream.
    hipError_t e = hipStreamCreateWithFlags(&gg.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(gg.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int c = 0; c < check_every; ++c) seg_iteration<T, Op>(gg.stream, op, B, c, dinv, x, r, z, p, q);
      e = hipStreamEndCapture(gg.stream, &gg.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "seg_pcg graph capture");
  }
  const hipGraphExec_t ge = gg.exec;
  int launched = 0;
  while (rc == FX_OK && !all_done() && launched < maxiter) {
    if (ge) {
      hipError_t e = hipGraphLaunch(ge, s);
      if (e != hipSuccess) {
        rc = hip_fail(e, "hipGraphLaunch");
        break;
      }
    } else {
      for (int c = 0; c < check_every; ++c) seg_iteration<T, Op>(s, op, B, launched + c, dinv, x, r, z, p, q);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        rc = hip_fail(e, "seg_pcg_iter");
        break;
      }
    }
    launched += check_every;
    hipError_t e = hipMemcpyAsync(h.data(), B.st, B.nseg * sizeof(PcgState), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "seg_pcg flags");
  }
  if (rc) return rc;
  for (int i = 0; i < B.nseg; ++i) {
    iters[i] = h[i].k;
    rs[i] = h[i].rs;
  }
  return FX_OK;
}

// Autotune cache: problem shape -> SpMV configuration.
std::mutex g_tune_mu;
std::map<std::tuple<int, long long, int, int, int>, SegCfg> g_tune;

// Times q = A x (5 launches after a warm-up) for lanes 8 / 16 / 32 x
// workgroups of 256 / 512 threads x map off / on (a map is built for each
// partition; candidates whose map is unusable are skipped) and returns the
// fastest.  Uses the table and map storage of the workspace.
template <typename V, typename T, bool DIFF>
int tune_cfg(hipStream_t s, int nseg, const int* seg_node_ptr, long long nse, const int* indptr, const int* bcol,
             const V* data, const T* xin, T* yout, BlockRange* dtab, SegMap map, SegCfg* out) {
  const auto key = std::make_tuple(seg_node_ptr[nseg], nse, nseg, static_cast<int>(sizeof(V)), DIFF ? 1 : 0);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      *out = it->second;
      return FX_OK;
    }
  }
  EventGuard evg;
  evg.ev.assign(2, nullptr);
  FX_HIP(hipEventCreate(&evg.ev[0]));
  FX_HIP(hipEventCreate(&evg.ev[1]));
  hipEvent_t e0 = evg.ev[0], e1 = evg.ev[1];
  SegCfg best{0, 0, false};
  float best_ms = 1e30f;
  std::vector<BlockRange> tab;
  std::vector<int> ptr;
  for (int lanes : {8, 16, 32}) {
    for (int bs : {256, 512}) {
      mv_table(nseg, seg_node_ptr, bs / lanes, tab, ptr);
      FX_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
      const int nb = static_cast<int>(tab.size());
      for (bool use_map : {false, true}) {
        if (use_map) {
          int rc = build_seg_map<T>(s, dtab, nb, indptr, bcol, map);
          if (rc) return rc;
          if (map.max_unique < 0) continue;
        }
        const SegCfg c{lanes, bs, use_map};
        auto run = [&](auto lc, auto bc, auto mc) -> int {
          const SegOp<V, T, decltype(lc)::value, DIFF, decltype(bc)::value, decltype(mc)::value> op{
              dtab, nb, indptr, bcol, data, map};
          op.apply(s, xin, yout);
          FX_HIP(hipEventRecord(e0, s));
          for (int rep = 0; rep < 5; ++rep) op.apply(s, xin, yout);
          FX_HIP(hipEventRecord(e1, s));
          FX_HIP(hipEventSynchronize(e1));
          FX_LAUNCHED("seg_spmv_tune");
          return FX_OK;
        };
        int rc = with_cfg(c, run);
        if (rc) return rc;
        float ms = 0.f;
        FX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
      FX_HIP(hipStreamSynchronize(s));  // `tab` is refilled next
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
  }
  *out = best;
  return FX_OK;
}

template <typename V, typename T, bool DIFF>
int run_batched(void* stream, int nseg, const int* seg_node_ptr, long long nse, int flags, const int* indptr,
                const int* bcol, const V* data, const T* dinv, const T* b, const T* x0, double tol, double atol,
                int maxiter, int check_every, T* x, int* iters, double* rs, int* tuned, void* workspace,
                size_t ws_bytes, SegRecord rec = SegRecord{nullptr, nullptr, 0, 0, 0}) {
  hipStream_t s = as_stream(stream);
  if (nseg <= 0 || !seg_node_ptr || !indptr || !bcol || !data || !b || !x || !iters || !rs || nse < 0)
    return fail(FX_EINVAL, "pcg_batched: bad arguments");
  if ((rec.P != nullptr) != (rec.pq != nullptr) || (rec.P && rec.rec_max <= 0))
    return fail(FX_EINVAL, "pcg_batched: a record needs P, pq and rec_max > 0");
  if (maxiter < 0) return fail(FX_EINVAL, "pcg_batched: maxiter < 0");
  for (int i = 0; i < nseg; ++i)
    if (seg_node_ptr[i + 1] < seg_node_ptr[i] || seg_node_ptr[0] != 0)
      return fail(FX_EINVAL, "pcg_batched: seg_node_ptr must start at 0 and be non-decreasing");
  const SegLayout Lay = seg_layout(nseg, seg_node_ptr, nse, sizeof(T));
  if (!workspace || ws_bytes < Lay.total)
    return fail(FX_EWORKSPACE, "pcg_batched: workspace %zu < %zu bytes", ws_bytes, Lay.total);
  if (Lay.n == 0) {  // no rows: nothing to launch
    for (int i = 0; i < nseg; ++i) {
      iters[i] = 0;
      rs[i] = 0.0;
    }
    return FX_OK;
  }
  char* w = static_cast<char*>(workspace);
  SegBuffers B;
  B.st = reinterpret_cast<PcgState*>(w + Lay.off_st);
  B.mvtab = reinterpret_cast<BlockRange*>(w + Lay.off_mvtab);
  B.mvptr = reinterpret_cast<int*>(w + Lay.off_mvptr);
  B.vtab = reinterpret_cast<BlockRange*>(w + Lay.off_vtab);
  B.vptr = reinterpret_cast<int*>(w + Lay.off_vptr);
  B.pmv = reinterpret_cast<double*>(w + Lay.off_pmv);
  B.prz = reinterpret_cast<double*>(w + Lay.off_pv);
  B.prr = B.prz + Lay.nv;
  B.pbb = B.prr + Lay.nv;
  B.nseg = nseg;
  B.rec = rec;
  B.rec.n = 3 * seg_node_ptr[nseg];
  B.rec.nseg = nseg;
  SegMap map{reinterpret_cast<uint16_t*>(w + Lay.off_lidx), reinterpret_cast<int*>(w + Lay.off_ucol),
             reinterpret_cast<int*>(w + Lay.off_ucnt), reinterpret_cast<int*>(w + Lay.off_info), -1};
  T* r = reinterpret_cast<T*>(w + Lay.off_vec);
  T* z = reinterpret_cast<T*>(w + Lay.off_vec + Lay.vec_bytes);
  T* p = reinterpret_cast<T*>(w + Lay.off_vec + 2 * Lay.vec_bytes);
  T* q = reinterpret_cast<T*>(w + Lay.off_vec + 3 * Lay.vec_bytes);

  // Default: FP64 values 32 lanes through bcol; FP32 values 16 lanes with
  // the map (falls back to bcol when the map is unusable).
  SegCfg cfg{sizeof(V) == 8 ? 32 : 16, 256, sizeof(V) == 4};
  if (flags & FX_PCG_SEG_FORCE) {
    if (!tuned) return fail(FX_EINVAL, "pcg_batched: FX_PCG_SEG_FORCE needs the configuration in tuned");
    cfg = SegCfg{tuned[0], tuned[1], tuned[2] != 0};
    if ((cfg.lanes != 8 && cfg.lanes != 16 && cfg.lanes != 32) || (cfg.bs != 256 && cfg.bs != 512))
      return fail(FX_EINVAL, "pcg_batched: configuration %d lanes x %d threads", cfg.lanes, cfg.bs);
  } else if (flags & FX_PCG_AUTOTUNE) {
    int rc = tune_cfg<V, T, DIFF>(s, nseg, seg_node_ptr, nse, indptr, bcol, data, b, q, B.mvtab, map, &cfg);
    if (rc) return rc;
  }
  std::vector<BlockRange> mvtab, vtab;
  std::vector<int> mvptr, vptr, rowptr;
  mv_table(nseg, seg_node_ptr, cfg.bs / cfg.lanes, mvtab, mvptr);
  v_table(nseg, seg_node_ptr, vtab, vptr, rowptr);
  if (static_cast<int>(mvtab.size()) > Lay.nmv || static_cast<int>(vtab.size()) > Lay.nv)
    return fail(FX_ELIMIT, "pcg_batched: block tables exceed the workspace layout");
  B.nv = static_cast<int>(vtab.size());
  const int nb = static_cast<int>(mvtab.size());
  FX_HIP(hipMemcpyAsync(B.mvtab, mvtab.data(), mvtab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(B.mvptr, mvptr.data(), mvptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(B.vtab, vtab.data(), vtab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(B.vptr, vptr.data(), vptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
  if (cfg.map) {
    int rc = build_seg_map<T>(s, B.mvtab, nb, indptr, bcol, map);  // synchronises the stream
    if (rc) return rc;
    if (map.max_unique < 0) cfg.map = false;
  }
  FX_HIP(hipStreamSynchronize(s));  // the host vectors above go out of scope
  if (tuned) {
    tuned[0] = cfg.lanes;
    tuned[1] = cfg.bs;
    tuned[2] = cfg.map ? 1 : 0;
  }
  const bool graph = !(flags & FX_PCG_NO_GRAPH);
  const int n = Lay.n;
  auto go = [&](auto lc, auto bc, auto mc) -> int {
    const SegOp<V, T, decltype(lc)::value, DIFF, decltype(bc)::value, decltype(mc)::value> op{B.mvtab, nb, indptr,
                                                                                             bcol, data, map};
    return seg_pcg_loop<T>(s, op, B, dinv, b, x0, x, r, z, p, q, n, tol, atol, maxiter, check_every, graph, iters,
                           rs);
  };
  return with_cfg(cfg, go);
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_pcg_batched_workspace_bytes(int32_t nseg, const int32_t* seg_node_ptr, int64_t nse,
                                      int32_t vec_elem_bytes) {
  if (nseg <= 0 || !seg_node_ptr || nse < 0 || (vec_elem_bytes != 4 && vec_elem_bytes != 8)) return 0;
  return seg_layout(nseg, seg_node_ptr, nse, static_cast<size_t>(vec_elem_bytes)).total;
}

int fx_pcg_batched_f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse, int32_t flags,
                       const int32_t* indptr, const int32_t* bcol, const double* data, const double* dinv,
                       const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                       int32_t check_every, double* x, int32_t* iters, double* rs, int32_t* tuned,
                       void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3) return fail(FX_EINVAL, "fx_pcg_batched_f64: FX_PCG_DIFF3 is for FP32 matrices");
  return run_batched<double, double, false>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                            tol, atol, maxiter, check_every, x, iters, rs, tuned, workspace,
                                            ws_bytes);
}

int fx_pcg_batched_f32f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse, int32_t flags,
                          const int32_t* indptr, const int32_t* bcol, const float* data, const double* dinv,
                          const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                          int32_t check_every, double* x, int32_t* iters, double* rs, int32_t* tuned,
                          void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3)
    return run_batched<float, double, true>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                            tol, atol, maxiter, check_every, x, iters, rs, tuned, workspace,
                                            ws_bytes);
  return run_batched<float, double, false>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                           tol, atol, maxiter, check_every, x, iters, rs, tuned, workspace,
                                           ws_bytes);
}

int fx_pcg_batched_record_f32f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse,
                                 int32_t flags, const int32_t* indptr, const int32_t* bcol, const float* data,
                                 const double* dinv, const double* b, const double* x0, double tol, double atol,
                                 int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
                                 int32_t* tuned, float* rec_p, double* rec_pq, int32_t rec_max,
                                 void* workspace, size_t ws_bytes) {
  if (!rec_p || !rec_pq || rec_max <= 0) return fail(FX_EINVAL, "fx_pcg_batched_record_f32f64: bad record");
  const SegRecord rec{rec_p, rec_pq, rec_max, 0, 0};
  if (flags & FX_PCG_DIFF3)
    return run_batched<float, double, true>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                            tol, atol, maxiter, check_every, x, iters, rs, tuned, workspace,
                                            ws_bytes, rec);
  return run_batched<float, double, false>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                           tol, atol, maxiter, check_every, x, iters, rs, tuned, workspace,
                                           ws_bytes, rec);
}

size_t fx_seg_project_workspace_bytes(int32_t nseg, const int32_t* seg_node_ptr, int32_t kmax) {
  if (nseg <= 0 || !seg_node_ptr || kmax < 0) return 0;
  const size_t nv = static_cast<size_t>(v_blocks(seg_node_ptr, nseg));
  return al256(nv * sizeof(BlockRange)) + 2 * al256((static_cast<size_t>(nseg) + 1) * sizeof(int)) +
         al256(static_cast<size_t>(std::max(kmax, 1)) * nv * sizeof(double)) +
         al256(static_cast<size_t>(std::max(kmax, 1)) * nseg * sizeof(double));
}

int fx_seg_project_f32(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int32_t kmax, const int32_t* counts,
                       const float* P, const double* pq, const double* r, double* x0, void* workspace,
                       size_t ws_bytes) {
  if (nseg <= 0 || !seg_node_ptr || kmax < 0 || !counts || !r || !x0 || (kmax > 0 && (!P || !pq)))
    return fail(FX_EINVAL, "fx_seg_project_f32: bad arguments");
  for (int s = 0; s < nseg; ++s)
    if (counts[s] < 0 || counts[s] > kmax || seg_node_ptr[s + 1] < seg_node_ptr[s] || seg_node_ptr[0] != 0)
      return fail(FX_EINVAL, "fx_seg_project_f32: counts must lie in [0, kmax], seg_node_ptr non-decreasing from 0");
  if (kmax > 65535) return fail(FX_ELIMIT, "fx_seg_project_f32: kmax %d > 65535", kmax);
  if (!workspace || ws_bytes < fx_seg_project_workspace_bytes(nseg, seg_node_ptr, kmax))
    return fail(FX_EWORKSPACE, "fx_seg_project_f32: workspace too small");
  hipStream_t s = as_stream(stream);
  std::vector<BlockRange> vtab;
  std::vector<int> vptr, rowptr;
  v_table(nseg, seg_node_ptr, vtab, vptr, rowptr);
  const int nv = static_cast<int>(vtab.size());
  const int n = 3 * seg_node_ptr[nseg];
  char* w = static_cast<char*>(workspace);
  BlockRange* dvtab = reinterpret_cast<BlockRange*>(w);
  w += al256(static_cast<size_t>(nv) * sizeof(BlockRange));
  int* dvptr = reinterpret_cast<int*>(w);
  w += al256((static_cast<size_t>(nseg) + 1) * sizeof(int));
  int* dcnt = reinterpret_cast<int*>(w);
  w += al256((static_cast<size_t>(nseg) + 1) * sizeof(int));
  double* partials = reinterpret_cast<double*>(w);
  w += al256(static_cast<size_t>(std::max(kmax, 1)) * nv * sizeof(double));
  double* c = reinterpret_cast<double*>(w);
  FX_HIP(hipMemcpyAsync(dvtab, vtab.data(), vtab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(dvptr, vptr.data(), vptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(dcnt, counts, static_cast<size_t>(nseg) * sizeof(int), hipMemcpyHostToDevice, s));
  if (kmax > 0) {
    hipLaunchKernelGGL(seg_proj_dot_kernel, dim3(nv, kmax), dim3(kSegBlock), 0, s, dvtab, n, nv, dcnt, P, r,
                       partials);
    const long nc = static_cast<long>(kmax) * nseg;
    hipLaunchKernelGGL(seg_proj_coef_kernel, dim3(static_cast<unsigned>((nc + kSegBlock - 1) / kSegBlock)),
                       dim3(kSegBlock), 0, s, nseg, kmax, nv, dvptr, dcnt, partials, pq, c);
  }
  hipLaunchKernelGGL(seg_proj_combine_kernel, dim3(nv), dim3(kSegBlock), 0, s, dvtab, n, nseg, dcnt, P, c, x0);
  FX_LAUNCHED("seg_proj kernels");
  FX_HIP(hipStreamSynchronize(s));  // the host tables above go out of scope
  return FX_OK;
}

int fx_seg_dot_f64(void* stream, int32_t nseg, const int32_t* row_ptr, const double* a, const double* b,
                   double* out) {
  if (nseg <= 0 || !row_ptr || !a || !b || !out) return fail(FX_EINVAL, "fx_seg_dot_f64: bad arguments");
  hipLaunchKernelGGL(seg_dot_kernel, dim3(nseg), dim3(kFinBlock), 0, as_stream(stream), row_ptr, a, b, out);
  FX_LAUNCHED("seg_dot_kernel");
  return FX_OK;
}

}  // extern "C"
