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
//    tables built on the host), so its partial sums belong to that instance
//    and the finalize kernels run one block per instance;
//  * each instance has its own alpha, beta, gamma, stopping test
//    (jax cg semantics) and done flag; converged instances' blocks exit at
//    their first instruction while the others iterate;
//  * a chunk of check_every iterations is captured once as a HIP graph and
//    replayed on the caller's stream until all instances are done (one host
//    read of the per-instance flags per chunk);
//  * FX_PCG_AUTOTUNE times the SpMV lane counts 8 / 16 / 32 on the
//    instance set and keeps the fastest per problem shape;
//  * optionally each instance's search directions p_k (FP32) and p_k.A p_k
//    are recorded at its own iteration count k, and fx_seg_project_f32
//    forms the per-instance Galerkin projection onto them (Krylov recycling
//    for the later inner solves of a mixed-precision defect correction).
// FP64 partial sums in a fixed order: bitwise repeatable, with or without
// the graph.
#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.h"
#include "pcg_common.h"
#include "reduce.h"
#include "spmv_common.h"

namespace fx {
namespace {

constexpr int kSegBlock = 256;  // threads per block
constexpr int kVecRows = 4096;  // rows per vector-kernel block

struct BlockRange {
  int seg, begin, end;  // instance, [begin, end) of nodes (SpMV) or rows (vector kernels)
};

// Sums of cnt partials starting at p[0..NV) (one finalize block, one
// block reduction for all NV quantities); valid in thread 0.
template <int NV>
__device__ __forceinline__ DVec<NV> seg_sums(const double* const (&p)[NV], int cnt) {
  DVec<NV> acc;
#pragma unroll
  for (int k = 0; k < NV; ++k) acc.v[k] = 0.0;
  for (int i = threadIdx.x; i < cnt; i += kFinBlock) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc.v[k] += p[k][i];
  }
  return block_sum<NV, kFinBlock>(acc);
}

template <typename V, typename T, int L, bool DIFF, bool DOT>
__global__ __launch_bounds__(kSegBlock) void seg_spmv_kernel(const BlockRange* __restrict__ tab,
                                                             const int* __restrict__ indptr,
                                                             const int* __restrict__ bcol, const V* __restrict__ data,
                                                             const T* __restrict__ x, T* __restrict__ y,
                                                             const PcgState* __restrict__ st, double* partials) {
  const BlockRange br = tab[blockIdx.x];
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
    const DVec<1> t = block_sum<1, kSegBlock>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = t.v[0];
  }
}

// Recording of search directions: P[k n + i] = p_k[i] (FP32) for the rows
// of an instance at its iteration k < rec_max, pq[k nseg + s] = p_k.A p_k.
struct SegRecord {
  float* P;
  double* pq;
  int rec_max, n, nseg;
};

__global__ __launch_bounds__(kFinBlock) void seg_fin_alpha(PcgState* st, const int* __restrict__ mv_ptr,
                                                           const double* __restrict__ partials, SegRecord rec) {
  const int s = blockIdx.x;
  if (st[s].done) return;
  const double* src[1] = {partials + mv_ptr[s]};
  const DVec<1> pq = seg_sums<1>(src, mv_ptr[s + 1] - mv_ptr[s]);
  if (threadIdx.x == 0) {
    st[s].alpha = st[s].gamma / pq.v[0];
    if (rec.pq && st[s].k < rec.rec_max) rec.pq[static_cast<size_t>(st[s].k) * rec.nseg + s] = pq.v[0];
  }
}

// x += alpha p; r -= alpha q; z = M r; partials of r.z and r.r.
template <typename T>
__global__ __launch_bounds__(kSegBlock) void seg_update_kernel(const BlockRange* __restrict__ vtab, T* __restrict__ x,
                                                               T* __restrict__ r, T* __restrict__ z,
                                                               const T* __restrict__ p, const T* __restrict__ q,
                                                               const T* __restrict__ dinv,
                                                               const PcgState* __restrict__ st, double* prz,
                                                               double* prr, SegRecord rec) {
  const BlockRange br = vtab[blockIdx.x];
  if (st[br.seg].done) return;
  const T alpha = static_cast<T>(st[br.seg].alpha);
  const int k = st[br.seg].k;
  float* rp = (rec.P && k < rec.rec_max) ? rec.P + static_cast<size_t>(k) * rec.n : nullptr;
  DVec<2> acc;
  acc.v[0] = acc.v[1] = 0.0;
  for (int i = br.begin + threadIdx.x; i < br.end; i += kSegBlock) {
    if (rp) rp[i] = static_cast<float>(p[i]);
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    z[i] = zi;
    acc.v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    acc.v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const DVec<2> t = block_sum<2, kSegBlock>(acc);
  if (threadIdx.x == 0) {
    prz[blockIdx.x] = t.v[0];
    prr[blockIdx.x] = t.v[1];
  }
}

__global__ __launch_bounds__(kFinBlock) void seg_fin_beta(PcgState* st, const int* __restrict__ v_ptr,
                                                          const double* __restrict__ prz,
                                                          const double* __restrict__ prr) {
  const int s = blockIdx.x;
  if (st[s].done) return;
  const int b0 = v_ptr[s], cnt = v_ptr[s + 1] - v_ptr[s];
  const double* src[2] = {prz + b0, prr + b0};
  const DVec<2> t = seg_sums<2>(src, cnt);
  if (threadIdx.x == 0) {
    PcgState& S = st[s];
    const double g_new = t.v[0];
    S.beta = g_new / S.gamma;
    S.gamma = g_new;
    S.rs = t.v[1];
    S.k += 1;
    S.done = !(S.rs > S.atol2 && S.k < S.maxiter);
  }
}

template <typename T>
__global__ __launch_bounds__(kSegBlock) void seg_dir_kernel(const BlockRange* __restrict__ vtab, T* __restrict__ p,
                                                            const T* __restrict__ z,
                                                            const PcgState* __restrict__ st) {
  const BlockRange br = vtab[blockIdx.x];
  if (st[br.seg].done) return;
  const T beta = static_cast<T>(st[br.seg].beta);
  for (int i = br.begin + threadIdx.x; i < br.end; i += kSegBlock) p[i] = z[i] + beta * p[i];
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
  const DVec<3> t = seg_sums<3>(src, cnt);
  if (threadIdx.x == 0) {
    PcgState& S = st[s];
    S.alpha = S.beta = 0.0;
    S.gamma = t.v[0];
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
  size_t off_st = 0, off_mvtab = 0, off_mvptr = 0, off_vtab = 0, off_vptr = 0, off_rowptr = 0, off_pmv = 0,
         off_pv = 0, off_vec = 0, vec_bytes = 0, total = 0;
};

// Worst-case SpMV block count over the lane counts used (L = 32: 8 nodes per block).
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

inline SegLayout seg_layout(int nseg, const int* seg_node_ptr, size_t elem) {
  SegLayout L;
  L.nseg = nseg;
  L.n = 3 * seg_node_ptr[nseg];
  L.nmv = mv_blocks_cap(seg_node_ptr, nseg);
  L.nv = v_blocks(seg_node_ptr, nseg);
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
  L.off_vec = o;
  L.vec_bytes = al256(static_cast<size_t>(L.n) * elem);
  o += 4 * L.vec_bytes;
  L.total = o;
  return L;
}

// SpMV block table for L lanes per node (kSegBlock / L nodes per block).
inline void mv_table(int nseg, const int* seg_node_ptr, int lanes, std::vector<BlockRange>& tab,
                     std::vector<int>& ptr) {
  const int gpb = kSegBlock / lanes;
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

template <typename V, typename T, int L, bool DIFF>
struct SegOp {
  const BlockRange* tab;
  int nblocks;
  const int *indptr, *bcol;
  const V* data;
  void apply(hipStream_t s, const T* x, T* y) const {
    hipLaunchKernelGGL((seg_spmv_kernel<V, T, L, DIFF, false>), dim3(nblocks), dim3(kSegBlock), 0, s, tab, indptr,
                       bcol, data, x, y, nullptr, nullptr);
  }
  void iter(hipStream_t s, const T* p, T* q, const PcgState* st, double* partials) const {
    hipLaunchKernelGGL((seg_spmv_kernel<V, T, L, DIFF, true>), dim3(nblocks), dim3(kSegBlock), 0, s, tab, indptr,
                       bcol, data, p, q, st, partials);
  }
};

struct SegBuffers {
  PcgState* st;
  BlockRange *mvtab, *vtab;
  int *mvptr, *vptr;
  double *pmv, *prz, *prr, *pbb;
  int nseg, nv;
  SegRecord rec;
};

template <typename T, typename Op>
void seg_iteration(hipStream_t s, const Op& op, const SegBuffers& B, const T* dinv, T* x, T* r, T* z, T* p, T* q) {
  op.iter(s, p, q, B.st, B.pmv);
  hipLaunchKernelGGL(seg_fin_alpha, dim3(B.nseg), dim3(kFinBlock), 0, s, B.st, B.mvptr, B.pmv, B.rec);
  hipLaunchKernelGGL((seg_update_kernel<T>), dim3(B.nv), dim3(kSegBlock), 0, s, B.vtab, x, r, z, p, q, dinv, B.st,
                     B.prz, B.prr, B.rec);
  hipLaunchKernelGGL(seg_fin_beta, dim3(B.nseg), dim3(kFinBlock), 0, s, B.st, B.vptr, B.prz, B.prr);
  hipLaunchKernelGGL((seg_dir_kernel<T>), dim3(B.nv), dim3(kSegBlock), 0, s, B.vtab, p, z, B.st);
}

template <typename T, typename Op>
int seg_pcg_loop(hipStream_t s, const Op& op, const SegBuffers& B, const T* dinv, const T* b, const T* x0, T* x,
                 T* r, T* z, T* p, T* q, int n, double tol, double atol, int maxiter, int check_every, bool graph,
                 int* iters, double* rs, int* chunks_out) {
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
      for (int c = 0; c < check_every; ++c) seg_iteration<T, Op>(gg.stream, op, B, dinv, x, r, z, p, q);
      e = hipStreamEndCapture(gg.stream, &gg.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "seg_pcg graph capture");
  }
  const hipGraphExec_t ge = gg.exec;
  int launched = 0, chunks = 0;
  while (rc == FX_OK && !all_done() && launched < maxiter) {
    if (ge) {
      hipError_t e = hipGraphLaunch(ge, s);
      if (e != hipSuccess) {
        rc = hip_fail(e, "hipGraphLaunch");
        break;
      }
    } else {
      for (int c = 0; c < check_every; ++c) seg_iteration<T, Op>(s, op, B, dinv, x, r, z, p, q);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        rc = hip_fail(e, "seg_pcg_iter");
        break;
      }
    }
    launched += check_every;
    ++chunks;
    hipError_t e = hipMemcpyAsync(h.data(), B.st, B.nseg * sizeof(PcgState), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "seg_pcg flags");
  }
  if (rc) return rc;
  for (int i = 0; i < B.nseg; ++i) {
    iters[i] = h[i].k;
    rs[i] = h[i].rs;
  }
  if (chunks_out) *chunks_out = chunks;
  return FX_OK;
}

// Autotune cache: problem shape -> lanes.
std::mutex g_tune_mu;
std::map<std::tuple<int, long long, int, int, int>, int> g_tune;

template <typename V, typename T, bool DIFF>
int tune_lanes(hipStream_t s, int nseg, const int* seg_node_ptr, long long nse, const int* indptr, const int* bcol,
               const V* data, const T* xin, T* yout, BlockRange* dtab, int* lanes_out) {
  const auto key = std::make_tuple(seg_node_ptr[nseg], nse, nseg, static_cast<int>(sizeof(V)), DIFF ? 1 : 0);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      *lanes_out = it->second;
      return FX_OK;
    }
  }
  EventGuard evg;
  evg.ev.assign(2, nullptr);
  FX_HIP(hipEventCreate(&evg.ev[0]));
  FX_HIP(hipEventCreate(&evg.ev[1]));
  hipEvent_t e0 = evg.ev[0], e1 = evg.ev[1];
  int best = 0;
  float best_ms = 1e30f;
  std::vector<BlockRange> tab;
  std::vector<int> ptr;
  for (int lanes : {8, 16, 32}) {
    mv_table(nseg, seg_node_ptr, lanes, tab, ptr);
    FX_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
    const int nb = static_cast<int>(tab.size());
    auto launch = [&]() {
      switch (lanes) {
        case 8:
          hipLaunchKernelGGL((seg_spmv_kernel<V, T, 8, DIFF, false>), dim3(nb), dim3(kSegBlock), 0, s, dtab, indptr,
                             bcol, data, xin, yout, nullptr, nullptr);
          break;
        case 16:
          hipLaunchKernelGGL((seg_spmv_kernel<V, T, 16, DIFF, false>), dim3(nb), dim3(kSegBlock), 0, s, dtab, indptr,
                             bcol, data, xin, yout, nullptr, nullptr);
          break;
        default:
          hipLaunchKernelGGL((seg_spmv_kernel<V, T, 32, DIFF, false>), dim3(nb), dim3(kSegBlock), 0, s, dtab, indptr,
                             bcol, data, xin, yout, nullptr, nullptr);
      }
    };
    launch();
    FX_HIP(hipEventRecord(e0, s));
    for (int rep = 0; rep < 5; ++rep) launch();
    FX_HIP(hipEventRecord(e1, s));
    FX_HIP(hipEventSynchronize(e1));
    FX_LAUNCHED("seg_spmv_tune");
    float ms = 0.f;
    FX_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = lanes;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
  }
  *lanes_out = best;
  return FX_OK;
}

template <typename V, typename T, bool DIFF>
int run_batched(void* stream, int nseg, const int* seg_node_ptr, long long nse, int flags, const int* indptr,
                const int* bcol, const V* data, const T* dinv, const T* b, const T* x0, double tol, double atol,
                int maxiter, int check_every, T* x, int* iters, double* rs, int* lanes_used, void* workspace,
                size_t ws_bytes, SegRecord rec = SegRecord{nullptr, nullptr, 0, 0, 0}) {
  hipStream_t s = as_stream(stream);
  if (nseg <= 0 || !seg_node_ptr || !indptr || !bcol || !data || !b || !x || !iters || !rs)
    return fail(FX_EINVAL, "pcg_batched: bad arguments");
  if ((rec.P != nullptr) != (rec.pq != nullptr) || (rec.P && rec.rec_max <= 0))
    return fail(FX_EINVAL, "pcg_batched: a record needs P, pq and rec_max > 0");
  if (maxiter < 0) return fail(FX_EINVAL, "pcg_batched: maxiter < 0");
  for (int i = 0; i < nseg; ++i)
    if (seg_node_ptr[i + 1] < seg_node_ptr[i] || seg_node_ptr[0] != 0)
      return fail(FX_EINVAL, "pcg_batched: seg_node_ptr must start at 0 and be non-decreasing");
  const SegLayout Lay = seg_layout(nseg, seg_node_ptr, sizeof(T));
  if (!workspace || ws_bytes < Lay.total)
    return fail(FX_EWORKSPACE, "pcg_batched: workspace %zu < %zu bytes", ws_bytes, Lay.total);
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
  T* r = reinterpret_cast<T*>(w + Lay.off_vec);
  T* z = reinterpret_cast<T*>(w + Lay.off_vec + Lay.vec_bytes);
  T* p = reinterpret_cast<T*>(w + Lay.off_vec + 2 * Lay.vec_bytes);
  T* q = reinterpret_cast<T*>(w + Lay.off_vec + 3 * Lay.vec_bytes);

  int lanes = sizeof(V) == 8 ? 32 : 16;
  if (flags & FX_PCG_AUTOTUNE) {
    int rc = tune_lanes<V, T, DIFF>(s, nseg, seg_node_ptr, nse, indptr, bcol, data, b, q, B.mvtab, &lanes);
    if (rc) return rc;
  }
  std::vector<BlockRange> mvtab, vtab;
  std::vector<int> mvptr, vptr, rowptr;
  mv_table(nseg, seg_node_ptr, lanes, mvtab, mvptr);
  v_table(nseg, seg_node_ptr, vtab, vptr, rowptr);
  if (static_cast<int>(mvtab.size()) > Lay.nmv || static_cast<int>(vtab.size()) > Lay.nv)
    return fail(FX_ELIMIT, "pcg_batched: block tables exceed the workspace layout");
  B.nv = static_cast<int>(vtab.size());
  FX_HIP(hipMemcpyAsync(B.mvtab, mvtab.data(), mvtab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(B.mvptr, mvptr.data(), mvptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(B.vtab, vtab.data(), vtab.size() * sizeof(BlockRange), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(B.vptr, vptr.data(), vptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
  FX_HIP(hipStreamSynchronize(s));  // the host vectors above go out of scope
  if (lanes_used) *lanes_used = lanes;
  const int nb = static_cast<int>(mvtab.size());
  const bool graph = !(flags & FX_PCG_NO_GRAPH);
  const int n = Lay.n;
  switch (lanes) {
    case 8:
      return seg_pcg_loop<T>(s, SegOp<V, T, 8, DIFF>{B.mvtab, nb, indptr, bcol, data}, B, dinv, b, x0, x, r, z, p, q,
                             n, tol, atol, maxiter, check_every, graph, iters, rs, nullptr);
    case 16:
      return seg_pcg_loop<T>(s, SegOp<V, T, 16, DIFF>{B.mvtab, nb, indptr, bcol, data}, B, dinv, b, x0, x, r, z, p,
                             q, n, tol, atol, maxiter, check_every, graph, iters, rs, nullptr);
    case 32:
      return seg_pcg_loop<T>(s, SegOp<V, T, 32, DIFF>{B.mvtab, nb, indptr, bcol, data}, B, dinv, b, x0, x, r, z, p,
                             q, n, tol, atol, maxiter, check_every, graph, iters, rs, nullptr);
    default:
      return fail(FX_EINVAL, "pcg_batched: lane count %d", lanes);
  }
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_pcg_batched_workspace_bytes(int32_t nseg, const int32_t* seg_node_ptr, int32_t vec_elem_bytes) {
  if (nseg <= 0 || !seg_node_ptr || (vec_elem_bytes != 4 && vec_elem_bytes != 8)) return 0;
  return seg_layout(nseg, seg_node_ptr, static_cast<size_t>(vec_elem_bytes)).total;
}

int fx_pcg_batched_f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse, int32_t flags,
                       const int32_t* indptr, const int32_t* bcol, const double* data, const double* dinv,
                       const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                       int32_t check_every, double* x, int32_t* iters, double* rs, int32_t* lanes_used,
                       void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3) return fail(FX_EINVAL, "fx_pcg_batched_f64: FX_PCG_DIFF3 is for FP32 matrices");
  return run_batched<double, double, false>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                            tol, atol, maxiter, check_every, x, iters, rs, lanes_used, workspace,
                                            ws_bytes);
}

int fx_pcg_batched_f32f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse, int32_t flags,
                          const int32_t* indptr, const int32_t* bcol, const float* data, const double* dinv,
                          const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                          int32_t check_every, double* x, int32_t* iters, double* rs, int32_t* lanes_used,
                          void* workspace, size_t ws_bytes) {
  if (flags & FX_PCG_DIFF3)
    return run_batched<float, double, true>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                            tol, atol, maxiter, check_every, x, iters, rs, lanes_used, workspace,
                                            ws_bytes);
  return run_batched<float, double, false>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                           tol, atol, maxiter, check_every, x, iters, rs, lanes_used, workspace,
                                           ws_bytes);
}

int fx_pcg_batched_record_f32f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse,
                                 int32_t flags, const int32_t* indptr, const int32_t* bcol, const float* data,
                                 const double* dinv, const double* b, const double* x0, double tol, double atol,
                                 int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
                                 int32_t* lanes_used, float* rec_p, double* rec_pq, int32_t rec_max,
                                 void* workspace, size_t ws_bytes) {
  if (!rec_p || !rec_pq || rec_max <= 0) return fail(FX_EINVAL, "fx_pcg_batched_record_f32f64: bad record");
  const SegRecord rec{rec_p, rec_pq, rec_max, 0, 0};
  if (flags & FX_PCG_DIFF3)
    return run_batched<float, double, true>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                            tol, atol, maxiter, check_every, x, iters, rs, lanes_used, workspace,
                                            ws_bytes, rec);
  return run_batched<float, double, false>(stream, nseg, seg_node_ptr, nse, flags, indptr, bcol, data, dinv, b, x0,
                                           tol, atol, maxiter, check_every, x, iters, rs, lanes_used, workspace,
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
