// Geometric multigrid preconditioned CG for HEX8 problems on structured
// grids (gfx950).
//
// Jacobi-PCG iteration counts grow with the mesh (951 at 1M DOF); a V-cycle
// on the grid hierarchy h, 2h, 4h, ... keeps them nearly mesh-independent.
// feax's NarrowBandCMG (feax/solvers/cmg.py:95-300) has the same structure
// (rediscretised coarse operators on cell averages, trilinear transfers,
// V-cycle preconditioned CG); this one is built from the block-3 SpMV of
// the Krylov path:
//
//   levels    the BC-eliminated CSR of each grid, assembled on the coarse
//             grid by the HEX8 assembly kernel (rediscretisation)
//   transfer  trilinear prolongation P between node grids (each direction
//             halved, n -> ceil(n/2) cells, or kept) and restriction
//             R = P^T, constrained DOFs of both grids masked
//   smoother  Chebyshev polynomial in D^-1 A on [lmax / ratio, 1.1 lmax]
//             (Saad, Iterative Methods, Alg. 12.1; lmax from a few Lanczos
//             steps, fx_gmg_eig_bounds_f64), the same polynomial
//             before and after the coarse correction: the V-cycle is a
//             fixed symmetric positive definite operator
//   coarsest  dense inverse (one wave per row), or a long Chebyshev on
//             [lmin, 1.1 lmax]
//   outer CG  jax cg semantics with z = V(r), in FP64; the V-cycle runs in
//             FP64 or FP32 (FX_GMG_F32: half the bytes per level sweep).
//             A chunk of check_every iterations is captured once as a HIP
//             graph and replayed; every kernel of an iteration returns at
//             once after convergence.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.h"
#include "feaxhip.h"
#include "pcg_common.h"
#include "reduce.h"
#include "spmv_common.h"

namespace fx {
namespace {

// Lanes per node of the block-3 SpMV: 27 blocks per interior node row.
template <typename T>
constexpr int lanes() {
  return sizeof(T) == 8 ? 32 : 16;
}
template <typename T>
constexpr int groups() {
  return kBlock / lanes<T>();
}

inline int nodes_of(const fx_gmg_level& L) { return (L.nx + 1) * (L.ny + 1) * (L.nz + 1); }
template <typename T>
inline int mv_grid(int nn) {
  return std::max(1, (nn + groups<T>() - 1) / groups<T>());
}
inline int v_grid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, kMaxGridBlocks)); }
inline int n_grid(int nn) { return std::max(1, (nn + kBlock - 1) / kBlock); }

__device__ __forceinline__ bool finished(const PcgState* st) { return st && st->done; }

// r = b - A x (rows 3g..3g+2).
template <typename T>
__global__ __launch_bounds__(kBlock) void resid_kernel(int nn, const int* __restrict__ indptr,
                                                       const int* __restrict__ bcol, const T* __restrict__ data,
                                                       const T* __restrict__ x, const T* __restrict__ b,
                                                       T* __restrict__ r, const PcgState* __restrict__ st) {
  if (finished(st)) return;
  constexpr int L = lanes<T>();
  const int g = blockIdx.x * groups<T>() + threadIdx.x / L, gl = threadIdx.x % L;
  if (g >= nn) return;  // whole groups leave together
  T y0, y1, y2;
  b3_rows<T, T, L, false>(indptr, bcol, data, x, g, gl, y0, y1, y2);
  if (gl == 0) {
    r[3 * g] = b[3 * g] - y0;
    r[3 * g + 1] = b[3 * g + 1] - y1;
    r[3 * g + 2] = b[3 * g + 2] - y2;
  }
}

// One Chebyshev step: d = cd d + cr dinv .* src, then x = d (assign, the
// first step from x = 0) or x += d.
template <typename T>
__global__ __launch_bounds__(kBlock) void cheb_kernel(int n, T cd, T cr, const T* __restrict__ dinv,
                                                      const T* __restrict__ src, T* __restrict__ d,
                                                      T* __restrict__ x, int assign,
                                                      const PcgState* __restrict__ st) {
  if (finished(st)) return;
  const bool with_d = !assign && cd != T(0);  // cd = 0: first step, d not read
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T di = cr * dinv[i] * src[i];
    if (with_d) di += cd * d[i];
    d[i] = di;
    x[i] = assign ? di : x[i] + di;
  }
}

// Node (I, J, K) of a grid with NY x NZ nodes in its last two directions;
// node id (I NY + J) NZ + K (feax/spgrid.py:28-41).
__device__ __forceinline__ void node_ijk(int id, int NY, int NZ, int& I, int& J, int& K) {
  K = id % NZ;
  const int t = id / NZ;
  J = t % NY;
  I = t / NY;
}

__device__ __forceinline__ bool free_dof(const uint8_t* __restrict__ mask, int i) { return !mask || !mask[i]; }

// Coarsening of one direction: kept (nc = nf, factor 1) or halved
// (nc = ceil(nf / 2), factor 2; for odd nf the last coarse cell reaches one
// fine cell past the fine grid, as in feax/solvers/cmg.py:82-87, 178-182).
__host__ __device__ __forceinline__ int cfactor(int nf, int nc) { return nf == nc ? 1 : 2; }
__host__ __device__ __forceinline__ bool coarsening_ok(int nf, int nc) {
  return nc >= 1 && (nc == nf || (nf >= 2 && nc == (nf + 1) / 2));
}

// Coarse nodes under fine index I along a direction coarsened by f (1 or
// 2): f = 2 takes floor(I/2) .. ceil(I/2) (one node for even I, two with
// weight 1/2 for odd I, trilinear); f = 1 takes I.
__device__ __forceinline__ void coarse_span(int I, int f, int& i0, int& n) {
  i0 = f == 2 ? I >> 1 : I;
  n = f == 2 ? 1 + (I & 1) : 1;
}

// x_f += P x_c on the unconstrained fine DOFs; the fine grid has
// (nxf, nyf, nzf) cells, the coarse grid (nxc, nyc, nzc), each direction
// halved or kept.
template <typename T>
__global__ __launch_bounds__(kBlock) void prolong_add_kernel(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                                             const T* __restrict__ xc, T* __restrict__ xf,
                                                             const uint8_t* __restrict__ mask_f,
                                                             const PcgState* __restrict__ st) {
  if (finished(st)) return;
  const int NYf = nyf + 1, NZf = nzf + 1, NYc = nyc + 1, NZc = nzc + 1;
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= (nxf + 1) * NYf * NZf) return;
  int I, J, K;
  node_ijk(f, NYf, NZf, I, J, K);
  int i0, j0, k0, ni, nj, nk;
  coarse_span(I, cfactor(nxf, nxc), i0, ni);
  coarse_span(J, cfactor(nyf, nyc), j0, nj);
  coarse_span(K, cfactor(nzf, nzc), k0, nk);
  const T w = T(1) / T(ni * nj * nk);
  T a0 = T(0), a1 = T(0), a2 = T(0);
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < nj; ++b)
      for (int c = 0; c < nk; ++c) {
        const int C = ((i0 + a) * NYc + (j0 + b)) * NZc + (k0 + c);
        a0 += xc[3 * C];
        a1 += xc[3 * C + 1];
        a2 += xc[3 * C + 2];
      }
  if (free_dof(mask_f, 3 * f)) xf[3 * f] += w * a0;
  if (free_dof(mask_f, 3 * f + 1)) xf[3 * f + 1] += w * a1;
  if (free_dof(mask_f, 3 * f + 2)) xf[3 * f + 2] += w * a2;
}

// b_c = P^T r_f, constrained fine and coarse DOFs masked: along a halved
// direction coarse node i gathers the fine nodes 2i + o, o in {-1, 0, 1},
// with weight (o ? 1/2 : 1); along a kept direction fine node i, weight 1.
template <typename T>
__global__ __launch_bounds__(kBlock) void restrict_kernel(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                                          const T* __restrict__ rf,
                                                          const uint8_t* __restrict__ mask_f, T* __restrict__ bc,
                                                          const uint8_t* __restrict__ mask_c,
                                                          const PcgState* __restrict__ st) {
  if (finished(st)) return;
  const int NXf = nxf + 1, NYf = nyf + 1, NZf = nzf + 1, NYc = nyc + 1, NZc = nzc + 1;
  const int fx = cfactor(nxf, nxc), fy = cfactor(nyf, nyc), fz = cfactor(nzf, nzc);
  const int C = blockIdx.x * kBlock + threadIdx.x;
  if (C >= (nxc + 1) * NYc * NZc) return;
  int i, j, k;
  node_ijk(C, NYc, NZc, i, j, k);
  T a0 = T(0), a1 = T(0), a2 = T(0);
  for (int oi = 1 - fx; oi <= fx - 1; ++oi) {
    const int I = fx * i + oi;
    if (I < 0 || I >= NXf) continue;
    for (int oj = 1 - fy; oj <= fy - 1; ++oj) {
      const int J = fy * j + oj;
      if (J < 0 || J >= NYf) continue;
      for (int ok = 1 - fz; ok <= fz - 1; ++ok) {
        const int K = fz * k + ok;
        if (K < 0 || K >= NZf) continue;
        const T w = T(1) / T((oi ? 2 : 1) * (oj ? 2 : 1) * (ok ? 2 : 1));
        const int f = (I * NYf + J) * NZf + K;
        if (free_dof(mask_f, 3 * f)) a0 += w * rf[3 * f];
        if (free_dof(mask_f, 3 * f + 1)) a1 += w * rf[3 * f + 1];
        if (free_dof(mask_f, 3 * f + 2)) a2 += w * rf[3 * f + 2];
      }
    }
  }
  bc[3 * C] = free_dof(mask_c, 3 * C) ? a0 : T(0);
  bc[3 * C + 1] = free_dof(mask_c, 3 * C + 1) ? a1 : T(0);
  bc[3 * C + 2] = free_dof(mask_c, 3 * C + 2) ? a2 : T(0);
}

// x = Ainv b (masked), one 64-lane wave per row, FP64 accumulation.
template <typename T>
__global__ __launch_bounds__(kBlock) void dense_kernel(int n, const T* __restrict__ ainv, const T* __restrict__ b,
                                                       T* __restrict__ x, const uint8_t* __restrict__ mask,
                                                       const PcgState* __restrict__ st) {
  if (finished(st)) return;
  const int row = blockIdx.x * (kBlock / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (row >= n) return;
  const T* a = ainv + static_cast<size_t>(row) * n;
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) acc += static_cast<double>(a[j]) * static_cast<double>(b[j]);
  acc = group_sum<double, 64>(acc);
  if (lane == 0) x[row] = free_dof(mask, row) ? static_cast<T>(acc) : T(0);
}

template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void convert_kernel(int n, const S* __restrict__ src, D* __restrict__ dst,
                                                         const PcgState* __restrict__ st) {
  if (finished(st)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) dst[i] = static_cast<D>(src[i]);
}

// ------------------------------------------------------------ outer CG
// q = A p (FP64) and block partials of p.q.
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(int nn, const int* __restrict__ indptr,
                                                         const int* __restrict__ bcol,
                                                         const double* __restrict__ data,
                                                         const double* __restrict__ p, double* __restrict__ q,
                                                         const PcgState* __restrict__ st, double* partials) {
  if (st->done) return;
  constexpr int L = lanes<double>();
  const int g = blockIdx.x * groups<double>() + threadIdx.x / L, gl = threadIdx.x % L;
  DVec<1> acc;
  acc.v[0] = 0.0;
  if (g < nn) {
    double y0, y1, y2;
    b3_rows<double, double, L, false>(indptr, bcol, data, p, g, gl, y0, y1, y2);
    if (gl == 0) {
      q[3 * g] = y0;
      q[3 * g + 1] = y1;
      q[3 * g + 2] = y2;
      acc.v[0] = p[3 * g] * y0 + p[3 * g + 1] * y1 + p[3 * g + 2] * y2;
    }
  }
  const DVec<1> t = block_sum<1, kBlock>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t.v[0];
}

// Contiguous slice of [0, n) of this block (XCD-aware block order).
__device__ __forceinline__ void slice(int n, int& lb, int& begin, int& end) {
  lb = xcd_block(blockIdx.x, gridDim.x);
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  begin = min(n, lb * chunk);
  end = min(n, begin + chunk);
}

__global__ __launch_bounds__(kFinBlock) void cg_fin_alpha(PcgState* st, const double* __restrict__ partials,
                                                          int np) {
  if (st->done) return;
  const DVec<1> t = sum_partials<1>(partials, np);
  if (threadIdx.x == 0) st->alpha = st->gamma / t.v[0];
}

// x += alpha p; r -= alpha q; partials of r.r in slot 0.
__global__ __launch_bounds__(kBlock) void cg_update_kernel(int n, double* __restrict__ x, double* __restrict__ r,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ q,
                                                           const PcgState* __restrict__ st, double* partials) {
  if (st->done) return;
  int lb, begin, end;
  slice(n, lb, begin, end);
  const double alpha = st->alpha;
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    acc.v[0] += ri * ri;
  }
  store_partials<1>(block_sum<1, kBlock>(acc), lb, partials);
}

// Partials of r.z in slot 1; with init also r.r (slot 0) and b.b (slot 2).
__global__ __launch_bounds__(kBlock) void cg_rz_kernel(int n, const double* __restrict__ r,
                                                       const double* __restrict__ z, const double* __restrict__ b,
                                                       const PcgState* __restrict__ st, int init,
                                                       double* partials) {
  if (!init && st->done) return;
  int lb, begin, end;
  slice(n, lb, begin, end);
  DVec<3> acc;
  acc.v[0] = acc.v[1] = acc.v[2] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    acc.v[1] += r[i] * z[i];
    if (init) {
      acc.v[0] += r[i] * r[i];
      acc.v[2] += b[i] * b[i];
    }
  }
  const DVec<3> t = block_sum<3, kBlock>(acc);
  if (threadIdx.x == 0) {
    partials[kMaxGridBlocks + lb] = t.v[1];
    if (init) {
      partials[lb] = t.v[0];
      partials[2 * kMaxGridBlocks + lb] = t.v[2];
    }
  }
}

__global__ __launch_bounds__(kFinBlock) void cg_fin_beta(PcgState* st, const double* __restrict__ partials, int np) {
  if (st->done) return;
  const DVec<2> t = sum_partials<2>(partials, np);  // r.r, r.z
  if (threadIdx.x == 0) {
    st->beta = t.v[1] / st->gamma;
    st->gamma = t.v[1];
    st->rs = t.v[0];
    st->k += 1;
    st->done = !(st->rs > st->atol2 && st->k < st->maxiter);
  }
}

__global__ __launch_bounds__(kFinBlock) void cg_fin_init(PcgState* st, const double* __restrict__ partials, int np,
                                                         double tol, double atol, int maxiter) {
  const DVec<3> t = sum_partials<3>(partials, np);  // r.r, r.z, b.b
  if (threadIdx.x == 0) {
    st->alpha = st->beta = 0.0;
    st->rs = t.v[0];
    st->gamma = t.v[1];
    st->bb = t.v[2];
    st->atol2 = fmax(tol * tol * t.v[2], atol * atol);
    st->k = 0;
    st->maxiter = maxiter;
    st->nparts = np;
    st->done = !(st->rs > st->atol2 && 0 < maxiter);
  }
}

// p = z + beta p (beta = 0 at init).
__global__ __launch_bounds__(kBlock) void cg_dir_kernel(int n, double* __restrict__ p, const double* __restrict__ z,
                                                        const PcgState* __restrict__ st, int init) {
  if (!init && st->done) return;
  const double beta = init ? 0.0 : st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = z[i] + beta * p[i];
}

// ------------------------------------------------------------ spectrum
// Lanczos through Jacobi-preconditioned CG on A x = v (v fixed): the CG
// coefficients give the Lanczos tridiagonal of D^-1 A, whose extreme
// eigenvalues (Ritz values) bound the spectrum from inside and converge to
// its ends in a few steps (Golub and Van Loan, Matrix Computations, 10.2).

// r = v = 1 + (i mod 7) / 4, p = dinv .* r; partials of r.p in slot 0.
__global__ __launch_bounds__(kBlock) void lz_init_kernel(int n, const double* __restrict__ dinv,
                                                         double* __restrict__ r, double* __restrict__ p,
                                                         double* partials) {
  int lb, begin, end;
  slice(n, lb, begin, end);
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    const double v = 1.0 + 0.25 * static_cast<double>(i % 7);
    r[i] = v;
    p[i] = dinv[i] * v;
    acc.v[0] += v * dinv[i] * v;
  }
  store_partials<1>(block_sum<1, kBlock>(acc), lb, partials);
}

__global__ __launch_bounds__(kFinBlock) void lz_fin_init(PcgState* st, const double* __restrict__ partials, int np) {
  const DVec<1> t = sum_partials<1>(partials, np);
  if (threadIdx.x == 0) {
    st->gamma = t.v[0];
    st->k = 0;
    st->done = 0;
  }
}

// r -= alpha q; partials of r.(dinv r) in slot 0.
__global__ __launch_bounds__(kBlock) void lz_update_kernel(int n, double* __restrict__ r,
                                                           const double* __restrict__ q,
                                                           const double* __restrict__ dinv,
                                                           const PcgState* __restrict__ st, double* partials) {
  int lb, begin, end;
  slice(n, lb, begin, end);
  const double alpha = st->alpha;
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    acc.v[0] += ri * dinv[i] * ri;
  }
  store_partials<1>(block_sum<1, kBlock>(acc), lb, partials);
}

// beta = rz / gamma; hist[2k] = alpha, hist[2k + 1] = beta.
__global__ __launch_bounds__(kFinBlock) void lz_fin_beta(PcgState* st, const double* __restrict__ partials, int np,
                                                         double* __restrict__ hist) {
  const DVec<1> t = sum_partials<1>(partials, np);
  if (threadIdx.x == 0) {
    st->beta = t.v[0] / st->gamma;
    st->gamma = t.v[0];
    hist[2 * st->k] = st->alpha;
    hist[2 * st->k + 1] = st->beta;
    st->k += 1;
  }
}

// p = dinv .* r + beta p.
__global__ __launch_bounds__(kBlock) void lz_dir_kernel(int n, double* __restrict__ p, const double* __restrict__ r,
                                                        const double* __restrict__ dinv,
                                                        const PcgState* __restrict__ st) {
  const double beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = dinv[i] * r[i] + beta * p[i];
}

__global__ __launch_bounds__(kBlock) void csr_dense_kernel(int n, const int* __restrict__ indptr,
                                                           const int* __restrict__ indices,
                                                           const double* __restrict__ data,
                                                           double* __restrict__ dense) {
  const int row = blockIdx.x;
  double* out = dense + static_cast<size_t>(row) * n;
  for (int j = threadIdx.x; j < n; j += kBlock) out[j] = 0.0;
  __syncthreads();
  for (int k = indptr[row] + threadIdx.x; k < indptr[row + 1]; k += kBlock) out[indices[k]] = data[k];
}

// Coarse-grid copy of a node field (ncomp values per node: the value at the
// same node of the fine grid) or of a cell field (mean over the child
// cells, the density averaging of feax/solvers/cmg.py:375-386, 426-431).
__global__ __launch_bounds__(kBlock) void coarsen_node_kernel(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                                              int ncomp, const double* __restrict__ fine,
                                                              double* __restrict__ coarse) {
  const int C = blockIdx.x * kBlock + threadIdx.x;
  if (C >= (nxc + 1) * (nyc + 1) * (nzc + 1)) return;
  int i, j, k;
  node_ijk(C, nyc + 1, nzc + 1, i, j, k);
  // a coarse node past the end of an odd fine direction takes the last fine node
  const int I = min(cfactor(nxf, nxc) * i, nxf), J = min(cfactor(nyf, nyc) * j, nyf),
            K = min(cfactor(nzf, nzc) * k, nzf);
  const int f = (I * (nyf + 1) + J) * (nzf + 1) + K;
  for (int c = 0; c < ncomp; ++c) coarse[ncomp * C + c] = fine[ncomp * f + c];
}

__global__ __launch_bounds__(kBlock) void coarsen_cell_kernel(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                                              const double* __restrict__ fine,
                                                              double* __restrict__ coarse) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nxc * nyc * nzc) return;
  const int fx = cfactor(nxf, nxc), fy = cfactor(nyf, nyc), fz = cfactor(nzf, nzc);
  const int ck = e % nzc, t = e / nzc;
  const int cj = t % nyc, ci = t / nyc;
  double s = 0.0;
  int cnt = 0;  // child cells inside the fine grid (cmg.py:375-386 counts them the same way)
  for (int a = 0; a < fx; ++a) {
    if (fx * ci + a >= nxf) continue;
    for (int b = 0; b < fy; ++b) {
      if (fy * cj + b >= nyf) continue;
      for (int c = 0; c < fz; ++c) {
        if (fz * ck + c >= nzf) continue;
        s += fine[((fx * ci + a) * nyf + fy * cj + b) * nzf + fz * ck + c];
        ++cnt;
      }
    }
  }
  coarse[e] = s / static_cast<double>(cnt);
}

// ------------------------------------------------------------ Galerkin coarsest
// Weight of coarse node index i in the prolongation row of fine index I
// along one direction (factor f): 1 (even I, i = I/2), 1/2 (odd I, i one of
// its two neighbours), or 1 (kept direction, i = I); else 0.
__device__ __forceinline__ double pweight(int I, int i, int f) {
  if (f == 1) return I == i ? 1.0 : 0.0;
  if (I & 1) return (i == (I >> 1) || i == (I >> 1) + 1) ? 0.5 : 0.0;
  return i == (I >> 1) ? 1.0 : 0.0;
}

// B = A P' (n_f x n_c, row-major), P' = M_f P M_c: one thread per (fine row
// a, coarse DOF J); sums the CSR row of a over its columns b of J's
// component that are free, times P[b][J].
__global__ __launch_bounds__(kBlock) void galerkin_ap_kernel(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                                             const int* __restrict__ indptr,
                                                             const int* __restrict__ indices,
                                                             const double* __restrict__ data,
                                                             const uint8_t* __restrict__ mask_f,
                                                             const uint8_t* __restrict__ mask_c,
                                                             double* __restrict__ B) {
  const int nf = 3 * (nxf + 1) * (nyf + 1) * (nzf + 1), nc = 3 * (nxc + 1) * (nyc + 1) * (nzc + 1);
  const long idx = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= static_cast<long>(nf) * nc) return;
  const int a = static_cast<int>(idx / nc), J = static_cast<int>(idx - static_cast<long>(a) * nc);
  double v = 0.0;
  if (free_dof(mask_c, J)) {
    const int cJ = J % 3;
    int ci, cj, ck;
    node_ijk(J / 3, nyc + 1, nzc + 1, ci, cj, ck);
    const int fx = cfactor(nxf, nxc), fy = cfactor(nyf, nyc), fz = cfactor(nzf, nzc);
    for (int k = indptr[a]; k < indptr[a + 1]; ++k) {
      const int b = indices[k];
      if (b % 3 != cJ || !free_dof(mask_f, b)) continue;
      int I, Jn, K;
      node_ijk(b / 3, nyf + 1, nzf + 1, I, Jn, K);
      const double w = pweight(I, ci, fx) * pweight(Jn, cj, fy) * pweight(K, ck, fz);
      if (w != 0.0) v += data[k] * w;
    }
  }
  B[idx] = v;
}

// Ac = P'^T B (n_c x n_c): coarse DOF I gathers the free fine DOFs of its
// restriction stencil; constrained coarse DOFs get identity rows/columns.
__global__ __launch_bounds__(kBlock) void galerkin_ptb_kernel(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                                              const double* __restrict__ B,
                                                              const uint8_t* __restrict__ mask_f,
                                                              const uint8_t* __restrict__ mask_c,
                                                              double* __restrict__ Ac) {
  const int nc = 3 * (nxc + 1) * (nyc + 1) * (nzc + 1);
  const long idx = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= static_cast<long>(nc) * nc) return;
  const int I = static_cast<int>(idx / nc), J = static_cast<int>(idx - static_cast<long>(I) * nc);
  if (!free_dof(mask_c, I) || !free_dof(mask_c, J)) {
    Ac[idx] = I == J ? 1.0 : 0.0;
    return;
  }
  const int NXf = nxf + 1, NYf = nyf + 1, NZf = nzf + 1;
  const int fx = cfactor(nxf, nxc), fy = cfactor(nyf, nyc), fz = cfactor(nzf, nzc);
  const int cI = I % 3;
  int i, j, k;
  node_ijk(I / 3, nyc + 1, nzc + 1, i, j, k);
  double v = 0.0;
  for (int oi = 1 - fx; oi <= fx - 1; ++oi) {
    const int X = fx * i + oi;
    if (X < 0 || X >= NXf) continue;
    for (int oj = 1 - fy; oj <= fy - 1; ++oj) {
      const int Y = fy * j + oj;
      if (Y < 0 || Y >= NYf) continue;
      for (int ok = 1 - fz; ok <= fz - 1; ++ok) {
        const int Z = fz * k + ok;
        if (Z < 0 || Z >= NZf) continue;
        const int a = 3 * ((X * NYf + Y) * NZf + Z) + cI;
        if (!free_dof(mask_f, a)) continue;
        const double w = (oi ? 0.5 : 1.0) * (oj ? 0.5 : 1.0) * (ok ? 0.5 : 1.0);
        v += w * B[static_cast<long>(a) * nc + J];
      }
    }
  }
  Ac[idx] = v;
}

// ------------------------------------------------------------ host side
struct Cheb {
  int degree;
  double theta, delta, sigma;
};

// Interval [lo, 1.1 lmax]: lo = lmax / ratio for a smoother, or a given
// lower bound (lmin > 0) for the Chebyshev solve of the coarsest level.
inline Cheb cheb_params(double lmax, double ratio, int degree, double lmin = 0.0) {
  const double hi = 1.1 * lmax, lo = lmin > 0.0 ? lmin : lmax / ratio;
  Cheb c;
  c.degree = degree;
  c.theta = 0.5 * (hi + lo);
  c.delta = 0.5 * (hi - lo);
  c.sigma = c.theta / c.delta;
  return c;
}

template <typename T>
struct VCycle {
  int nlev;
  const fx_gmg_level* lev;
  std::vector<Cheb> cheb;
  const PcgState* st;  // kernels return at once when st->done (nullptr: never)

  // Chebyshev iteration for A x = b on level L (x from 0 when zero_init).
  void smooth(hipStream_t s, const fx_gmg_level& L, const Cheb& c, const T* b, T* x, bool zero_init) const {
    const int nn = nodes_of(L), n = 3 * nn;
    const int gv = v_grid(n), gm = mv_grid<T>(nn);
    const T* data = static_cast<const T*>(L.data);
    const T* dinv = static_cast<const T*>(L.dinv);
    T* r = static_cast<T*>(L.r);
    T* d = static_cast<T*>(L.d);
    const T* src = b;
    if (!zero_init) {
      hipLaunchKernelGGL(resid_kernel<T>, dim3(gm), dim3(kBlock), 0, s, nn, L.indptr, L.bcol, data, x, b, r, st);
      src = r;
    }
    hipLaunchKernelGGL(cheb_kernel<T>, dim3(gv), dim3(kBlock), 0, s, n, T(0), static_cast<T>(1.0 / c.theta), dinv,
                       src, d, x, zero_init ? 1 : 0, st);
    double rho_prev = 1.0 / c.sigma;
    for (int k = 1; k < c.degree; ++k) {
      hipLaunchKernelGGL(resid_kernel<T>, dim3(gm), dim3(kBlock), 0, s, nn, L.indptr, L.bcol, data, x, b, r, st);
      const double rho = 1.0 / (2.0 * c.sigma - rho_prev);
      hipLaunchKernelGGL(cheb_kernel<T>, dim3(gv), dim3(kBlock), 0, s, n, static_cast<T>(rho * rho_prev),
                         static_cast<T>(2.0 * rho / c.delta), dinv, r, d, x, 0, st);
      rho_prev = rho;
    }
  }

  // x0 = V(b0) on level 0.
  void run(hipStream_t s, const T* b0, T* x0) const {
    for (int l = 0; l < nlev; ++l) {
      const fx_gmg_level& L = lev[l];
      const T* b = l == 0 ? b0 : static_cast<const T*>(L.b);
      T* x = l == 0 ? x0 : static_cast<T*>(L.x);
      if (l == nlev - 1) {
        if (L.dense_inv) {
          const int n = 3 * nodes_of(L);
          hipLaunchKernelGGL(dense_kernel<T>, dim3((n + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, s, n,
                             static_cast<const T*>(L.dense_inv), b, x, L.bc_mask, st);
        } else {
          smooth(s, L, cheb[l], b, x, true);
        }
        break;
      }
      smooth(s, L, cheb[l], b, x, true);
      const int nn = nodes_of(L);
      T* r = static_cast<T*>(L.r);
      hipLaunchKernelGGL(resid_kernel<T>, dim3(mv_grid<T>(nn)), dim3(kBlock), 0, s, nn, L.indptr, L.bcol,
                         static_cast<const T*>(L.data), x, b, r, st);
      const fx_gmg_level& C = lev[l + 1];
      hipLaunchKernelGGL(restrict_kernel<T>, dim3(n_grid(nodes_of(C))), dim3(kBlock), 0, s, L.nx, L.ny, L.nz, C.nx,
                         C.ny, C.nz, r, L.bc_mask, static_cast<T*>(C.b), C.bc_mask, st);
    }
    for (int l = nlev - 2; l >= 0; --l) {
      const fx_gmg_level& L = lev[l];
      const fx_gmg_level& C = lev[l + 1];
      const T* b = l == 0 ? b0 : static_cast<const T*>(L.b);
      T* x = l == 0 ? x0 : static_cast<T*>(L.x);
      hipLaunchKernelGGL(prolong_add_kernel<T>, dim3(n_grid(nodes_of(L))), dim3(kBlock), 0, s, L.nx, L.ny, L.nz,
                         C.nx, C.ny, C.nz, static_cast<const T*>(C.x), x, L.bc_mask, st);
      smooth(s, L, cheb[l], b, x, false);
    }
  }
};

int check_levels(int flags, int nlev, const fx_gmg_level* lev, int smooth_degree, int coarse_degree,
                 double ratio) {
  if (nlev < 1 || !lev) return fail(FX_EINVAL, "gmg: no levels");
  if (smooth_degree < 1 || coarse_degree < 1 || !(ratio > 1.0))
    return fail(FX_EINVAL, "gmg: smooth_degree %d, coarse_degree %d must be >= 1 and cheb_ratio %g > 1",
                smooth_degree, coarse_degree, ratio);
  for (int l = 0; l < nlev; ++l) {
    const fx_gmg_level& L = lev[l];
    if (L.nx < 1 || L.ny < 1 || L.nz < 1) return fail(FX_EINVAL, "gmg: level %d has no cells", l);
    if (static_cast<long long>(L.nx + 1) * (L.ny + 1) * (L.nz + 1) > (1 << 29) / 3)
      return fail(FX_ELIMIT, "gmg: level %d too large for int32 DOF indices", l);
    const bool dense_coarsest = l == nlev - 1 && L.dense_inv;  // uses only dense_inv, x, b, bc_mask
    if (!dense_coarsest && (!L.indptr || !L.bcol || !L.data || !L.dinv || !L.r || !L.d))
      return fail(FX_EINVAL, "gmg: level %d lacks its CSR, dinv or scratch vectors", l);
    if ((l > 0 || (flags & FX_GMG_F32)) && (!L.x || !L.b))
      return fail(FX_EINVAL, "gmg: level %d needs x and b", l);
    if (l + 1 < nlev) {
      const fx_gmg_level& C = lev[l + 1];
      if (!coarsening_ok(L.nx, C.nx) || !coarsening_ok(L.ny, C.ny) || !coarsening_ok(L.nz, C.nz) ||
          (L.nx == C.nx && L.ny == C.ny && L.nz == C.nz))
        return fail(FX_EINVAL,
                    "gmg: level %d grid (%d,%d,%d) -> level %d (%d,%d,%d): each direction must be halved or kept, "
                    "one at least halved",
                    l, L.nx, L.ny, L.nz, l + 1, C.nx, C.ny, C.nz);
      if (!(L.lmax > 0.0)) return fail(FX_EINVAL, "gmg: level %d lmax must be positive", l);
    } else if (!L.dense_inv && !(L.lmax > 0.0)) {
      return fail(FX_EINVAL, "gmg: coarsest level needs dense_inv or a positive lmax");
    }
  }
  return FX_OK;
}

template <typename T>
VCycle<T> make_vcycle(int nlev, const fx_gmg_level* lev, int smooth_degree, int coarse_degree, double ratio,
                      const PcgState* st) {
  VCycle<T> V;
  V.nlev = nlev;
  V.lev = lev;
  V.st = st;
  for (int l = 0; l < nlev; ++l) {
    const bool coarsest = l == nlev - 1;
    V.cheb.push_back(cheb_params(lev[l].lmax > 0.0 ? lev[l].lmax : 1.0, ratio,
                                 coarsest ? coarse_degree : smooth_degree, coarsest ? lev[l].lmin : 0.0));
  }
  return V;
}

// z = V(r) for FP64 r, z: directly (FP64 levels) or through level 0's
// FP32 x and b.
template <typename T>
void apply_vcycle(hipStream_t s, const VCycle<T>& V, const double* r, double* z, int n) {
  if constexpr (sizeof(T) == 8) {
    V.run(s, r, z);
  } else {
    const fx_gmg_level& F = V.lev[0];
    T* b0 = static_cast<T*>(F.b);
    T* x0 = static_cast<T*>(F.x);
    hipLaunchKernelGGL((convert_kernel<double, T>), dim3(v_grid(n)), dim3(kBlock), 0, s, n, r, b0, V.st);
    V.run(s, b0, x0);
    hipLaunchKernelGGL((convert_kernel<T, double>), dim3(v_grid(n)), dim3(kBlock), 0, s, n, x0, z, V.st);
  }
}

inline size_t al256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Number of eigenvalues below x of the symmetric tridiagonal (diag d,
// off-diagonal e) by the signs of its LDL^T pivots (Sturm count).
int sturm_below(const std::vector<double>& d, const std::vector<double>& e, double x) {
  int count = 0;
  double q = 1.0;
  for (size_t j = 0; j < d.size(); ++j) {
    q = d[j] - x - (j ? e[j - 1] * e[j - 1] / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++count;
  }
  return count;
}

// Extreme eigenvalues of the Lanczos tridiagonal built from m CG steps
// (alpha_j, beta_j): T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1},
// T_{j,j+1} = sqrt(beta_j) / alpha_j; bisection on the Sturm count.
void ritz_bounds(const std::vector<double>& hist, int m, double& lmin, double& lmax) {
  std::vector<double> d(m), e(m > 1 ? m - 1 : 0);
  for (int j = 0; j < m; ++j) {
    d[j] = 1.0 / hist[2 * j] + (j ? hist[2 * j - 1] / hist[2 * j - 2] : 0.0);
    if (j + 1 < m) e[j] = std::sqrt(hist[2 * j + 1]) / hist[2 * j];
  }
  double lo = d[0], hi = d[0];
  for (int j = 0; j < m; ++j) {
    const double rad = (j ? std::fabs(e[j - 1]) : 0.0) + (j + 1 < m ? std::fabs(e[j]) : 0.0);
    lo = std::min(lo, d[j] - rad);
    hi = std::max(hi, d[j] + rad);
  }
  double a = lo, b = hi;  // lmax: largest x with fewer than m eigenvalues below
  for (int it = 0; it < 200 && b - a > 1e-15 * std::fabs(b); ++it) {
    const double c = 0.5 * (a + b);
    (sturm_below(d, e, c) < m ? a : b) = c;
  }
  lmax = 0.5 * (a + b);
  a = lo;
  b = hi;  // lmin: smallest x with one eigenvalue below
  for (int it = 0; it < 200 && b - a > 1e-15 * std::fabs(b); ++it) {
    const double c = 0.5 * (a + b);
    (sturm_below(d, e, c) >= 1 ? b : a) = c;
  }
  lmin = 0.5 * (a + b);
}

size_t eig_ws_bytes(int nn, int steps) {
  const size_t nparts = static_cast<size_t>(std::max(mv_grid<double>(nn), kMaxGridBlocks));
  return al256(sizeof(PcgState)) + al256(nparts * sizeof(double)) + al256(2 * static_cast<size_t>(steps) * 8) +
         3 * al256(3 * static_cast<size_t>(nn) * sizeof(double));
}

template <typename T>
int pcg_gmg(hipStream_t s, int flags, int nlev, const fx_gmg_level* levels, int smooth_degree, int coarse_degree,
            double ratio, const double* data, const double* b, const double* x0, double tol, double atol,
            int maxiter, int check_every, double* x, int* iters, double* rs, char* w) {
  const fx_gmg_level& F = levels[0];
  const int nn = nodes_of(F), n = 3 * nn;
  const int gv = v_grid(n), gm = mv_grid<double>(nn);
  PcgState* st = reinterpret_cast<PcgState*>(w);
  const size_t nparts = static_cast<size_t>(std::max(gm, 3 * kMaxGridBlocks));
  double* partials = reinterpret_cast<double*>(w + al256(sizeof(PcgState)));
  const size_t vb = al256(static_cast<size_t>(n) * sizeof(double));
  char* vbase = w + al256(sizeof(PcgState)) + al256(nparts * sizeof(double));
  double* r = reinterpret_cast<double*>(vbase);
  double* z = reinterpret_cast<double*>(vbase + vb);
  double* p = reinterpret_cast<double*>(vbase + 2 * vb);
  double* q = reinterpret_cast<double*>(vbase + 3 * vb);
  // Init (V-cycle not gated: the state may hold a previous solve's flag).
  const VCycle<T> V0 = make_vcycle<T>(nlev, levels, smooth_degree, coarse_degree, ratio, nullptr);
  if (x0) {
    if (x0 != x)
      FX_HIP(hipMemcpyAsync(x, x0, static_cast<size_t>(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
  } else {
    FX_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * sizeof(double), s));
  }
  hipLaunchKernelGGL(resid_kernel<double>, dim3(gm), dim3(kBlock), 0, s, nn, F.indptr, F.bcol, data, x, b, r,
                     nullptr);
  apply_vcycle<T>(s, V0, r, z, n);
  hipLaunchKernelGGL(cg_rz_kernel, dim3(gv), dim3(kBlock), 0, s, n, r, z, b, st, 1, partials);
  hipLaunchKernelGGL(cg_fin_init, dim3(1), dim3(kFinBlock), 0, s, st, partials, gv, tol, atol, maxiter);
  hipLaunchKernelGGL(cg_dir_kernel, dim3(gv), dim3(kBlock), 0, s, n, p, z, st, 1);
  FX_LAUNCHED("fx_pcg_gmg_f64 init");
  const VCycle<T> V = make_vcycle<T>(nlev, levels, smooth_degree, coarse_degree, ratio, st);
  auto iteration = [&](hipStream_t t) {
    hipLaunchKernelGGL(cg_spmv_kernel, dim3(gm), dim3(kBlock), 0, t, nn, F.indptr, F.bcol, data, p, q, st,
                       partials);
    hipLaunchKernelGGL(cg_fin_alpha, dim3(1), dim3(kFinBlock), 0, t, st, partials, gm);
    hipLaunchKernelGGL(cg_update_kernel, dim3(gv), dim3(kBlock), 0, t, n, x, r, p, q, st, partials);
    apply_vcycle<T>(t, V, r, z, n);
    hipLaunchKernelGGL(cg_rz_kernel, dim3(gv), dim3(kBlock), 0, t, n, r, z, b, st, 0, partials);
    hipLaunchKernelGGL(cg_fin_beta, dim3(1), dim3(kFinBlock), 0, t, st, partials, gv);
    hipLaunchKernelGGL(cg_dir_kernel, dim3(gv), dim3(kBlock), 0, t, n, p, z, st, 0);
  };
  struct {
    int k, done;
  } h = {0, 1};
  FX_HIP(hipMemcpyAsync(&h, &st->k, sizeof(h), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  GraphGuard gg;
  if (!(flags & FX_PCG_NO_GRAPH) && !h.done) {
    hipError_t e = hipStreamCreateWithFlags(&gg.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(gg.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int c = 0; c < check_every; ++c) iteration(gg.stream);
      e = hipStreamEndCapture(gg.stream, &gg.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "fx_pcg_gmg_f64 graph capture");
  }
  int launched = 0;
  while (!h.done && launched < maxiter) {
    if (gg.exec) {
      FX_HIP(hipGraphLaunch(gg.exec, s));
    } else {
      for (int c = 0; c < check_every; ++c) iteration(s);
      FX_LAUNCHED("fx_pcg_gmg_f64 iteration");
    }
    launched += check_every;
    FX_HIP(hipMemcpyAsync(&h, &st->k, sizeof(h), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
  }
  double hrs = 0.0;
  FX_HIP(hipMemcpyAsync(&hrs, &st->rs, sizeof(double), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  *iters = h.k;
  *rs = hrs;
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_gmg_eig_bounds_workspace_bytes(int32_t num_nodes, int32_t steps) {
  if (num_nodes < 1 || steps < 1) return 0;
  return eig_ws_bytes(num_nodes, steps);
}

int fx_gmg_eig_bounds_f64(void* stream, int32_t num_nodes, const int32_t* indptr, const int32_t* bcol,
                          const double* data, const double* dinv, int32_t steps, double* lmin, double* lmax,
                          void* workspace, size_t ws_bytes) {
  if (num_nodes < 1 || steps < 1 || steps > 1000 || !indptr || !bcol || !data || !dinv || !lmin || !lmax)
    return fail(FX_EINVAL, "fx_gmg_eig_bounds_f64: bad arguments");
  if (static_cast<long long>(num_nodes) * 3 > (1LL << 31) - 1)
    return fail(FX_ELIMIT, "fx_gmg_eig_bounds_f64: %d nodes exceed int32 DOF indices", num_nodes);
  const size_t need = eig_ws_bytes(num_nodes, steps);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_gmg_eig_bounds_f64: workspace %zu < %zu", ws_bytes, need);
  const int nn = num_nodes, n = 3 * nn;
  const int gv = v_grid(n), gm = mv_grid<double>(nn);
  char* w = static_cast<char*>(workspace);
  PcgState* st = reinterpret_cast<PcgState*>(w);
  const size_t nparts = static_cast<size_t>(std::max(gm, kMaxGridBlocks));
  double* partials = reinterpret_cast<double*>(w + al256(sizeof(PcgState)));
  double* hist = reinterpret_cast<double*>(w + al256(sizeof(PcgState)) + al256(nparts * sizeof(double)));
  char* vb = reinterpret_cast<char*>(hist) + al256(2 * static_cast<size_t>(steps) * 8);
  const size_t vs = al256(static_cast<size_t>(n) * sizeof(double));
  double* r = reinterpret_cast<double*>(vb);
  double* p = reinterpret_cast<double*>(vb + vs);
  double* q = reinterpret_cast<double*>(vb + 2 * vs);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(lz_init_kernel, dim3(gv), dim3(kBlock), 0, s, n, dinv, r, p, partials);
  hipLaunchKernelGGL(lz_fin_init, dim3(1), dim3(kFinBlock), 0, s, st, partials, gv);
  for (int k = 0; k < steps; ++k) {
    hipLaunchKernelGGL(cg_spmv_kernel, dim3(gm), dim3(kBlock), 0, s, nn, indptr, bcol, data, p, q, st, partials);
    hipLaunchKernelGGL(cg_fin_alpha, dim3(1), dim3(kFinBlock), 0, s, st, partials, gm);
    hipLaunchKernelGGL(lz_update_kernel, dim3(gv), dim3(kBlock), 0, s, n, r, q, dinv, st, partials);
    hipLaunchKernelGGL(lz_fin_beta, dim3(1), dim3(kFinBlock), 0, s, st, partials, gv, hist);
    hipLaunchKernelGGL(lz_dir_kernel, dim3(gv), dim3(kBlock), 0, s, n, p, r, dinv, st);
  }
  FX_LAUNCHED("fx_gmg_eig_bounds_f64");
  std::vector<double> h(2 * static_cast<size_t>(steps));
  FX_HIP(hipMemcpyAsync(h.data(), hist, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  // Steps up to the first breakdown (exact convergence on a small level).
  int m = 0;
  while (m < steps && std::isfinite(h[2 * m]) && h[2 * m] > 0.0 && std::isfinite(h[2 * m + 1]) &&
         h[2 * m + 1] > 0.0)
    ++m;
  if (m < steps && std::isfinite(h[2 * m]) && h[2 * m] > 0.0) ++m;  // its alpha is still valid
  if (m == 0) return fail(FX_EINVAL, "fx_gmg_eig_bounds_f64: operator not positive definite along the start vector");
  ritz_bounds(h, m, *lmin, *lmax);
  return FX_OK;
}

int fx_gmg_vcycle(void* stream, int32_t flags, int32_t nlev, const fx_gmg_level* levels, int32_t smooth_degree,
                  int32_t coarse_degree, double cheb_ratio, const double* r, double* z) {
  int rc = check_levels(flags, nlev, levels, smooth_degree, coarse_degree, cheb_ratio);
  if (rc) return rc;
  if (!r || !z) return fail(FX_EINVAL, "fx_gmg_vcycle: null vector");
  hipStream_t s = as_stream(stream);
  const int n = 3 * nodes_of(levels[0]);
  if (flags & FX_GMG_F32) {
    apply_vcycle<float>(s, make_vcycle<float>(nlev, levels, smooth_degree, coarse_degree, cheb_ratio, nullptr), r,
                        z, n);
  } else {
    apply_vcycle<double>(s, make_vcycle<double>(nlev, levels, smooth_degree, coarse_degree, cheb_ratio, nullptr),
                         r, z, n);
  }
  FX_LAUNCHED("fx_gmg_vcycle");
  return FX_OK;
}

size_t fx_pcg_gmg_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  const size_t nparts = static_cast<size_t>(std::max(mv_grid<double>(n / 3), 3 * kMaxGridBlocks));
  return al256(sizeof(PcgState)) + al256(nparts * sizeof(double)) +
         4 * al256(static_cast<size_t>(n) * sizeof(double));
}

int fx_pcg_gmg_f64(void* stream, int32_t flags, int32_t nlev, const fx_gmg_level* levels, int32_t smooth_degree,
                   int32_t coarse_degree, double cheb_ratio, const double* data, const double* b,
                   const double* x0, double tol, double atol, int32_t maxiter, int32_t check_every, double* x,
                   int32_t* iters, double* rs, void* workspace, size_t ws_bytes) {
  int rc = check_levels(flags, nlev, levels, smooth_degree, coarse_degree, cheb_ratio);
  if (rc) return rc;
  if (!data || !b || !x || !iters || !rs || maxiter < 0) return fail(FX_EINVAL, "fx_pcg_gmg_f64: bad arguments");
  const int n = 3 * nodes_of(levels[0]);
  const size_t need = fx_pcg_gmg_workspace_bytes(n);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_pcg_gmg_f64: workspace %zu < %zu", ws_bytes, need);
  if (check_every <= 0) check_every = 4;
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(workspace);
  if (flags & FX_GMG_F32)
    return pcg_gmg<float>(s, flags, nlev, levels, smooth_degree, coarse_degree, cheb_ratio, data, b, x0, tol, atol,
                          maxiter, check_every, x, iters, rs, w);
  return pcg_gmg<double>(s, flags, nlev, levels, smooth_degree, coarse_degree, cheb_ratio, data, b, x0, tol, atol,
                         maxiter, check_every, x, iters, rs, w);
}

size_t fx_gmg_galerkin_workspace_bytes(int32_t nxf, int32_t nyf, int32_t nzf, int32_t nxc, int32_t nyc,
                                       int32_t nzc) {
  const size_t nf = 3 * static_cast<size_t>(nxf + 1) * (nyf + 1) * (nzf + 1);
  const size_t nc = 3 * static_cast<size_t>(nxc + 1) * (nyc + 1) * (nzc + 1);
  return nf * nc * sizeof(double);
}

int fx_gmg_galerkin_dense_f64(void* stream, int32_t nxf, int32_t nyf, int32_t nzf, int32_t nxc, int32_t nyc,
                              int32_t nzc, const int32_t* indptr, const int32_t* indices, const double* data,
                              const uint8_t* mask_f, const uint8_t* mask_c, double* Ac, void* workspace,
                              size_t ws_bytes) {
  if (!coarsening_ok(nxf, nxc) || !coarsening_ok(nyf, nyc) || !coarsening_ok(nzf, nzc) || !indptr || !indices ||
      !data || !Ac)
    return fail(FX_EINVAL, "fx_gmg_galerkin_dense_f64: grids (%d,%d,%d) -> (%d,%d,%d) or null pointers", nxf, nyf,
                nzf, nxc, nyc, nzc);
  const size_t need = fx_gmg_galerkin_workspace_bytes(nxf, nyf, nzf, nxc, nyc, nzc);
  const size_t nc = 3 * static_cast<size_t>(nxc + 1) * (nyc + 1) * (nzc + 1);
  if (nc > 20000) return fail(FX_ELIMIT, "fx_gmg_galerkin_dense_f64: %zu coarse DOFs is too many for a dense matrix", nc);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_gmg_galerkin_dense_f64: workspace %zu < %zu", ws_bytes, need);
  hipStream_t s = as_stream(stream);
  double* B = static_cast<double*>(workspace);
  const size_t nb = need / sizeof(double);
  hipLaunchKernelGGL(galerkin_ap_kernel, dim3(static_cast<unsigned>((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                     nxf, nyf, nzf, nxc, nyc, nzc, indptr, indices, data, mask_f, mask_c, B);
  hipLaunchKernelGGL(galerkin_ptb_kernel, dim3(static_cast<unsigned>((nc * nc + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, s, nxf, nyf, nzf, nxc, nyc, nzc, B, mask_f, mask_c, Ac);
  FX_LAUNCHED("fx_gmg_galerkin_dense_f64");
  return FX_OK;
}

int fx_gmg_coarsen_f64(void* stream, int32_t loc, int32_t ncomp, int32_t nxf, int32_t nyf, int32_t nzf, int32_t nxc,
                       int32_t nyc, int32_t nzc, const double* fine, double* coarse) {
  if (!coarsening_ok(nxf, nxc) || !coarsening_ok(nyf, nyc) || !coarsening_ok(nzf, nzc) || !fine || !coarse)
    return fail(FX_EINVAL, "fx_gmg_coarsen_f64: grids (%d,%d,%d) -> (%d,%d,%d) or null pointers", nxf, nyf, nzf, nxc,
                nyc, nzc);
  hipStream_t s = as_stream(stream);
  if (loc == FX_PARAM_NODE) {
    if (ncomp < 1) return fail(FX_EINVAL, "fx_gmg_coarsen_f64: ncomp %d", ncomp);
    hipLaunchKernelGGL(coarsen_node_kernel, dim3(n_grid((nxc + 1) * (nyc + 1) * (nzc + 1))), dim3(kBlock), 0, s, nxf,
                       nyf, nzf, nxc, nyc, nzc, ncomp, fine, coarse);
  } else if (loc == FX_PARAM_CELL) {
    if (ncomp != 1) return fail(FX_EINVAL, "fx_gmg_coarsen_f64: cell fields have one component");
    hipLaunchKernelGGL(coarsen_cell_kernel, dim3(n_grid(nxc * nyc * nzc)), dim3(kBlock), 0, s, nxf, nyf, nzf, nxc,
                       nyc, nzc, fine, coarse);
  } else {
    return fail(FX_EINVAL, "fx_gmg_coarsen_f64: loc %d", loc);
  }
  FX_LAUNCHED("fx_gmg_coarsen_f64");
  return FX_OK;
}

int fx_csr_to_dense_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                        const double* data, double* dense) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !data || !dense)))
    return fail(FX_EINVAL, "fx_csr_to_dense_f64: bad arguments");
  if (n == 0) return FX_OK;
  hipLaunchKernelGGL(csr_dense_kernel, dim3(n), dim3(kBlock), 0, as_stream(stream), n, indptr, indices, data, dense);
  FX_LAUNCHED("fx_csr_to_dense_f64");
  return FX_OK;
}

}  // extern "C"
