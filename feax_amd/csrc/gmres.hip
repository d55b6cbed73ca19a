// Restarted GMRES on the device (gfx950): the solver feax picks for a
// general operator ("auto" -> gmres, feax/solvers/options.py:837-889) and
// one a caller may name for the non-symmetric operator of
// symmetric_elimination=False.  Iteration and stopping rule of
// jax.scipy.sparse.linalg.gmres with its default solve_method='batched'
// (called at feax/solvers/common.py:254-265), left-preconditioned with
// M = Jacobi or the identity:
//
//   ptol = ||M b|| * min(1, max(tol ||b||, atol) / ||b||)
//   v0, rn = normalise(M (b - A x0))          (||.|| <= eps counts as 0)
//   while cycles < maxiter and rn > ptol:
//     H = eye(R, R+1)
//     for j < R until breakdown:
//       w = M A v_j;  w0 = ||w||
//       h = V^T w (one classical Gram-Schmidt pass);  w -= V h
//       v_{j+1}, w1 = normalise(w, threshold eps * w0);  H[j, :] = h, H[j, j+1] = w1
//       breakdown: w1 = 0
//     y = (H H^T)^{-1} H (rn e0)      (normal equations, Cholesky)
//     x += V[:, :R] y;  v0, rn = normalise(M (b - A x))
//
// One Arnoldi step = an SpMV with M and ||w||^2 fused in, a multi-column
// dot kernel (8 basis vectors per pass over w), the orthogonalisation with
// ||w||^2 fused in, two one-block finalizes and the normalisation.  Every
// scalar stays on the device; reductions are fixed-order (bitwise
// repeatable); chunks of whole cycles are replayed as a HIP graph.  The
// Krylov basis V is (R+1) x n FP64 in HBM (R = restart <= 64).
#include <algorithm>
#include <cstdint>

#include "common.h"
#include "feaxhip.h"
#include "reduce.h"
#include "spmv_common.h"

namespace fx {
namespace {

constexpr int kL = 32;                // lanes per row / node
constexpr int kGroups = kBlock / kL;  // rows / nodes per SpMV block
constexpr int kMaxRestart = 64;
constexpr int kCols = 8;              // basis vectors per dot pass
constexpr double kEps = 2.220446049250313e-16;

struct GmState {
  double rn, ptol, w0, w1;
  int k, done, stop, m;
};

inline int vgrid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, kMaxGridBlocks)); }
inline int mgrid(int n, bool b3) {
  const int groups = b3 ? n / 3 : n;
  return std::max(1, (groups + kGroups - 1) / kGroups);
}
inline size_t al(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

__device__ __forceinline__ void slice(int n, int& lb, int& begin, int& end) {
  lb = xcd_block(blockIdx.x, gridDim.x);
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  begin = min(n, lb * chunk);
  end = min(n, begin + chunk);
}

// Norm with jax's safe-normalise threshold: ||v|| if above thresh, else 0.
__device__ __forceinline__ double safe_norm(double ss, double thresh) {
  const double nrm = sqrt(ss);
  return nrm > thresh ? nrm : 0.0;
}

// RESID: y = M (b - A v), else y = M A v (skipped once the cycle broke
// down).  Partials of y.y, one per block.
template <bool B3, bool RESID>
__global__ __launch_bounds__(kBlock) void gm_spmv_kernel(int n, const int* __restrict__ indptr,
                                                         const int* __restrict__ idx,
                                                         const double* __restrict__ data,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ dinv,
                                                         const double* __restrict__ b, double* __restrict__ y,
                                                         const GmState* __restrict__ st, double* partials) {
  if (st->done || (!RESID && st->stop)) return;
  const int g = blockIdx.x * kGroups + threadIdx.x / kL, gl = threadIdx.x % kL;
  double out[3];
  int rows = 0;
  if constexpr (B3) {
    if (g < n / 3) {
      b3_rows<double, double, kL, false>(indptr, idx, data, v, g, gl, out[0], out[1], out[2]);
      rows = 3;
    }
  } else {
    if (g < n) {
      out[0] = csr_row<double, double, kL>(indptr, idx, data, v, g, gl);
      rows = 1;
    }
  }
  DVec<1> acc;
  acc.v[0] = 0.0;
  if (gl == 0) {
    const int r0 = B3 ? 3 * g : g;
    for (int i = 0; i < rows; ++i) {
      double yi = RESID ? b[r0 + i] - out[i] : out[i];
      if (dinv) yi = dinv[r0 + i] * yi;
      y[r0 + i] = yi;
      acc.v[0] += yi * yi;
    }
  }
  const DVec<1> t = block_sum<1, kBlock>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t.v[0];
}

// Partials of v_c . w for the basis vectors c in [kCols*blockIdx.y,
// min(kCols*(blockIdx.y+1), ncols)): partials[c * kMaxGridBlocks + block].
__global__ __launch_bounds__(kBlock) void gm_dots_kernel(int n, const double* __restrict__ V, size_t ld, int ncols,
                                                         const double* __restrict__ w,
                                                         const GmState* __restrict__ st, double* partials) {
  if (st->done || st->stop) return;
  int lb, begin, end;
  slice(n, lb, begin, end);
  const int c0 = blockIdx.y * kCols;
  const int nc = min(kCols, ncols - c0);
  DVec<kCols> acc;
#pragma unroll
  for (int c = 0; c < kCols; ++c) acc.v[c] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    const double wi = w[i];
#pragma unroll
    for (int c = 0; c < kCols; ++c)
      if (c < nc) acc.v[c] += V[(c0 + c) * ld + i] * wi;
  }
  const DVec<kCols> t = block_sum<kCols, kBlock>(acc);
  if (threadIdx.x == 0) {
    for (int c = 0; c < nc; ++c) partials[(c0 + c) * kMaxGridBlocks + lb] = t.v[c];
  }
}

// w0 = ||M A v_j|| (thresh eps) from the SpMV partials; h_c for c < ncols
// from the dot partials, one wave per basis vector (fixed order).
__global__ __launch_bounds__(kFinBlock) void gm_fin_h(GmState* st, const double* __restrict__ spart, int nsp,
                                                      const double* __restrict__ dpart, int ndp, int ncols,
                                                      double* __restrict__ h) {
  if (st->done || st->stop) return;
  const DVec<1> t = sum_partials<1>(spart, nsp);
  if (threadIdx.x == 0) st->w0 = safe_norm(t.v[0], kEps);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int c = wid; c < ncols; c += kFinBlock / 64) {
    const double* src = dpart + static_cast<size_t>(c) * kMaxGridBlocks;
    double a = 0.0;
    for (int i = lane; i < ndp; i += 64) a += src[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) h[c] = a;
  }
}

// w -= sum_c h_c v_c (c < ncols); partials of w.w.
__global__ __launch_bounds__(kBlock) void gm_orth_kernel(int n, const double* __restrict__ V, size_t ld, int ncols,
                                                         const double* __restrict__ h, double* __restrict__ w,
                                                         const GmState* __restrict__ st, double* partials) {
  if (st->done || st->stop) return;
  int lb, begin, end;
  slice(n, lb, begin, end);
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    double s = 0.0;
    for (int c = 0; c < ncols; ++c) s += V[c * ld + i] * h[c];
    const double wi = w[i] - s;
    w[i] = wi;
    acc.v[0] += wi * wi;
  }
  store_partials<1>(block_sum<1, kBlock>(acc), lb, partials);
}

// w1 = ||w|| (thresh eps * w0); row j of H; breakdown stops the cycle.
__global__ __launch_bounds__(kFinBlock) void gm_fin_norm(GmState* st, const double* __restrict__ partials, int np,
                                                         int j, int R, const double* __restrict__ h,
                                                         double* __restrict__ H) {
  if (st->done || st->stop) return;
  const DVec<1> t = sum_partials<1>(partials, np);
  __shared__ double w1s;
  if (threadIdx.x == 0) {
    const double w1 = safe_norm(t.v[0], kEps * st->w0);
    st->w1 = w1;
    st->m = j + 1;
    if (w1 == 0.0) st->stop = 1;
    w1s = w1;
  }
  __syncthreads();
  for (int c = threadIdx.x; c <= R; c += kFinBlock)
    H[j * (R + 1) + c] = c <= j ? h[c] : (c == j + 1 ? w1s : 0.0);
}

// out = w / nrm, or 0 when nrm = 0 (the normalised basis vector).  ARNOLDI
// reads w1 and skips after a breakdown; otherwise the residual norm.
template <bool ARNOLDI>
__global__ __launch_bounds__(kBlock) void gm_scale_kernel(int n, const double* __restrict__ w,
                                                          double* __restrict__ out,
                                                          const GmState* __restrict__ st) {
  if (st->done || (ARNOLDI && st->stop)) return;
  const double nrm = ARNOLDI ? st->w1 : st->rn;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    out[i] = nrm != 0.0 ? w[i] / nrm : 0.0;
}

// y = (H H^T)^{-1} H (rn e0) over the m steps taken (the rows of H past m
// are rows of the identity, so y is zero there): one wave, Cholesky of the
// m x m normal matrix in LDS.
__global__ __launch_bounds__(64) void gm_lstsq(GmState* st, int R, const double* __restrict__ H,
                                               double* __restrict__ y) {
  if (st->done) return;
  __shared__ double a[kMaxRestart][kMaxRestart + 1];
  __shared__ double rhs[kMaxRestart];
  const int m = st->m, lane = threadIdx.x;
  const double rn = st->rn;
  // a = H H^T (rows p, q < m), rhs = H[:, 0] rn
  for (int p = 0; p < m; ++p) {
    for (int q = lane; q <= p; q += 64) {
      double s = 0.0;
      for (int c = 0; c <= R; ++c) s += H[p * (R + 1) + c] * H[q * (R + 1) + c];
      a[p][q] = s;
    }
  }
  for (int p = lane; p < m; p += 64) rhs[p] = H[p * (R + 1)] * rn;
  __syncthreads();
  // Right-looking Cholesky, lower triangle: lanes own rows i > k.
  for (int k = 0; k < m; ++k) {
    const double d = sqrt(a[k][k]);
    __syncthreads();
    if (lane == 0) a[k][k] = d;
    for (int i = k + 1 + lane; i < m; i += 64) a[i][k] /= d;
    __syncthreads();
    for (int i = k + 1 + lane; i < m; i += 64) {
      const double lik = a[i][k];
      for (int jj = k + 1; jj <= i; ++jj) a[i][jj] -= lik * a[jj][k];
    }
    __syncthreads();
  }
  if (lane == 0) {
    for (int i = 0; i < m; ++i) {  // L z = rhs
      double s = rhs[i];
      for (int k = 0; k < i; ++k) s -= a[i][k] * rhs[k];
      rhs[i] = s / a[i][i];
    }
    for (int i = m - 1; i >= 0; --i) {  // L^T y = z
      double s = rhs[i];
      for (int k = i + 1; k < m; ++k) s -= a[k][i] * y[k];
      y[i] = s / a[i][i];
    }
  }
}

// x += sum_{c < m} y_c v_c.
__global__ __launch_bounds__(kBlock) void gm_update_kernel(int n, const double* __restrict__ V, size_t ld,
                                                           const double* __restrict__ y, double* __restrict__ x,
                                                           const GmState* __restrict__ st) {
  if (st->done) return;
  const int m = st->m;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double s = 0.0;
    for (int c = 0; c < m; ++c) s += V[c * ld + i] * y[c];
    x[i] += s;
  }
}

// Partials of b.b and (M b).(M b).
__global__ __launch_bounds__(kBlock) void gm_bnorm_kernel(int n, const double* __restrict__ b,
                                                          const double* __restrict__ dinv, double* partials) {
  int lb, begin, end;
  slice(n, lb, begin, end);
  DVec<2> acc;
  acc.v[0] = acc.v[1] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    const double bi = b[i], mb = dinv ? dinv[i] * bi : bi;
    acc.v[0] += bi * bi;
    acc.v[1] += mb * mb;
  }
  store_partials<2>(block_sum<2, kBlock>(acc), lb, partials);
}

__global__ __launch_bounds__(kFinBlock) void gm_fin_ptol(GmState* st, const double* __restrict__ partials, int np,
                                                         double tol, double atol) {
  const DVec<2> t = sum_partials<2>(partials, np);
  if (threadIdx.x == 0) {
    const double bn = sqrt(t.v[0]), mbn = sqrt(t.v[1]);
    const double ratio = fmax(tol * bn, atol) / bn;  // inf for b = 0, NaN when atol = 0 too
    st->ptol = isnan(ratio) ? ratio : mbn * (ratio < 1.0 ? ratio : 1.0);
    st->k = 0;
    st->done = 0;
  }
}

// rn = ||M (b - A x)|| (thresh eps); COUNT: one more cycle; stop test; a
// fresh H = eye(R, R+1) for the next cycle.
template <bool COUNT>
__global__ __launch_bounds__(kFinBlock) void gm_fin_res(GmState* st, const double* __restrict__ partials, int np,
                                                        int maxiter, int R, double* __restrict__ H) {
  if (st->done) return;
  const DVec<1> t = sum_partials<1>(partials, np);
  if (threadIdx.x == 0) {
    st->rn = safe_norm(t.v[0], kEps);
    if (COUNT) st->k += 1;
    st->done = !(st->k < maxiter && st->rn > st->ptol);
    st->stop = 0;
    st->m = 0;
  }
  for (int e = threadIdx.x; e < R * (R + 1); e += kFinBlock) H[e] = (e / (R + 1) == e % (R + 1)) ? 1.0 : 0.0;
}

struct Layout {
  GmState* st;
  double *H, *h, *y, *vpart, *spart, *V, *w;
  size_t ld;
};

inline Layout layout(char* w, int n, int R, bool b3) {
  Layout L;
  const size_t nsp = static_cast<size_t>(std::max(mgrid(n, b3), kMaxGridBlocks));
  L.st = reinterpret_cast<GmState*>(w);
  char* p = w + al(sizeof(GmState));
  L.H = reinterpret_cast<double*>(p);
  p += al(static_cast<size_t>(R) * (R + 1) * 8);
  L.h = reinterpret_cast<double*>(p);
  p += al(static_cast<size_t>(R + 1) * 8);
  L.y = reinterpret_cast<double*>(p);
  p += al(static_cast<size_t>(R) * 8);
  L.vpart = reinterpret_cast<double*>(p);
  p += al(static_cast<size_t>(std::max(R + 1, 2)) * kMaxGridBlocks * 8);
  L.spart = reinterpret_cast<double*>(p);
  p += al(nsp * 8);
  L.ld = al(static_cast<size_t>(n) * 8) / 8;
  L.V = reinterpret_cast<double*>(p);
  p += static_cast<size_t>(R + 1) * L.ld * 8;
  L.w = reinterpret_cast<double*>(p);
  return L;
}

inline size_t workspace_bytes(int n, int R, bool b3) {
  const size_t nsp = static_cast<size_t>(std::max(mgrid(n, b3), kMaxGridBlocks));
  const size_t vec = al(static_cast<size_t>(n) * 8);
  return al(sizeof(GmState)) + al(static_cast<size_t>(R) * (R + 1) * 8) + al(static_cast<size_t>(R + 1) * 8) +
         al(static_cast<size_t>(R) * 8) + al(static_cast<size_t>(std::max(R + 1, 2)) * kMaxGridBlocks * 8) +
         al(nsp * 8) + static_cast<size_t>(R + 2) * vec;
}

template <bool B3>
int gmres(hipStream_t s, int n, int flags, const int* indptr, const int* idx, const double* data,
          const double* dinv, const double* b, const double* x0, double tol, double atol, int R, int maxiter,
          int check_every, double* x, int* cycles, double* rn_out, char* wsp) {
  const int gv = vgrid(n), gm = mgrid(n, B3);
  const Layout L = layout(wsp, n, R, B3);
  if (x0) {
    if (x0 != x) FX_HIP(hipMemcpyAsync(x, x0, static_cast<size_t>(n) * 8, hipMemcpyDeviceToDevice, s));
  } else {
    FX_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * 8, s));
  }
  // Cleared state (a previous solve may have left done set); fin_ptol and
  // the first fin_res write every field the cycle reads.
  FX_HIP(hipMemsetAsync(L.st, 0, sizeof(GmState), s));
  hipLaunchKernelGGL(gm_bnorm_kernel, dim3(gv), dim3(kBlock), 0, s, n, b, dinv, L.vpart);
  hipLaunchKernelGGL(gm_fin_ptol, dim3(1), dim3(kFinBlock), 0, s, L.st, L.vpart, gv, tol, atol);
  hipLaunchKernelGGL((gm_spmv_kernel<B3, true>), dim3(gm), dim3(kBlock), 0, s, n, indptr, idx, data, x, dinv, b,
                     L.w, L.st, L.spart);
  hipLaunchKernelGGL((gm_fin_res<false>), dim3(1), dim3(kFinBlock), 0, s, L.st, L.spart, gm, maxiter, R, L.H);
  hipLaunchKernelGGL(gm_scale_kernel<false>, dim3(gv), dim3(kBlock), 0, s, n, L.w, L.V, L.st);
  FX_LAUNCHED("fx_gmres_f64 init");
  auto cycle = [&](hipStream_t u) {
    for (int j = 0; j < R; ++j) {
      const double* vj = L.V + static_cast<size_t>(j) * L.ld;
      double* vj1 = L.V + static_cast<size_t>(j + 1) * L.ld;
      const int ncols = j + 1;
      hipLaunchKernelGGL((gm_spmv_kernel<B3, false>), dim3(gm), dim3(kBlock), 0, u, n, indptr, idx, data, vj, dinv,
                         b, L.w, L.st, L.spart);
      hipLaunchKernelGGL(gm_dots_kernel, dim3(gv, (ncols + kCols - 1) / kCols), dim3(kBlock), 0, u, n, L.V, L.ld,
                         ncols, L.w, L.st, L.vpart);
      hipLaunchKernelGGL(gm_fin_h, dim3(1), dim3(kFinBlock), 0, u, L.st, L.spart, gm, L.vpart, gv, ncols, L.h);
      hipLaunchKernelGGL(gm_orth_kernel, dim3(gv), dim3(kBlock), 0, u, n, L.V, L.ld, ncols, L.h, L.w, L.st,
                         L.vpart);
      hipLaunchKernelGGL(gm_fin_norm, dim3(1), dim3(kFinBlock), 0, u, L.st, L.vpart, gv, j, R, L.h, L.H);
      hipLaunchKernelGGL(gm_scale_kernel<true>, dim3(gv), dim3(kBlock), 0, u, n, L.w, vj1, L.st);
    }
    hipLaunchKernelGGL(gm_lstsq, dim3(1), dim3(64), 0, u, L.st, R, L.H, L.y);
    hipLaunchKernelGGL(gm_update_kernel, dim3(gv), dim3(kBlock), 0, u, n, L.V, L.ld, L.y, x, L.st);
    hipLaunchKernelGGL((gm_spmv_kernel<B3, true>), dim3(gm), dim3(kBlock), 0, u, n, indptr, idx, data, x, dinv, b,
                       L.w, L.st, L.spart);
    hipLaunchKernelGGL((gm_fin_res<true>), dim3(1), dim3(kFinBlock), 0, u, L.st, L.spart, gm, maxiter, R, L.H);
    hipLaunchKernelGGL(gm_scale_kernel<false>, dim3(gv), dim3(kBlock), 0, u, n, L.w, L.V, L.st);
  };
  struct {
    int k, done;
  } hs = {0, 1};
  FX_HIP(hipMemcpyAsync(&hs, &L.st->k, sizeof(hs), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  GraphGuard gg;
  if (!(flags & FX_PCG_NO_GRAPH) && !hs.done) {
    hipError_t e = hipStreamCreateWithFlags(&gg.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(gg.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int c = 0; c < check_every; ++c) cycle(gg.stream);
      e = hipStreamEndCapture(gg.stream, &gg.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "fx_gmres_f64 graph capture");
  }
  int launched = 0;
  while (!hs.done && launched < maxiter) {
    if (gg.exec) {
      FX_HIP(hipGraphLaunch(gg.exec, s));
    } else {
      for (int c = 0; c < check_every; ++c) cycle(s);
      FX_LAUNCHED("fx_gmres_f64 cycle");
    }
    launched += check_every;
    FX_HIP(hipMemcpyAsync(&hs, &L.st->k, sizeof(hs), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
  }
  double hrn = 0.0;
  FX_HIP(hipMemcpyAsync(&hrn, &L.st->rn, sizeof(double), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  *cycles = hs.k;
  *rn_out = hrn;
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_gmres_workspace_bytes(int32_t n, int32_t restart, int32_t flags) {
  if (n < 0 || restart < 1 || restart > kMaxRestart) return 0;
  const int R = std::max(1, std::min(restart, n));
  return workspace_bytes(n, R, (flags & FX_PCG_BLOCK3) != 0);
}

int fx_gmres_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const int32_t* indices,
                 const int32_t* bcol, const double* data, const double* dinv, const double* b, const double* x0,
                 double tol, double atol, int32_t restart, int32_t maxiter, int32_t check_every, double* x,
                 int32_t* cycles, double* rn, void* workspace, size_t ws_bytes) {
  const bool b3 = (flags & FX_PCG_BLOCK3) != 0;
  if (n < 0 || nse < 0 || !indptr || !data || !b || !x || !cycles || !rn || maxiter < 0 ||
      (b3 && (!bcol || n % 3)) || (!b3 && !indices))
    return fail(FX_EINVAL, "fx_gmres_f64: bad arguments");
  if (restart < 1 || restart > kMaxRestart)
    return fail(FX_EINVAL, "fx_gmres_f64: restart %d outside [1, %d]", restart, kMaxRestart);
  const size_t need = fx_gmres_workspace_bytes(n, restart, flags);
  if (!workspace || ws_bytes < need) return fail(FX_EWORKSPACE, "fx_gmres_f64: workspace %zu < %zu", ws_bytes, need);
  if (n == 0) {
    *cycles = 0;
    *rn = 0.0;
    return FX_OK;
  }
  if (check_every <= 0) check_every = 1;
  const int R = std::min(restart, n);
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(workspace);
  if (b3)
    return gmres<true>(s, n, flags, indptr, bcol, data, dinv, b, x0, tol, atol, R, maxiter, check_every, x, cycles,
                       rn, w);
  return gmres<false>(s, n, flags, indptr, indices, data, dinv, b, x0, tol, atol, R, maxiter, check_every, x, cycles,
                      rn, w);
}

}  // extern "C"
