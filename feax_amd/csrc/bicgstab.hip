// BiCGSTAB on the device (gfx950): the solver feax picks for a symmetric
// but indefinite or a general operator ("auto" -> bicgstab / gmres,
// feax/solvers/options.py:837-889) and the one a caller names for the
// non-symmetric operator of symmetric_elimination=False.  Iteration and
// stopping rule of jax.scipy.sparse.linalg.bicgstab (called at
// feax/solvers/common.py:241-252), right-preconditioned with M = Jacobi or
// the identity:
//
//   r = b - A x0, rhat = r, rho = alpha = omega = 1, p = q = r
//   while r.r > max(tol^2 b.b, atol^2) and k < maxiter and no breakdown:
//     rho' = rhat.r;  beta = rho'/rho * alpha/omega
//     p = r + beta (p - omega q);  phat = M p;  q = A phat
//     alpha = rho' / rhat.q;  s = r - alpha q
//     if s.s < atol2:  x += alpha phat;  r = s                  (early exit)
//     else:            shat = M s;  t = A shat;  omega = t.s / t.t
//                      x += alpha phat + omega shat;  r = s - omega t
//     breakdown: alpha = 0 (also when rho' = 0) or omega = 0 ends the loop
//
// One iteration = 2 SpMVs (block-3 or scalar CSR, FP64) with their dot
// products fused in, 3 vector kernels and 4 one-block finalizes that keep
// every scalar on the device; fixed-order reductions (bitwise repeatable);
// chunks of check_every iterations replayed as a HIP graph.
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

struct BcgState {
  double rho, rho_new, alpha, omega, beta, rs, atol2, ts, tt;
  int k, done, early, pad;
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

// y = A v for the rows of this group (block-3: rows 3g..3g+2).
template <bool B3>
__device__ __forceinline__ int spmv_rows(int n, const int* __restrict__ indptr, const int* __restrict__ idx,
                                         const double* __restrict__ data, const double* __restrict__ v, int g,
                                         int gl, double (&y)[3]) {
  if constexpr (B3) {
    if (g >= n / 3) return 0;
    b3_rows<double, double, kL, false>(indptr, idx, data, v, g, gl, y[0], y[1], y[2]);
    return 3;
  } else {
    if (g >= n) return 0;
    y[0] = csr_row<double, double, kL>(indptr, idx, data, v, g, gl);
    return 1;
  }
}

// q = A phat; partials of rhat.q.  t = A shat; partials of t.s and t.t
// (SECOND).  Skipped once done (and the second after an early exit).
template <bool B3, bool SECOND>
__global__ __launch_bounds__(kBlock) void bcg_spmv_kernel(int n, const int* __restrict__ indptr,
                                                          const int* __restrict__ idx,
                                                          const double* __restrict__ data,
                                                          const double* __restrict__ v, double* __restrict__ y,
                                                          const double* __restrict__ w,
                                                          const BcgState* __restrict__ st, double* partials,
                                                          int nparts) {
  if (st->done || (SECOND && st->early)) return;
  const int g = blockIdx.x * kGroups + threadIdx.x / kL, gl = threadIdx.x % kL;
  double out[3];
  const int rows = spmv_rows<B3>(n, indptr, idx, data, v, g, gl, out);
  DVec<2> acc;
  acc.v[0] = acc.v[1] = 0.0;
  if (gl == 0) {
    const int r0 = B3 ? 3 * g : g;
    for (int i = 0; i < rows; ++i) {
      y[r0 + i] = out[i];
      acc.v[0] += w[r0 + i] * out[i];  // rhat.q, or t.s
      acc.v[1] += out[i] * out[i];     // t.t
    }
  }
  const DVec<2> t = block_sum<2, kBlock>(acc);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t.v[0];
    if (SECOND) partials[nparts + blockIdx.x] = t.v[1];
  }
}

// r = b - q (q = A x0); rhat = p = q = r; partials of r.r and b.b.
__global__ __launch_bounds__(kBlock) void bcg_init_kernel(int n, const double* __restrict__ b,
                                                          double* __restrict__ q, double* __restrict__ r,
                                                          double* __restrict__ rhat, double* __restrict__ p,
                                                          double* partials) {
  int lb, begin, end;
  slice(n, lb, begin, end);
  DVec<2> acc;
  acc.v[0] = acc.v[1] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    const double ri = b[i] - q[i];
    r[i] = ri;
    rhat[i] = ri;
    p[i] = ri;
    q[i] = ri;
    acc.v[0] += ri * ri;
    acc.v[1] += b[i] * b[i];
  }
  store_partials<2>(block_sum<2, kBlock>(acc), lb, partials);
}

__global__ __launch_bounds__(kFinBlock) void bcg_fin_init(BcgState* st, const double* __restrict__ partials,
                                                          int np, double tol, double atol, int maxiter) {
  const DVec<2> t = sum_partials<2>(partials, np);  // r.r, b.b
  if (threadIdx.x == 0) {
    st->rho = st->alpha = st->omega = 1.0;
    st->rs = t.v[0];
    st->rho_new = t.v[0];  // rhat.r with rhat = r
    st->beta = st->rho_new / st->rho * st->alpha / st->omega;
    st->atol2 = fmax(tol * tol * t.v[1], atol * atol);
    st->k = 0;
    st->early = 0;
    st->done = !(st->rs > st->atol2 && 0 < maxiter);
  }
}

// p = r + beta (p - omega q); phat = M p.
__global__ __launch_bounds__(kBlock) void bcg_dir_kernel(int n, const double* __restrict__ r,
                                                         const double* __restrict__ q, double* __restrict__ p,
                                                         double* __restrict__ phat, const double* __restrict__ dinv,
                                                         const BcgState* __restrict__ st) {
  if (st->done) return;
  const double beta = st->beta, omega = st->omega;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double pi = r[i] + beta * (p[i] - omega * q[i]);
    p[i] = pi;
    phat[i] = dinv ? dinv[i] * pi : pi;
  }
}

// alpha = rho' / rhat.q
__global__ __launch_bounds__(kFinBlock) void bcg_fin_alpha(BcgState* st, const double* __restrict__ partials,
                                                           int np) {
  if (st->done) return;
  const DVec<1> t = sum_partials<1>(partials, np);
  if (threadIdx.x == 0) st->alpha = st->rho_new / t.v[0];
}

// s = r - alpha q; shat = M s; partials of s.s.
__global__ __launch_bounds__(kBlock) void bcg_s_kernel(int n, const double* __restrict__ r,
                                                       const double* __restrict__ q, double* __restrict__ s,
                                                       double* __restrict__ shat, const double* __restrict__ dinv,
                                                       const BcgState* __restrict__ st, double* partials) {
  if (st->done) return;
  int lb, begin, end;
  slice(n, lb, begin, end);
  const double alpha = st->alpha;
  DVec<1> acc;
  acc.v[0] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    const double si = r[i] - alpha * q[i];
    s[i] = si;
    shat[i] = dinv ? dinv[i] * si : si;
    acc.v[0] += si * si;
  }
  store_partials<1>(block_sum<1, kBlock>(acc), lb, partials);
}

__global__ __launch_bounds__(kFinBlock) void bcg_fin_s(BcgState* st, const double* __restrict__ partials, int np) {
  if (st->done) return;
  const DVec<1> t = sum_partials<1>(partials, np);
  if (threadIdx.x == 0) st->early = t.v[0] < st->atol2;
}

// Fixed-order sums of two partial arrays of np entries each, the second at
// partials + stride (the SpMV grid may exceed kMaxGridBlocks, the stride of
// sum_partials); valid in thread 0.  One block reduction.
__device__ __forceinline__ DVec<2> sum2_strided(const double* __restrict__ partials, int stride, int np) {
  DVec<2> acc;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double* src = partials + static_cast<size_t>(k) * stride;
    double a = 0.0, b = 0.0;
    int i = threadIdx.x;
    for (; i + kFinBlock < np; i += 2 * kFinBlock) {
      a += src[i];
      b += src[i + kFinBlock];
    }
    if (i < np) a += src[i];
    acc.v[k] = a + b;
  }
  return block_sum<2, kFinBlock>(acc);
}

// omega = t.s / t.t
__global__ __launch_bounds__(kFinBlock) void bcg_fin_omega(BcgState* st, const double* __restrict__ partials,
                                                           int np, int stride) {
  if (st->done || st->early) return;
  const DVec<2> t = sum2_strided(partials, stride, np);
  if (threadIdx.x == 0) {
    st->ts = t.v[0];
    st->tt = t.v[1];
    st->omega = t.v[0] / t.v[1];
  }
}

// x += alpha phat (+ omega shat); r = s (- omega t); partials of r.r and
// rhat.r.
__global__ __launch_bounds__(kBlock) void bcg_x_kernel(int n, double* __restrict__ x, double* __restrict__ r,
                                                       const double* __restrict__ s, const double* __restrict__ t,
                                                       const double* __restrict__ phat,
                                                       const double* __restrict__ shat,
                                                       const double* __restrict__ rhat,
                                                       const BcgState* __restrict__ st, double* partials) {
  if (st->done) return;
  int lb, begin, end;
  slice(n, lb, begin, end);
  const double alpha = st->alpha, omega = st->omega;
  const bool early = st->early;
  DVec<2> acc;
  acc.v[0] = acc.v[1] = 0.0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) {
    double ri;
    if (early) {
      x[i] += alpha * phat[i];
      ri = s[i];
    } else {
      x[i] += alpha * phat[i] + omega * shat[i];
      ri = s[i] - omega * t[i];
    }
    r[i] = ri;
    acc.v[0] += ri * ri;
    acc.v[1] += rhat[i] * ri;
  }
  store_partials<2>(block_sum<2, kBlock>(acc), lb, partials);
}

// rs = r.r; next rho' = rhat.r and beta; count, breakdown (alpha = 0, or
// omega = 0 after a full step) and stop test.
__global__ __launch_bounds__(kFinBlock) void bcg_fin_r(BcgState* st, const double* __restrict__ partials, int np,
                                                       int maxiter) {
  if (st->done) return;
  const DVec<2> t = sum_partials<2>(partials, np);
  if (threadIdx.x == 0) {
    const bool broke = st->alpha == 0.0 || (!st->early && st->omega == 0.0);
    st->rs = t.v[0];
    st->k += 1;
    st->rho = st->rho_new;
    st->rho_new = t.v[1];
    st->beta = st->rho_new / st->rho * st->alpha / st->omega;
    st->done = (broke || !(st->rs > st->atol2 && st->k < maxiter)) ? 1 : 0;
  }
}

template <bool B3>
int bicgstab(hipStream_t s, int n, int flags, const int* indptr, const int* idx, const double* data,
             const double* dinv, const double* b, const double* x0, double tol, double atol, int maxiter,
             int check_every, double* x, int* iters, double* rs, char* w) {
  const int gv = vgrid(n), gm = mgrid(n, B3);
  BcgState* st = reinterpret_cast<BcgState*>(w);
  const size_t np = static_cast<size_t>(std::max(gm, kMaxGridBlocks));
  double* vpart = reinterpret_cast<double*>(w + al(sizeof(BcgState)));       // 2 kMax: vector kernels
  double* mpart = reinterpret_cast<double*>(reinterpret_cast<char*>(vpart) + al(2 * kMaxGridBlocks * 8));
  char* vb = reinterpret_cast<char*>(mpart) + al(2 * np * 8);
  const size_t vs = al(static_cast<size_t>(n) * 8);
  double* r = reinterpret_cast<double*>(vb);
  double* rhat = reinterpret_cast<double*>(vb + vs);
  double* p = reinterpret_cast<double*>(vb + 2 * vs);
  double* q = reinterpret_cast<double*>(vb + 3 * vs);
  double* phat = reinterpret_cast<double*>(vb + 4 * vs);
  double* sv = reinterpret_cast<double*>(vb + 5 * vs);
  double* shat = reinterpret_cast<double*>(vb + 6 * vs);
  double* t = reinterpret_cast<double*>(vb + 7 * vs);
  if (x0) {
    if (x0 != x) FX_HIP(hipMemcpyAsync(x, x0, static_cast<size_t>(n) * 8, hipMemcpyDeviceToDevice, s));
  } else {
    FX_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * 8, s));
  }
  // q = A x0 through the first SpMV kernel, with the state cleared (a
  // previous solve may have left done set); the init finalize writes every
  // field.
  FX_HIP(hipMemsetAsync(st, 0, sizeof(BcgState), s));
  hipLaunchKernelGGL((bcg_spmv_kernel<B3, false>), dim3(gm), dim3(kBlock), 0, s, n, indptr, idx, data, x, q, b, st,
                     mpart, static_cast<int>(np));
  hipLaunchKernelGGL(bcg_init_kernel, dim3(gv), dim3(kBlock), 0, s, n, b, q, r, rhat, p, vpart);
  hipLaunchKernelGGL(bcg_fin_init, dim3(1), dim3(kFinBlock), 0, s, st, vpart, gv, tol, atol, maxiter);
  FX_LAUNCHED("fx_bicgstab_f64 init");
  auto iteration = [&](hipStream_t u) {
    hipLaunchKernelGGL(bcg_dir_kernel, dim3(gv), dim3(kBlock), 0, u, n, r, q, p, phat, dinv, st);
    hipLaunchKernelGGL((bcg_spmv_kernel<B3, false>), dim3(gm), dim3(kBlock), 0, u, n, indptr, idx, data, phat, q,
                       rhat, st, mpart, static_cast<int>(np));
    hipLaunchKernelGGL(bcg_fin_alpha, dim3(1), dim3(kFinBlock), 0, u, st, mpart, gm);
    hipLaunchKernelGGL(bcg_s_kernel, dim3(gv), dim3(kBlock), 0, u, n, r, q, sv, shat, dinv, st, vpart);
    hipLaunchKernelGGL(bcg_fin_s, dim3(1), dim3(kFinBlock), 0, u, st, vpart, gv);
    hipLaunchKernelGGL((bcg_spmv_kernel<B3, true>), dim3(gm), dim3(kBlock), 0, u, n, indptr, idx, data, shat, t,
                       sv, st, mpart, static_cast<int>(np));
    hipLaunchKernelGGL(bcg_fin_omega, dim3(1), dim3(kFinBlock), 0, u, st, mpart, gm, static_cast<int>(np));
    hipLaunchKernelGGL(bcg_x_kernel, dim3(gv), dim3(kBlock), 0, u, n, x, r, sv, t, phat, shat, rhat, st, vpart);
    hipLaunchKernelGGL(bcg_fin_r, dim3(1), dim3(kFinBlock), 0, u, st, vpart, gv, maxiter);
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
    if (e != hipSuccess) return hip_fail(e, "fx_bicgstab_f64 graph capture");
  }
  int launched = 0;
  while (!h.done && launched < maxiter) {
    if (gg.exec) {
      FX_HIP(hipGraphLaunch(gg.exec, s));
    } else {
      for (int c = 0; c < check_every; ++c) iteration(s);
      FX_LAUNCHED("fx_bicgstab_f64 iteration");
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

size_t fx_bicgstab_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  const size_t np = static_cast<size_t>(std::max(mgrid(n, false), kMaxGridBlocks));
  return al(sizeof(BcgState)) + al(2 * kMaxGridBlocks * 8) + al(2 * np * 8) + 8 * al(static_cast<size_t>(n) * 8);
}

int fx_bicgstab_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                    const int32_t* indices, const int32_t* bcol, const double* data, const double* dinv,
                    const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                    int32_t check_every, double* x, int32_t* iters, double* rs, void* workspace, size_t ws_bytes) {
  const bool b3 = (flags & FX_PCG_BLOCK3) != 0;
  if (n < 0 || nse < 0 || !indptr || !data || !b || !x || !iters || !rs || maxiter < 0 || (b3 && (!bcol || n % 3)) ||
      (!b3 && !indices))
    return fail(FX_EINVAL, "fx_bicgstab_f64: bad arguments");
  const size_t need = fx_bicgstab_workspace_bytes(n);
  if (!workspace || ws_bytes < need) return fail(FX_EWORKSPACE, "fx_bicgstab_f64: workspace %zu < %zu", ws_bytes, need);
  if (n == 0) {
    *iters = 0;
    *rs = 0.0;
    return FX_OK;
  }
  if (check_every <= 0) check_every = 16;
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(workspace);
  if (b3)
    return bicgstab<true>(s, n, flags, indptr, bcol, data, dinv, b, x0, tol, atol, maxiter, check_every, x, iters,
                          rs, w);
  return bicgstab<false>(s, n, flags, indptr, indices, data, dinv, b, x0, tol, atol, maxiter, check_every, x, iters,
                         rs, w);
}

}  // extern "C"
