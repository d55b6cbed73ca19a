// Direct solve on the device: banded Cholesky of the SPD J_bc (symmetric
// Dirichlet elimination), the MI355X stand-in for feax's GPU direct solver
// (DirectSolverOptions "cudss" / "auto", feax/solvers/common.py:116-210,
// feax/solvers/linear.py:308-417).
//
// Storage: LAPACK lower band layout, column j at ab + j * ldab, entry
// (i, j), 0 <= i - j < ldab, at offset i - j.  Then entry (i, j) sits at
// ab + i + j * (ldab - 1): every block of the band is an ordinary
// column-major matrix with leading dimension ldab - 1 as long as the
// blocks's entries keep 0 <= i - j < ldab.  With ldab = bw + 2 nb the
// blocked right-looking factorisation stays inside that range:
//
//   for each block column [c0, c0 + w), w <= nb:
//     L11 = potrf(A11)                    (rocSOLVER, w x w)
//     L21 = A21 L11^-T                    (rocBLAS trsm, h x w, h <= bw)
//     A22 -= L21 L21^T  (lower triangle)  (rocBLAS syrk, h x h)
//
// The entries of L21 past the bandwidth are stored zeros and stay zero
// (Cholesky fills only inside the envelope).  The triangular solves are
// hand-written: one launch per 64-column block, the block's diagonal
// triangle solved by one wave from LDS, the coupling to the next block
// (forward) or the previous bandwidth rows (backward) as coalesced
// multiply-adds.  Rows are in the caller's ordering `perm` (new -> old),
// chosen to reduce the bandwidth.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "common.h"
#include "feaxhip.h"

namespace fx {
namespace {

constexpr int kFactorNb = 128;  // block columns of the factorisation
constexpr int kSolveNb = 64;    // block of the triangular solves (one wave)
constexpr int kThreads = 256;

__device__ __forceinline__ size_t boff(int64_t i, int64_t j, int64_t ldab) {
  return static_cast<size_t>(i - j) + static_cast<size_t>(j) * static_cast<size_t>(ldab);
}

// ab[(i' - j') + j' ldab] = data[s] for the lower-triangle slots (i' >= j'
// in the new numbering, i' = inv[i]).  One wave per CSR row.  *bad = 1 if a
// slot lies past the bandwidth (the caller's bw does not fit the pattern).
__global__ __launch_bounds__(kThreads) void band_scatter_kernel(int n, const int* __restrict__ indptr,
                                                                const int* __restrict__ indices,
                                                                const double* __restrict__ data,
                                                                const int* __restrict__ inv, int bw, int64_t ldab,
                                                                double* __restrict__ ab, int* __restrict__ bad) {
  const int row = blockIdx.x * (kThreads / 64) + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (row >= n) return;
  const int ip = inv ? inv[row] : row;
  const int end = indptr[row + 1];
  for (int s = indptr[row] + lane; s < end; s += 64) {
    const int col = indices[s];
    const int jp = inv ? inv[col] : col;
    if (jp > ip) continue;
    if (ip - jp > bw) {
      bad[0] = 1;
      continue;
    }
    ab[boff(ip, jp, ldab)] = data[s];
  }
}

// y[i'] = b[perm[i']]  /  x[perm[i']] = y[i']
__global__ __launch_bounds__(kThreads) void permute_kernel(int n, const int* __restrict__ perm,
                                                           const double* __restrict__ src, double* __restrict__ dst,
                                                           int scatter) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int p = perm ? perm[i] : i;
    if (scatter)
      dst[p] = src[i];
    else
      dst[i] = src[p];
  }
}

// Diagonal block [c, c + w) of L into LDS: Ld[i][j] = L(c + i, c + j), i >= j.
__device__ __forceinline__ void load_diag(const double* __restrict__ ab, int64_t ldab, int c, int w,
                                          double (*Ld)[kSolveNb + 1]) {
  const int lane = threadIdx.x & 63;
  for (int j = 0; j < w; ++j)
    if (lane >= j && lane < w) Ld[lane][j] = ab[boff(c + lane, c + j, ldab)];
}

// Forward step.  On entry y[cK, cK + wK) is final (cK < 0: nothing yet).
// Block 0: rows of the next block [c1, c1 + w1) take the coupling to block
// K, then wave 0 solves the next diagonal triangle.  Other blocks: rows
// [c1 + w1, min(n, cK + wK + bw)) take the coupling to block K.
__global__ __launch_bounds__(kThreads) void band_fwd_kernel(int n, int bw, int64_t ldab,
                                                            const double* __restrict__ ab, double* __restrict__ y,
                                                            int cK, int wK) {
  __shared__ double yk[kSolveNb];
  __shared__ double Ld[kSolveNb][kSolveNb + 1];
  const int c1 = cK < 0 ? 0 : cK + wK;
  const int w1 = min(kSolveNb, n - c1);
  if (cK >= 0 && threadIdx.x < wK) yk[threadIdx.x] = y[cK + threadIdx.x];
  __syncthreads();
  if (blockIdx.x == 0) {
    const int t = threadIdx.x;
    double v = 0.0;
    if (t < 64) {
      load_diag(ab, ldab, c1, w1, Ld);
      if (t < w1) {
        v = y[c1 + t];
        if (cK >= 0)
          for (int j = 0; j < wK; ++j) v -= ab[boff(c1 + t, cK + j, ldab)] * yk[j];
      }
    }
    __syncthreads();
    if (t >= 64) return;
    for (int jj = 0; jj < w1; ++jj) {
      const double xj = __shfl(v, jj, 64) / Ld[jj][jj];
      if (t == jj)
        v = xj;
      else if (t > jj && t < w1)
        v -= Ld[t][jj] * xj;
    }
    if (t < w1) y[c1 + t] = v;
    return;
  }
  const int r = c1 + w1 + (blockIdx.x - 1) * kThreads + threadIdx.x;
  const int rend = min(n, cK + wK + bw);
  if (r >= rend) return;
  double s = 0.0;
  for (int j = 0; j < wK; ++j) s += ab[boff(r, cK + j, ldab)] * yk[j];
  y[r] -= s;
}

// Backward step (L^T x = y).  On entry y[cK, cK + wK) holds final x (the
// last block's launch has cK = n: nothing yet).  Rows j of the previous
// bandwidth take sum_{r in K} L(r, j) x_r (one wave per j, coalesced in
// r); block 0 owns the previous block [c1, c1 + 64) and then solves its
// transposed triangle with wave 0.
__global__ __launch_bounds__(kThreads) void band_bwd_kernel(int n, int bw, int64_t ldab,
                                                            const double* __restrict__ ab, double* __restrict__ y,
                                                            int cK, int wK, int c1, int w1) {
  __shared__ double xk[kSolveNb];
  __shared__ double Ld[kSolveNb][kSolveNb + 1];
  __shared__ double upd[kSolveNb];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x < wK) xk[threadIdx.x] = y[cK + threadIdx.x];
  __syncthreads();
  if (blockIdx.x == 0) {
    // coupling of the rows of block [c1, c1 + w1) to block K
    for (int j = wid; j < w1; j += kThreads / 64) {
      double s = 0.0;
      if (lane < wK) s = ab[boff(cK + lane, c1 + j, ldab)] * xk[lane];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) upd[j] = s;
    }
    if (wid == 0) load_diag(ab, ldab, c1, w1, Ld);
    __syncthreads();
    if (wid != 0) return;
    const int t = lane;
    double v = t < w1 ? y[c1 + t] - (wK > 0 ? upd[t] : 0.0) : 0.0;
    for (int jj = w1 - 1; jj >= 0; --jj) {
      const double xj = __shfl(v, jj, 64) / Ld[jj][jj];
      if (t == jj)
        v = xj;
      else if (t < jj)
        v -= Ld[jj][t] * xj;
    }
    if (t < w1) y[c1 + t] = v;
    return;
  }
  if (wK == 0) return;
  const int jlo = max(0, cK - bw);
  const int j = c1 - 1 - ((blockIdx.x - 1) * (kThreads / 64) + wid);
  if (j < jlo) return;
  double s = 0.0;
  if (lane < wK) s = ab[boff(cK + lane, j, ldab)] * xk[lane];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[j] -= s;
}

// One rocBLAS handle per device for the process (rocSOLVER runs on it).
rocblas_handle blas_handle(int& err) {
  static std::mutex mu;
  static std::vector<rocblas_handle> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    err = 1;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (static_cast<int>(handles.size()) <= dev) handles.resize(dev + 1, nullptr);
  if (!handles[dev] && rocblas_create_handle(&handles[dev]) != rocblas_status_success) {
    handles[dev] = nullptr;
    err = 1;
  }
  return handles[dev];
}

inline int blas_fail(rocblas_status st, const char* where) {
  return fail(FX_ELIB, "%s: rocBLAS/rocSOLVER status %d", where, static_cast<int>(st));
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

int64_t fx_band_ldab(int32_t bw) { return bw < 0 ? -1 : static_cast<int64_t>(bw) + 2 * kFactorNb; }

size_t fx_band_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  const size_t nb = (static_cast<size_t>(n) + kFactorNb - 1) / kFactorNb;
  return 256 + ((nb * sizeof(int) + 255) & ~static_cast<size_t>(255)) + static_cast<size_t>(n) * 8;
}

int fx_band_from_csr_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                         const int32_t* inv_perm, int32_t bw, int64_t ldab, double* ab, void* workspace,
                         size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || (n > 0 && (!indptr || !indices || !data || !ab)))
    return fail(FX_EINVAL, "fx_band_from_csr_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n))
    return fail(FX_EWORKSPACE, "fx_band_from_csr_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n));
  if (n == 0) return FX_OK;
  hipStream_t s = as_stream(stream);
  int* bad = static_cast<int*>(workspace);
  FX_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  FX_HIP(hipMemsetAsync(ab, 0, static_cast<size_t>(n) * static_cast<size_t>(ldab) * 8, s));
  const int rows_per_block = kThreads / 64;
  hipLaunchKernelGGL(band_scatter_kernel, dim3((n + rows_per_block - 1) / rows_per_block), dim3(kThreads), 0, s, n,
                     indptr, indices, data, inv_perm, bw, ldab, ab, bad);
  FX_LAUNCHED("fx_band_from_csr_f64");
  int hbad = 0;
  FX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  if (hbad) return fail(FX_EINVAL, "fx_band_from_csr_f64: an entry lies past bandwidth %d", bw);
  return FX_OK;
}

int fx_band_cholesky_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, double* ab, int32_t* info,
                         void* workspace, size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || !info || (n > 0 && !ab))
    return fail(FX_EINVAL, "fx_band_cholesky_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n))
    return fail(FX_EWORKSPACE, "fx_band_cholesky_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n));
  *info = 0;
  if (n == 0) return FX_OK;
  if (ldab - 1 > INT32_MAX) return fail(FX_ELIMIT, "fx_band_cholesky_f64: ldab %lld too large", (long long)ldab);
  int err = 0;
  rocblas_handle h = blas_handle(err);
  if (err) return fail(FX_ELIB, "fx_band_cholesky_f64: rocblas_create_handle failed");
  hipStream_t s = as_stream(stream);
  rocblas_status st = rocblas_set_stream(h, s);
  if (st != rocblas_status_success) return blas_fail(st, "rocblas_set_stream");
  const int nblocks = (n + kFactorNb - 1) / kFactorNb;
  int* dinfo = reinterpret_cast<int*>(static_cast<char*>(workspace) + 256);
  const int ld = static_cast<int>(ldab - 1);
  auto A = [&](int64_t i, int64_t j) { return ab + i + j * static_cast<int64_t>(ld); };
  const double one = 1.0, minus_one = -1.0;
  for (int k = 0; k < nblocks; ++k) {
    const int c0 = k * kFactorNb;
    const int w = std::min(kFactorNb, n - c0);
    const int hh = std::min(bw, n - c0 - w);
    st = rocsolver_dpotrf(h, rocblas_fill_lower, w, A(c0, c0), ld, dinfo + k);
    if (st != rocblas_status_success) return blas_fail(st, "rocsolver_dpotrf");
    if (hh > 0) {
      st = rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                         rocblas_diagonal_non_unit, hh, w, &one, A(c0, c0), ld, A(c0 + w, c0), ld);
      if (st != rocblas_status_success) return blas_fail(st, "rocblas_dtrsm");
      st = rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, hh, w, &minus_one, A(c0 + w, c0), ld, &one,
                         A(c0 + w, c0 + w), ld);
      if (st != rocblas_status_success) return blas_fail(st, "rocblas_dsyrk");
    }
  }
  std::vector<int> hinfo(nblocks);
  FX_HIP(hipMemcpyAsync(hinfo.data(), dinfo, nblocks * sizeof(int), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < nblocks; ++k) {
    if (hinfo[k] != 0) {
      *info = k * kFactorNb + hinfo[k];
      break;
    }
  }
  return FX_OK;
}

int fx_band_solve_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, const double* ab, const int32_t* perm,
                      const double* b, double* x, void* workspace, size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || (n > 0 && (!ab || !b || !x)))
    return fail(FX_EINVAL, "fx_band_solve_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n))
    return fail(FX_EWORKSPACE, "fx_band_solve_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n));
  if (n == 0) return FX_OK;
  hipStream_t s = as_stream(stream);
  const size_t nblk = (static_cast<size_t>(n) + kFactorNb - 1) / kFactorNb;
  double* y = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256 +
                                        ((nblk * sizeof(int) + 255) & ~static_cast<size_t>(255)));
  const int gv = std::max(1, std::min((n + kThreads - 1) / kThreads, 2048));
  hipLaunchKernelGGL(permute_kernel, dim3(gv), dim3(kThreads), 0, s, n, perm, b, y, 0);
  // L y = P b
  const int nsb = (n + kSolveNb - 1) / kSolveNb;
  for (int k = -1; k < nsb - 1; ++k) {
    const int cK = k < 0 ? -1 : k * kSolveNb;
    const int wK = k < 0 ? 0 : kSolveNb;
    const int c1 = k < 0 ? 0 : cK + wK;
    const int w1 = std::min(kSolveNb, n - c1);
    const int rend = k < 0 ? 0 : std::min(n, cK + wK + bw);
    const int others = std::max(0, rend - (c1 + w1));
    hipLaunchKernelGGL(band_fwd_kernel, dim3(1 + (others + kThreads - 1) / kThreads), dim3(kThreads), 0, s, n, bw,
                       ldab, ab, y, cK, wK);
  }
  FX_LAUNCHED("fx_band_solve_f64 forward");
  // L^T x = y, from the last block
  for (int k = nsb; k >= 1; --k) {
    const int cK = k < nsb ? k * kSolveNb : n;
    const int wK = k < nsb ? std::min(kSolveNb, n - cK) : 0;
    const int c1 = (k - 1) * kSolveNb;
    const int w1 = std::min(kSolveNb, n - c1);
    const int jlo = std::max(0, cK - bw);
    const int others = wK > 0 ? std::max(0, c1 - jlo) : 0;
    const int wpb = kThreads / 64;
    hipLaunchKernelGGL(band_bwd_kernel, dim3(1 + (others + wpb - 1) / wpb), dim3(kThreads), 0, s, n, bw, ldab, ab, y,
                       cK, wK, c1, w1);
  }
  FX_LAUNCHED("fx_band_solve_f64 backward");
  hipLaunchKernelGGL(permute_kernel, dim3(gv), dim3(kThreads), 0, s, n, perm, y, x, 1);
  FX_LAUNCHED("fx_band_solve_f64 permute");
  return FX_OK;
}

}  // extern "C"
