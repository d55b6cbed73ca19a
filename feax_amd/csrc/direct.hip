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
//     L11 = potrf(A11)                    (chol_block_kernel, w x w in LDS)
//     L21 = A21 L11^-T                    (rocBLAS trsm, h x w, h <= bw)
//     A22 -= L21 L21^T  (lower triangle)  (rocBLAS syrk, h x h)
//
// The entries of L21 past the bandwidth are stored zeros and stay zero
// (Cholesky fills only inside the envelope).  The triangular solves are
// hand-written: per 256-row block, a wide launch forms the block's coupling
// to the block solved just before it (16 workgroups forward, one wave per
// 4 columns backward), then one launch whose workgroup 0 subtracts it and
// solves the diagonal triangle in 64-row sub-blocks (one wave from LDS,
// then all threads update the rows below / above), while its other
// workgroups apply the previous block to the rest of the bandwidth rows
// (forward) or columns (backward) as coalesced multiply-adds; only entries
// within the bandwidth are read.  Rows are in the caller's ordering `perm`
// (new -> old), chosen to reduce the bandwidth.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "common.h"
#include "feaxhip.h"

namespace fx {
namespace {

constexpr int kFactorNb = 128;  // block columns of the factorisation
constexpr int kSolveNb = 256;   // block of the triangular solves (one launch each)
constexpr int kSub = 64;        // diagonal sub-block solved by one wave
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

// Diagonal triangle [c, c + w) (w <= kSub) of L into LDS, all threads:
// Ld[i][j] = L(c + i, c + j) for i > j (0 past the bandwidth) and
// Ld[j][j] = 1 / L(c + j, c + j), so the sequential steps multiply.
// unit: L has a unit diagonal (LDL^T; the stored diagonal is D).
__device__ __forceinline__ void load_diag(const double* __restrict__ ab, int64_t ldab, int bw, int c, int w,
                                          bool unit, double (*Ld)[kSub + 1]) {
#pragma unroll
  for (int q = 0; q < kSub * kSub / kThreads; ++q) {
    const int e = threadIdx.x + q * kThreads;
    const int i = e % kSub, j = e / kSub;
    if (j < w && i >= j && i < w) {
      if (i == j) {
        Ld[i][j] = unit ? 1.0 : 1.0 / ab[boff(c + i, c + j, ldab)];
      } else {
        Ld[i][j] = (i - j <= bw) ? ab[boff(c + i, c + j, ldab)] : 0.0;
      }
    }
  }
}

// sum_{j in [j0, w)} L(r, c + j) v[j] for this thread's row r; for one j
// the threads' consecutive rows read consecutive addresses (coalesced).
// kBatch loads are issued before the first is used (latency, not
// bandwidth, bounds these short dots).
constexpr int kBatch = 32;
__device__ __forceinline__ double row_dot(const double* __restrict__ ab, int64_t ldab, int r, int c, int j0, int w,
                                          const double* __restrict__ v) {
  double acc = 0.0;
  for (int base = j0; base < w; base += kBatch) {
    double l[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) l[u] = base + u < w ? ab[boff(r, c + base + u, ldab)] : 0.0;
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (base + u < w) acc += l[u] * v[base + u];
  }
  return acc;
}

// Wave-wide sums s[q] = sum_{r < wr, c + r - j_q <= bw} L(c + r, j_q) v[r]
// for NJ columns j_q = j0 + q * stride (j_q < jend), lanes over r
// (coalesced); all loads issued before the reductions.  Valid in every lane.
template <int NJ>
__device__ __forceinline__ void col_dots(const double* __restrict__ ab, int64_t ldab, int bw, int c, int wr, int j0,
                                         int stride, int jend, const double* __restrict__ v, double (&s)[NJ]) {
  const int lane = threadIdx.x & 63;
  constexpr int kQ = kSolveNb / 64;
  double l[NJ][kQ];
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = j0 + q * stride;
#pragma unroll
    for (int h = 0; h < kQ; ++h) {
      const int r = h * 64 + lane;
      l[q][h] = (j < jend && r < wr && c + r - j <= bw) ? ab[boff(c + r, j, ldab)] : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    double a = 0.0;
#pragma unroll
    for (int h = 0; h < kQ; ++h) {
      const int r = h * 64 + lane;
      if (r < wr) a += l[q][h] * v[r];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    s[q] = a;
  }
}

// Coupling of the next block to the block just solved, spread over many
// workgroups so that it leaves the serial path of the single workgroup
// that solves the next block's triangle.  Forward: rows [c, c + w) of
// block k against the kSolveNb columns from cK (block k - 1, final y):
// P[q kSolveNb + r] = sum_{u < kChunk} L(c + r, cK + q kChunk + u) y, one
// workgroup per column chunk q, one thread per row (for one column the
// threads read consecutive addresses).  Entries past the bandwidth count 0.
constexpr int kChunk = 16;
constexpr int kChunks = kSolveNb / kChunk;

// Workspace of the triangular solves: y (n values, padded to 32) and the
// coupling partials (kChunks x kSolveNb), 256-byte aligned.
inline size_t solve_region_bytes(int n) {
  return ((((static_cast<size_t>(n) + 31) & ~static_cast<size_t>(31)) + kChunks * kSolveNb) * 8 + 255) &
         ~static_cast<size_t>(255);
}
__global__ __launch_bounds__(kThreads) void band_fwd_couple_kernel(int bw, int64_t ldab, const double* __restrict__ ab,
                                                                   const double* __restrict__ y, int cK, int c, int w,
                                                                   double* __restrict__ P) {
  __shared__ double yk[kChunk];
  const int q = blockIdx.x, t = threadIdx.x;
  const int j0 = cK + q * kChunk;
  if (t < kChunk) yk[t] = y[j0 + t];
  __syncthreads();
  double acc = 0.0;
  if (t < w) {
    const int r = c + t;
    double l[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) l[u] = (r - (j0 + u) <= bw) ? ab[boff(r, j0 + u, ldab)] : 0.0;
#pragma unroll
    for (int u = 0; u < kChunk; ++u) acc += l[u] * yk[u];
  }
  P[q * kSolveNb + t] = acc;
}

// Forward step (L y = b, in place) for block [c1, c1 + w1).  On entry
// y[cK, cK + wK) is final (wK = 0: the first block) and, when wK > 0,
// band_fwd_couple_kernel has left the block's coupling to it in P.
// Block 0: subtracts that coupling (chunks in a fixed order), then solves
// the triangle in kSub sub-blocks (wave 0 solves a sub-block from LDS, all
// threads update the rows below it).  Other blocks: rows [c1 + w1,
// min(n, cK + wK + bw)) take the coupling to block K.  Only entries within
// the bandwidth are read.
__global__ __launch_bounds__(kThreads) void band_fwd_kernel(int n, int bw, int64_t ldab,
                                                            const double* __restrict__ ab, double* __restrict__ y,
                                                            int cK, int wK, int c1, int unit,
                                                            const double* __restrict__ P) {
  __shared__ double yk[kSolveNb];
  __shared__ double yb[kSolveNb];
  __shared__ double Ld[kSub][kSub + 1];
  const int t = threadIdx.x;
  const int w1 = min(kSolveNb, n - c1);
  if (blockIdx.x == 0) {
    if (t < w1) {
      double v = y[c1 + t];
      if (wK > 0) {
        double cpl = 0.0;
#pragma unroll
        for (int q = 0; q < kChunks; ++q) cpl += P[q * kSolveNb + t];
        v -= cpl;
      }
      yb[t] = v;
    }
    for (int s0 = 0; s0 < w1; s0 += kSub) {
      const int sw = min(kSub, w1 - s0);
      load_diag(ab, ldab, bw, c1 + s0, sw, unit, Ld);
      __syncthreads();
      if (t < 64) {
        double vv = t < sw ? yb[s0 + t] : 0.0;
        for (int jj = 0; jj < sw; ++jj) {
          const double xj = __shfl(vv, jj, 64) * Ld[jj][jj];
          if (t == jj)
            vv = xj;
          else if (t > jj && t < sw)
            vv -= Ld[t][jj] * xj;
        }
        if (t < sw) yb[s0 + t] = vv;
      }
      __syncthreads();
      if (t >= s0 + sw && t < w1)
        yb[t] -= row_dot(ab, ldab, c1 + t, c1 + s0, max(0, t - bw - s0), sw, yb + s0);
      __syncthreads();
    }
    if (t < w1) y[c1 + t] = yb[t];
    return;
  }
  if (t < wK) yk[t] = y[cK + t];
  __syncthreads();
  const int r = c1 + w1 + (blockIdx.x - 1) * kThreads + t;
  const int rend = min(n, cK + wK + bw);
  if (r >= rend) return;
  y[r] -= row_dot(ab, ldab, r, cK, max(0, r - bw - cK), wK, yk);
}

// Backward coupling: C[j] = sum over the wK rows from cK (block k + 1,
// final x) of L(r, c1 + j) x_r for the w1 columns of block k.  One wave per
// kJc columns, lanes over the rows (coalesced), wave reductions.
constexpr int kJc = 4;
__global__ __launch_bounds__(kThreads) void band_bwd_couple_kernel(int bw, int64_t ldab, const double* __restrict__ ab,
                                                                   const double* __restrict__ y, int cK, int wK,
                                                                   int c1, int w1, double* __restrict__ C) {
  __shared__ double xk[kSolveNb];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t < wK) xk[t] = y[cK + t];
  __syncthreads();
  const int jj = (blockIdx.x * (kThreads / 64) + wid) * kJc;
  if (jj >= w1) return;  // uniform over the wave
  double sv[kJc];
  col_dots<kJc>(ab, ldab, bw, cK, wK, c1 + jj, 1, c1 + w1, xk, sv);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kJc; ++q)
      if (jj + q < w1) C[jj + q] = sv[q];
  }
}

// Backward step (L^T x = y, in place).  On entry y[cK, cK + wK) holds final
// x (the last block's launch has wK = 0).  Other blocks: rows j of the
// previous bandwidth take sum_{r in K} L(r, j) x_r, one wave per j.  Block
// 0 does the same for its block [c1, c1 + w1), then solves the transposed
// triangle in kSub sub-blocks from the bottom.
__global__ __launch_bounds__(kThreads) void band_bwd_kernel(int n, int bw, int64_t ldab,
                                                            const double* __restrict__ ab, double* __restrict__ y,
                                                            int cK, int wK, int c1, int w1, int unit,
                                                            const double* __restrict__ C) {
  __shared__ double xk[kSolveNb];
  __shared__ double xb[kSolveNb];
  __shared__ double Ld[kSub][kSub + 1];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  constexpr int kWaves = kThreads / 64;
  if (blockIdx.x == 0) {
    constexpr int kJ = 8;
    if (t < w1) xb[t] = wK > 0 ? y[c1 + t] - C[t] : y[c1 + t];
    __syncthreads();
    for (int s0 = ((w1 - 1) / kSub) * kSub; s0 >= 0; s0 -= kSub) {
      const int sw = min(kSub, w1 - s0);
      load_diag(ab, ldab, bw, c1 + s0, sw, unit, Ld);
      __syncthreads();
      if (t < 64) {
        double vv = t < sw ? xb[s0 + t] : 0.0;
        for (int jj = sw - 1; jj >= 0; --jj) {
          const double xj = __shfl(vv, jj, 64) * Ld[jj][jj];
          if (t == jj)
            vv = xj;
          else if (t < jj)
            vv -= Ld[jj][t] * xj;
        }
        if (t < sw) xb[s0 + t] = vv;
      }
      __syncthreads();
      for (int jj = wid; jj < s0; jj += kWaves * kJ) {
        double s[kJ];
        col_dots<kJ>(ab, ldab, bw, c1 + s0, sw, c1 + jj, kWaves, c1 + s0, xb + s0, s);
        if (lane == 0) {
#pragma unroll
          for (int q = 0; q < kJ; ++q)
            if (jj + q * kWaves < s0) xb[jj + q * kWaves] -= s[q];
        }
      }
      __syncthreads();
    }
    if (t < w1) y[c1 + t] = xb[t];
    return;
  }
  if (wK == 0) return;
  if (t < wK) xk[t] = y[cK + t];
  __syncthreads();
  const int jlo = max(0, cK - bw);
  const int j = c1 - 1 - ((blockIdx.x - 1) * kWaves + wid);
  if (j < jlo) return;
  double s[1];
  col_dots<1>(ab, ldab, bw, cK, wK, j, 0, j + 1, xk, s);
  if (lane == 0) y[j] -= s[0];
}

// y[i] /= D[i] (the stored diagonal of an LDL^T factor).
__global__ __launch_bounds__(kThreads) void diag_scale_kernel(int n, int64_t ldab, const double* __restrict__ ab,
                                                              double* __restrict__ y) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) y[i] /= ab[boff(i, i, ldab)];
}

// Mirror the lower triangle of the w x w diagonal block into the upper
// triangle of its dense view (leading dimension ldab - 1), which LU reads:
// view entry (i, j), i < j, is the unused tail of band column c0 + j - 1
// at distance ldab - (j - i) > bw + kFactorNb, never read elsewhere.
__global__ __launch_bounds__(kThreads) void mirror_block_kernel(int64_t ldab, double* __restrict__ ab, int c0,
                                                                int w) {
  const int64_t ld = ldab - 1;
  for (int e = blockIdx.x * kThreads + threadIdx.x; e < w * w; e += gridDim.x * kThreads) {
    const int i = e % w, j = e / w;
    if (i < j) ab[(c0 + i) + (c0 + j) * ld] = ab[(c0 + j) + (c0 + i) * ld];
  }
}

// LDL^T panel of block column [c0, c0 + w): the band holds W = A21 L11^-T
// (h x w); copy W to wt (column-major, leading dimension h) and overwrite
// the band with L21 = W D1^-1.
__global__ __launch_bounds__(kThreads) void ldlt_panel_kernel(int64_t ldab, double* __restrict__ ab, int c0, int w,
                                                              int h, double* __restrict__ wt) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (r >= h || j >= w) return;
  const size_t o = boff(c0 + w + r, c0 + j, ldab);
  const double v = ab[o];
  wt[static_cast<size_t>(j) * h + r] = v;
  ab[o] = v / ab[boff(c0 + j, c0 + j, ldab)];
}

// Cholesky of the w x w (w <= kFactorNb) diagonal block at (c0, c0) of the
// band's dense view (leading dimension ld), in place, lower triangle: one
// workgroup, the block in LDS.  Right-looking in panels of kPanel columns:
// wave 0 factors a panel from registers (lane l holds rows p0 + l and
// p0 + 64 + l of the panel's columns; pivots and the column entries of the
// panel rows travel by lane shuffles, no barrier per column), then all
// waves apply the panel to the trailing triangle (each entry one
// kPanel-term dot from LDS; 4 x 4 register tiles measured slower).  The
// block is loaded with all of a thread's loads in flight at once.  A pivot
// that is not positive (or NaN) stops
// the factorisation and sets *info to its column + 1, as LAPACK's potrf
// does; else *info = 0.
constexpr int kPanel = 16;
constexpr int kCholThreads = 1024;
static_assert(kFactorNb <= 128, "two panel rows per lane");
__global__ __launch_bounds__(kCholThreads) void chol_block_kernel(double* __restrict__ ab, int64_t ld, int c0, int w,
                                                                  int* __restrict__ info) {
  __shared__ double A[kFactorNb][kFactorNb + 1];
  __shared__ int s_failed;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  double* blk = ab + c0 + static_cast<int64_t>(c0) * ld;
  {
    constexpr int kLd = kFactorNb * kFactorNb / kCholThreads;
    double v[kLd];
#pragma unroll
    for (int q = 0; q < kLd; ++q) {
      const int e = t + q * kCholThreads;
      const int i = e % kFactorNb, j = e / kFactorNb;
      v[q] = (i >= j && i < w) ? blk[i + static_cast<int64_t>(j) * ld] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kLd; ++q) {
      const int e = t + q * kCholThreads;
      A[e % kFactorNb][e / kFactorNb] = v[q];
    }
  }
  if (t == 0) s_failed = 0;
  __syncthreads();
  for (int p0 = 0; p0 < w; p0 += kPanel) {
    const int pw = min(kPanel, w - p0), pe = p0 + pw;
    if (wid == 0) {
      double a[2][kPanel];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int i = p0 + lane + 64 * r;
#pragma unroll
        for (int c = 0; c < kPanel; ++c) a[r][c] = (i < w && c < pw && i >= p0 + c) ? A[i][p0 + c] : 0.0;
      }
      int failed = 0;
#pragma unroll
      for (int c = 0; c < kPanel; ++c) {
        if (c < pw && !failed) {  // uniform
          const double d = __shfl(a[0][c], c, 64);
          if (!(d > 0.0)) {
            failed = p0 + c + 1;
          } else {
            const double sq = sqrt(d), inv = 1.0 / sq;
#pragma unroll
            for (int r = 0; r < 2; ++r) {
              const int i = p0 + lane + 64 * r;
              a[r][c] = i == p0 + c ? sq : a[r][c] * inv;
            }
#pragma unroll
            for (int k = c + 1; k < kPanel; ++k) {
              const double lkc = __shfl(a[0][c], k, 64);  // L(p0 + k, p0 + c)
#pragma unroll
              for (int r = 0; r < 2; ++r) a[r][k] -= a[r][c] * lkc;
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int i = p0 + lane + 64 * r;
#pragma unroll
        for (int c = 0; c < kPanel; ++c)
          if (i < w && c < pw && i >= p0 + c) A[i][p0 + c] = a[r][c];
      }
      if (lane == 0 && failed) s_failed = failed;
    }
    __syncthreads();
    if (s_failed) break;  // uniform
    const int m = w - pe;  // trailing triangle rows / columns [pe, w)
    for (int e = t; e < m * m; e += kCholThreads) {
      const int i = pe + e % m, k = pe + e / m;
      if (i >= k) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < kPanel; ++j)
          if (j < pw) acc += A[i][p0 + j] * A[k][p0 + j];
        A[i][k] -= acc;
      }
    }
    __syncthreads();
  }
  for (int e = t; e < kFactorNb * kFactorNb; e += kCholThreads) {
    const int i = e % kFactorNb, j = e / kFactorNb;
    if (i >= j && i < w) blk[i + static_cast<int64_t>(j) * ld] = A[i][j];
  }
  if (t == 0) *info = s_failed;
}

// ------------------------------------------------------------------
// Triangular solves with the inverses of the diagonal blocks.
//
// The solve of block k (NBS rows) is two launches per direction instead of
// a chain of serial sub-block solves: (1) a wide launch forms the block's
// coupling to the rest of the band as partial sums (forward: the rows of
// block k against the <= bw solved columns before it, left-looking;
// backward: the columns of block k against the <= bw solved rows after it),
// 64 columns (rows) per workgroup; (2) a launch of NBS / 64 workgroups
// subtracts the partials (fixed order) from the right-hand side and
// multiplies by the precomputed inverse of the diagonal block (a dense
// lower-triangular GEMV), so no step waits on a serial triangle.  The
// inverses (rocBLAS trtri, strided batched) are formed once per
// factorisation.  Only entries within the bandwidth are read.
constexpr int kPart = 64;        // columns (forward) / rows (backward) per partial workgroup
constexpr int kDiagThreads = 1024;

// Largest solve block below ldab (the diagonal block must be a plain
// column-major matrix inside the band storage), at most 512.
inline int inv_block(int bw) {
  const int64_t ldab = static_cast<int64_t>(bw) + 2 * kFactorNb;
  int nbs = 512;
  while (nbs > 64 && nbs >= ldab) nbs /= 2;
  return nbs;
}
inline int part_chunks(int bw) { return (bw + kPart - 1) / kPart + 1; }

// Forward partials: P[q NBS + r] = sum_{j in chunk q} L(c1 + r, j) y_j over
// the columns [jlo + 128 q, min(c1, jlo + 128 q + 128)), jlo = max(0,
// c1 - bw).  Workgroup (q, row tile): 64 rows, its 4 waves take 32 columns
// each (lanes over the rows: consecutive rows read consecutive addresses,
// 32 loads in flight per lane), then a fixed-order sum over the waves.  A
// block's coupling is only ~21 MB at 1M DOF, so the work is spread thin
// over many short-lived waves.
constexpr int kFwdCols = 128;
template <int NBS>
__global__ __launch_bounds__(256) void inv_fwd_partial_kernel(int bw, int64_t ldab, const double* __restrict__ ab,
                                                              const double* __restrict__ y, int c1, int w1, int jlo,
                                                              double* __restrict__ P) {
  __shared__ double ys[kFwdCols];
  __shared__ double red[4][64];
  const int q = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int j0 = jlo + q * kFwdCols;
  const int jn = min(kFwdCols, c1 - j0);
  if (t < kFwdCols) ys[t] = t < jn ? y[j0 + t] : 0.0;
  __syncthreads();
  const int rr = blockIdx.y * 64 + lane;
  double acc = 0.0;
  if (rr < w1) {
    const int r = c1 + rr;
    double l[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) {
      const int jj = wv * 32 + u, j = j0 + jj;
      l[u] = (jj < jn && r - j <= bw) ? ab[boff(r, j, ldab)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 32; ++u) acc += l[u] * ys[wv * 32 + u];
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && blockIdx.y * 64 + lane < NBS)
    P[q * NBS + blockIdx.y * 64 + lane] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Backward partials: P[q NBS + j] = sum_{i in chunk q} L(i, c1 + j) x_i over
// the rows [i0, i0 + 64), i0 = c1 + w1 + 64 q, below the block (i < ihi).
// Workgroup (q, tile): a 64 x 64 tile of L staged in LDS (coalesced: rows
// consecutive for one column), then column sums.
template <int NBS>
__global__ __launch_bounds__(256) void inv_bwd_partial_kernel(int bw, int64_t ldab, const double* __restrict__ ab,
                                                              const double* __restrict__ x, int c1, int w1, int ihi,
                                                              double* __restrict__ P) {
  __shared__ double T[kPart][kPart + 1];
  __shared__ double xs[kPart];
  __shared__ double red[4][kPart];
  const int q = blockIdx.x, tile = blockIdx.y, t = threadIdx.x;
  const int i0 = c1 + w1 + q * kPart;
  const int cj = c1 + tile * kPart;
  if (t < kPart) xs[t] = (i0 + t < ihi) ? x[i0 + t] : 0.0;
#pragma unroll
  for (int e0 = 0; e0 < kPart * kPart; e0 += 256) {
    const int e = e0 + t;
    const int row = e % kPart, col = e / kPart;
    const int i = i0 + row, j = cj + col;
    T[row][col] = (i < ihi && j < c1 + w1 && i - j <= bw) ? ab[boff(i, j, ldab)] : 0.0;
  }
  __syncthreads();
  const int col = t % kPart, part = t / kPart;
  double s = 0.0;
#pragma unroll
  for (int rr = 0; rr < kPart / 4; ++rr) {
    const int row = part * (kPart / 4) + rr;
    s += T[row][col] * xs[row];
  }
  red[part][col] = s;
  __syncthreads();
  if (t < kPart && tile * kPart + t < NBS) P[q * NBS + tile * kPart + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// v_r = rhs[c1 + r] - sum_q P[q NBS + r] for r < w1, into LDS (every
// workgroup of the diagonal launch forms all of v).  Fixed order: each of
// the kDiagThreads / NBS thread groups sums the chunks q = h, h + H, ...
// (32 loads in flight, added in q order); the groups' sums are added in
// group order.
template <int NBS>
__device__ __forceinline__ void inv_rhs(const double* __restrict__ rhs, int c1, int w1, int nq,
                                        const double* __restrict__ P, double* v, double (*grp)[NBS]) {
  constexpr int H = kDiagThreads / NBS;
  constexpr int kB = 32;  // partials loaded before any is added
  const int t = threadIdx.x;
  const int r = t % NBS, h = t / NBS;
  double a = 0.0;
  if (r < w1) {
    for (int q0 = h; q0 < nq; q0 += kB * H) {
      double pv[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int q = q0 + u * H;
        pv[u] = q < nq ? P[q * NBS + r] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) a += pv[u];
    }
  }
  grp[h][r] = a;
  __syncthreads();
  if (h == 0) {
    // v past the block (the last block of the matrix) is 0, so that a zero
    // of the inverse never multiplies stale LDS
    double sum = 0.0;
#pragma unroll
    for (int g = 0; g < H; ++g) sum += grp[g][r];
    v[r] = r < w1 ? rhs[c1 + r] - sum : 0.0;
  }
  __syncthreads();
}

// Forward diagonal step: y[c1 + r] = sum_{j <= r} Linv(r, j) v_j for the 64
// rows of this workgroup (Linv column-major, leading dimension NBS); 16
// waves take NBS / 16 columns each, lanes the rows (coalesced), then a
// fixed-order sum over the waves.
template <int NBS>
__global__ __launch_bounds__(kDiagThreads) void inv_fwd_diag_kernel(const double* __restrict__ rhs,
                                                                    double* __restrict__ y, int c1, int w1, int nq,
                                                                    const double* __restrict__ P,
                                                                    const double* __restrict__ Linv) {
  constexpr int H = kDiagThreads / NBS;
  constexpr int kW = kDiagThreads / 64, kCols = NBS / kW;
  __shared__ double v[NBS];
  __shared__ double grp[H][NBS];
  __shared__ double red[kW][64];
  inv_rhs<NBS>(rhs, c1, w1, nq, P, v, grp);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (r < w1) {
    double l[kCols];
#pragma unroll
    for (int u = 0; u < kCols; ++u) {
      const int j = wv * kCols + u;
      l[u] = (j <= r) ? Linv[r + static_cast<size_t>(NBS) * j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kCols; ++u) acc += l[u] * v[wv * kCols + u];
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && r < w1) {
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < kW; ++w) sum += red[w][lane];
    y[c1 + r] = sum;
  }
}

// Backward diagonal step: x[c1 + r] = sum_{j >= r} Linv(j, r) v_j (column r
// of Linv, contiguous): 16 waves, each takes 4 of the workgroup's 64 rows,
// lanes over j, wave sums.
template <int NBS>
__global__ __launch_bounds__(kDiagThreads) void inv_bwd_diag_kernel(const double* __restrict__ rhs,
                                                                    double* __restrict__ x, int c1, int w1, int nq,
                                                                    const double* __restrict__ P,
                                                                    const double* __restrict__ Linv) {
  constexpr int H = kDiagThreads / NBS;
  constexpr int kW = kDiagThreads / 64, kPer = NBS / 64;
  __shared__ double v[NBS];
  __shared__ double grp[H][NBS];
  inv_rhs<NBS>(rhs, c1, w1, nq, P, v, grp);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double l[4][kPer];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int r = blockIdx.x * 64 + wv * 4 + k;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int j = lane + 64 * u;
      l[k][u] = (r < w1 && j >= r && j < w1) ? Linv[j + static_cast<size_t>(NBS) * r] : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int r = blockIdx.x * 64 + wv * 4 + k;
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int j = lane + 64 * u;
      a += l[k][u] * (j < w1 ? v[j] : 0.0);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0 && r < w1) x[c1 + r] = a;
  }
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

size_t fx_band_workspace_bytes(int32_t n, int32_t bw) {
  if (n < 0 || bw < 0) return 0;
  const size_t nb = (static_cast<size_t>(n) + kFactorNb - 1) / kFactorNb;
  // info flags, the solves' region, the LDL^T panel copy (bw x kFactorNb)
  return 256 + ((nb * sizeof(int) + 255) & ~static_cast<size_t>(255)) + solve_region_bytes(n) +
         static_cast<size_t>(bw) * kFactorNb * 8;
}

int fx_band_from_csr_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                         const int32_t* inv_perm, int32_t bw, int64_t ldab, double* ab, void* workspace,
                         size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || (n > 0 && (!indptr || !indices || !data || !ab)))
    return fail(FX_EINVAL, "fx_band_from_csr_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n, bw))
    return fail(FX_EWORKSPACE, "fx_band_from_csr_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n, bw));
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

}  // extern "C"

namespace fx {
namespace {

// Blocked right-looking band factorisation, Cholesky (L L^T) or, ldlt,
// L D L^T without pivoting (the diagonal block, mirrored to full, by
// rocSOLVER's LU without pivoting: U11 = D1 L11^T; W = A21 L11^-T with unit
// L11; L21 = W D1^-1; A22 -= L21 W^T by syrkx).
int band_factor(const char* name, bool ldlt, hipStream_t s, int n, int bw, int64_t ldab, double* ab, int* info,
                char* wsp) {
  *info = 0;
  if (n == 0) return FX_OK;
  if (ldab - 1 > INT32_MAX) return fail(FX_ELIMIT, "%s: ldab %lld too large", name, (long long)ldab);
  int err = 0;
  rocblas_handle h = blas_handle(err);
  if (err) return fail(FX_ELIB, "%s: rocblas_create_handle failed", name);
  rocblas_status st = rocblas_set_stream(h, s);
  if (st != rocblas_status_success) return blas_fail(st, "rocblas_set_stream");
  const int nblocks = (n + kFactorNb - 1) / kFactorNb;
  int* dinfo = reinterpret_cast<int*>(wsp + 256);
  double* wt = reinterpret_cast<double*>(wsp + 256 + ((nblocks * sizeof(int) + 255) & ~static_cast<size_t>(255)) +
                                         solve_region_bytes(n));
  const int ld = static_cast<int>(ldab - 1);
  auto A = [&](int64_t i, int64_t j) { return ab + i + j * static_cast<int64_t>(ld); };
  const double one = 1.0, minus_one = -1.0;
  for (int k = 0; k < nblocks; ++k) {
    const int c0 = k * kFactorNb;
    const int w = std::min(kFactorNb, n - c0);
    const int hh = std::min(bw, n - c0 - w);
    if (ldlt) {
      hipLaunchKernelGGL(mirror_block_kernel, dim3((w * w + kThreads - 1) / kThreads), dim3(kThreads), 0, s, ldab,
                         ab, c0, w);
      FX_LAUNCHED("mirror_block_kernel");
      st = rocsolver_dgetrf_npvt(h, w, w, A(c0, c0), ld, dinfo + k);
      if (st != rocblas_status_success) return blas_fail(st, "rocsolver_dgetrf_npvt");
    } else {
      hipLaunchKernelGGL(chol_block_kernel, dim3(1), dim3(kCholThreads), 0, s, ab, static_cast<int64_t>(ld), c0, w,
                         dinfo + k);
      FX_LAUNCHED("chol_block_kernel");
    }
    if (hh <= 0) continue;
    st = rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                       ldlt ? rocblas_diagonal_unit : rocblas_diagonal_non_unit, hh, w, &one, A(c0, c0), ld,
                       A(c0 + w, c0), ld);
    if (st != rocblas_status_success) return blas_fail(st, "rocblas_dtrsm");
    if (ldlt) {
      hipLaunchKernelGGL(ldlt_panel_kernel, dim3((hh + kThreads - 1) / kThreads, w), dim3(kThreads), 0, s, ldab, ab,
                         c0, w, hh, wt);
      FX_LAUNCHED("ldlt_panel_kernel");
      st = rocblas_dsyrkx(h, rocblas_fill_lower, rocblas_operation_none, hh, w, &minus_one, A(c0 + w, c0), ld, wt,
                          hh, &one, A(c0 + w, c0 + w), ld);
      if (st != rocblas_status_success) return blas_fail(st, "rocblas_dsyrkx");
    } else {
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

}  // namespace
}  // namespace fx

extern "C" {

int fx_band_cholesky_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, double* ab, int32_t* info,
                         void* workspace, size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || !info || (n > 0 && !ab))
    return fail(FX_EINVAL, "fx_band_cholesky_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n, bw))
    return fail(FX_EWORKSPACE, "fx_band_cholesky_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n, bw));
  return band_factor("fx_band_cholesky_f64", false, as_stream(stream), n, bw, ldab, ab, info,
                     static_cast<char*>(workspace));
}

int fx_band_ldlt_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, double* ab, int32_t* info, void* workspace,
                     size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || !info || (n > 0 && !ab))
    return fail(FX_EINVAL, "fx_band_ldlt_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n, bw))
    return fail(FX_EWORKSPACE, "fx_band_ldlt_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n, bw));
  return band_factor("fx_band_ldlt_f64", true, as_stream(stream), n, bw, ldab, ab, info,
                     static_cast<char*>(workspace));
}

int32_t fx_band_inv_block(int32_t bw) { return bw < 0 ? -1 : inv_block(bw); }

size_t fx_band_inv_bytes(int32_t n, int32_t bw) {
  if (n < 0 || bw < 0) return 0;
  const size_t nbs = static_cast<size_t>(inv_block(bw));
  return ((static_cast<size_t>(n) + nbs - 1) / nbs) * nbs * nbs * sizeof(double);
}

size_t fx_band_solve_inv_workspace_bytes(int32_t n, int32_t bw) {
  if (n < 0 || bw < 0) return 0;
  const size_t vec = ((static_cast<size_t>(n) * 8 + 255) & ~static_cast<size_t>(255));
  const size_t parts = static_cast<size_t>(part_chunks(bw)) * inv_block(bw) * 8;
  return 2 * vec + ((parts + 255) & ~static_cast<size_t>(255));
}

int fx_band_invert_diag_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, int32_t flags, const double* ab,
                            double* inv) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || (flags & ~FX_BAND_LDLT) || (n > 0 && (!ab || !inv)))
    return fail(FX_EINVAL, "fx_band_invert_diag_f64: bad arguments");
  if (n == 0) return FX_OK;
  if (ldab - 1 > INT32_MAX) return fail(FX_ELIMIT, "fx_band_invert_diag_f64: ldab %lld too large", (long long)ldab);
  int err = 0;
  rocblas_handle h = blas_handle(err);
  if (err) return fail(FX_ELIB, "fx_band_invert_diag_f64: rocblas_create_handle failed");
  hipStream_t s = as_stream(stream);
  rocblas_status st = rocblas_set_stream(h, s);
  if (st != rocblas_status_success) return blas_fail(st, "rocblas_set_stream");
  const int nbs = inv_block(bw);
  const int nfull = n / nbs, rest = n - nfull * nbs;
  const rocblas_diagonal dg = (flags & FX_BAND_LDLT) ? rocblas_diagonal_unit : rocblas_diagonal_non_unit;
  const int ld = static_cast<int>(ldab - 1);
  if (nfull > 0) {
    st = rocblas_dtrtri_strided_batched(h, rocblas_fill_lower, dg, nbs, ab, ld, static_cast<rocblas_stride>(nbs) * ldab,
                                        inv, nbs, static_cast<rocblas_stride>(nbs) * nbs, nfull);
    if (st != rocblas_status_success) return blas_fail(st, "rocblas_dtrtri_strided_batched");
  }
  if (rest > 0) {
    st = rocblas_dtrtri(h, rocblas_fill_lower, dg, rest, ab + static_cast<size_t>(nfull) * nbs * ldab, ld,
                        inv + static_cast<size_t>(nfull) * nbs * nbs, nbs);
    if (st != rocblas_status_success) return blas_fail(st, "rocblas_dtrtri");
  }
  return FX_OK;
}

int fx_band_solve_inv_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, int32_t flags, const double* ab,
                          const double* inv, const int32_t* perm, const double* b, double* x, void* workspace,
                          size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || (flags & ~FX_BAND_LDLT) || (n > 0 && (!ab || !inv || !b || !x)))
    return fail(FX_EINVAL, "fx_band_solve_inv_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_solve_inv_workspace_bytes(n, bw))
    return fail(FX_EWORKSPACE, "fx_band_solve_inv_f64: workspace %zu < %zu", ws_bytes,
                fx_band_solve_inv_workspace_bytes(n, bw));
  if (n == 0) return FX_OK;
  hipStream_t s = as_stream(stream);
  const size_t vec = ((static_cast<size_t>(n) * 8 + 255) & ~static_cast<size_t>(255));
  double* yb = static_cast<double*>(workspace);                                  // P b, then x (permuted)
  double* y = reinterpret_cast<double*>(static_cast<char*>(workspace) + vec);   // L^-1 P b
  double* P = reinterpret_cast<double*>(static_cast<char*>(workspace) + 2 * vec);
  const int gv = std::max(1, std::min((n + kThreads - 1) / kThreads, 2048));
  const int unit = (flags & FX_BAND_LDLT) ? 1 : 0;
  const int nbs = inv_block(bw);
  const int nsb = (n + nbs - 1) / nbs;
  hipLaunchKernelGGL(permute_kernel, dim3(gv), dim3(kThreads), 0, s, n, perm, b, yb, 0);
  // The 2 x 2 x nsb launches of the two directions touch only ab, inv and
  // the workspace: they are captured once as a HIP graph per (device, ab,
  // inv, workspace, n, bw, ldab, flags) and replayed (kept until the process
  // ends; the graph reads the current contents of those buffers, so a new
  // factorisation in the same storage reuses it).
  static std::mutex gmu;
  static std::map<std::tuple<int, const void*, const void*, const void*, int, int, int64_t, int>, hipGraphExec_t>
      graphs;
  int dev = 0;
  FX_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, static_cast<const void*>(ab), static_cast<const void*>(inv),
                                   static_cast<const void*>(workspace), static_cast<int>(n), static_cast<int>(bw),
                                   static_cast<int64_t>(ldab), static_cast<int>(flags));
  auto launch = [&](hipStream_t s, auto nbc) -> int {
    constexpr int NBS = decltype(nbc)::value;
    // L y = P b, block by block from the first
    for (int k = 0; k < nsb; ++k) {
      const int c1 = k * NBS, w1 = std::min(NBS, n - c1);
      const int jlo = std::max(0, c1 - bw);
      const int nq = (c1 - jlo + kFwdCols - 1) / kFwdCols;
      if (nq > 0)
        hipLaunchKernelGGL((inv_fwd_partial_kernel<NBS>), dim3(nq, (w1 + 63) / 64), dim3(256), 0, s, bw, ldab, ab, y,
                           c1, w1, jlo, P);
      hipLaunchKernelGGL((inv_fwd_diag_kernel<NBS>), dim3((w1 + 63) / 64), dim3(kDiagThreads), 0, s, yb, y, c1, w1, nq,
                         P, inv + static_cast<size_t>(k) * NBS * NBS);
    }
    FX_LAUNCHED("fx_band_solve_inv_f64 forward");
    if (unit) hipLaunchKernelGGL(diag_scale_kernel, dim3(gv), dim3(kThreads), 0, s, n, ldab, ab, y);
    // L^T x = y, block by block from the last (x into yb)
    for (int k = nsb - 1; k >= 0; --k) {
      const int c1 = k * NBS, w1 = std::min(NBS, n - c1);
      const int ihi = std::min(n, c1 + w1 + bw);
      const int nq = (ihi - (c1 + w1) + kPart - 1) / kPart;
      if (nq > 0)
        hipLaunchKernelGGL((inv_bwd_partial_kernel<NBS>), dim3(nq, (w1 + kPart - 1) / kPart), dim3(256), 0, s, bw,
                           ldab, ab, yb, c1, w1, ihi, P);
      hipLaunchKernelGGL((inv_bwd_diag_kernel<NBS>), dim3((w1 + 63) / 64), dim3(kDiagThreads), 0, s, y, yb, c1, w1,
                         std::max(nq, 0), P, inv + static_cast<size_t>(k) * NBS * NBS);
    }
    FX_LAUNCHED("fx_band_solve_inv_f64 backward");
    return FX_OK;
  };
  auto launch_all = [&](hipStream_t cs) -> int {
    switch (nbs) {
      case 512: return launch(cs, std::integral_constant<int, 512>{});
      case 256: return launch(cs, std::integral_constant<int, 256>{});
      case 128: return launch(cs, std::integral_constant<int, 128>{});
      default: return launch(cs, std::integral_constant<int, 64>{});
    }
  };
  hipGraphExec_t exec = nullptr;
  {
    std::lock_guard<std::mutex> lock(gmu);
    auto it = graphs.find(key);
    if (it != graphs.end()) exec = it->second;
  }
  if (!exec) {
    GraphGuard gg;
    hipError_t e = hipStreamCreateWithFlags(&gg.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(gg.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "fx_band_solve_inv_f64 capture");
    const int rc = launch_all(gg.stream);
    e = hipStreamEndCapture(gg.stream, &gg.graph);
    if (rc) return rc;
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "fx_band_solve_inv_f64 graph");
    exec = gg.exec;
    gg.exec = nullptr;  // kept in the cache
    std::lock_guard<std::mutex> lock(gmu);
    graphs[key] = exec;
  }
  FX_HIP(hipGraphLaunch(exec, s));
  hipLaunchKernelGGL(permute_kernel, dim3(gv), dim3(kThreads), 0, s, n, perm, yb, x, 1);
  FX_LAUNCHED("fx_band_solve_inv_f64 permute");
  return FX_OK;
}

int fx_band_solve_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, int32_t flags, const double* ab,
                      const int32_t* perm, const double* b, double* x, void* workspace, size_t ws_bytes) {
  if (n < 0 || bw < 0 || ldab < fx_band_ldab(bw) || (flags & ~FX_BAND_LDLT) || (n > 0 && (!ab || !b || !x)))
    return fail(FX_EINVAL, "fx_band_solve_f64: bad arguments");
  if (!workspace || ws_bytes < fx_band_workspace_bytes(n, bw))
    return fail(FX_EWORKSPACE, "fx_band_solve_f64: workspace %zu < %zu", ws_bytes, fx_band_workspace_bytes(n, bw));
  if (n == 0) return FX_OK;
  hipStream_t s = as_stream(stream);
  const size_t nblk = (static_cast<size_t>(n) + kFactorNb - 1) / kFactorNb;
  double* y = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256 +
                                        ((nblk * sizeof(int) + 255) & ~static_cast<size_t>(255)));
  const int gv = std::max(1, std::min((n + kThreads - 1) / kThreads, 2048));
  const int unit = (flags & FX_BAND_LDLT) ? 1 : 0;
  hipLaunchKernelGGL(permute_kernel, dim3(gv), dim3(kThreads), 0, s, n, perm, b, y, 0);
  // L y = P b
  double* cp = y + ((static_cast<size_t>(n) + 31) & ~static_cast<size_t>(31));  // coupling partials
  const int nsb = (n + kSolveNb - 1) / kSolveNb;
  for (int k = 0; k < nsb; ++k) {
    const int c1 = k * kSolveNb;
    const int w1 = std::min(kSolveNb, n - c1);
    const int cK = k > 0 ? c1 - kSolveNb : 0;
    const int wK = k > 0 ? kSolveNb : 0;
    if (k > 0)
      hipLaunchKernelGGL(band_fwd_couple_kernel, dim3(kChunks), dim3(kThreads), 0, s, bw, ldab, ab, y, cK, c1, w1, cp);
    const int rend = k > 0 ? std::min(n, cK + wK + bw) : 0;
    const int others = std::max(0, rend - (c1 + w1));
    hipLaunchKernelGGL(band_fwd_kernel, dim3(1 + (others + kThreads - 1) / kThreads), dim3(kThreads), 0, s, n, bw,
                       ldab, ab, y, cK, wK, c1, unit, cp);
  }
  FX_LAUNCHED("fx_band_solve_f64 forward");
  if (unit) hipLaunchKernelGGL(diag_scale_kernel, dim3(gv), dim3(kThreads), 0, s, n, ldab, ab, y);
  // L^T x = y, from the last block
  for (int k = nsb; k >= 1; --k) {
    const int cK = k < nsb ? k * kSolveNb : n;
    const int wK = k < nsb ? std::min(kSolveNb, n - cK) : 0;
    const int c1 = (k - 1) * kSolveNb;
    const int w1 = std::min(kSolveNb, n - c1);
    if (wK > 0) {
      const int waves = (w1 + kJc - 1) / kJc;
      hipLaunchKernelGGL(band_bwd_couple_kernel, dim3((waves + kThreads / 64 - 1) / (kThreads / 64)), dim3(kThreads),
                         0, s, bw, ldab, ab, y, cK, wK, c1, w1, cp);
    }
    const int jlo = std::max(0, cK - bw);
    const int others = wK > 0 ? std::max(0, c1 - jlo) : 0;
    const int wpb = kThreads / 64;
    hipLaunchKernelGGL(band_bwd_kernel, dim3(1 + (others + wpb - 1) / wpb), dim3(kThreads), 0, s, n, bw, ldab, ab, y,
                       cK, wK, c1, w1, unit, cp);
  }
  FX_LAUNCHED("fx_band_solve_f64 backward");
  hipLaunchKernelGGL(permute_kernel, dim3(gv), dim3(kThreads), 0, s, n, perm, y, x, 1);
  FX_LAUNCHED("fx_band_solve_f64 permute");
  return FX_OK;
}

}  // extern "C"
