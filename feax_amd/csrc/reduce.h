// Deterministic FP64 reductions.
//
// A reduction is split across a kernel boundary: every block of the
// producing kernel writes one partial sum per quantity (plain stores), and
// a single-block finalize kernel adds the partials in a fixed order.  The
// kernel boundary gives visibility across XCDs (each XCD has its own L2),
// so no device-scope fences are issued inside the streaming kernels (on
// MI355X such fences write back / invalidate the L2 and evict the gathered
// vector).  No float atomics: results are bitwise repeatable.
#pragma once

#include <hip/hip_runtime.h>

namespace fx {

constexpr int kBlock = 256;          // threads per block for vector kernels
constexpr int kMaxGridBlocks = 2048; // partial slots per reduced quantity
constexpr int kFinBlock = 1024;      // threads of a finalize kernel

template <int NV>
struct DVec {
  double v[NV];
};

// Sum NV values over a block of BS threads; result valid in thread 0 (the
// other threads get zeros: a defined value, so the compiler cannot treat
// their use of the result as undefined and specialise the caller for
// thread 0).  Fixed order: butterflies inside each wave, then a serial sum
// over waves.
// Call it once per kernel with all quantities (NV > 1): a second call would
// reuse the LDS slots while thread 0 may still read the first call's.
template <int NV, int BS>
__device__ __forceinline__ DVec<NV> block_sum(DVec<NV> a) {
  __shared__ double red[BS / 64][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = a.v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    a.v[k] = x;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid][k] = a.v[k];
  }
  __syncthreads();
  DVec<NV> out;
#pragma unroll
  for (int k = 0; k < NV; ++k) out.v[k] = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = red[0][k];
      for (int w = 1; w < BS / 64; ++w) s += red[w][k];
      out.v[k] = s;
    }
  }
  return out;
}

// Block partials of a producing kernel: partials[k * kMaxGridBlocks + b].
template <int NV>
__device__ __forceinline__ void store_partials(DVec<NV> part, int lblock, double* partials) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * kMaxGridBlocks + lblock] = part.v[k];
  }
}

// In a finalize kernel (one block of kFinBlock threads): fixed-order sum of
// the nblocks partials of each quantity; valid in thread 0.
template <int NV>
__device__ __forceinline__ DVec<NV> sum_partials(const double* __restrict__ partials, int nblocks) {
  // Eight independent loads in flight per thread: a long partial array
  // (one entry per SpMV block) is latency-bound, not bandwidth-bound.
  constexpr int kU = 8;
  DVec<NV> acc;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double* src = partials + k * kMaxGridBlocks;
    double s[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) s[u] = 0.0;
    int i = threadIdx.x;
    for (; i + (kU - 1) * kFinBlock < nblocks; i += kU * kFinBlock) {
#pragma unroll
      for (int u = 0; u < kU; ++u) s[u] += src[i + u * kFinBlock];
    }
    for (; i < nblocks; i += kFinBlock) s[0] += src[i];
    acc.v[k] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
  }
  return block_sum<NV, kFinBlock>(acc);
}

}  // namespace fx
