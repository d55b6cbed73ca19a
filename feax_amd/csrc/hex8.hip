// HEX8 element kernels for gfx950: CSR Jacobian assembly and residual.
//
// feax evaluates element Jacobians with 24 forward-mode JVPs per element
// inside a 256-cell lax.scan and accumulates them with a sorted
// segment_sum into the CSR value array (feax/assembler.py:808-818,
// 1253-1318); geometry (shape gradients, JxW) is precomputed and stored per
// element (feax/fe.py:96-129), ~3.2 KB/element.
//
// Here one 64-lane wavefront owns one mesh node n and produces its three
// CSR rows 3n..3n+2 (contiguous in the CSR value array, 9*deg(n) values):
//   for each element e of n, in ascending element order (= feax's
//   summation order for a slot):
//     lane (q, b) = (lane >> 3, lane & 7) recomputes the geometry at Gauss
//     point q from the element's SoA coordinates (nothing per element is
//     stored), forms the 3x3 block K_ab^q of node pair (a, b) with a the
//     local index of n, and the 8 quadrature lanes of each b are summed
//     with a fixed butterfly;
//     lanes q == 0 add the block into an LDS accumulator at the slot given
//     by elem_rank[e][a][b] (position of node b in n's neighbour list);
//   the accumulator is written once, coalesced, with the Dirichlet
//   elimination of feax/DCboundary.py:147-199 fused into the store.
// No float atomics; bitwise repeatable; CSR values written exactly once.
//
// Given a workspace (576 values per element; 324 for the linear laws,
// whose exactly symmetric element matrices are stored as their upper block
// triangle), the same sums run in two passes instead: one wave per element
// forms the element matrix once, then the per-node waves gather it
// (elem_matrix_lin_kernel / elem_matrix_kernel, gather_rows_kernel).  The
// residual has the same two forms.
#include <cstdint>

#include "cell_common.h"
#include "hex8_common.h"
#include "pcg_common.h"
#include "reduce.h"

namespace fx {


// ------------------------------------------------------------------ assembly
// NH: Neo-Hookean tangent at the state u (the Gauss-point deformation
// gradient is formed from the element's 8 nodal displacements across the
// 8 lanes of each Gauss-point group).
template <typename T, typename C, bool NH>
__global__ __launch_bounds__(64 * kNodesPerBlock) void assemble_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
    const uint8_t* __restrict__ rank, const int* __restrict__ indptr, const int* __restrict__ indices,
    MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    const uint8_t* __restrict__ bc_mask, int symmetric, int diff_form, double jacobi_shift,
    double* __restrict__ dinv, T* __restrict__ data) {
  __shared__ C acc_all[kNodesPerBlock][9 * kMaxDeg];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = blockIdx.x * kNodesPerBlock + wid;
  if (nd >= num_nodes) return;
  C* acc = acc_all[wid];
  const int row0 = indptr[3 * nd];
  const int len = indptr[3 * nd + 1] - row0;  // 3 * deg
  const int nval = 3 * len;
  if (len > 3 * kMaxDeg) return;  // host guarantees deg <= kMaxDeg
  for (int i = lane; i < nval; i += 64) acc[i] = C(0);
  __builtin_amdgcn_wave_barrier();

  const int q = lane >> 3, b = lane & 7;
  const C x0 = static_cast<C>(gauss_pt(q >> 2)), x1 = static_cast<C>(gauss_pt((q >> 1) & 1)),
          x2 = static_cast<C>(gauss_pt(q & 1));
  const C nu = static_cast<C>(mat.nu);
  const C lam0 = nu / ((C(1) + nu) * (C(1) - C(2) * nu)), mu0 = C(1) / (C(2) * (C(1) + nu));

  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    const int e = pk >> 3, a = pk & 7;
    const int* cn = cells + 8 * static_cast<long>(e);
    C X[8][3];
    load_coords<C>(x, y, z, cn, X);
    C Ji[3][3];
    const C det = jac_inv<C>(X, x0, x1, x2, Ji);
    C N, d0, d1, d2, ga[3], gb[3];
    shape<C>(a, x0, x1, x2, N, d0, d1, d2);
    phys_grad<C>(Ji, d0, d1, d2, ga);
    shape<C>(b, x0, x1, x2, N, d0, d1, d2);
    phys_grad<C>(Ji, d0, d1, d2, gb);
    const C w = det * C(0.125);  // JxW, Gauss weights 1/8 on [0,1]^3
    C K[9];
    if constexpr (NH) {
      // grad u at this Gauss point: lane b holds u_b (x) grad N_b, summed
      // over the 8 lanes of the point's group.
      const int nb = cn[b];
      C ug[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          C v = static_cast<C>(u[3 * nb + i]) * gb[d];
          v = xor_sum(v, 1);
          v = xor_sum(v, 2);
          v = xor_sum(v, 4);
          ug[i][d] = v;
        }
      C FiT[3][3];
      const C J = inv_transpose_F<C>(ug, FiT);
      const C mu = param_at<C>(mat, param, e, cn, x0, x1, x2);
      const C kap = static_cast<C>(mat.kappa);
      // dP/dF = mu d_ik d_JL + kappa (2J-1) J F^-T_iJ F^-T_kL - c F^-T_iL F^-T_kJ,
      // c = -mu + kappa J (J-1); contracted with grad N_a (J) and grad N_b (L).
      const C c1 = kap * (C(2) * J - C(1)) * J, c2 = -mu + kap * J * (J - C(1));
      C fa[3], fb[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        fa[i] = FiT[i][0] * ga[0] + FiT[i][1] * ga[1] + FiT[i][2] * ga[2];
        fb[i] = FiT[i][0] * gb[0] + FiT[i][1] * gb[1] + FiT[i][2] * gb[2];
      }
      const C gg = mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          K[3 * i + j] = w * ((i == j ? gg : C(0)) + c1 * fa[i] * fb[j] - c2 * fb[i] * fa[j]);
    } else {
      const C E = modulus_at<C>(mat, param, e, cn, x0, x1, x2);
      const C wl = w * E * lam0, wm = w * E * mu0;
      const C gg = wm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          K[3 * i + j] = wl * ga[i] * gb[j] + wm * ga[j] * gb[i] + (i == j ? gg : C(0));
    }
    // sum over the 8 Gauss points (lanes b, b+8, ..., b+56)
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      C v = K[t];
      v = xor_sum(v, 8);
      v = xor_sum(v, 16);
      v = xor_sum(v, 32);
      K[t] = v;
    }
    if (q == 0) {
      const int kb = rank[64 * static_cast<long>(e) + 8 * a + b];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i * len + 3 * kb + j] += K[3 * i + j];
    }
    __builtin_amdgcn_wave_barrier();
  }
  finish_rows<T, C>(nd, lane, acc, len, row0, nval, inc_ptr, inc, rank, indices, bc_mask, symmetric, diff_form,
                    jacobi_shift, dinv, data);
}

// ------------------------------------------------- two-pass assembly
// With a workspace the assembly runs in two passes:
//   1. elem_matrix_kernel, one wave per element: lanes (q, c) form the
//      Jacobian inverse and grad N_c at Gauss point q once (LDS), lanes
//      (a, b) then sum the 3x3 block K_ab over q = 0..7; the 576 values
//      are staged in LDS and stored coalesced to ke[e][a][b][i][j];
//   2. gather_wide_rows_kernel, one wave per node (as assemble_kernel):
//      the blocks of the node's elements, ascending element order, are
//      added into the LDS rows at the slots of elem_rank and stored by
//      finish_rows.
// Each element's geometry is computed once instead of once per node, and
// no wave waits on a chain of dependent loads per element.  The arithmetic
// and its order are those of assemble_kernel (values agree to the rounding
// of an operation: the compiler may contract a multiply-add differently).
constexpr int kElemPerBlock = 4;  // one wave per element

template <typename C, bool NH>
__global__ __launch_bounds__(64 * kElemPerBlock) void elem_matrix_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    C* __restrict__ ke) {
  __shared__ C G_all[kElemPerBlock][8][8][3];  // grad N_c at Gauss point q
  __shared__ C Q_all[kElemPerBlock][8][13];    // per Gauss point: JxW and material factors
  __shared__ C K_all[kElemPerBlock][576];      // the element matrix, staged for the store
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e0 = blockIdx.x * kElemPerBlock + wid;
  const bool live = e0 < num_cells;
  const int e = live ? e0 : num_cells - 1;  // a dead wave repeats the last element and stores nothing
  C(*G)[8][3] = G_all[wid];
  C(*Q)[13] = Q_all[wid];
  C* KS = K_all[wid];
  const int* cn = cells + 8 * static_cast<long>(e);
  {
    const int q = lane >> 3, c = lane & 7;
    const C x0 = static_cast<C>(gauss_pt(q >> 2)), x1 = static_cast<C>(gauss_pt((q >> 1) & 1)),
            x2 = static_cast<C>(gauss_pt(q & 1));
    C X[8][3];
    load_coords<C>(x, y, z, cn, X);
    C Ji[3][3];
    const C det = jac_inv<C>(X, x0, x1, x2, Ji);
    C N, d0, d1, d2, g[3];
    shape<C>(c, x0, x1, x2, N, d0, d1, d2);
    phys_grad<C>(Ji, d0, d1, d2, g);
    G[q][c][0] = g[0];
    G[q][c][1] = g[1];
    G[q][c][2] = g[2];
    const C w = det * C(0.125);  // JxW, Gauss weights 1/8 on [0,1]^3
    if constexpr (NH) {
      C ug[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          C v = static_cast<C>(u[3 * cn[c] + i]) * g[d];
          v = xor_sum(v, 1);
          v = xor_sum(v, 2);
          v = xor_sum(v, 4);
          ug[i][d] = v;
        }
      C FiT[3][3];
      const C J = inv_transpose_F<C>(ug, FiT);
      const C mu = param_at<C>(mat, param, e, cn, x0, x1, x2);
      const C kap = static_cast<C>(mat.kappa);
      if (c == 0) {
        Q[q][0] = w;
        Q[q][1] = mu;
        Q[q][2] = kap * (C(2) * J - C(1)) * J;
        Q[q][3] = -mu + kap * J * (J - C(1));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) Q[q][4 + 3 * i + j] = FiT[i][j];
      }
    } else {
      const C nu = static_cast<C>(mat.nu);
      const C lam0 = nu / ((C(1) + nu) * (C(1) - C(2) * nu)), mu0 = C(1) / (C(2) * (C(1) + nu));
      const C E = modulus_at<C>(mat, param, e, cn, x0, x1, x2);
      if (c == 0) {
        Q[q][0] = w * E * lam0;
        Q[q][1] = w * E * mu0;
      }
    }
  }
  __syncthreads();
  {
    // Block (a, b) at Gauss point q, with the arithmetic of assemble_kernel.
    const int a = lane >> 3, b = lane & 7;
    auto block = [&](int q, C (&K)[9]) {
      const C ga[3] = {G[q][a][0], G[q][a][1], G[q][a][2]};
      const C gb[3] = {G[q][b][0], G[q][b][1], G[q][b][2]};
      if constexpr (NH) {
        C fa[3], fb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          fa[i] = Q[q][4 + 3 * i] * ga[0] + Q[q][5 + 3 * i] * ga[1] + Q[q][6 + 3 * i] * ga[2];
          fb[i] = Q[q][4 + 3 * i] * gb[0] + Q[q][5 + 3 * i] * gb[1] + Q[q][6 + 3 * i] * gb[2];
        }
        const C w = Q[q][0], c1 = Q[q][2], c2 = Q[q][3];
        const C gg = Q[q][1] * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            K[3 * i + j] = w * ((i == j ? gg : C(0)) + c1 * fa[i] * fb[j] - c2 * fb[i] * fa[j]);
      } else {
        const C wl = Q[q][0], wm = Q[q][1];
        const C gg = wm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) K[3 * i + j] = wl * ga[i] * gb[j] + wm * ga[j] * gb[i] + (i == j ? gg : C(0));
      }
    };
    // Sum over q in the order of assemble_kernel's butterfly (lane offsets
    // 8, 16, 32): ((K0 + K1) + (K2 + K3)) + ((K4 + K5) + (K6 + K7)).
    C half[2][9];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      C pair[2][9];
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        C K0[9], K1[9];
        block(4 * h + 2 * p, K0);
        block(4 * h + 2 * p + 1, K1);
#pragma unroll
        for (int t = 0; t < 9; ++t) pair[p][t] = K0[t] + K1[t];
      }
#pragma unroll
      for (int t = 0; t < 9; ++t) half[h][t] = pair[0][t] + pair[1][t];
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) KS[9 * lane + t] = half[0][t] + half[1][t];
  }
  __syncthreads();
  if (live) {
    C* dst = ke + 576 * static_cast<long>(e);
    for (int i = lane; i < 576; i += 64) dst[i] = KS[i];
  }
}

// Linear laws (linear elastic, SIMP): each wave forms the matrices of 8
// elements.  Geometry: lane (el, q) = (lane >> 3, lane & 7) builds the
// Jacobian of element el at Gauss point q once and writes grad N_c for all
// 8 nodes c and the weights JxW E lam0 / JxW E mu0 to LDS (one lane per
// point instead of eight).  Blocks: only the upper block triangle is formed
// and stored (36 blocks, 324 values per element; the gather reads block
// (b, a) as the transpose of block (a, b), so K_ba = K_ab^T holds bitwise).
// The wave's 8 x 36 blocks are dealt to its lanes in storage order (block g
// = 36 el + packed_block(a, b) in 4.5 passes, instead of 8 passes of all 64
// blocks (a, b)), and block g's 9 values go to ke[324 ebase + 9 g ..].
// Per Gauss point a lane adds wl ga gb^T and wm ga gb^T to two 3x3 sums
// (18 fused multiply-adds); the block is then
//   K_ab = sum_q wl ga gb^T + (sum_q wm ga gb^T)^T + trace(sum_q wm ga gb^T) I,
// which is sum_q wl ga gb^T + wm gb ga^T + wm (ga . gb) I.  A diagonal
// block takes its upper triangle for its lower one, so that it is exactly
// symmetric too (CG iteration counts react to a non-symmetric rounding of
// the matrix).
constexpr int kLinWaves = 2;  // waves per block
constexpr int kLinElems = 8;  // elements per wave
constexpr int kPacked = 324;  // values of the upper block triangle of a HEX8 element matrix (36 blocks)

template <typename C>
__global__ __launch_bounds__(64 * kLinWaves) void elem_matrix_lin_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, C* __restrict__ ke) {
  __shared__ C G_all[kLinWaves][kLinElems][8][8][3];  // grad N_c of element el at Gauss point q
  __shared__ C W_all[kLinWaves][kLinElems][8][2];     // JxW E lam0, JxW E mu0
  __shared__ C S_all[kLinWaves][9 * 32];               // 32 blocks, staged for the store
  const int wid = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int ebase = (blockIdx.x * kLinWaves + wid) * kLinElems;
  if (ebase >= num_cells) return;
  C(*G)[8][8][3] = G_all[wid];
  C(*W)[8][2] = W_all[wid];
  C* S = S_all[wid];
  {
    const int el = lane >> 3, q = lane & 7;
    const int e = min(ebase + el, num_cells - 1);  // a lane past the end repeats the last element
    const int* cn = cells + 8 * static_cast<long>(e);
    const C x0 = static_cast<C>(gauss_pt(q >> 2)), x1 = static_cast<C>(gauss_pt((q >> 1) & 1)),
            x2 = static_cast<C>(gauss_pt(q & 1));
    C X[8][3];
    load_coords<C>(x, y, z, cn, X);
    C Ji[3][3];
    const C det = jac_inv<C>(X, x0, x1, x2, Ji);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      C N, d0, d1, d2, g[3];
      shape<C>(c, x0, x1, x2, N, d0, d1, d2);
      phys_grad<C>(Ji, d0, d1, d2, g);
      G[el][q][c][0] = g[0];
      G[el][q][c][1] = g[1];
      G[el][q][c][2] = g[2];
    }
    const C nu = static_cast<C>(mat.nu);
    const C lam0 = nu / ((C(1) + nu) * (C(1) - C(2) * nu)), mu0 = C(1) / (C(2) * (C(1) + nu));
    const C wE = det * C(0.125) * modulus_at<C>(mat, param, e, cn, x0, x1, x2);  // JxW E
    W[el][q][0] = wE * lam0;
    W[el][q][1] = wE * mu0;
  }
  __builtin_amdgcn_wave_barrier();
  const int nblk = 36 * min(kLinElems, num_cells - ebase);
  C* dst = ke + kPacked * static_cast<long>(ebase);
  for (int g0 = 0; g0 < nblk; g0 += 64) {
    const int g = g0 + lane;
    C K[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) K[t] = C(0);
    if (g < nblk) {
      const int el = g / 36, pb = g - 36 * el;
      // block row a of the packed triangle starts at a (17 - a) / 2
      const int a = (pb >= 8) + (pb >= 15) + (pb >= 21) + (pb >= 26) + (pb >= 30) + (pb >= 33) + (pb >= 35);
      const int b = pb - a * (17 - a) / 2 + a;
      C K1[9], K2[9];
#pragma unroll
      for (int t = 0; t < 9; ++t) K1[t] = K2[t] = C(0);
#pragma unroll 2
      for (int q = 0; q < 8; ++q) {
        const C wl = W[el][q][0], wm = W[el][q][1];
        const C gb0 = G[el][q][b][0], gb1 = G[el][q][b][1], gb2 = G[el][q][b][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const C ga = G[el][q][a][i];
          const C al = wl * ga, am = wm * ga;
          K1[3 * i + 0] = fma_c(al, gb0, K1[3 * i + 0]);
          K1[3 * i + 1] = fma_c(al, gb1, K1[3 * i + 1]);
          K1[3 * i + 2] = fma_c(al, gb2, K1[3 * i + 2]);
          K2[3 * i + 0] = fma_c(am, gb0, K2[3 * i + 0]);
          K2[3 * i + 1] = fma_c(am, gb1, K2[3 * i + 1]);
          K2[3 * i + 2] = fma_c(am, gb2, K2[3 * i + 2]);
        }
      }
      const C tr = K2[0] + K2[4] + K2[8];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const C v = K1[3 * i + j] + K2[3 * j + i];
          K[3 * i + j] = i == j ? v + tr : v;
        }
      if (a == b) {
        K[3] = K[1];
        K[6] = K[2];
        K[7] = K[5];
      }
    }
    // Stores through LDS, 32 blocks (288 contiguous values) at a time, so
    // that each store instruction writes consecutive addresses.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int gh = g0 + 32 * h;
      if (gh >= nblk) break;
      __builtin_amdgcn_wave_barrier();
      if ((lane >> 5) == h) {
#pragma unroll
        for (int t = 0; t < 9; ++t) S[9 * (lane & 31) + t] = K[t];
      }
      __builtin_amdgcn_wave_barrier();
      const int nv = 9 * min(32, nblk - gh);
      for (int i = lane; i < nv; i += 64) dst[9 * gh + i] = S[i];
    }
  }
}

// ------------------------------------------------------------------ residual
__global__ __launch_bounds__(64 * kNodesPerBlock) void residual_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, const int* __restrict__ inc_ptr, const int* __restrict__ inc, MatParams mat,
    const double* __restrict__ param, const double* __restrict__ u, const int* __restrict__ faces,
    const int* __restrict__ finc_ptr, const int* __restrict__ finc, const double* __restrict__ svar,
    const double* __restrict__ fdir, int negate, double* __restrict__ res) {
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = blockIdx.x * kNodesPerBlock + wid;
  if (nd >= num_nodes) return;
  const int q = lane >> 3, c = lane & 7;
  const double x0 = gauss_pt(q >> 2), x1 = gauss_pt((q >> 1) & 1), x2 = gauss_pt(q & 1);
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    const int e = pk >> 3, a = pk & 7;
    const int* cn = cells + 8 * static_cast<long>(e);
    double X[8][3];
    load_coords<double>(x, y, z, cn, X);
    double Ji[3][3];
    const double det = jac_inv<double>(X, x0, x1, x2, Ji);
    double N, d0, d1, d2, gc[3], ga[3];
    shape<double>(c, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, gc);
    shape<double>(a, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, ga);
    const int nc = cn[c];
    double ug[3][3];  // du_i/dx_d, summed over the 8 nodes of this Gauss point
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double v = u[3 * nc + i] * gc[d];
        v = xor_sum(v, 1);
        v = xor_sum(v, 2);
        v = xor_sum(v, 4);
        ug[i][d] = v;
      }
    double sig[3][3];
    const double pq = param_at<double>(mat, param, e, cn, x0, x1, x2);
    material_stress<double>(mat, pq, ug, sig);
    const double kth = thermal_shift<double>(mat, pq, cn, x0, x1, x2);
    sig[0][0] -= kth;
    sig[1][1] -= kth;
    sig[2][2] -= kth;
    const double w = det * 0.125;
    double ri[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ri[i] = (sig[i][0] * ga[0] + sig[i][1] * ga[1] + sig[i][2] * ga[2]) * w;
    // every lane of a Gauss-point group holds the same value: sum over groups
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double v = ri[i];
      v = xor_sum(v, 8);
      v = xor_sum(v, 16);
      v = xor_sum(v, 32);
      ri[i] = v;
    }
    r0 += ri[0];
    r1 += ri[1];
    r2 += ri[2];
  }
  // Dead surface traction t = dir * svar[f, q] on the loaded faces of this
  // node, ascending face order (feax/assembler.py:494-535, + sign).
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> 3, a = pk & 7;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      double v = 0.0;
      if (lane < 4) {
        double p0, p1, p2;
        face_point(lf, lane, p0, p1, p2);
        double X[8][3];
        load_coords<double>(x, y, z, cells + 8 * static_cast<long>(e), X);
        const double nanson = face_scale(X, lf, p0, p1, p2);
        double N, d0, d1, d2;
        shape<double>(a, p0, p1, p2, N, d0, d1, d2);
        v = svar[4 * f + lane] * N * nanson;
      }
      v = xor_sum(v, 1);
      v = xor_sum(v, 2);
      r0 += fdir[3 * f] * v;
      r1 += fdir[3 * f + 1] * v;
      r2 += fdir[3 * f + 2] * v;
    }
  }
  if (lane == 0) {
    const double s = negate ? -1.0 : 1.0;
    res[3 * nd] = s * r0;
    res[3 * nd + 1] = s * r1;
    res[3 * nd + 2] = s * r2;
  }
}

// ------------------------------------------------- two-pass residual
// With a workspace: elem_residual_kernel (one wave per element, lanes
// (q, c): the same Gauss-point arithmetic and butterfly over q as
// residual_kernel, once per element) writes re[e][c][i]; node_residual_kernel
// (one thread per node) adds them in ascending element order, then the
// loaded faces.  The sums and their order are those of residual_kernel.
__global__ __launch_bounds__(64 * kElemPerBlock) void elem_residual_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    double* __restrict__ re) {
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e = blockIdx.x * kElemPerBlock + wid;
  if (e >= num_cells) return;
  const int q = lane >> 3, c = lane & 7;
  const double x0 = gauss_pt(q >> 2), x1 = gauss_pt((q >> 1) & 1), x2 = gauss_pt(q & 1);
  const int* cn = cells + 8 * static_cast<long>(e);
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double Ji[3][3];
  const double det = jac_inv<double>(X, x0, x1, x2, Ji);
  double N, d0, d1, d2, gc[3];
  shape<double>(c, x0, x1, x2, N, d0, d1, d2);
  phys_grad<double>(Ji, d0, d1, d2, gc);
  const int nc = cn[c];
  double ug[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = u[3 * nc + i] * gc[d];
      v = xor_sum(v, 1);
      v = xor_sum(v, 2);
      v = xor_sum(v, 4);
      ug[i][d] = v;
    }
  double sig[3][3];
  const double pq = param_at<double>(mat, param, e, cn, x0, x1, x2);
  material_stress<double>(mat, pq, ug, sig);
  const double kth = thermal_shift<double>(mat, pq, cn, x0, x1, x2);
  sig[0][0] -= kth;
  sig[1][1] -= kth;
  sig[2][2] -= kth;
  const double w = det * 0.125;
  double ri[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double v = (sig[i][0] * gc[0] + sig[i][1] * gc[1] + sig[i][2] * gc[2]) * w;
    v = xor_sum(v, 8);
    v = xor_sum(v, 16);
    v = xor_sum(v, 32);
    ri[i] = v;
  }
  if (q == 0) {
    double* dst = re + 24 * static_cast<long>(e) + 3 * c;
    dst[0] = ri[0];
    dst[1] = ri[1];
    dst[2] = ri[2];
  }
}

__global__ __launch_bounds__(256) void node_residual_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
    const double* __restrict__ re, const int* __restrict__ faces, const int* __restrict__ finc_ptr,
    const int* __restrict__ finc, const double* __restrict__ svar, const double* __restrict__ fdir, int negate,
    double* __restrict__ res) {
  const int nd = blockIdx.x * 256 + threadIdx.x;
  if (nd >= num_nodes) return;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    const double* src = re + 24 * static_cast<long>(pk >> 3) + 3 * (pk & 7);
    r0 += src[0];
    r1 += src[1];
    r2 += src[2];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> 3, a = pk & 7;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      double X[8][3];
      load_coords<double>(x, y, z, cells + 8 * static_cast<long>(e), X);
      double vq[4];
      for (int qq = 0; qq < 4; ++qq) {
        double p0, p1, p2, N, d0, d1, d2;
        face_point(lf, qq, p0, p1, p2);
        const double nanson = face_scale(X, lf, p0, p1, p2);
        shape<double>(a, p0, p1, p2, N, d0, d1, d2);
        vq[qq] = svar[4 * f + qq] * N * nanson;
      }
      const double v = (vq[0] + vq[1]) + (vq[2] + vq[3]);  // the 4-lane butterfly's order
      r0 += fdir[3 * f] * v;
      r1 += fdir[3 * f + 1] * v;
      r2 += fdir[3 * f + 2] * v;
    }
  }
  const double sgn = negate ? -1.0 : 1.0;
  res[3 * nd] = sgn * r0;
  res[3 * nd + 1] = sgn * r1;
  res[3 * nd + 2] = sgn * r2;
}

// ------------------------------------------------- matrix-free J_bc(u) v
// feax's Krylov operator (feax/assembler.py:1747-1778):
//   J_bc v = where(bc, v, JVP_R(u)[v * !bc]).
// Same wave-per-node layout as the residual: lanes (q, c) form grad v (and
// grad u for the Neo-Hookean law) at Gauss point q by a butterfly over the
// node lanes c, contract the linearised stress dP = A(grad u) : grad v and
// test it with grad N_a.
template <bool NH, bool DOT>
__global__ __launch_bounds__(64 * kNodesPerBlock) void jvp_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, const int* __restrict__ inc_ptr, const int* __restrict__ inc, MatParams mat,
    const double* __restrict__ param, const double* __restrict__ u, const double* __restrict__ v,
    const uint8_t* __restrict__ bc_mask, double* __restrict__ out, const PcgState* __restrict__ st,
    double* partials) {
  if constexpr (DOT) {
    if (st->done) return;  // the whole grid leaves together
  }
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = blockIdx.x * kNodesPerBlock + wid;
  const bool live = nd < num_nodes;
  const int q = lane >> 3, c = lane & 7;
  const double x0 = gauss_pt(q >> 2), x1 = gauss_pt((q >> 1) & 1), x2 = gauss_pt(q & 1);
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  const int k_end = live ? inc_ptr[nd + 1] : 0;
  for (int k = live ? inc_ptr[nd] : 0; k < k_end; ++k) {
    const int pk = inc[k];
    const int e = pk >> 3, a = pk & 7;
    const int* cn = cells + 8 * static_cast<long>(e);
    double X[8][3];
    load_coords<double>(x, y, z, cn, X);
    double Ji[3][3];
    const double det = jac_inv<double>(X, x0, x1, x2, Ji);
    double N, d0, d1, d2, gc[3], ga[3];
    shape<double>(c, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, gc);
    shape<double>(a, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, ga);
    const int nc = cn[c];
    double vg[3][3], ug[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double vv = (bc_mask && bc_mask[3 * nc + i]) ? 0.0 : v[3 * nc + i];
      const double uu = NH ? u[3 * nc + i] : 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double t = vv * gc[d];
        t = xor_sum(t, 1);
        t = xor_sum(t, 2);
        t = xor_sum(t, 4);
        vg[i][d] = t;
        if constexpr (NH) {
          double s = uu * gc[d];
          s = xor_sum(s, 1);
          s = xor_sum(s, 2);
          s = xor_sum(s, 4);
          ug[i][d] = s;
        }
      }
    }
    const double p = param_at<double>(mat, param, e, cn, x0, x1, x2);
    double dP[3][3];
    if constexpr (NH) {
      // dP_iJ = A_iJkL vg_kL = mu vg_iJ + c1 F^-T_iJ (F^-T : vg) - c2 (F^-T vg^T F^-T)_iJ
      double FiT[3][3];
      const double J = inv_transpose_F<double>(ug, FiT);
      const double kap = mat.kappa;
      const double c1 = kap * (2.0 * J - 1.0) * J, c2 = -p + kap * J * (J - 1.0);
      double t = 0.0, M[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) t += FiT[i][d] * vg[i][d];
#pragma unroll
      for (int L = 0; L < 3; ++L)
#pragma unroll
        for (int Jc = 0; Jc < 3; ++Jc) M[L][Jc] = vg[0][L] * FiT[0][Jc] + vg[1][L] * FiT[1][Jc] + vg[2][L] * FiT[2][Jc];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int Jc = 0; Jc < 3; ++Jc)
          dP[i][Jc] = p * vg[i][Jc] + c1 * FiT[i][Jc] * t -
                      c2 * (FiT[i][0] * M[0][Jc] + FiT[i][1] * M[1][Jc] + FiT[i][2] * M[2][Jc]);
    } else {
      material_stress<double>(mat, p, vg, dP);  // linear in grad v
    }
    const double w = det * 0.125;
    double ri[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ri[i] = (dP[i][0] * ga[0] + dP[i][1] * ga[1] + dP[i][2] * ga[2]) * w;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double t = ri[i];
      t = xor_sum(t, 8);
      t = xor_sum(t, 16);
      t = xor_sum(t, 32);
      ri[i] = t;
    }
    r0 += ri[0];
    r1 += ri[1];
    r2 += ri[2];
  }
  DVec<1> dot;
  dot.v[0] = 0.0;
  if (live && lane < 3) {
    const int r = 3 * nd + lane;
    const double val = (bc_mask && bc_mask[r]) ? v[r] : (lane == 0 ? r0 : (lane == 1 ? r1 : r2));
    out[r] = val;
    if constexpr (DOT) dot.v[0] = v[r] * val;
  }
  if constexpr (DOT) store_partials<1>(block_sum<1, 64 * kNodesPerBlock>(dot), blockIdx.x, partials);
}

int hex8_op_blocks(int num_nodes) { return (num_nodes + kNodesPerBlock - 1) / kNodesPerBlock; }

void launch_hex8_op(hipStream_t s, const Hex8Op& op, const double* v, double* out, const PcgState* st,
                    double* partials) {
  const int grid = hex8_op_blocks(op.num_nodes);
  const bool nh = op.mat.kind == FX_MAT_NEO_HOOKEAN;
#define FX_JVP_LAUNCH(NHV, DOTV)                                                                                \
  hipLaunchKernelGGL((jvp_kernel<NHV, DOTV>), dim3(grid), dim3(64 * kNodesPerBlock), 0, s, op.num_nodes, op.x, \
                     op.y, op.z, op.cells, op.inc_ptr, op.inc, op.mat, op.param, op.u, v, op.bc_mask, out, st,    \
                     partials)
  if (nh) {
    if (st) FX_JVP_LAUNCH(true, true);
    else FX_JVP_LAUNCH(true, false);
  } else {
    if (st) FX_JVP_LAUNCH(false, true);
    else FX_JVP_LAUNCH(false, false);
  }
#undef FX_JVP_LAUNCH
}

// Bytes of the element-matrix workspace: 576 values per element, FP64
// unless FP32 values are computed in FP32 (FX_ASM_COMPUTE32).
inline size_t elem_matrix_bytes(int num_cells, int value_bytes, int flags) {
  const size_t c = (value_bytes == 4 && (flags & FX_ASM_COMPUTE32)) ? 4 : 8;
  return 576 * c * static_cast<size_t>(num_cells);
}

template <typename T>
int assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y, const double* z,
             const int* cells, const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr,
             const int* indices, const fx_material* mat, const double* param, const double* u,
             const uint8_t* bc_mask, int symmetric, int flags, double jacobi_shift, double* dinv, T* data,
             void* workspace, size_t ws_bytes, const int* node_order, const int* up_ptr) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !rank || !indptr ||
      !indices || !mat || !param || !data)
    return fail(FX_EINVAL, "fx_hex8_assemble: bad arguments");
  const bool upper = (flags & FX_ASM_UPPER) != 0;
  if (upper && (sizeof(T) != 4 || !(flags & FX_ASM_DIFF3) || !symmetric || !workspace || !up_ptr ||
                mat->kind == FX_MAT_NEO_HOOKEAN))
    return fail(FX_EINVAL, "fx_hex8_assemble: FX_ASM_UPPER stores the FP32 difference form of a linear law with "
                           "symmetric elimination; it needs the workspace and up_ptr");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_assemble: unknown material %d", mat->kind);
  const bool nh = mat->kind == FX_MAT_NEO_HOOKEAN;
  if (nh && !u) return fail(FX_EINVAL, "fx_hex8_assemble: the Neo-Hookean tangent needs u");
  const int diff = (flags & FX_ASM_DIFF3) != 0;
  const int grid = (num_nodes + kNodesPerBlock - 1) / kNodesPerBlock;
  const MatParams mp = to_params(mat);
  hipStream_t s = as_stream(stream);
  // Element matrices are formed in FP64 and rounded once at the store, unless
  // FX_ASM_COMPUTE32 asks for FP32 arithmetic as well.
  const bool c32 = sizeof(T) == 4 && (flags & FX_ASM_COMPUTE32);
  if (workspace) {
    const size_t need = elem_matrix_bytes(num_cells, sizeof(T), flags);
    if (ws_bytes < need) return fail(FX_EWORKSPACE, "fx_hex8_assemble: workspace %zu < %zu", ws_bytes, need);
  }
  const int egrid = (num_cells + kElemPerBlock - 1) / kElemPerBlock;
  const int lgrid = (num_cells + kLinWaves * kLinElems - 1) / (kLinWaves * kLinElems);
#define FX_ASM_LAUNCH(C, NHV)                                                                                      \
  do {                                                                                                             \
    if (workspace) {                                                                                               \
      C* ke = static_cast<C*>(workspace);                                                                          \
      if (NHV)                                                                                                     \
        hipLaunchKernelGGL((elem_matrix_kernel<C, NHV>), dim3(egrid), dim3(64 * kElemPerBlock), 0, s, num_cells,   \
                           x, y, z, cells, mp, param, u, ke);                                                      \
      else                                                                                                         \
        hipLaunchKernelGGL((elem_matrix_lin_kernel<C>), dim3(lgrid), dim3(64 * kLinWaves), 0, s, num_cells, x, y,  \
                           z, cells, mp, param, ke);                                                               \
      hipLaunchKernelGGL((gather_wide_rows_kernel<T, C, 8, 3, 3, kMaxDeg, !NHV>), dim3(grid),                       \
                         dim3(64 * kNodesPerBlock), 0, s, num_nodes, node_order, inc_ptr, inc, rank, indptr,       \
                         indices, ke, bc_mask, symmetric, diff, jacobi_shift, dinv, data,                          \
                         upper ? up_ptr : nullptr);                                                                \
    } else {                                                                                                       \
      hipLaunchKernelGGL((assemble_kernel<T, C, NHV>), dim3(grid), dim3(64 * kNodesPerBlock), 0, s, num_nodes, x,  \
                         y, z, cells, inc_ptr, inc, rank, indptr, indices, mp, param, u, bc_mask, symmetric, diff, \
                         jacobi_shift, dinv, data);                                                                \
    }                                                                                                              \
  } while (0)
  if (nh) {
    if (c32) FX_ASM_LAUNCH(float, true);
    else FX_ASM_LAUNCH(double, true);
  } else {
    if (c32) FX_ASM_LAUNCH(float, false);
    else FX_ASM_LAUNCH(double, false);
  }
#undef FX_ASM_LAUNCH
  FX_LAUNCHED("assemble_kernel");
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_hex8_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                         const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                         int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                         void* workspace, size_t ws_bytes, const int32_t* node_order, const int32_t* up_ptr) {
  return assemble<double>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                          mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes,
                          node_order, up_ptr);
}

int fx_hex8_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                         const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                         int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                         void* workspace, size_t ws_bytes, const int32_t* node_order, const int32_t* up_ptr) {
  return assemble<float>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                         mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes,
                         node_order, up_ptr);
}

size_t fx_hex8_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  if (num_cells < 0 || (value_bytes != 4 && value_bytes != 8)) return 0;
  return elem_matrix_bytes(num_cells, value_bytes, flags);
}

size_t fx_hex8_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 24 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex8_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                         const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                         const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                         int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !u ||
      !res)
    return fail(FX_EINVAL, "fx_hex8_residual_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_residual_f64: unknown material %d", mat->kind);
  if (mat->temperature && mat->kind == FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_residual_f64: thermal strain is built for the linear / SIMP laws");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_hex8_residual_f64: face arrays missing");
  hipStream_t s = as_stream(stream);
  const bool has_faces = num_faces > 0;
  if (workspace) {
    const size_t need = 24 * sizeof(double) * static_cast<size_t>(num_cells);
    if (ws_bytes < need) return fail(FX_EWORKSPACE, "fx_hex8_residual_f64: workspace %zu < %zu", ws_bytes, need);
    double* re = static_cast<double*>(workspace);
    hipLaunchKernelGGL(elem_residual_kernel, dim3((num_cells + kElemPerBlock - 1) / kElemPerBlock),
                       dim3(64 * kElemPerBlock), 0, s, num_cells, x, y, z, cells, to_params(mat), param, u, re);
    hipLaunchKernelGGL(node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, z,
                       cells, inc_ptr, inc, re, has_faces ? faces : nullptr, has_faces ? finc_ptr : nullptr,
                       has_faces ? finc : nullptr, svar, fdir, negate, res);
  } else {
    const int grid = (num_nodes + kNodesPerBlock - 1) / kNodesPerBlock;
    hipLaunchKernelGGL(residual_kernel, dim3(grid), dim3(64 * kNodesPerBlock), 0, s, num_nodes, x, y, z, cells,
                       inc_ptr, inc, to_params(mat), param, u, has_faces ? faces : nullptr,
                       has_faces ? finc_ptr : nullptr, has_faces ? finc : nullptr, svar, fdir, negate, res);
  }
  FX_LAUNCHED("residual_kernel");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_hex8_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

int fx_hex8_jvp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                    const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                    const fx_material* mat, const double* param, const double* u, const double* v,
                    const uint8_t* bc_mask, double* out) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !v ||
      !out)
    return fail(FX_EINVAL, "fx_hex8_jvp_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_jvp_f64: unknown material %d", mat->kind);
  const bool nh = mat->kind == FX_MAT_NEO_HOOKEAN;
  if (nh && !u) return fail(FX_EINVAL, "fx_hex8_jvp_f64: the Neo-Hookean tangent needs u");
  if (out == v) return fail(FX_EINVAL, "fx_hex8_jvp_f64: out must not alias v");
  const Hex8Op op{num_nodes, x, y, z, cells, inc_ptr, inc, to_params(mat), param, nh ? u : nullptr, bc_mask};
  launch_hex8_op(as_stream(stream), op, v, out, nullptr, nullptr);
  FX_LAUNCHED("jvp_kernel");
  return FX_OK;
}

}  // extern "C"
