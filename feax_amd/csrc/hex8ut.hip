// A monolithic two-variable HEX8 problem: displacement u (vec 3) and
// temperature T (vec 1) on one mesh, in one CSR (feax's multi-variable
// Problem(mesh=[m, m], vec=[3, 1]), feax/problem.py:155-212,
// feax/assembler.py:591-647).
//
// DOF layout (feax offsets): u of node n at 3 n + i, T of node n at 3 N + n.
// Weak form per Gauss point (grad terms only):
//   variable 0: sigma = E (lam0 tr eps I + 2 mu0 eps) - E alpha (T - t_ref) / (1 - 2 nu) I
//   variable 1: k grad T
// with E = the volume variable (node or cell; SIMP law optional), nu,
// alpha, t_ref and the conductivity k (fx_material.kappa) constants.  The
// Jacobian blocks per element pair (a, b): K_uu (the HEX8 elasticity
// block), K_uT = -sum_q E alpha / (1 - 2 nu) N_b dN_a/dx_i JxW, K_Tu = 0,
// K_TT = sum_q k grad N_a . grad N_b JxW.  Surface loads: per face
// fdir[4 f + i] (i < 3) times the surface value on u, fdir[4 f + 3] times
// it on T (+ sign, feax/assembler.py:494-535).
//
// CSR rows (the dedupe of all element DOF pairs, columns ascending): every
// row of node n holds 4 deg entries, the u columns 3 m + j of its deg
// neighbours m (ascending) and then their T columns 3 N + m.
//
// Two passes: one wave per element forms the 32 x 32 element matrix as 64
// 4x4 blocks (a, b) (rows u_a0..2, T_a; columns u_b0..2, T_b) or its 32
// residual entries in a workspace; then one wave (thread) per node gathers
// them in ascending element order.
#include <cstdint>

#include "cell_common.h"
#include "common.h"
#include "feaxhip.h"
#include "hex8_common.h"

namespace fx {
namespace {

constexpr int kUtElemPerBlock = 4;  // one wave per element
constexpr int kUtMaxDeg = 64;

// One wave per element: lane (q, c) forms grad N_c and N_c at Gauss point q;
// lane (a, b) then the 4x4 block (a, b) at ke[1024 e + 16 (8 a + b)].
__global__ __launch_bounds__(64 * kUtElemPerBlock) void ut_elem_matrix_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, double* __restrict__ ke) {
  __shared__ double G_all[kUtElemPerBlock][8][8][3];
  __shared__ double N_all[kUtElemPerBlock][8][8];
  __shared__ double Q_all[kUtElemPerBlock][8][2];  // E JxW, JxW
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e0 = blockIdx.x * kUtElemPerBlock + wid;
  const bool live = e0 < num_cells;
  const int e = live ? e0 : num_cells - 1;  // a dead wave repeats the last element and stores nothing
  double(*G)[8][3] = G_all[wid];
  double(*Nq)[8] = N_all[wid];
  double(*Q)[2] = Q_all[wid];
  const int* cn = cells + 8 * static_cast<long>(e);
  {
    const int q = lane >> 3, c = lane & 7;
    double x0, x1, x2;
    gauss_point(q, x0, x1, x2);
    double X[8][3], Ji[3][3];
    load_coords<double>(x, y, z, cn, X);
    const double det = jac_inv<double>(X, x0, x1, x2, Ji);
    double N, d0, d1, d2, g[3];
    shape<double>(c, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, g);
    G[q][c][0] = g[0];
    G[q][c][1] = g[1];
    G[q][c][2] = g[2];
    Nq[q][c] = N;
    if (c == 0) {
      const double w = det * 0.125;
      Q[q][0] = w * modulus_at<double>(mat, param, e, cn, x0, x1, x2);
      Q[q][1] = w;
    }
  }
  __syncthreads();
  const int a = lane >> 3, b = lane & 7;
  const double nu = mat.nu;
  const double lam0 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu0 = 1.0 / (2.0 * (1.0 + nu));
  const double th = mat.alpha / (1.0 - 2.0 * nu), k = mat.kappa;
  double K[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) K[t] = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double ga[3] = {G[q][a][0], G[q][a][1], G[q][a][2]};
    const double gb[3] = {G[q][b][0], G[q][b][1], G[q][b][2]};
    const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    const double Ew = Q[q][0];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j)
        K[4 * i + j] += Ew * (lam0 * ga[i] * gb[j] + mu0 * ga[j] * gb[i] + (i == j ? mu0 * gg : 0.0));
      K[4 * i + 3] -= Ew * th * Nq[q][b] * ga[i];
    }
    K[15] += k * Q[q][1] * gg;
  }
  if (live) {
    double* dst = ke + 1024 * static_cast<long>(e) + 16 * lane;
#pragma unroll
    for (int t = 0; t < 16; ++t) dst[t] = K[t];
  }
}

// One wave per node: its 4 rows (u 3 n + r, T 3 N + n) from the blocks of
// its elements in ascending order; Jacobi dinv; the Dirichlet elimination
// of feax/DCboundary.py:147-199 (BC row: zero, diagonal 1; symmetric: BC
// columns zeroed).
template <typename T>
__global__ __launch_bounds__(64 * kNodesPerBlock) void ut_gather_rows_kernel(
    int num_nodes, const int* __restrict__ inc_ptr, const int* __restrict__ inc, const uint8_t* __restrict__ rank,
    const int* __restrict__ indptr, const int* __restrict__ indices, const double* __restrict__ ke,
    const uint8_t* __restrict__ bc_mask, int symmetric, double jacobi_shift, double* __restrict__ dinv,
    T* __restrict__ data) {
  __shared__ double acc_all[kNodesPerBlock][16 * kUtMaxDeg];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = blockIdx.x * kNodesPerBlock + wid;
  if (nd >= num_nodes) return;
  double* acc = acc_all[wid];
  const int len = indptr[3 * nd + 1] - indptr[3 * nd];  // 4 deg
  const int deg = len / 4;
  if (deg > kUtMaxDeg) return;  // host guarantees deg <= kUtMaxDeg
  for (int i = lane; i < 4 * len; i += 64) acc[i] = 0.0;
  __builtin_amdgcn_wave_barrier();
  for (int kk = inc_ptr[nd]; kk < inc_ptr[nd + 1]; ++kk) {
    const int pk = inc[kk];
    const int e = pk >> 3, a = pk & 7;
    const double* src = ke + 1024 * static_cast<long>(e) + 128 * a;
    const uint8_t* rk = rank + 64 * static_cast<long>(e) + 8 * a;
    for (int idx = lane; idx < 128; idx += 64) {
      const int b = idx >> 4, t = idx & 15;
      const int r = t >> 2, c = t & 3;
      const int col = c < 3 ? 3 * rk[b] + c : 3 * deg + rk[b];
      acc[r * len + col] += src[idx];
    }
    __builtin_amdgcn_wave_barrier();
  }
  int kself;
  {
    const int pk = inc[inc_ptr[nd]];
    kself = rank[64 * static_cast<long>(pk >> 3) + 9 * (pk & 7)];
  }
  const int N3 = 3 * num_nodes;
  if (dinv && lane < 4) {
    const int row = lane < 3 ? 3 * nd + lane : N3 + nd;
    const int pos = lane < 3 ? 3 * kself + lane : 3 * deg + kself;
    const double d = (bc_mask && bc_mask[row]) ? 1.0 : acc[lane * len + pos];
    dinv[row] = 1.0 / (d + jacobi_shift);
  }
  for (int idx = lane; idx < 4 * len; idx += 64) {
    const int r = idx / len, p = idx - len * (idx / len);
    const int row = r < 3 ? 3 * nd + r : N3 + nd;
    const int slot = indptr[row] + p;
    double v = acc[idx];
    if (bc_mask) {
      const int c = indices[slot];
      const bool zr = bc_mask[row] != 0;
      if (zr || (symmetric && bc_mask[c] != 0)) v = 0.0;
      if (zr && row == c) v = 1.0;
    }
    data[slot] = static_cast<T>(v);
  }
}

// One wave per element: re[32 e + 3 a + i] = sum_q sigma_q,id dN_a/dx_d JxW,
// re[32 e + 24 + a] = sum_q k grad T_q . grad N_a JxW.
__global__ __launch_bounds__(64 * kUtElemPerBlock) void ut_elem_residual_kernel(
    int num_cells, int num_nodes, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int* __restrict__ cells, MatParams mat, const double* __restrict__ param,
    const double* __restrict__ w, double* __restrict__ re) {
  __shared__ double G_all[kUtElemPerBlock][8][8][3];
  __shared__ double S_all[kUtElemPerBlock][8][12];  // sigma JxW (9), k grad T JxW (3)
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e0 = blockIdx.x * kUtElemPerBlock + wid;
  const bool live = e0 < num_cells;
  const int e = live ? e0 : num_cells - 1;
  double(*G)[8][3] = G_all[wid];
  double(*S)[12] = S_all[wid];
  const int* cn = cells + 8 * static_cast<long>(e);
  const double* Tn = w + 3 * static_cast<long>(num_nodes);
  {
    const int q = lane >> 3, c = lane & 7;
    double x0, x1, x2;
    gauss_point(q, x0, x1, x2);
    double X[8][3], Ji[3][3];
    load_coords<double>(x, y, z, cn, X);
    const double det = jac_inv<double>(X, x0, x1, x2, Ji);
    double N, d0, d1, d2, g[3];
    shape<double>(c, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, g);
    G[q][c][0] = g[0];
    G[q][c][1] = g[1];
    G[q][c][2] = g[2];
    // sums over the 8 lanes c of this Gauss point
    double ug[3][3], gT[3];
    const double tc = Tn[cn[c]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double v = w[3 * cn[c] + i] * g[d];
        v = xor_sum(v, 1);
        v = xor_sum(v, 2);
        v = xor_sum(v, 4);
        ug[i][d] = v;
      }
      double v = tc * g[d];
      v = xor_sum(v, 1);
      v = xor_sum(v, 2);
      v = xor_sum(v, 4);
      gT[d] = v;
    }
    double tq = tc * N;
    tq = xor_sum(tq, 1);
    tq = xor_sum(tq, 2);
    tq = xor_sum(tq, 4);
    if (c == 0) {
      const double jw = det * 0.125;
      const double E = modulus_at<double>(mat, param, e, cn, x0, x1, x2);
      const double nu = mat.nu;
      const double lam0 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu0 = 1.0 / (2.0 * (1.0 + nu));
      const double tr = ug[0][0] + ug[1][1] + ug[2][2];
      const double sh = mat.alpha * (tq - mat.tref) / (1.0 - 2.0 * nu);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          S[q][3 * i + d] = jw * E * (mu0 * (ug[i][d] + ug[d][i]) + (i == d ? lam0 * tr - sh : 0.0));
#pragma unroll
      for (int d = 0; d < 3; ++d) S[q][9 + d] = jw * mat.kappa * gT[d];
    }
  }
  __syncthreads();
  if (live && lane < 32) {
    double r = 0.0;
    if (lane < 24) {
      const int a = lane / 3, i = lane - 3 * (lane / 3);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        r += S[q][3 * i] * G[q][a][0] + S[q][3 * i + 1] * G[q][a][1] + S[q][3 * i + 2] * G[q][a][2];
    } else {
      const int a = lane - 24;
#pragma unroll
      for (int q = 0; q < 8; ++q) r += S[q][9] * G[q][a][0] + S[q][10] * G[q][a][1] + S[q][11] * G[q][a][2];
    }
    re[32 * static_cast<long>(e) + lane] = r;
  }
}

// One thread per node: element residuals in ascending order, then the loaded
// faces (4 corner nodes each; incidence entries f * 8 + a).
__global__ __launch_bounds__(256) void ut_node_residual_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
    const double* __restrict__ re, const int* __restrict__ faces, const int* __restrict__ finc_ptr,
    const int* __restrict__ finc, const double* __restrict__ svar, const double* __restrict__ fdir, int negate,
    double* __restrict__ res) {
  const int nd = blockIdx.x * 256 + threadIdx.x;
  if (nd >= num_nodes) return;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    const double* src = re + 32 * static_cast<long>(pk >> 3);
    const int a = pk & 7;
    r[0] += src[3 * a];
    r[1] += src[3 * a + 1];
    r[2] += src[3 * a + 2];
    r[3] += src[24 + a];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> 3, a = pk & 7;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      double X[8][3];
      const int* cn = cells + 8 * static_cast<long>(e);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        X[c][0] = x[cn[c]];
        X[c][1] = y[cn[c]];
        X[c][2] = z[cn[c]];
      }
      double v = 0.0;
      for (int q = 0; q < 4; ++q) {
        double p0, p1, p2, N, d0, d1, d2;
        face_point(lf, q, p0, p1, p2);
        shape<double>(a, p0, p1, p2, N, d0, d1, d2);
        v += svar[4 * f + q] * N * face_scale(X, lf, p0, p1, p2);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] += fdir[4 * f + i] * v;
    }
  }
  const double sgn = negate ? -1.0 : 1.0;
  res[3 * nd] = sgn * r[0];
  res[3 * nd + 1] = sgn * r[1];
  res[3 * nd + 2] = sgn * r[2];
  res[3 * static_cast<long>(num_nodes) + nd] = sgn * r[3];
}

inline bool ut_law(const fx_material* m) { return m->kind == FX_MAT_LINEAR_ELASTIC || m->kind == FX_MAT_SIMP; }

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_hex8ut_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  (void)value_bytes;
  (void)flags;
  return num_cells < 0 ? 0 : 1024 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex8ut_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                           const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                           int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                           void* workspace, size_t ws_bytes) {
  (void)u;  // the weak form is linear: the Jacobian does not depend on the state
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !elem_rank || !indptr ||
      !indices || !mat || !param || !data)
    return fail(FX_EINVAL, "fx_hex8ut_assemble_f64: bad arguments");
  if (!ut_law(mat)) return fail(FX_EINVAL, "fx_hex8ut_assemble_f64: needs a linear elastic or SIMP modulus law");
  if (flags) return fail(FX_EINVAL, "fx_hex8ut_assemble_f64: no assembly flags (FP64 CSR only)");
  const size_t need = fx_hex8ut_assemble_workspace_bytes(num_cells, 8, 0);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8ut_assemble_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  double* ke = static_cast<double*>(workspace);
  hipLaunchKernelGGL(ut_elem_matrix_kernel, dim3((num_cells + kUtElemPerBlock - 1) / kUtElemPerBlock),
                     dim3(64 * kUtElemPerBlock), 0, s, num_cells, x, y, z, cells, to_params(mat), param, ke);
  FX_LAUNCHED("ut_elem_matrix_kernel");
  hipLaunchKernelGGL(ut_gather_rows_kernel<double>, dim3((num_nodes + kNodesPerBlock - 1) / kNodesPerBlock),
                     dim3(64 * kNodesPerBlock), 0, s, num_nodes, inc_ptr, inc, elem_rank, indptr, indices, ke,
                     bc_mask, symmetric, jacobi_shift, dinv, data);
  FX_LAUNCHED("ut_gather_rows_kernel");
  return FX_OK;
}

size_t fx_hex8ut_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 32 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex8ut_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                           const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                           const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                           int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !u ||
      !res)
    return fail(FX_EINVAL, "fx_hex8ut_residual_f64: bad arguments");
  if (!ut_law(mat)) return fail(FX_EINVAL, "fx_hex8ut_residual_f64: needs a linear elastic or SIMP modulus law");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_hex8ut_residual_f64: face arrays missing");
  const size_t need = fx_hex8ut_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8ut_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool has_faces = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(ut_elem_residual_kernel, dim3((num_cells + kUtElemPerBlock - 1) / kUtElemPerBlock),
                     dim3(64 * kUtElemPerBlock), 0, s, num_cells, num_nodes, x, y, z, cells, to_params(mat), param, u,
                     re);
  hipLaunchKernelGGL(ut_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, z,
                     cells, inc_ptr, inc, re, has_faces ? faces : nullptr, has_faces ? finc_ptr : nullptr,
                     has_faces ? finc : nullptr, svar, fdir, negate, res);
  FX_LAUNCHED("fx_hex8ut_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_hex8ut_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

}  // extern "C"
