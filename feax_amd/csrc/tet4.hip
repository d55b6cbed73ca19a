// TET4 element kernels for gfx950: CSR Jacobian assembly and residual.
//
// feax's TET4 (feax/basis.py:78-83): basix's P1 tetrahedron, vertex order
// kept, N0 = 1 - x - y - z, N1 = x, N2 = y, N3 = z on the reference cell,
// one Gauss point (the centroid, weight 1/6); facet i is the triangle
// opposite vertex i with one point (weight 1/2 times the facet Jacobian
// size).  Geometry as for HEX8 (feax/fe.py:96-129, 175-239): J_ij =
// sum_a x_a,i dN_a/dxi_j = (x_{j+1} - x_0)_i, grad N = grad_xi N J^-1,
// JxW = det J / 6 (no abs), Nanson scale ||n_ref J^-1|| det J w_face.
//
// Two passes, as the HEX8 two-pass kernels: one thread per element block
// (a, b) forms the 3x3 block K_ab of the element matrix (144 values per
// element) or one thread per element its residual (12 values); then one
// wave (thread) per node gathers them in ascending element order into the
// node's CSR rows (cell_common.h, NPE = 4) or residual entries.
#include <cstdint>

#include "cell_common.h"
#include "common.h"
#include "feaxhip.h"
#include "hex8_common.h"
#include "reduce.h"

namespace fx {
namespace {

constexpr int kTetThreads = 256;

// Shape-function gradients g[a][d] and det J of element cn (C arithmetic).
template <typename C>
__device__ __forceinline__ C tet_grads(const double* __restrict__ x, const double* __restrict__ y,
                                       const double* __restrict__ z, const int* __restrict__ cn, C (&g)[4][3]) {
  C X[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    X[a][0] = static_cast<C>(x[cn[a]]);
    X[a][1] = static_cast<C>(y[cn[a]]);
    X[a][2] = static_cast<C>(z[cn[a]]);
  }
  C J[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) J[i][j] = X[j + 1][i] - X[0][i];
  const C c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const C c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const C c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const C det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const C id = C(1) / det;
  C Ji[3][3];
  Ji[0][0] = c00 * id;
  Ji[1][0] = c01 * id;
  Ji[2][0] = c02 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  // grad_xi N_{k+1} = e_k, grad_xi N_0 = -(1, 1, 1)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g[1][d] = Ji[0][d];
    g[2][d] = Ji[1][d];
    g[3][d] = Ji[2][d];
    g[0][d] = -((Ji[0][d] + Ji[1][d]) + Ji[2][d]);
  }
  return det;
}

// The material parameter at the centroid: sum_c N_c p_c with N_c = 1/4
// (node variable), or the cell value.
template <typename C>
__device__ __forceinline__ C tet_param(const MatParams& m, const double* __restrict__ param, int e,
                                       const int* __restrict__ cn) {
  if (m.loc == FX_PARAM_NODE) {
    C p = C(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) p += C(0.25) * static_cast<C>(param[cn[c]]);
    return p;
  }
  return static_cast<C>(param[e]);
}

// du_i/dX_d = sum_c u[3 cn[c] + i] g[c][d].
template <typename C>
__device__ __forceinline__ void tet_ugrad(const double* __restrict__ u, const int* __restrict__ cn,
                                          const C (&g)[4][3], C (&ug)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      C v = C(0);
#pragma unroll
      for (int c = 0; c < 4; ++c) v += static_cast<C>(u[3 * cn[c] + i]) * g[c][d];
      ug[i][d] = v;
    }
}

// One thread per (element, a, b): block K_ab (row-major 3x3) at
// ke[144 e + 36 a + 9 b], the HEX8 block formula at the single Gauss point.
template <typename C, bool NH>
__global__ __launch_bounds__(kTetThreads) void tet_elem_matrix_kernel(int num_cells, const double* __restrict__ x,
                                                                      const double* __restrict__ y,
                                                                      const double* __restrict__ z,
                                                                      const int* __restrict__ cells, MatParams mat,
                                                                      const double* __restrict__ param,
                                                                      const double* __restrict__ u,
                                                                      C* __restrict__ ke) {
  const long t = static_cast<long>(blockIdx.x) * kTetThreads + threadIdx.x;
  const long e = t >> 4;
  if (e >= num_cells) return;
  const int a = static_cast<int>((t >> 2) & 3), b = static_cast<int>(t & 3);
  const int* cn = cells + 4 * e;
  C g[4][3];
  const C w = tet_grads<C>(x, y, z, cn, g) / C(6);
  const C ga[3] = {g[a][0], g[a][1], g[a][2]};
  const C gb[3] = {g[b][0], g[b][1], g[b][2]};
  C K[9];
  if constexpr (NH) {
    C ug[3][3], FiT[3][3];
    tet_ugrad<C>(u, cn, g, ug);
    const C J = inv_transpose_F<C>(ug, FiT);
    const C mu = tet_param<C>(mat, param, static_cast<int>(e), cn);
    const C kap = static_cast<C>(mat.kappa);
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
      for (int j = 0; j < 3; ++j) K[3 * i + j] = w * ((i == j ? gg : C(0)) + c1 * fa[i] * fb[j] - c2 * fb[i] * fa[j]);
  } else {
    const C nu = static_cast<C>(mat.nu);
    const C lam0 = nu / ((C(1) + nu) * (C(1) - C(2) * nu)), mu0 = C(1) / (C(2) * (C(1) + nu));
    const C E = modulus_of<C>(mat, tet_param<C>(mat, param, static_cast<int>(e), cn));
    const C wl = w * E * lam0, wm = w * E * mu0;
    const C gg = wm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) K[3 * i + j] = wl * ga[i] * gb[j] + wm * ga[j] * gb[i] + (i == j ? gg : C(0));
  }
  C* dst = ke + 144 * e + 36 * a + 9 * b;
#pragma unroll
  for (int k = 0; k < 9; ++k) dst[k] = K[k];
}

// One thread per element: r_a,i = sum_d sigma_id g_a,d JxW at re[12 e + 3 a + i].
__global__ __launch_bounds__(kTetThreads) void tet_elem_residual_kernel(int num_cells, const double* __restrict__ x,
                                                                        const double* __restrict__ y,
                                                                        const double* __restrict__ z,
                                                                        const int* __restrict__ cells, MatParams mat,
                                                                        const double* __restrict__ param,
                                                                        const double* __restrict__ u,
                                                                        double* __restrict__ re) {
  const int e = blockIdx.x * kTetThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double g[4][3];
  const double w = tet_grads<double>(x, y, z, cn, g) / 6.0;
  double ug[3][3], sig[3][3];
  tet_ugrad<double>(u, cn, g, ug);
  material_stress<double>(mat, tet_param<double>(mat, param, e, cn), ug, sig);
  double* dst = re + 12 * static_cast<long>(e);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) dst[3 * a + i] = (sig[i][0] * g[a][0] + sig[i][1] * g[a][1] + sig[i][2] * g[a][2]) * w;
}

// basix facet data of the reference tetrahedron: outward normals and
// w_face = 1/2 * |facet Jacobian| (facet 0 is the slanted face).
__device__ __forceinline__ void tet_facet(int lf, double (&n)[3], double& wf) {
  const double r3 = 0.57735026918962576451;  // 1 / sqrt(3)
  n[0] = n[1] = n[2] = 0.0;
  if (lf == 0) {
    n[0] = n[1] = n[2] = r3;
    wf = 0.5 * 1.73205080756887729353;  // sqrt(3) / 2
  } else {
    n[lf - 1] = -1.0;
    wf = 0.5;
  }
}

// One thread per node: the element residuals of its elements in ascending
// order, then its loaded faces: t = fdir[f] * svar[f] at the facet's
// centroid, N_a = 1/3 on the facet's vertices, times the Nanson scale.
__global__ __launch_bounds__(256) void tet_node_residual_kernel(
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
    const double* src = re + 12 * static_cast<long>(pk >> 3) + 3 * (pk & 7);
    r0 += src[0];
    r1 += src[1];
    r2 += src[2];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> 3;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      double g[4][3];
      const double det = tet_grads<double>(x, y, z, cells + 4 * static_cast<long>(e), g);
      // n_ref J^-1 from the gradients: grad N_{k+1} are the rows of J^-1
      double n[3], wf;
      tet_facet(lf, n, wf);
      double m[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) m[d] = n[0] * g[1][d] + n[1] * g[2][d] + n[2] * g[3][d];
      const double nanson = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) * det * wf;
      const double v = svar[f] * (1.0 / 3.0) * nanson;
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

// ------------------------------------------------------------- adjoint
// The HEX8 pullbacks (adjoint.hip) at the single Gauss point: per element
// w_e = dR_e/dp . lambda_e (SELF: lambda = u, the compliance case), with
//   linear / SIMP: sigma_1(grad u) : grad lambda JxW dE/dp,
//   Neo-Hookean:   (F - F^-T) : grad lambda JxW;
// a cell variable gets scale * w_e, a node variable scale * sum over its
// elements of N_a(centroid) w_e = w_e / 4.  lambda is read as 0 on the rows
// with bc_mask set.
template <bool SELF>
__global__ __launch_bounds__(kTetThreads) void tet_param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    const double* __restrict__ lam, const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell,
    double* __restrict__ we) {
  const int e = blockIdx.x * kTetThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double g[4][3];
  const double jxw = tet_grads<double>(x, y, z, cn, g) / 6.0;
  double ug[3][3], lg[3][3];
  tet_ugrad<double>(u, cn, g, ug);
  if constexpr (SELF) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) lg[i][d] = ug[i][d];
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int r = 3 * cn[c] + i;
          v += ((bc_mask && bc_mask[r]) ? 0.0 : lam[r]) * g[c][d];
        }
        lg[i][d] = v;
      }
  }
  double dens = 0.0;
  if (mat.kind == FX_MAT_NEO_HOOKEAN) {
    double FiT[3][3];
    inv_transpose_F<double>(ug, FiT);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) dens += (ug[i][d] + (i == d ? 1.0 : 0.0) - FiT[i][d]) * lg[i][d];
  } else {
    const double nu = mat.nu;
    const double lam0 = nu / ((1 + nu) * (1 - 2 * nu)), mu0 = 1 / (2 * (1 + nu));
    const double tr = ug[0][0] + ug[1][1] + ug[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) dens += (mu0 * (ug[i][d] + ug[d][i]) + (i == d ? lam0 * tr : 0.0)) * lg[i][d];
    double dEdp;
    modulus_of<double>(mat, tet_param<double>(mat, param, e, cn), &dEdp);
    dens *= dEdp;
  }
  const double w = dens * jxw;
  if (out_cell)
    out_cell[e] = scale * w;
  else
    we[e] = w;
}

__global__ __launch_bounds__(256) void tet_param_vjp_node_kernel(int num_nodes, const int* __restrict__ inc_ptr,
                                                                 const int* __restrict__ inc,
                                                                 const double* __restrict__ we, double scale,
                                                                 double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num_nodes) return;
  double s = 0.0;
  for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) s += 0.25 * we[inc[k] >> 3];
  out[i] = scale * s;
}

// grad_svar[f] = scale * nanson * sum over the facet's vertices a of
// (dir_f . lambda_a) / 3 (facet lf is opposite vertex lf).
__global__ __launch_bounds__(256) void tet_traction_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ z,
                                                               const int* __restrict__ cells,
                                                               const int* __restrict__ faces,
                                                               const double* __restrict__ fdir,
                                                               const double* __restrict__ lam,
                                                               const uint8_t* __restrict__ bc_mask, double scale,
                                                               double* __restrict__ out) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= num_faces) return;
  const int e = faces[2 * f], lf = faces[2 * f + 1];
  const int* cn = cells + 4 * static_cast<long>(e);
  double g[4][3];
  const double det = tet_grads<double>(x, y, z, cn, g);
  double n[3], wf;
  tet_facet(lf, n, wf);
  double m[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) m[d] = n[0] * g[1][d] + n[1] * g[2][d] + n[2] * g[3][d];
  const double nanson = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) * det * wf;
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (a == lf) continue;
    double dl = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int r = 3 * cn[a] + i;
      dl += fdir[3 * f + i] * ((bc_mask && bc_mask[r]) ? 0.0 : lam[r]);
    }
    s += (1.0 / 3.0) * dl;
  }
  out[f] = scale * s * nanson;
}

// Neo-Hookean strain energy psi = mu/2 (I1 - 3) - mu ln J + kappa/2 (J - 1)^2
// (feax/assembler.py:1866-1961) at the single Gauss point: per element
// psi JxW and its partial d/dmu_e, block partial sums of the total.
__global__ __launch_bounds__(kTetThreads) void tet_energy_kernel(int num_cells, const double* __restrict__ x,
                                                                 const double* __restrict__ y,
                                                                 const double* __restrict__ z,
                                                                 const int* __restrict__ cells, MatParams mat,
                                                                 const double* __restrict__ param,
                                                                 const double* __restrict__ u,
                                                                 double* __restrict__ dpsi_dmu,
                                                                 double* __restrict__ partials) {
  const int e = blockIdx.x * kTetThreads + threadIdx.x;
  double psi = 0.0;
  if (e < num_cells) {
    const int* cn = cells + 4 * static_cast<long>(e);
    double g[4][3], ug[3][3], FiT[3][3];
    const double jxw = tet_grads<double>(x, y, z, cn, g) / 6.0;
    tet_ugrad<double>(u, cn, g, ug);
    const double J = inv_transpose_F<double>(ug, FiT);
    double I1 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double F = ug[i][d] + (i == d ? 1.0 : 0.0);
        I1 += F * F;
      }
    const double mu = tet_param<double>(mat, param, e, cn);
    const double dmu = (0.5 * (I1 - 3.0) - log(J)) * jxw;
    psi = mu * dmu + 0.5 * mat.kappa * (J - 1.0) * (J - 1.0) * jxw;
    if (dpsi_dmu) dpsi_dmu[e] = dmu;
  }
  DVec<1> acc;
  acc.v[0] = psi;
  const DVec<1> t = block_sum<1, kTetThreads>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t.v[0];
}

__global__ __launch_bounds__(kFinBlock) void tet_energy_fin(const double* __restrict__ partials, int nparts,
                                                            double* __restrict__ out) {
  const DVec<1> t = sum_partials<1>(partials, nparts);
  if (threadIdx.x == 0) *out = t.v[0];
}

inline size_t tet_matrix_bytes(int num_cells, int value_bytes, int flags) {
  const size_t c = (value_bytes == 4 && (flags & FX_ASM_COMPUTE32)) ? 4 : 8;
  return 144 * c * static_cast<size_t>(num_cells);
}

template <typename T>
int tet_assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y, const double* z,
                 const int* cells, const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr,
                 const int* indices, const fx_material* mat, const double* param, const double* u,
                 const uint8_t* bc_mask, int symmetric, int flags, double jacobi_shift, double* dinv, T* data,
                 void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !rank || !indptr ||
      !indices || !mat || !param || !data)
    return fail(FX_EINVAL, "fx_tet4_assemble: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_tet4_assemble: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_tet4_assemble: thermal strain is built for HEX8");
  const bool nh = mat->kind == FX_MAT_NEO_HOOKEAN;
  if (nh && !u) return fail(FX_EINVAL, "fx_tet4_assemble: the Neo-Hookean tangent needs u");
  const size_t need = tet_matrix_bytes(num_cells, sizeof(T), flags);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4_assemble: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  const int diff = (flags & FX_ASM_DIFF3) != 0;
  const bool c32 = sizeof(T) == 4 && (flags & FX_ASM_COMPUTE32);
  const MatParams mp = to_params(mat);
  hipStream_t s = as_stream(stream);
  const int egrid = static_cast<int>((16L * num_cells + kTetThreads - 1) / kTetThreads);
  const int grid = (num_nodes + kNodesPerBlock - 1) / kNodesPerBlock;
#define FX_TET_LAUNCH(C, NHV)                                                                                      \
  do {                                                                                                             \
    C* ke = static_cast<C*>(workspace);                                                                            \
    hipLaunchKernelGGL((tet_elem_matrix_kernel<C, NHV>), dim3(egrid), dim3(kTetThreads), 0, s, num_cells, x, y, z, \
                       cells, mp, param, u, ke);                                                                   \
    hipLaunchKernelGGL((gather_wide_rows_kernel<T, C, 4>), dim3(grid), dim3(64 * kNodesPerBlock), 0, s, num_nodes, \
                       nullptr, inc_ptr, inc, rank, indptr, indices, ke, bc_mask, symmetric, diff, jacobi_shift,   \
                       dinv, data, nullptr);                                                                       \
  } while (0)
  if (nh) {
    if (c32) FX_TET_LAUNCH(float, true);
    else FX_TET_LAUNCH(double, true);
  } else {
    if (c32) FX_TET_LAUNCH(float, false);
    else FX_TET_LAUNCH(double, false);
  }
#undef FX_TET_LAUNCH
  FX_LAUNCHED("fx_tet4_assemble");
  return FX_OK;
}

// ---------------------------------------------------------------- scalar TET4
// vec = 1 conduction on tets: K_ab = k grad N_a . grad N_b JxW, R_a =
// k grad N_a . grad T JxW at the single Gauss point, facet flux q / 3 per
// facet vertex times the Nanson scale; k = p or its SIMP interpolation.
__device__ __forceinline__ double ts_conductivity(const MatParams& m, double p, double* dkdp) {
  if (m.kind == FX_MAT_SIMP_CONDUCTION) {
    const double span = m.E0 - m.Emin;
    if (dkdp) *dkdp = m.penal * pow(p, m.penal - 1.0) * span;
    return m.Emin + pow(p, m.penal) * span;
  }
  if (dkdp) *dkdp = 1.0;
  return p;
}

__device__ __forceinline__ void ts_grad(const double* __restrict__ f, const uint8_t* __restrict__ mask,
                                        const int* __restrict__ cn, const double (&g)[4][3], double (&G)[3]) {
  G[0] = G[1] = G[2] = 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double v = (mask && mask[cn[c]]) ? 0.0 : f[cn[c]];
    G[0] += v * g[c][0];
    G[1] += v * g[c][1];
    G[2] += v * g[c][2];
  }
}

__device__ __forceinline__ double ts_facet_nanson(const double (&g)[4][3], double det, int lf) {
  double n[3], wf;
  tet_facet(lf, n, wf);
  double m[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) m[d] = n[0] * g[1][d] + n[1] * g[2][d] + n[2] * g[3][d];
  return sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) * det * wf;
}

// One thread per (element, a, b): ke[16 e + 4 a + b].
__global__ __launch_bounds__(kTetThreads) void ts_elem_matrix_kernel(int num_cells, const double* __restrict__ x,
                                                                     const double* __restrict__ y,
                                                                     const double* __restrict__ z,
                                                                     const int* __restrict__ cells, MatParams mat,
                                                                     const double* __restrict__ param,
                                                                     double* __restrict__ ke) {
  const long t = static_cast<long>(blockIdx.x) * kTetThreads + threadIdx.x;
  const long e = t >> 4;
  if (e >= num_cells) return;
  const int a = static_cast<int>((t >> 2) & 3), b = static_cast<int>(t & 3);
  const int* cn = cells + 4 * e;
  double g[4][3];
  const double w = tet_grads<double>(x, y, z, cn, g) / 6.0;
  const double k = ts_conductivity(mat, tet_param<double>(mat, param, static_cast<int>(e), cn), nullptr);
  ke[16 * e + 4 * a + b] = k * w * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
}

// One thread per element: re[4 e + a].
__global__ __launch_bounds__(kTetThreads) void ts_elem_residual_kernel(int num_cells, const double* __restrict__ x,
                                                                       const double* __restrict__ y,
                                                                       const double* __restrict__ z,
                                                                       const int* __restrict__ cells, MatParams mat,
                                                                       const double* __restrict__ param,
                                                                       const double* __restrict__ u,
                                                                       double* __restrict__ re) {
  const int e = blockIdx.x * kTetThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double g[4][3], G[3];
  const double w = tet_grads<double>(x, y, z, cn, g) / 6.0;
  ts_grad(u, nullptr, cn, g, G);
  const double kw = ts_conductivity(mat, tet_param<double>(mat, param, e, cn), nullptr) * w;
#pragma unroll
  for (int a = 0; a < 4; ++a) re[4 * static_cast<long>(e) + a] = kw * (G[0] * g[a][0] + G[1] * g[a][1] + G[2] * g[a][2]);
}

// One thread per node: element residuals in ascending order, then the
// loaded facets: flux = fdir[f] * svar[f] at the facet centroid.
__global__ __launch_bounds__(256) void ts_node_residual_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
    const double* __restrict__ re, const int* __restrict__ faces, const int* __restrict__ finc_ptr,
    const int* __restrict__ finc, const double* __restrict__ svar, const double* __restrict__ fdir, int negate,
    double* __restrict__ res) {
  const int nd = blockIdx.x * 256 + threadIdx.x;
  if (nd >= num_nodes) return;
  double r = 0.0;
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    r += re[4 * static_cast<long>(pk >> 3) + (pk & 7)];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int f = finc[k] >> 3;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      double g[4][3];
      const double det = tet_grads<double>(x, y, z, cells + 4 * static_cast<long>(e), g);
      r += fdir[f] * (svar[f] * (1.0 / 3.0) * ts_facet_nanson(g, det, lf));
    }
  }
  res[nd] = negate ? -r : r;
}

// w_e = dk/dp grad lambda . grad T JxW (SELF: lambda = T); cell variables:
// scale * w_e; node variables keep we[e] (tet_param_vjp_node_kernel).
template <bool SELF>
__global__ __launch_bounds__(kTetThreads) void ts_param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    const double* __restrict__ lam, const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell,
    double* __restrict__ we) {
  const int e = blockIdx.x * kTetThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double g[4][3], G[3], L[3];
  const double w = tet_grads<double>(x, y, z, cn, g) / 6.0;
  ts_grad(u, nullptr, cn, g, G);
  if constexpr (SELF) {
    L[0] = G[0];
    L[1] = G[1];
    L[2] = G[2];
  } else {
    ts_grad(lam, bc_mask, cn, g, L);
  }
  double dkdp;
  ts_conductivity(mat, tet_param<double>(mat, param, e, cn), &dkdp);
  const double v = dkdp * w * (G[0] * L[0] + G[1] * L[1] + G[2] * L[2]);
  if (out_cell) {
    out_cell[e] = scale * v;
  } else {
    we[e] = v;
  }
}

// grad_svar[f] = scale * fdir[f] * nanson * sum over the facet's vertices of lambda_a / 3.
__global__ __launch_bounds__(256) void ts_flux_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                          const double* __restrict__ y, const double* __restrict__ z,
                                                          const int* __restrict__ cells,
                                                          const int* __restrict__ faces,
                                                          const double* __restrict__ fdir,
                                                          const double* __restrict__ lam,
                                                          const uint8_t* __restrict__ bc_mask, double scale,
                                                          double* __restrict__ out) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= num_faces) return;
  const int e = faces[2 * f], lf = faces[2 * f + 1];
  const int* cn = cells + 4 * static_cast<long>(e);
  double g[4][3];
  const double det = tet_grads<double>(x, y, z, cn, g);
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (a == lf) continue;
    s += (1.0 / 3.0) * ((bc_mask && bc_mask[cn[a]]) ? 0.0 : lam[cn[a]]);
  }
  out[f] = scale * fdir[f] * s * ts_facet_nanson(g, det, lf);
}

inline bool ts_law(const fx_material* m) {
  return (m->kind == FX_MAT_CONDUCTION || m->kind == FX_MAT_SIMP_CONDUCTION) && !m->temperature;
}

template <typename T>
int ts_assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y, const double* z,
                const int* cells, const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr,
                const int* indices, const fx_material* mat, const double* param, const uint8_t* bc_mask,
                int symmetric, int flags, double jacobi_shift, double* dinv, T* data, void* workspace,
                size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !rank || !indptr ||
      !indices || !mat || !param || !data)
    return fail(FX_EINVAL, "fx_tet4s_assemble: bad arguments");
  if (!ts_law(mat)) return fail(FX_EINVAL, "fx_tet4s_assemble: needs a conduction law");
  if (flags & (FX_ASM_DIFF3 | FX_ASM_COMPUTE32))
    return fail(FX_EINVAL, "fx_tet4s_assemble: the difference form and FP32 arithmetic are built for vec = 3");
  const size_t need = 16 * sizeof(double) * static_cast<size_t>(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4s_assemble: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  double* ke = static_cast<double*>(workspace);
  hipLaunchKernelGGL(ts_elem_matrix_kernel, dim3(static_cast<int>((16L * num_cells + kTetThreads - 1) / kTetThreads)),
                     dim3(kTetThreads), 0, s, num_cells, x, y, z, cells, to_params(mat), param, ke);
  hipLaunchKernelGGL((gather_wide_rows_kernel<T, double, 4, 1>),
                     dim3((num_nodes + kNodesPerBlock - 1) / kNodesPerBlock),
                     dim3(64 * kNodesPerBlock), 0, s, num_nodes, nullptr, inc_ptr, inc, rank, indptr, indices, ke,
                     bc_mask, symmetric, 0, jacobi_shift, dinv, data, nullptr);
  FX_LAUNCHED("fx_tet4s_assemble");
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_tet4_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  return num_cells < 0 ? 0 : tet_matrix_bytes(num_cells, value_bytes, flags);
}

int fx_tet4_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                         const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                         int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                         void* workspace, size_t ws_bytes) {
  return tet_assemble<double>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr,
                              indices, mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace,
                              ws_bytes);
}

int fx_tet4_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                         const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                         int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                         void* workspace, size_t ws_bytes) {
  return tet_assemble<float>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                             mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace,
                             ws_bytes);
}

size_t fx_tet4_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 12 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_tet4_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                         const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                         const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                         int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !u ||
      !res)
    return fail(FX_EINVAL, "fx_tet4_residual_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_tet4_residual_f64: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_tet4_residual_f64: thermal strain is built for HEX8");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_tet4_residual_f64: face arrays missing");
  const size_t need = fx_tet4_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool has_faces = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(tet_elem_residual_kernel, dim3((num_cells + kTetThreads - 1) / kTetThreads), dim3(kTetThreads),
                     0, s, num_cells, x, y, z, cells, to_params(mat), param, u, re);
  hipLaunchKernelGGL(tet_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, z,
                     cells, inc_ptr, inc, re, has_faces ? faces : nullptr, has_faces ? finc_ptr : nullptr,
                     has_faces ? finc : nullptr, svar, fdir, negate, res);
  FX_LAUNCHED("fx_tet4_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_tet4_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

size_t fx_tet4_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? static_cast<size_t>(num_cells) * sizeof(double) : 0;
}

int fx_tet4_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, const double* lam,
                          const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_tet4_param_vjp_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_tet4_param_vjp_f64: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_tet4_param_vjp_f64: thermal strain is built for HEX8");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_tet4_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  const dim3 eg((num_cells + kTetThreads - 1) / kTetThreads);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(tet_param_vjp_kernel<false>, eg, dim3(kTetThreads), 0, s, num_cells, x, y, z, cells, mp,
                       param, u, lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("tet_param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_tet4_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_tet4_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* we = static_cast<double*>(workspace);
  hipLaunchKernelGGL(tet_param_vjp_kernel<false>, eg, dim3(kTetThreads), 0, s, num_cells, x, y, z, cells, mp, param,
                     u, lam, bc_mask, 1.0, nullptr, we);
  hipLaunchKernelGGL(tet_param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc, we, scale, grad);
  FX_LAUNCHED("tet_param_vjp_node_kernel");
  return FX_OK;
}

int fx_tet4_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                     const double* z, const int32_t* cells, const fx_material* mat,
                                     const double* param, const double* u, double* dc) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_tet4_simp_compliance_sens_f64: bad arguments");
  if (mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_LINEAR_ELASTIC)
    return fail(FX_EINVAL, "fx_tet4_simp_compliance_sens_f64: needs a linear law (SIMP or linear elastic)");
  if (mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_tet4_simp_compliance_sens_f64: needs a cell variable");
  hipLaunchKernelGGL(tet_param_vjp_kernel<true>, dim3((num_cells + kTetThreads - 1) / kTetThreads),
                     dim3(kTetThreads), 0, as_stream(stream), num_cells, x, y, z, cells, to_params(mat), param, u,
                     nullptr, nullptr, -1.0, dc, nullptr);
  FX_LAUNCHED("tet_param_vjp_kernel");
  return FX_OK;
}

int fx_tet4_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                             const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                             const uint8_t* bc_mask, double scale, double* grad_svar) {
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !z || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_tet4_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  hipLaunchKernelGGL(tet_traction_vjp_kernel, dim3((num_faces + 255) / 256), dim3(256), 0, as_stream(stream),
                     num_faces, x, y, z, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("tet_traction_vjp_kernel");
  return FX_OK;
}

size_t fx_tet4_energy_workspace_bytes(int32_t num_cells) {
  if (num_cells < 0) return 0;
  return static_cast<size_t>((num_cells + kTetThreads - 1) / kTetThreads) * sizeof(double);
}

int fx_tet4_energy_f64(void* stream, int32_t num_cells, const double* x, const double* y, const double* z,
                       const int32_t* cells, const fx_material* mat, const double* param, const double* u,
                       double* energy, double* dpsi_dmu, void* workspace, size_t ws_bytes) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !energy)
    return fail(FX_EINVAL, "fx_tet4_energy_f64: bad arguments");
  if (mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_tet4_energy_f64: strain energy is built for the Neo-Hookean law");
  if (dpsi_dmu && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_tet4_energy_f64: dpsi/dmu per element needs a cell variable");
  const size_t need = fx_tet4_energy_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4_energy_f64: workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = as_stream(stream);
  const int grid = (num_cells + kTetThreads - 1) / kTetThreads;
  double* partials = static_cast<double*>(workspace);
  hipLaunchKernelGGL(tet_energy_kernel, dim3(grid), dim3(kTetThreads), 0, s, num_cells, x, y, z, cells,
                     to_params(mat), param, u, dpsi_dmu, partials);
  hipLaunchKernelGGL(tet_energy_fin, dim3(1), dim3(kFinBlock), 0, s, partials, grid, energy);
  FX_LAUNCHED("tet_energy_kernel");
  return FX_OK;
}

size_t fx_tet4s_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  (void)value_bytes;
  (void)flags;
  return num_cells < 0 ? 0 : 16 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_tet4s_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes) {
  (void)u;
  return ts_assemble<double>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                             mat, param, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes);
}

int fx_tet4s_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes) {
  (void)u;
  return ts_assemble<float>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                            mat, param, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes);
}

size_t fx_tet4s_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 4 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_tet4s_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !u ||
      !res)
    return fail(FX_EINVAL, "fx_tet4s_residual_f64: bad arguments");
  if (!ts_law(mat)) return fail(FX_EINVAL, "fx_tet4s_residual_f64: needs a conduction law");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_tet4s_residual_f64: face arrays missing");
  const size_t need = fx_tet4s_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4s_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool hf = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(ts_elem_residual_kernel, dim3((num_cells + kTetThreads - 1) / kTetThreads), dim3(kTetThreads),
                     0, s, num_cells, x, y, z, cells, to_params(mat), param, u, re);
  hipLaunchKernelGGL(ts_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, z,
                     cells, inc_ptr, inc, re, hf ? faces : nullptr, hf ? finc_ptr : nullptr, hf ? finc : nullptr, svar,
                     fdir, negate, res);
  FX_LAUNCHED("fx_tet4s_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_tet4s_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

size_t fx_tet4s_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? static_cast<size_t>(num_cells) * sizeof(double) : 0;
}

int fx_tet4s_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_tet4s_param_vjp_f64: bad arguments");
  if (!ts_law(mat)) return fail(FX_EINVAL, "fx_tet4s_param_vjp_f64: needs a conduction law");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_tet4s_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  const dim3 eg((num_cells + kTetThreads - 1) / kTetThreads);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(ts_param_vjp_kernel<false>, eg, dim3(kTetThreads), 0, s, num_cells, x, y, z, cells, mp, param,
                       u, lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("ts_param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_tet4s_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_tet4s_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_tet4s_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* we = static_cast<double*>(workspace);
  hipLaunchKernelGGL(ts_param_vjp_kernel<false>, eg, dim3(kTetThreads), 0, s, num_cells, x, y, z, cells, mp, param, u,
                     lam, bc_mask, 1.0, nullptr, we);
  hipLaunchKernelGGL(tet_param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc, we, scale, grad);
  FX_LAUNCHED("tet_param_vjp_node_kernel");
  return FX_OK;
}

int fx_tet4s_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_tet4s_simp_compliance_sens_f64: bad arguments");
  if (!ts_law(mat) || mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_tet4s_simp_compliance_sens_f64: needs a conduction law with a cell variable");
  hipLaunchKernelGGL(ts_param_vjp_kernel<true>, dim3((num_cells + kTetThreads - 1) / kTetThreads), dim3(kTetThreads),
                     0, as_stream(stream), num_cells, x, y, z, cells, to_params(mat), param, u, nullptr, nullptr, -1.0,
                     dc, nullptr);
  FX_LAUNCHED("ts_param_vjp_kernel");
  return FX_OK;
}

int fx_tet4s_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar) {
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !z || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_tet4s_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  hipLaunchKernelGGL(ts_flux_vjp_kernel, dim3((num_faces + 255) / 256), dim3(256), 0, as_stream(stream), num_faces, x,
                     y, z, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("ts_flux_vjp_kernel");
  return FX_OK;
}

}  // extern "C"
