// Scalar HEX8 kernels for gfx950 (vec = 1): steady conduction
// div(k grad T) = 0 -- CSR Jacobian assembly, residual with a surface flux,
// and the adjoint pullbacks.
//
// feax writes a scalar problem as a Problem with vec = 1 whose tensor map is
// the flux, e.g. ``lambda grad_T, k: k * grad_T`` (feax's
// examples/basic/thermo_mech_coupling.py:64-69, tests/test_params_bridge.py:
// 131-133); the element kernels are those of any tensor map
// (feax/assembler.py:406-448, 808-818) with one component: R_a = sum_q
// k_q grad N_a . grad T_q JxW_q, K_ab = sum_q k_q grad N_a . grad N_b JxW_q.
// A surface map with one component is a prescribed flux, added with a +
// sign like a traction (feax/assembler.py:494-535).  HEX8 geometry and the
// 2x2x2 / 2x2 Gauss rules are those of hex8_common.h.
//
// k is the parameter (FX_MAT_CONDUCTION) or its SIMP interpolation
// k = kmin + rho^p (k0 - kmin) (FX_MAT_SIMP_CONDUCTION, the E0 / Emin /
// penal fields), a node or a cell variable.  Two passes as for the other
// elements: one thread per element pair (a, b) sums its entry over the
// Gauss points (64 values per element), one wave per node gathers
// (cell_common.h with NPE = 8, V = 1).
#include <cstdint>

#include "cell_common.h"
#include "common.h"
#include "feaxhip.h"
#include "hex8_common.h"

namespace fx {
namespace {

constexpr int kSThreads = 256;

__device__ __forceinline__ double conductivity(const MatParams& m, double p, double* dkdp) {
  if (m.kind == FX_MAT_SIMP_CONDUCTION) {
    const double span = m.E0 - m.Emin;
    if (dkdp) *dkdp = m.penal * pow(p, m.penal - 1.0) * span;
    return m.Emin + pow(p, m.penal) * span;
  }
  if (dkdp) *dkdp = 1.0;
  return p;
}

// Physical shape-function gradients g[a] and JxW at Gauss point q.
__device__ __forceinline__ double grads_at(const double (&X)[8][3], int q, double& x0, double& x1, double& x2,
                                           double (&g)[8][3]) {
  gauss_point(q, x0, x1, x2);
  double Ji[3][3];
  const double det = jac_inv<double>(X, x0, x1, x2, Ji);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    double N, d0, d1, d2;
    shape<double>(a, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, g[a]);
  }
  return det * 0.125;  // Gauss weights 1/8 on [0,1]^3
}

// grad f at the Gauss point for a nodal scalar f (read as 0 on masked rows).
__device__ __forceinline__ void scalar_grad(const double* __restrict__ f, const uint8_t* __restrict__ mask,
                                            const int* __restrict__ cn, const double (&g)[8][3], double (&G)[3]) {
  G[0] = G[1] = G[2] = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int r = cn[c];
    const double v = (mask && mask[r]) ? 0.0 : f[r];
    G[0] += v * g[c][0];
    G[1] += v * g[c][1];
    G[2] += v * g[c][2];
  }
}

// One thread per (element, a, b): ke[64 e + 8 a + b], summed over the
// Gauss points in the order ((q0 + q1) + (q2 + q3)) + ((q4 + q5) + (q6 + q7)).
template <typename C>
__global__ __launch_bounds__(kSThreads) void s_elem_matrix_kernel(int num_cells, const double* __restrict__ x,
                                                                  const double* __restrict__ y,
                                                                  const double* __restrict__ z,
                                                                  const int* __restrict__ cells, MatParams mat,
                                                                  const double* __restrict__ param,
                                                                  C* __restrict__ ke) {
  const long t = static_cast<long>(blockIdx.x) * kSThreads + threadIdx.x;
  const long e = t >> 6;
  if (e >= num_cells) return;
  const int a = static_cast<int>((t >> 3) & 7), b = static_cast<int>(t & 7);
  const int* cn = cells + 8 * e;
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double kq[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    double x0, x1, x2, g[8][3];
    const double w = grads_at(X, q, x0, x1, x2, g);
    const double k = conductivity(mat, param_at<double>(mat, param, static_cast<int>(e), cn, x0, x1, x2), nullptr);
    kq[q] = k * w * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
  }
  ke[64 * e + 8 * a + b] = static_cast<C>(((kq[0] + kq[1]) + (kq[2] + kq[3])) + ((kq[4] + kq[5]) + (kq[6] + kq[7])));
}

// One thread per element: re[8 e + a].
__global__ __launch_bounds__(kSThreads) void s_elem_residual_kernel(int num_cells, const double* __restrict__ x,
                                                                    const double* __restrict__ y,
                                                                    const double* __restrict__ z,
                                                                    const int* __restrict__ cells, MatParams mat,
                                                                    const double* __restrict__ param,
                                                                    const double* __restrict__ u,
                                                                    double* __restrict__ re) {
  const int e = blockIdx.x * kSThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 8 * static_cast<long>(e);
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double rq[8][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    double x0, x1, x2, g[8][3], G[3];
    const double w = grads_at(X, q, x0, x1, x2, g);
    scalar_grad(u, nullptr, cn, g, G);
    const double kw = conductivity(mat, param_at<double>(mat, param, e, cn, x0, x1, x2), nullptr) * w;
#pragma unroll
    for (int a = 0; a < 8; ++a) rq[q][a] = kw * (G[0] * g[a][0] + G[1] * g[a][1] + G[2] * g[a][2]);
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
    re[8 * static_cast<long>(e) + a] =
        ((rq[0][a] + rq[1][a]) + (rq[2][a] + rq[3][a])) + ((rq[4][a] + rq[5][a]) + (rq[6][a] + rq[7][a]));
}

// One thread per node: element residuals in ascending element order, then
// the loaded faces: flux(f, q) = fdir[f] * svar[4 f + q].
__global__ __launch_bounds__(256) void s_node_residual_kernel(
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
    r += re[8 * static_cast<long>(pk >> 3) + (pk & 7)];
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
      r += fdir[f] * ((vq[0] + vq[1]) + (vq[2] + vq[3]));
    }
  }
  res[nd] = negate ? -r : r;
}

// Pullbacks per (element, Gauss point): w_eq = dR/dp . lambda = dk/dp
// grad lambda . grad T JxW (SELF: lambda = T).  Cell variables: scale *
// sum_q w_eq; node variables keep w_eq (we[8 e + q]).
template <bool SELF>
__global__ __launch_bounds__(kSThreads) void s_param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    const double* __restrict__ lam, const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell,
    double* __restrict__ we) {
  const int e = blockIdx.x * kSThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 8 * static_cast<long>(e);
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double wq[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    double x0, x1, x2, g[8][3], G[3], L[3];
    const double w = grads_at(X, q, x0, x1, x2, g);
    scalar_grad(u, nullptr, cn, g, G);
    if constexpr (SELF) {
      L[0] = G[0];
      L[1] = G[1];
      L[2] = G[2];
    } else {
      scalar_grad(lam, bc_mask, cn, g, L);
    }
    double dkdp;
    conductivity(mat, param_at<double>(mat, param, e, cn, x0, x1, x2), &dkdp);
    wq[q] = dkdp * w * (G[0] * L[0] + G[1] * L[1] + G[2] * L[2]);
  }
  if (out_cell) {
    out_cell[e] = scale * (((wq[0] + wq[1]) + (wq[2] + wq[3])) + ((wq[4] + wq[5]) + (wq[6] + wq[7])));
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) we[8 * static_cast<long>(e) + q] = wq[q];
  }
}

__global__ __launch_bounds__(256) void s_param_vjp_node_kernel(int num_nodes, const int* __restrict__ inc_ptr,
                                                               const int* __restrict__ inc,
                                                               const double* __restrict__ we, double scale,
                                                               double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num_nodes) return;
  double acc = 0.0;
  for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
    const int pk = inc[k];
    const long e = pk >> 3;
    const int a = pk & 7;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double x0, x1, x2, N, d0, d1, d2;
      gauss_point(q, x0, x1, x2);
      shape<double>(a, x0, x1, x2, N, d0, d1, d2);
      acc += N * we[8 * e + q];
    }
  }
  out[i] = scale * acc;
}

// grad_svar[4 f + q] = scale * nanson_q * fdir[f] * sum over the face's
// vertices of N_a(q) lambda_a.
__global__ __launch_bounds__(256) void s_flux_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double* __restrict__ z,
                                                         const int* __restrict__ cells,
                                                         const int* __restrict__ faces,
                                                         const double* __restrict__ fdir,
                                                         const double* __restrict__ lam,
                                                         const uint8_t* __restrict__ bc_mask, double scale,
                                                         double* __restrict__ out) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long f = t >> 2;
  const int qq = static_cast<int>(t & 3);
  if (f >= num_faces) return;
  const int e = faces[2 * f], lf = faces[2 * f + 1];
  const int* cn = cells + 8 * static_cast<long>(e);
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double p0, p1, p2;
  face_point(lf, qq, p0, p1, p2);
  const double nanson = face_scale(X, lf, p0, p1, p2);
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    double N, d0, d1, d2;
    shape<double>(a, p0, p1, p2, N, d0, d1, d2);
    const int r = cn[a];
    acc += N * ((bc_mask && bc_mask[r]) ? 0.0 : lam[r]);
  }
  out[4 * f + qq] = scale * fdir[f] * acc * nanson;
}

inline bool conduction_law(const fx_material* m) {
  return m->kind == FX_MAT_CONDUCTION || m->kind == FX_MAT_SIMP_CONDUCTION;
}

template <typename T>
int s_assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y, const double* z,
               const int* cells, const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr,
               const int* indices, const fx_material* mat, const double* param, const uint8_t* bc_mask,
               int symmetric, int flags, double jacobi_shift, double* dinv, T* data, void* workspace,
               size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !rank || !indptr ||
      !indices || !mat || !param || !data)
    return fail(FX_EINVAL, "fx_hex8s_assemble: bad arguments");
  if (!conduction_law(mat)) return fail(FX_EINVAL, "fx_hex8s_assemble: needs a conduction law (got %d)", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_hex8s_assemble: thermal strain is for the elastic laws");
  if (flags & (FX_ASM_DIFF3 | FX_ASM_COMPUTE32))
    return fail(FX_EINVAL, "fx_hex8s_assemble: the difference form and FP32 arithmetic are built for vec = 3");
  const size_t need = 64 * sizeof(double) * static_cast<size_t>(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8s_assemble: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  double* ke = static_cast<double*>(workspace);
  const int egrid = static_cast<int>((64L * num_cells + kSThreads - 1) / kSThreads);
  hipLaunchKernelGGL((s_elem_matrix_kernel<double>), dim3(egrid), dim3(kSThreads), 0, s, num_cells, x, y, z, cells,
                     to_params(mat), param, ke);
  const int grid = (num_nodes + kNodesPerBlock - 1) / kNodesPerBlock;
  hipLaunchKernelGGL((gather_wide_rows_kernel<T, double, 8, 1>), dim3(grid), dim3(64 * kNodesPerBlock), 0, s,
                     num_nodes, nullptr, inc_ptr, inc, rank, indptr, indices, ke, bc_mask, symmetric, 0, jacobi_shift,
                     dinv, data, nullptr);
  FX_LAUNCHED("fx_hex8s_assemble");
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_hex8s_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  (void)value_bytes;
  (void)flags;
  return num_cells < 0 ? 0 : 64 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex8s_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes) {
  (void)u;  // linear law: the Jacobian does not depend on the state
  return s_assemble<double>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                            mat, param, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes);
}

int fx_hex8s_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes) {
  (void)u;
  return s_assemble<float>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                           mat, param, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes);
}

size_t fx_hex8s_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 8 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex8s_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !u ||
      !res)
    return fail(FX_EINVAL, "fx_hex8s_residual_f64: bad arguments");
  if (!conduction_law(mat)) return fail(FX_EINVAL, "fx_hex8s_residual_f64: needs a conduction law (got %d)", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_hex8s_residual_f64: thermal strain is for the elastic laws");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_hex8s_residual_f64: face arrays missing");
  const size_t need = fx_hex8s_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8s_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool has_faces = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(s_elem_residual_kernel, dim3((num_cells + kSThreads - 1) / kSThreads), dim3(kSThreads), 0, s,
                     num_cells, x, y, z, cells, to_params(mat), param, u, re);
  hipLaunchKernelGGL(s_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, z,
                     cells, inc_ptr, inc, re, has_faces ? faces : nullptr, has_faces ? finc_ptr : nullptr,
                     has_faces ? finc : nullptr, svar, fdir, negate, res);
  FX_LAUNCHED("fx_hex8s_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_hex8s_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

size_t fx_hex8s_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? static_cast<size_t>(num_cells) * 8 * sizeof(double) : 0;
}

int fx_hex8s_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_hex8s_param_vjp_f64: bad arguments");
  if (!conduction_law(mat))
    return fail(FX_EINVAL, "fx_hex8s_param_vjp_f64: needs a conduction law (got %d)", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_hex8s_param_vjp_f64: thermal strain is for the elastic laws");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_hex8s_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  const dim3 eg((num_cells + kSThreads - 1) / kSThreads);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(s_param_vjp_kernel<false>, eg, dim3(kSThreads), 0, s, num_cells, x, y, z, cells, mp, param, u,
                       lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("s_param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_hex8s_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_hex8s_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8s_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* we = static_cast<double*>(workspace);
  hipLaunchKernelGGL(s_param_vjp_kernel<false>, eg, dim3(kSThreads), 0, s, num_cells, x, y, z, cells, mp, param, u,
                     lam, bc_mask, 1.0, nullptr, we);
  hipLaunchKernelGGL(s_param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc, we, scale, grad);
  FX_LAUNCHED("s_param_vjp_node_kernel");
  return FX_OK;
}

int fx_hex8s_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_hex8s_simp_compliance_sens_f64: bad arguments");
  if (!conduction_law(mat))
    return fail(FX_EINVAL, "fx_hex8s_simp_compliance_sens_f64: needs a conduction law (got %d)", mat->kind);
  if (mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_hex8s_simp_compliance_sens_f64: needs a cell variable");
  hipLaunchKernelGGL(s_param_vjp_kernel<true>, dim3((num_cells + kSThreads - 1) / kSThreads), dim3(kSThreads), 0,
                     as_stream(stream), num_cells, x, y, z, cells, to_params(mat), param, u, nullptr, nullptr, -1.0,
                     dc, nullptr);
  FX_LAUNCHED("s_param_vjp_kernel");
  return FX_OK;
}

int fx_hex8s_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar) {
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !z || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_hex8s_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  const long threads = 4L * num_faces;
  hipLaunchKernelGGL(s_flux_vjp_kernel, dim3(static_cast<int>((threads + 255) / 256)), dim3(256), 0,
                     as_stream(stream), num_faces, x, y, z, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("s_flux_vjp_kernel");
  return FX_OK;
}

}  // extern "C"
