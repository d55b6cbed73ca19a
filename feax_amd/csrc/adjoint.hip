// Reverse-mode (adjoint) kernels of the HEX8 residual for gfx950.
//
// feax differentiates its solve with jax.custom_vjp: the backward pass
// solves J_bc^T lambda = v and pulls lambda back through the residual with
// jax.vjp of R_bc with respect to the traced parameters
// (feax/solvers/linear.py:426-481).  The pullbacks here are written by hand:
//
//   param_vjp      g_p = sum_e sum_q (dR_e/dp)^T lambda_e at Gauss point q
//                  for the volume variable p (cell or node), where, with
//                  grad u and grad lambda at the Gauss point and JxW,
//                    linear / SIMP:  dR/dE . lambda = sigma_1(grad u) : grad lambda JxW
//                                    (sigma_1 = stress at unit modulus), times dE/dp;
//                    Neo-Hookean:    dR/dmu . lambda = (F - F^-T) : grad lambda JxW.
//                  Node variables enter through the shape functions
//                  (p_q = sum_c N_c(q) p_c), so a node gets
//                  sum_{e, q} N_a(q) w_eq over its incident elements.
//   compliance     the self-adjoint special case lambda = u of a linear law
//                  (C = f.u, dC/dp = -u^T dK/dp u, feax/solvers/cmg.py:312-323).
//   traction_vjp   dR/d t(f, q) . lambda = sum_a N_a(q) (dir . lambda_a) nanson_q
//                  for the dead traction t = dir * svar (+ sign,
//                  feax/assembler.py:494-535).
//   energy         Pi = sum_e sum_q psi(grad u) JxW and its partial
//                  dPi/dmu_e (feax/assembler.py:1866-1961).
//
// Layout: 8 consecutive lanes own one element, one lane per Gauss point;
// each lane recomputes the geometry at its point from the SoA coordinates
// (nothing per element is stored) and the 8 lanes reduce with a fixed
// butterfly.  Node sums run over the node->element incidence in ascending
// element order.  No atomics: bitwise repeatable.
#include <cstdint>

#include "hex8_common.h"
#include "reduce.h"

namespace fx {

constexpr int kElemBlock = 256;  // 32 elements x 8 Gauss points

__device__ __forceinline__ double sum8(double v) {
  v = xor_sum(v, 1);
  v = xor_sum(v, 2);
  v = xor_sum(v, 4);
  return v;
}

// Gradients at Gauss point (x0, x1, x2) of element cn of the nodal fields u
// (and lambda unless SELF; lambda is read as 0 on rows with bc_mask set):
// G[i][d] = d f_i / d x_d.  Returns JxW.
template <bool SELF>
__device__ __forceinline__ double field_grads(const double (&X)[8][3], const int* __restrict__ cn, double x0,
                                              double x1, double x2, const double* __restrict__ u,
                                              const double* __restrict__ lam, const uint8_t* __restrict__ bc_mask,
                                              double (&ug)[3][3], double (&lg)[3][3]) {
  double Ji[3][3];
  const double det = jac_inv<double>(X, x0, x1, x2, Ji);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) ug[i][d] = lg[i][d] = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double N, d0, d1, d2, g[3];
    shape<double>(c, x0, x1, x2, N, d0, d1, d2);
    phys_grad<double>(Ji, d0, d1, d2, g);
    const int nd = cn[c];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double uv = u[3 * nd + i];
#pragma unroll
      for (int d = 0; d < 3; ++d) ug[i][d] += uv * g[d];
      if constexpr (!SELF) {
        const double lv = (bc_mask && bc_mask[3 * nd + i]) ? 0.0 : lam[3 * nd + i];
#pragma unroll
        for (int d = 0; d < 3; ++d) lg[i][d] += lv * g[d];
      }
    }
  }
  if constexpr (SELF) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) lg[i][d] = ug[i][d];
  }
  return det * 0.125;  // Gauss weights 1/8 on [0,1]^3
}

// F^-T of F = I + ug; returns det F.
__device__ __forceinline__ double inv_transpose(const double (&ug)[3][3], double (&FiT)[3][3]) {
  double F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) F[i][d] = ug[i][d] + (i == d ? 1.0 : 0.0);
  const double c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
  const double c01 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
  const double c02 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
  const double J = F[0][0] * c00 + F[0][1] * c01 + F[0][2] * c02;
  const double id = 1.0 / J;
  // (F^-1)^T = cofactor(F) / det F
  FiT[0][0] = c00 * id;
  FiT[0][1] = c01 * id;
  FiT[0][2] = c02 * id;
  FiT[1][0] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * id;
  FiT[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * id;
  FiT[1][2] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * id;
  FiT[2][0] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * id;
  FiT[2][1] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * id;
  FiT[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * id;
  return J;
}

// w_eq = (dR_e/dp_q . lambda_e) at Gauss point q (before the shape-function
// weights of node variables).
template <bool SELF>
__global__ __launch_bounds__(kElemBlock) void param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    const double* __restrict__ lam, const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell,
    double* __restrict__ wq) {
  const long t = static_cast<long>(blockIdx.x) * kElemBlock + threadIdx.x;
  const long e = t >> 3;
  const int q = static_cast<int>(t & 7);
  const bool live = e < num_cells;
  double w = 0.0;
  if (live) {
    const int* cn = cells + 8 * e;
    double X[8][3];
    load_coords<double>(x, y, z, cn, X);
    double x0, x1, x2;
    gauss_point(q, x0, x1, x2);
    double ug[3][3], lg[3][3];
    const double jxw = field_grads<SELF>(X, cn, x0, x1, x2, u, lam, bc_mask, ug, lg);
    double dens;
    if (mat.kind == FX_MAT_NEO_HOOKEAN) {
      double FiT[3][3];
      inv_transpose(ug, FiT);
      dens = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) dens += (ug[i][d] + (i == d ? 1.0 : 0.0) - FiT[i][d]) * lg[i][d];
      // mu = p: dmu/dp = 1
    } else {
      const double nu = mat.nu;
      const double lam0 = nu / ((1 + nu) * (1 - 2 * nu)), mu0 = 1 / (2 * (1 + nu));
      const double tr = ug[0][0] + ug[1][1] + ug[2][2];
      dens = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          dens += (mu0 * (ug[i][d] + ug[d][i]) + (i == d ? lam0 * tr : 0.0)) * lg[i][d];
      const double pq = param_at<double>(mat, param, static_cast<int>(e), cn, x0, x1, x2);
      if (mat.temp) {
        // thermal stress -E per_E I: d/dE contributes -per_E tr(grad lambda)
        double per_E;
        thermal_shift<double>(mat, pq, cn, x0, x1, x2, &per_E);
        dens -= per_E * (lg[0][0] + lg[1][1] + lg[2][2]);
      }
      double dEdp;
      modulus_of<double>(mat, pq, &dEdp);
      dens *= dEdp;
    }
    w = dens * jxw;
  }
  if (out_cell) {
    w = sum8(w);  // every lane of the element group participates
    if (live && q == 0) out_cell[e] = scale * w;
  } else if (live) {
    wq[8 * e + q] = w;
  }
}

// Node variables: g_a = scale * sum_{e in inc(a)} sum_q N_a(q) w_eq.
// Pullback onto the temperature node variable of the thermal strain, per
// (element, Gauss point): the thermal part of R_ai is -kth d_i N_a with
// kth = E c3 alpha (T_q - t_ref), so lambda . dR/dT_q = -E c3 alpha
// div(lambda) JxW (T_q = sum_b N_b(q) T_b: the node kernel spreads it).
__global__ __launch_bounds__(kElemBlock) void thermal_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ lam,
    const uint8_t* __restrict__ bc_mask, double* __restrict__ wq) {
  const long t = static_cast<long>(blockIdx.x) * kElemBlock + threadIdx.x;
  const long e = t >> 3;
  const int q = static_cast<int>(t & 7);
  if (e >= num_cells) return;
  const int* cn = cells + 8 * e;
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double x0, x1, x2;
  gauss_point(q, x0, x1, x2);
  double ug[3][3], lg[3][3];
  const double jxw = field_grads<false>(X, cn, x0, x1, x2, lam, lam, bc_mask, ug, lg);
  const double E = modulus_of<double>(mat, param_at<double>(mat, param, static_cast<int>(e), cn, x0, x1, x2));
  const double nu = mat.nu;
  const double c3 = 3.0 * nu / ((1 + nu) * (1 - 2 * nu)) + 1.0 / (1 + nu);
  wq[8 * e + q] = -E * c3 * mat.alpha * (lg[0][0] + lg[1][1] + lg[2][2]) * jxw;
}

__global__ __launch_bounds__(256) void param_vjp_node_kernel(int num_nodes, const int* __restrict__ inc_ptr,
                                                             const int* __restrict__ inc,
                                                             const double* __restrict__ wq, double scale,
                                                             double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num_nodes) return;
  double s = 0.0;
  for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
    const int pk = inc[k];
    const long e = pk >> 3;
    const int a = pk & 7;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double x0, x1, x2, N, d0, d1, d2;
      gauss_point(q, x0, x1, x2);
      shape<double>(a, x0, x1, x2, N, d0, d1, d2);
      s += N * wq[8 * e + q];
    }
  }
  out[i] = scale * s;
}

// grad_svar[f, q] = scale * nanson_q * sum_a N_a(q) (dir_f . lambda_a).
__global__ __launch_bounds__(256) void traction_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ z,
                                                           const int* __restrict__ cells,
                                                           const int* __restrict__ faces,
                                                           const double* __restrict__ fdir,
                                                           const double* __restrict__ lam,
                                                           const uint8_t* __restrict__ bc_mask, double scale,
                                                           double* __restrict__ out) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long f = t >> 2;
  const int q = static_cast<int>(t & 3);
  if (f >= num_faces) return;
  const long e = faces[2 * f];
  const int lf = faces[2 * f + 1];
  const int* cn = cells + 8 * e;
  double X[8][3];
  load_coords<double>(x, y, z, cn, X);
  double p0, p1, p2;
  face_point(lf, q, p0, p1, p2);
  const double nanson = face_scale(X, lf, p0, p1, p2);
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    double N, d0, d1, d2;
    shape<double>(a, p0, p1, p2, N, d0, d1, d2);
    const int nd = cn[a];
    double dl = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double lv = (bc_mask && bc_mask[3 * nd + i]) ? 0.0 : lam[3 * nd + i];
      dl += fdir[3 * f + i] * lv;
    }
    s += N * dl;
  }
  out[4 * f + q] = scale * s * nanson;
}

// Neo-Hookean strain energy psi = mu/2 (I1 - 3) - mu ln J + kappa/2 (J - 1)^2
// per element (sum over Gauss points of psi JxW), its partial d/dmu_e, and
// block partial sums of the total.
__global__ __launch_bounds__(kElemBlock) void energy_kernel(int num_cells, const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ z,
                                                            const int* __restrict__ cells, MatParams mat,
                                                            const double* __restrict__ param,
                                                            const double* __restrict__ u,
                                                            double* __restrict__ dpsi_dmu, double* partials) {
  const long t = static_cast<long>(blockIdx.x) * kElemBlock + threadIdx.x;
  const long e = t >> 3;
  const int q = static_cast<int>(t & 7);
  const bool live = e < num_cells;
  double psi = 0.0, dmu = 0.0;
  if (live) {
    const int* cn = cells + 8 * e;
    double X[8][3];
    load_coords<double>(x, y, z, cn, X);
    double x0, x1, x2;
    gauss_point(q, x0, x1, x2);
    double ug[3][3], lg[3][3];
    const double jxw = field_grads<true>(X, cn, x0, x1, x2, u, nullptr, nullptr, ug, lg);
    double FiT[3][3];
    const double J = inv_transpose(ug, FiT);
    double I1 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double F = ug[i][d] + (i == d ? 1.0 : 0.0);
        I1 += F * F;
      }
    const double mu = param_at<double>(mat, param, static_cast<int>(e), cn, x0, x1, x2);
    const double lnJ = log(J);
    dmu = (0.5 * (I1 - 3.0) - lnJ) * jxw;
    psi = mu * dmu + 0.5 * mat.kappa * (J - 1.0) * (J - 1.0) * jxw;
  }
  const double dsum = sum8(dmu);
  if (dpsi_dmu && live && q == 0) dpsi_dmu[e] = dsum;
  DVec<1> acc;
  acc.v[0] = psi;
  store_partials<1>(block_sum<1, kElemBlock>(acc), blockIdx.x, partials);
}

__global__ __launch_bounds__(kFinBlock) void energy_fin(const double* __restrict__ partials, int nparts,
                                                        double* __restrict__ out) {
  DVec<1> t = sum_partials<1>(partials, nparts);
  if (threadIdx.x == 0) *out = t.v[0];
}

inline int elem_grid(long num_cells) { return static_cast<int>((8 * num_cells + kElemBlock - 1) / kElemBlock); }

}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_hex8_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? static_cast<size_t>(num_cells) * 8 * sizeof(double) : 0;
}

int fx_hex8_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, const double* lam,
                          const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_hex8_param_vjp_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_param_vjp_f64: unknown material %d", mat->kind);
  if (mat->temperature && mat->kind == FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_param_vjp_f64: thermal strain is built for the linear / SIMP laws");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_hex8_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(param_vjp_kernel<false>, dim3(elem_grid(num_cells)), dim3(kElemBlock), 0, s, num_cells, x,
                       y, z, cells, mp, param, u, lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_hex8_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_hex8_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* wq = static_cast<double*>(workspace);
  hipLaunchKernelGGL(param_vjp_kernel<false>, dim3(elem_grid(num_cells)), dim3(kElemBlock), 0, s, num_cells, x, y,
                     z, cells, mp, param, u, lam, bc_mask, 1.0, nullptr, wq);
  FX_LAUNCHED("param_vjp_kernel");
  hipLaunchKernelGGL(param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr, inc,
                     wq, scale, grad);
  FX_LAUNCHED("param_vjp_node_kernel");
  return FX_OK;
}

int fx_hex8_thermal_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                            const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                            const fx_material* mat, const double* param, const double* lam, const uint8_t* bc_mask,
                            double scale, double* grad, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !lam ||
      !grad)
    return fail(FX_EINVAL, "fx_hex8_thermal_vjp_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP)
    return fail(FX_EINVAL, "fx_hex8_thermal_vjp_f64: thermal strain is built for the linear / SIMP laws");
  if (!mat->temperature) return fail(FX_EINVAL, "fx_hex8_thermal_vjp_f64: the material has no temperature field");
  const size_t need = fx_hex8_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8_thermal_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = as_stream(stream);
  double* wq = static_cast<double*>(workspace);
  hipLaunchKernelGGL(thermal_vjp_kernel, dim3(elem_grid(num_cells)), dim3(kElemBlock), 0, s, num_cells, x, y, z,
                     cells, to_params(mat), param, lam, bc_mask, wq);
  hipLaunchKernelGGL(param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr, inc,
                     wq, scale, grad);
  FX_LAUNCHED("thermal_vjp_kernel");
  return FX_OK;
}

int fx_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y, const double* z,
                                const int32_t* cells, const fx_material* mat, const double* param, const double* u,
                                double* dc) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_simp_compliance_sens_f64: bad arguments");
  if (mat->temperature)
    return fail(FX_EINVAL, "fx_simp_compliance_sens_f64: a thermal load depends on E; use the adjoint");
  if (mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_LINEAR_ELASTIC)
    return fail(FX_EINVAL, "fx_simp_compliance_sens_f64: needs a linear law (SIMP or linear elastic)");
  if (mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_simp_compliance_sens_f64: needs a cell variable");
  hipLaunchKernelGGL(param_vjp_kernel<true>, dim3(elem_grid(num_cells)), dim3(kElemBlock), 0, as_stream(stream),
                     num_cells, x, y, z, cells, to_params(mat), param, u, nullptr, nullptr, -1.0, dc, nullptr);
  FX_LAUNCHED("param_vjp_kernel");
  return FX_OK;
}

int fx_hex8_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                             const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                             const uint8_t* bc_mask, double scale, double* grad_svar) {
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !z || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_hex8_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  const long threads = 4L * num_faces;
  hipLaunchKernelGGL(traction_vjp_kernel, dim3(static_cast<int>((threads + 255) / 256)), dim3(256), 0,
                     as_stream(stream), num_faces, x, y, z, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("traction_vjp_kernel");
  return FX_OK;
}

size_t fx_hex8_energy_workspace_bytes(int32_t num_cells) {
  if (num_cells < 0) return 0;
  return static_cast<size_t>(elem_grid(num_cells)) * sizeof(double);
}

int fx_hex8_energy_f64(void* stream, int32_t num_cells, const double* x, const double* y, const double* z,
                       const int32_t* cells, const fx_material* mat, const double* param, const double* u,
                       double* energy, double* dpsi_dmu, void* workspace, size_t ws_bytes) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !energy)
    return fail(FX_EINVAL, "fx_hex8_energy_f64: bad arguments");
  if (mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_hex8_energy_f64: strain energy is built for the Neo-Hookean law");
  if (dpsi_dmu && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_hex8_energy_f64: dpsi/dmu per element needs a cell variable");
  const int grid = elem_grid(num_cells);
  if (grid > kMaxGridBlocks * 64)
    return fail(FX_ELIMIT, "fx_hex8_energy_f64: %d cells exceed the reduction limit", num_cells);
  const size_t need = fx_hex8_energy_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex8_energy_f64: workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = as_stream(stream);
  double* partials = static_cast<double*>(workspace);
  hipLaunchKernelGGL(energy_kernel, dim3(grid), dim3(kElemBlock), 0, s, num_cells, x, y, z, cells, to_params(mat),
                     param, u, dpsi_dmu, partials);
  FX_LAUNCHED("energy_kernel");
  hipLaunchKernelGGL(energy_fin, dim3(1), dim3(kFinBlock), 0, s, partials, grid, energy);
  FX_LAUNCHED("energy_fin");
  return FX_OK;
}

}  // extern "C"
