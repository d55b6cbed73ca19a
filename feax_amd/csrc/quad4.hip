// QUAD4 element kernels for gfx950 (dim = 2, vec = 2): CSR Jacobian
// assembly, residual and adjoint pullbacks.
//
// feax's QUAD4 (feax/basis.py:90-95): basix's P1 quadrilateral through the
// re-order [0, 1, 3, 2], i.e. local nodes at (0,0), (1,0), (1,1), (0,1) of
// [0,1]^2, 2x2 Gauss points (weights 1/4), edges bottom (0,1), left (0,3),
// right (1,2), top (3,2) with 2 Gauss points each (weights 1/2 times the
// edge Jacobian size, 1 here).  Geometry as in 3D (feax/fe.py:96-129,
// 175-239): J = sum_a x_a (x) grad_xi N_a (2x2), JxW = det J * w,
// Nanson scale ||n_ref J^-1|| det J w.
//
// The material laws are the 3-D ones of hex8_common.h applied to the
// plane-strain embedding (du_z = 0, d/dz = 0): the in-plane 2x2 blocks of
// stress, tangent and pullbacks are those of feax's 2-D tensor maps written
// with the same Lame constants.  Two passes as for HEX8 / TET4: one thread
// per element block (a, b) sums its 2x2 block over the Gauss points
// (64 values per element), one wave per node gathers (cell_common.h,
// NPE = 4, V = 2).
#include <cstdint>

#include "cell_common.h"
#include "common.h"
#include "feaxhip.h"
#include "hex8_common.h"

namespace fx {
namespace {

constexpr int kQuadThreads = 256;

__device__ __forceinline__ void quad_ref(int a, double& xa, double& ya) {
  xa = (a == 1 || a == 2) ? 1.0 : 0.0;
  ya = (a >= 2) ? 1.0 : 0.0;
}

// N_a and its reference gradient at (s, t).
__device__ __forceinline__ void quad_shape(int a, double s, double t, double& N, double& ds, double& dt) {
  double xa, ya;
  quad_ref(a, xa, ya);
  const double fs = xa > 0.5 ? s : 1.0 - s, ft = ya > 0.5 ? t : 1.0 - t;
  N = fs * ft;
  ds = (xa > 0.5 ? 1.0 : -1.0) * ft;
  dt = fs * (ya > 0.5 ? 1.0 : -1.0);
}

// Gradients g[a][0..1] (g[a][2] = 0) and det J at (s, t).
__device__ __forceinline__ double quad_grads(const double* __restrict__ x, const double* __restrict__ y,
                                             const int* __restrict__ cn, double s, double t, double (&g)[4][3]) {
  double dN[4][2];
  double J00 = 0.0, J01 = 0.0, J10 = 0.0, J11 = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double N;
    quad_shape(a, s, t, N, dN[a][0], dN[a][1]);
    const double xa = x[cn[a]], ya = y[cn[a]];
    J00 += xa * dN[a][0];
    J01 += xa * dN[a][1];
    J10 += ya * dN[a][0];
    J11 += ya * dN[a][1];
  }
  const double det = J00 * J11 - J01 * J10;
  const double i00 = J11 / det, i01 = -J01 / det, i10 = -J10 / det, i11 = J00 / det;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    g[a][0] = dN[a][0] * i00 + dN[a][1] * i10;
    g[a][1] = dN[a][0] * i01 + dN[a][1] * i11;
    g[a][2] = 0.0;
  }
  return det;
}

__device__ __forceinline__ void quad_gauss(int q, double& s, double& t) {
  s = gauss_pt(q >> 1);
  t = gauss_pt(q & 1);
}

__device__ __forceinline__ double quad_param(const MatParams& m, const double* __restrict__ param, int e,
                                             const int* __restrict__ cn, double s, double t) {
  if (m.loc == FX_PARAM_NODE) {
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double N, ds, dt;
      quad_shape(c, s, t, N, ds, dt);
      p += N * param[cn[c]];
    }
    return p;
  }
  return param[e];
}

// Plane-strain embedding of grad f for a 2-component nodal field f
// (read as 0 on rows with mask set when mask is given).
__device__ __forceinline__ void quad_fgrad(const double* __restrict__ f, const uint8_t* __restrict__ mask,
                                           const int* __restrict__ cn, const double (&g)[4][3], double (&G)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) G[i][d] = 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double v = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int r = 2 * cn[c] + i;
        v += ((mask && mask[r]) ? 0.0 : f[r]) * g[c][d];
      }
      G[i][d] = v;
    }
}

// The 3x3 tangent block K_ab of the 3-D laws at one Gauss point (weight w
// folded in); only its in-plane 2x2 part is used.
template <bool NH>
__device__ __forceinline__ void quad_block(const MatParams& mat, double p, const double (&ug)[3][3],
                                           const double (&ga)[3], const double (&gb)[3], double w, double (&K)[9]) {
  if constexpr (NH) {
    double FiT[3][3];
    const double J = inv_transpose_F<double>(ug, FiT);
    const double mu = p, kap = mat.kappa;
    const double c1 = kap * (2.0 * J - 1.0) * J, c2 = -mu + kap * J * (J - 1.0);
    double fa[3], fb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      fa[i] = FiT[i][0] * ga[0] + FiT[i][1] * ga[1] + FiT[i][2] * ga[2];
      fb[i] = FiT[i][0] * gb[0] + FiT[i][1] * gb[1] + FiT[i][2] * gb[2];
    }
    const double gg = mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) K[3 * i + j] = w * ((i == j ? gg : 0.0) + c1 * fa[i] * fb[j] - c2 * fb[i] * fa[j]);
  } else {
    const double nu = mat.nu;
    const double lam0 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu0 = 1.0 / (2.0 * (1.0 + nu));
    const double E = modulus_of<double>(mat, p);
    const double wl = w * E * lam0, wm = w * E * mu0;
    const double gg = wm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) K[3 * i + j] = wl * ga[i] * gb[j] + wm * ga[j] * gb[i] + (i == j ? gg : 0.0);
  }
}

// One thread per (element, a, b): ke[64 e + 16 a + 4 b + 2 i + j], the sum
// over the Gauss points in the order ((q0 + q1) + (q2 + q3)).
template <typename C, bool NH>
__global__ __launch_bounds__(kQuadThreads) void quad_elem_matrix_kernel(int num_cells, const double* __restrict__ x,
                                                                        const double* __restrict__ y,
                                                                        const int* __restrict__ cells, MatParams mat,
                                                                        const double* __restrict__ param,
                                                                        const double* __restrict__ u,
                                                                        C* __restrict__ ke) {
  const long t = static_cast<long>(blockIdx.x) * kQuadThreads + threadIdx.x;
  const long e = t >> 4;
  if (e >= num_cells) return;
  const int a = static_cast<int>((t >> 2) & 3), b = static_cast<int>(t & 3);
  const int* cn = cells + 4 * e;
  double Kq[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double s, tt, g[4][3];
    quad_gauss(q, s, tt);
    const double w = quad_grads(x, y, cn, s, tt, g) * 0.25;
    double ug[3][3];
    if constexpr (NH) {
      quad_fgrad(u, nullptr, cn, g, ug);
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) ug[i][d] = 0.0;
    }
    double K[9];
    quad_block<NH>(mat, quad_param(mat, param, static_cast<int>(e), cn, s, tt), ug, g[a], g[b], w, K);
    Kq[q][0] = K[0];
    Kq[q][1] = K[1];
    Kq[q][2] = K[3];
    Kq[q][3] = K[4];
  }
  C* dst = ke + 64 * e + 16 * a + 4 * b;
#pragma unroll
  for (int k = 0; k < 4; ++k) dst[k] = static_cast<C>((Kq[0][k] + Kq[1][k]) + (Kq[2][k] + Kq[3][k]));
}

// One thread per element: re[8 e + 2 a + i].
__global__ __launch_bounds__(kQuadThreads) void quad_elem_residual_kernel(int num_cells, const double* __restrict__ x,
                                                                          const double* __restrict__ y,
                                                                          const int* __restrict__ cells, MatParams mat,
                                                                          const double* __restrict__ param,
                                                                          const double* __restrict__ u,
                                                                          double* __restrict__ re) {
  const int e = blockIdx.x * kQuadThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double rq[4][8];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double s, t, g[4][3], ug[3][3], sig[3][3];
    quad_gauss(q, s, t);
    const double w = quad_grads(x, y, cn, s, t, g) * 0.25;
    quad_fgrad(u, nullptr, cn, g, ug);
    material_stress<double>(mat, quad_param(mat, param, e, cn, s, t), ug, sig);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 2; ++i) rq[q][2 * a + i] = (sig[i][0] * g[a][0] + sig[i][1] * g[a][1]) * w;
  }
  double* dst = re + 8 * static_cast<long>(e);
#pragma unroll
  for (int k = 0; k < 8; ++k) dst[k] = (rq[0][k] + rq[1][k]) + (rq[2][k] + rq[3][k]);
}

// Edge lf of the reference quad: first / second vertex and outward normal.
__device__ __forceinline__ void quad_edge(int lf, int& v0, int& v1, double& nx, double& ny) {
  const int E0[4] = {0, 0, 1, 3}, E1[4] = {1, 3, 2, 2};
  v0 = E0[lf];
  v1 = E1[lf];
  nx = lf == 1 ? -1.0 : (lf == 2 ? 1.0 : 0.0);
  ny = lf == 0 ? -1.0 : (lf == 3 ? 1.0 : 0.0);
}

// Gauss point qq of edge lf: reference coordinates (s, t) and the Nanson
// scale ||n_ref J^-1|| det J w (w = 1/2 per point).
__device__ __forceinline__ double quad_edge_point(const double* __restrict__ x, const double* __restrict__ y,
                                                  const int* __restrict__ cn, int lf, int qq, double& s, double& t) {
  int v0, v1;
  double nx, ny;
  quad_edge(lf, v0, v1, nx, ny);
  double x0, y0, x1, y1;
  quad_ref(v0, x0, y0);
  quad_ref(v1, x1, y1);
  const double tp = gauss_pt(qq);
  s = x0 + tp * (x1 - x0);
  t = y0 + tp * (y1 - y0);
  double J00 = 0.0, J01 = 0.0, J10 = 0.0, J11 = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double N, ds, dt;
    quad_shape(a, s, t, N, ds, dt);
    J00 += x[cn[a]] * ds;
    J01 += x[cn[a]] * dt;
    J10 += y[cn[a]] * ds;
    J11 += y[cn[a]] * dt;
  }
  const double det = J00 * J11 - J01 * J10;
  const double i00 = J11 / det, i01 = -J01 / det, i10 = -J10 / det, i11 = J00 / det;
  const double m0 = nx * i00 + ny * i10, m1 = nx * i01 + ny * i11;
  return sqrt(m0 * m0 + m1 * m1) * det * 0.5;
}

// One thread per node: element residuals in ascending element order, then
// the loaded edges: t(f, q) = fdir[f] * svar[2 f + q].
__global__ __launch_bounds__(256) void quad_node_residual_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const int* __restrict__ cells,
    const int* __restrict__ inc_ptr, const int* __restrict__ inc, const double* __restrict__ re,
    const int* __restrict__ faces, const int* __restrict__ finc_ptr, const int* __restrict__ finc,
    const double* __restrict__ svar, const double* __restrict__ fdir, int negate, double* __restrict__ res) {
  const int nd = blockIdx.x * 256 + threadIdx.x;
  if (nd >= num_nodes) return;
  double r0 = 0.0, r1 = 0.0;
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    const double* src = re + 8 * static_cast<long>(pk >> 3) + 2 * (pk & 7);
    r0 += src[0];
    r1 += src[1];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> 3, a = pk & 7;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      const int* cn = cells + 4 * static_cast<long>(e);
      double vq[2];
      for (int qq = 0; qq < 2; ++qq) {
        double s, t;
        const double nanson = quad_edge_point(x, y, cn, lf, qq, s, t);
        double N, ds, dt;
        quad_shape(a, s, t, N, ds, dt);
        vq[qq] = svar[2 * f + qq] * N * nanson;
      }
      const double v = vq[0] + vq[1];
      r0 += fdir[2 * f] * v;
      r1 += fdir[2 * f + 1] * v;
    }
  }
  const double sgn = negate ? -1.0 : 1.0;
  res[2 * nd] = sgn * r0;
  res[2 * nd + 1] = sgn * r1;
}

// Pullbacks per (element, Gauss point): w_eq = dR/dp . lambda (SELF:
// lambda = u).  Cell variables: scale * sum_q w_eq; node variables keep
// w_eq (we[4 e + q]) for quad_param_vjp_node_kernel.
template <bool SELF>
__global__ __launch_bounds__(kQuadThreads) void quad_param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const int* __restrict__ cells,
    MatParams mat, const double* __restrict__ param, const double* __restrict__ u, const double* __restrict__ lam,
    const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell, double* __restrict__ we) {
  const int e = blockIdx.x * kQuadThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double wq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double s, t, g[4][3], ug[3][3], lg[3][3];
    quad_gauss(q, s, t);
    const double jxw = quad_grads(x, y, cn, s, t, g) * 0.25;
    quad_fgrad(u, nullptr, cn, g, ug);
    if constexpr (SELF) {
      quad_fgrad(u, nullptr, cn, g, lg);
    } else {
      quad_fgrad(lam, bc_mask, cn, g, lg);
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
      modulus_of<double>(mat, quad_param(mat, param, e, cn, s, t), &dEdp);
      dens *= dEdp;
    }
    wq[q] = dens * jxw;
  }
  if (out_cell) {
    out_cell[e] = scale * ((wq[0] + wq[1]) + (wq[2] + wq[3]));
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) we[4 * static_cast<long>(e) + q] = wq[q];
  }
}

__global__ __launch_bounds__(256) void quad_param_vjp_node_kernel(int num_nodes, const int* __restrict__ inc_ptr,
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
    for (int q = 0; q < 4; ++q) {
      double s, t, N, ds, dt;
      quad_gauss(q, s, t);
      quad_shape(a, s, t, N, ds, dt);
      acc += N * we[4 * e + q];
    }
  }
  out[i] = scale * acc;
}

// grad_svar[2 f + q] = scale * nanson_q * sum over the edge's vertices of
// N_a(q) (dir_f . lambda_a).
__global__ __launch_bounds__(256) void quad_traction_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                                const double* __restrict__ y,
                                                                const int* __restrict__ cells,
                                                                const int* __restrict__ faces,
                                                                const double* __restrict__ fdir,
                                                                const double* __restrict__ lam,
                                                                const uint8_t* __restrict__ bc_mask, double scale,
                                                                double* __restrict__ out) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long f = t >> 1;
  const int qq = static_cast<int>(t & 1);
  if (f >= num_faces) return;
  const int e = faces[2 * f], lf = faces[2 * f + 1];
  const int* cn = cells + 4 * static_cast<long>(e);
  double s, tt;
  const double nanson = quad_edge_point(x, y, cn, lf, qq, s, tt);
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double N, ds, dt;
    quad_shape(a, s, tt, N, ds, dt);
    double dl = 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int r = 2 * cn[a] + i;
      dl += fdir[2 * f + i] * ((bc_mask && bc_mask[r]) ? 0.0 : lam[r]);
    }
    acc += N * dl;
  }
  out[2 * f + qq] = scale * acc * nanson;
}

// ---------------------------------------------------------------- scalar QUAD4
// vec = 1 conduction in the plane (the scalar Problem of feax with flux
// k grad T on QUAD4): K_ab = sum_q k_q grad N_a . grad N_b JxW_q, R_a =
// sum_q k_q grad N_a . grad T_q JxW_q, edge flux + sum_q q N_a nanson;
// k = p or the SIMP interpolation (FX_MAT_SIMP_CONDUCTION).
__device__ __forceinline__ double qs_conductivity(const MatParams& m, double p, double* dkdp) {
  if (m.kind == FX_MAT_SIMP_CONDUCTION) {
    const double span = m.E0 - m.Emin;
    if (dkdp) *dkdp = m.penal * pow(p, m.penal - 1.0) * span;
    return m.Emin + pow(p, m.penal) * span;
  }
  if (dkdp) *dkdp = 1.0;
  return p;
}

__device__ __forceinline__ void qs_grad(const double* __restrict__ f, const uint8_t* __restrict__ mask,
                                        const int* __restrict__ cn, const double (&g)[4][3], double& G0, double& G1) {
  G0 = G1 = 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double v = (mask && mask[cn[c]]) ? 0.0 : f[cn[c]];
    G0 += v * g[c][0];
    G1 += v * g[c][1];
  }
}

// One thread per (element, a, b): ke[16 e + 4 a + b].
__global__ __launch_bounds__(kQuadThreads) void qs_elem_matrix_kernel(int num_cells, const double* __restrict__ x,
                                                                      const double* __restrict__ y,
                                                                      const int* __restrict__ cells, MatParams mat,
                                                                      const double* __restrict__ param,
                                                                      double* __restrict__ ke) {
  const long t = static_cast<long>(blockIdx.x) * kQuadThreads + threadIdx.x;
  const long e = t >> 4;
  if (e >= num_cells) return;
  const int a = static_cast<int>((t >> 2) & 3), b = static_cast<int>(t & 3);
  const int* cn = cells + 4 * e;
  double kq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double s, tt, g[4][3];
    quad_gauss(q, s, tt);
    const double w = quad_grads(x, y, cn, s, tt, g) * 0.25;
    const double k = qs_conductivity(mat, quad_param(mat, param, static_cast<int>(e), cn, s, tt), nullptr);
    kq[q] = k * w * (g[a][0] * g[b][0] + g[a][1] * g[b][1]);
  }
  ke[16 * e + 4 * a + b] = (kq[0] + kq[1]) + (kq[2] + kq[3]);
}

// One thread per element: re[4 e + a].
__global__ __launch_bounds__(kQuadThreads) void qs_elem_residual_kernel(int num_cells, const double* __restrict__ x,
                                                                        const double* __restrict__ y,
                                                                        const int* __restrict__ cells, MatParams mat,
                                                                        const double* __restrict__ param,
                                                                        const double* __restrict__ u,
                                                                        double* __restrict__ re) {
  const int e = blockIdx.x * kQuadThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double rq[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double s, t, g[4][3], G0, G1;
    quad_gauss(q, s, t);
    const double w = quad_grads(x, y, cn, s, t, g) * 0.25;
    qs_grad(u, nullptr, cn, g, G0, G1);
    const double kw = qs_conductivity(mat, quad_param(mat, param, e, cn, s, t), nullptr) * w;
#pragma unroll
    for (int a = 0; a < 4; ++a) rq[q][a] = kw * (G0 * g[a][0] + G1 * g[a][1]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) re[4 * static_cast<long>(e) + a] = (rq[0][a] + rq[1][a]) + (rq[2][a] + rq[3][a]);
}

// One thread per node: element residuals in ascending order, then the
// loaded edges: flux(f, q) = fdir[f] * svar[2 f + q].
__global__ __launch_bounds__(256) void qs_node_residual_kernel(
    int num_nodes, const double* __restrict__ x, const double* __restrict__ y, const int* __restrict__ cells,
    const int* __restrict__ inc_ptr, const int* __restrict__ inc, const double* __restrict__ re,
    const int* __restrict__ faces, const int* __restrict__ finc_ptr, const int* __restrict__ finc,
    const double* __restrict__ svar, const double* __restrict__ fdir, int negate, double* __restrict__ res) {
  const int nd = blockIdx.x * 256 + threadIdx.x;
  if (nd >= num_nodes) return;
  double r = 0.0;
  for (int k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k) {
    const int pk = inc[k];
    r += re[4 * static_cast<long>(pk >> 3) + (pk & 7)];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> 3, a = pk & 7;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      const int* cn = cells + 4 * static_cast<long>(e);
      double vq[2];
      for (int qq = 0; qq < 2; ++qq) {
        double s, t, N, ds, dt;
        const double nanson = quad_edge_point(x, y, cn, lf, qq, s, t);
        quad_shape(a, s, t, N, ds, dt);
        vq[qq] = svar[2 * f + qq] * N * nanson;
      }
      r += fdir[f] * (vq[0] + vq[1]);
    }
  }
  res[nd] = negate ? -r : r;
}

// Pullbacks per (element, Gauss point) w_eq = dk/dp grad lambda . grad T
// JxW (SELF: lambda = T); cell variables: scale * sum_q, else we[4 e + q].
template <bool SELF>
__global__ __launch_bounds__(kQuadThreads) void qs_param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const int* __restrict__ cells,
    MatParams mat, const double* __restrict__ param, const double* __restrict__ u, const double* __restrict__ lam,
    const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell, double* __restrict__ we) {
  const int e = blockIdx.x * kQuadThreads + threadIdx.x;
  if (e >= num_cells) return;
  const int* cn = cells + 4 * static_cast<long>(e);
  double wq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double s, t, g[4][3], G0, G1, L0, L1;
    quad_gauss(q, s, t);
    const double w = quad_grads(x, y, cn, s, t, g) * 0.25;
    qs_grad(u, nullptr, cn, g, G0, G1);
    if constexpr (SELF) {
      L0 = G0;
      L1 = G1;
    } else {
      qs_grad(lam, bc_mask, cn, g, L0, L1);
    }
    double dkdp;
    qs_conductivity(mat, quad_param(mat, param, e, cn, s, t), &dkdp);
    wq[q] = dkdp * w * (G0 * L0 + G1 * L1);
  }
  if (out_cell) {
    out_cell[e] = scale * ((wq[0] + wq[1]) + (wq[2] + wq[3]));
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) we[4 * static_cast<long>(e) + q] = wq[q];
  }
}

// grad_svar[2 f + q] = scale * nanson_q * fdir[f] * sum_a N_a(q) lambda_a.
__global__ __launch_bounds__(256) void qs_flux_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                          const double* __restrict__ y, const int* __restrict__ cells,
                                                          const int* __restrict__ faces,
                                                          const double* __restrict__ fdir,
                                                          const double* __restrict__ lam,
                                                          const uint8_t* __restrict__ bc_mask, double scale,
                                                          double* __restrict__ out) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long f = t >> 1;
  const int qq = static_cast<int>(t & 1);
  if (f >= num_faces) return;
  const int e = faces[2 * f], lf = faces[2 * f + 1];
  const int* cn = cells + 4 * static_cast<long>(e);
  double s, tt;
  const double nanson = quad_edge_point(x, y, cn, lf, qq, s, tt);
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double N, ds, dt;
    quad_shape(a, s, tt, N, ds, dt);
    acc += N * ((bc_mask && bc_mask[cn[a]]) ? 0.0 : lam[cn[a]]);
  }
  out[2 * f + qq] = scale * fdir[f] * acc * nanson;
}

inline bool qs_law(const fx_material* m) {
  return m->kind == FX_MAT_CONDUCTION || m->kind == FX_MAT_SIMP_CONDUCTION;
}

inline size_t quad_matrix_bytes(int num_cells, int value_bytes, int flags) {
  const size_t c = (value_bytes == 4 && (flags & FX_ASM_COMPUTE32)) ? 4 : 8;
  return 64 * c * static_cast<size_t>(num_cells);
}

template <typename T>
int quad_assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y, const int* cells,
                  const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr, const int* indices,
                  const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                  int symmetric, int flags, double jacobi_shift, double* dinv, T* data, void* workspace,
                  size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !cells || !inc_ptr || !inc || !rank || !indptr || !indices ||
      !mat || !param || !data)
    return fail(FX_EINVAL, "fx_quad4_assemble: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_quad4_assemble: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_quad4_assemble: thermal strain is built for HEX8");
  if (flags & (FX_ASM_DIFF3 | FX_ASM_COMPUTE32))
    return fail(FX_EINVAL, "fx_quad4_assemble: the difference form and FP32 arithmetic are built for vec = 3");
  const bool nh = mat->kind == FX_MAT_NEO_HOOKEAN;
  if (nh && !u) return fail(FX_EINVAL, "fx_quad4_assemble: the Neo-Hookean tangent needs u");
  const size_t need = quad_matrix_bytes(num_cells, sizeof(T), flags);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_quad4_assemble: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  const MatParams mp = to_params(mat);
  hipStream_t s = as_stream(stream);
  const int egrid = static_cast<int>((16L * num_cells + kQuadThreads - 1) / kQuadThreads);
  const int grid = (num_nodes + kNodesPerBlock - 1) / kNodesPerBlock;
  double* ke = static_cast<double*>(workspace);
  if (nh)
    hipLaunchKernelGGL((quad_elem_matrix_kernel<double, true>), dim3(egrid), dim3(kQuadThreads), 0, s, num_cells, x,
                       y, cells, mp, param, u, ke);
  else
    hipLaunchKernelGGL((quad_elem_matrix_kernel<double, false>), dim3(egrid), dim3(kQuadThreads), 0, s, num_cells,
                       x, y, cells, mp, param, u, ke);
  hipLaunchKernelGGL((gather_wide_rows_kernel<T, double, 4, 2>), dim3(grid), dim3(64 * kNodesPerBlock), 0, s,
                     num_nodes, nullptr, inc_ptr, inc, rank, indptr, indices, ke, bc_mask, symmetric, 0, jacobi_shift,
                     dinv, data, nullptr);
  FX_LAUNCHED("fx_quad4_assemble");
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_quad4_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  return num_cells < 0 ? 0 : quad_matrix_bytes(num_cells, value_bytes, flags);
}

int fx_quad4_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes) {
  (void)z;
  return quad_assemble<double>(stream, num_nodes, num_cells, x, y, cells, inc_ptr, inc, elem_rank, indptr, indices,
                               mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace,
                               ws_bytes);
}

int fx_quad4_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes) {
  (void)z;
  return quad_assemble<float>(stream, num_nodes, num_cells, x, y, cells, inc_ptr, inc, elem_rank, indptr, indices,
                              mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace,
                              ws_bytes);
}

size_t fx_quad4_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 8 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_quad4_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  (void)z;
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !cells || !inc_ptr || !inc || !mat || !param || !u || !res)
    return fail(FX_EINVAL, "fx_quad4_residual_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_quad4_residual_f64: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_quad4_residual_f64: thermal strain is built for HEX8");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_quad4_residual_f64: face arrays missing");
  const size_t need = fx_quad4_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_quad4_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool has_faces = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(quad_elem_residual_kernel, dim3((num_cells + kQuadThreads - 1) / kQuadThreads),
                     dim3(kQuadThreads), 0, s, num_cells, x, y, cells, to_params(mat), param, u, re);
  hipLaunchKernelGGL(quad_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y,
                     cells, inc_ptr, inc, re, has_faces ? faces : nullptr, has_faces ? finc_ptr : nullptr,
                     has_faces ? finc : nullptr, svar, fdir, negate, res);
  FX_LAUNCHED("fx_quad4_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_quad4_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

size_t fx_quad4_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? static_cast<size_t>(num_cells) * 4 * sizeof(double) : 0;
}

int fx_quad4_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  (void)z;
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_quad4_param_vjp_f64: bad arguments");
  if (mat->kind != FX_MAT_LINEAR_ELASTIC && mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_NEO_HOOKEAN)
    return fail(FX_EINVAL, "fx_quad4_param_vjp_f64: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_quad4_param_vjp_f64: thermal strain is built for HEX8");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_quad4_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  const dim3 eg((num_cells + kQuadThreads - 1) / kQuadThreads);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(quad_param_vjp_kernel<false>, eg, dim3(kQuadThreads), 0, s, num_cells, x, y, cells, mp, param,
                       u, lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("quad_param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_quad4_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_quad4_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_quad4_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* we = static_cast<double*>(workspace);
  hipLaunchKernelGGL(quad_param_vjp_kernel<false>, eg, dim3(kQuadThreads), 0, s, num_cells, x, y, cells, mp, param, u,
                     lam, bc_mask, 1.0, nullptr, we);
  hipLaunchKernelGGL(quad_param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc, we, scale, grad);
  FX_LAUNCHED("quad_param_vjp_node_kernel");
  return FX_OK;
}

int fx_quad4_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc) {
  (void)z;
  if (num_cells <= 0 || !x || !y || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_quad4_simp_compliance_sens_f64: bad arguments");
  if (mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_LINEAR_ELASTIC)
    return fail(FX_EINVAL, "fx_quad4_simp_compliance_sens_f64: needs a linear law (SIMP or linear elastic)");
  if (mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_quad4_simp_compliance_sens_f64: needs a cell variable");
  hipLaunchKernelGGL(quad_param_vjp_kernel<true>, dim3((num_cells + kQuadThreads - 1) / kQuadThreads),
                     dim3(kQuadThreads), 0, as_stream(stream), num_cells, x, y, cells, to_params(mat), param, u,
                     nullptr, nullptr, -1.0, dc, nullptr);
  FX_LAUNCHED("quad_param_vjp_kernel");
  return FX_OK;
}

int fx_quad4_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar) {
  (void)z;
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_quad4_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  const long threads = 2L * num_faces;
  hipLaunchKernelGGL(quad_traction_vjp_kernel, dim3(static_cast<int>((threads + 255) / 256)), dim3(256), 0,
                     as_stream(stream), num_faces, x, y, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("quad_traction_vjp_kernel");
  return FX_OK;
}

size_t fx_quad4s_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  (void)value_bytes;
  (void)flags;
  return num_cells < 0 ? 0 : 16 * sizeof(double) * static_cast<size_t>(num_cells);
}

}  // extern "C"

namespace fx {
namespace {

template <typename T>
int qs_assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y,
                       const int* cells, const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr,
                       const int* indices, const fx_material* mat, const double* param, const uint8_t* bc_mask,
                       int symmetric, int flags, double jacobi_shift, double* dinv, T* data, void* workspace,
                       size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !cells || !inc_ptr || !inc || !rank || !indptr || !indices ||
      !mat || !param || !data)
    return fail(FX_EINVAL, "fx_quad4s_assemble: bad arguments");
  if (!qs_law(mat) || mat->temperature) return fail(FX_EINVAL, "fx_quad4s_assemble: needs a conduction law");
  if (flags & (FX_ASM_DIFF3 | FX_ASM_COMPUTE32))
    return fail(FX_EINVAL, "fx_quad4s_assemble: the difference form and FP32 arithmetic are built for vec = 3");
  const size_t need = fx_quad4s_assemble_workspace_bytes(num_cells, sizeof(T), flags);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_quad4s_assemble: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  double* ke = static_cast<double*>(workspace);
  hipLaunchKernelGGL(qs_elem_matrix_kernel, dim3(static_cast<int>((16L * num_cells + kQuadThreads - 1) / kQuadThreads)),
                     dim3(kQuadThreads), 0, s, num_cells, x, y, cells, to_params(mat), param, ke);
  hipLaunchKernelGGL((gather_wide_rows_kernel<T, double, 4, 1>),
                     dim3((num_nodes + kNodesPerBlock - 1) / kNodesPerBlock),
                     dim3(64 * kNodesPerBlock), 0, s, num_nodes, nullptr, inc_ptr, inc, rank, indptr, indices, ke,
                     bc_mask, symmetric, 0, jacobi_shift, dinv, data, nullptr);
  FX_LAUNCHED("fx_quad4s_assemble");
  return FX_OK;
}

}  // namespace
}  // namespace fx

extern "C" {

int fx_quad4s_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                           const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                           int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                           void* workspace, size_t ws_bytes) {
  (void)z;
  (void)u;
  return qs_assemble<double>(stream, num_nodes, num_cells, x, y, cells, inc_ptr, inc, elem_rank, indptr, indices, mat,
                             param, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes);
}

int fx_quad4s_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                           const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                           int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                           void* workspace, size_t ws_bytes) {
  (void)z;
  (void)u;
  return qs_assemble<float>(stream, num_nodes, num_cells, x, y, cells, inc_ptr, inc, elem_rank, indptr, indices, mat,
                            param, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace, ws_bytes);
}

size_t fx_quad4s_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 4 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_quad4s_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                           const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                           const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                           int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  (void)z;
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !cells || !inc_ptr || !inc || !mat || !param || !u || !res)
    return fail(FX_EINVAL, "fx_quad4s_residual_f64: bad arguments");
  if (!qs_law(mat) || mat->temperature) return fail(FX_EINVAL, "fx_quad4s_residual_f64: needs a conduction law");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_quad4s_residual_f64: face arrays missing");
  const size_t need = fx_quad4s_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_quad4s_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool hf = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(qs_elem_residual_kernel, dim3((num_cells + kQuadThreads - 1) / kQuadThreads), dim3(kQuadThreads),
                     0, s, num_cells, x, y, cells, to_params(mat), param, u, re);
  hipLaunchKernelGGL(qs_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, cells,
                     inc_ptr, inc, re, hf ? faces : nullptr, hf ? finc_ptr : nullptr, hf ? finc : nullptr, svar, fdir,
                     negate, res);
  FX_LAUNCHED("fx_quad4s_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_quad4s_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

size_t fx_quad4s_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? static_cast<size_t>(num_cells) * 4 * sizeof(double) : 0;
}

int fx_quad4s_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                            const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                            const fx_material* mat, const double* param, const double* u, const double* lam,
                            const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  (void)z;
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_quad4s_param_vjp_f64: bad arguments");
  if (!qs_law(mat) || mat->temperature) return fail(FX_EINVAL, "fx_quad4s_param_vjp_f64: needs a conduction law");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_quad4s_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  const dim3 eg((num_cells + kQuadThreads - 1) / kQuadThreads);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(qs_param_vjp_kernel<false>, eg, dim3(kQuadThreads), 0, s, num_cells, x, y, cells, mp, param, u,
                       lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("qs_param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_quad4s_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_quad4s_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_quad4s_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* we = static_cast<double*>(workspace);
  hipLaunchKernelGGL(qs_param_vjp_kernel<false>, eg, dim3(kQuadThreads), 0, s, num_cells, x, y, cells, mp, param, u,
                     lam, bc_mask, 1.0, nullptr, we);
  hipLaunchKernelGGL(quad_param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc, we, scale, grad);
  FX_LAUNCHED("quad_param_vjp_node_kernel");
  return FX_OK;
}

int fx_quad4s_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                       const double* z, const int32_t* cells, const fx_material* mat,
                                       const double* param, const double* u, double* dc) {
  (void)z;
  if (num_cells <= 0 || !x || !y || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_quad4s_simp_compliance_sens_f64: bad arguments");
  if (!qs_law(mat) || mat->temperature || mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_quad4s_simp_compliance_sens_f64: needs a conduction law with a cell variable");
  hipLaunchKernelGGL(qs_param_vjp_kernel<true>, dim3((num_cells + kQuadThreads - 1) / kQuadThreads),
                     dim3(kQuadThreads), 0, as_stream(stream), num_cells, x, y, cells, to_params(mat), param, u,
                     nullptr, nullptr, -1.0, dc, nullptr);
  FX_LAUNCHED("qs_param_vjp_kernel");
  return FX_OK;
}

int fx_quad4s_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                               const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                               const uint8_t* bc_mask, double scale, double* grad_svar) {
  (void)z;
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_quad4s_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  hipLaunchKernelGGL(qs_flux_vjp_kernel, dim3(static_cast<int>((2L * num_faces + 255) / 256)), dim3(256), 0,
                     as_stream(stream), num_faces, x, y, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("qs_flux_vjp_kernel");
  return FX_OK;
}

}  // extern "C"
