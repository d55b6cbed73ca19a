// HEX20 element kernels for gfx950: CSR Jacobian assembly, residual and the
// adjoint pullbacks.
//
// feax's HEX20 (feax/basis.py:71-77): basix's serendipity element of degree
// 2 on the hexahedron in the meshio / Abaqus node order (8 corners, then the
// 12 edge nodes of the edges (0,1) (1,2) (2,3) (3,0) (4,5) (5,6) (6,7)
// (7,4) (0,4) (1,5) (2,6) (3,7)).  basix takes as degrees of freedom the
// corner values and the mean of the function over each edge, so with the
// 20-node nodal serendipity functions psi the basis is
//   phi_edge = 3/2 psi_edge,  phi_corner = psi_corner - 1/4 sum_{edges at it} psi_edge.
// Quadrature: basix.make_quadrature(hexahedron, 2), the 2x2x2 Gauss rule
// (weights 1/8 on [0,1]^3); faces: 2x2 Gauss on each facet (weights 1/4),
// the HEX8 facets.  Geometry as feax/fe.py:96-129, 175-239.
//
// Two passes like the TET4 / HEX8 two-pass kernels: one wave per element
// forms the 60x60 element matrix (3x3 blocks (a, b) at 9 (20 a + b)) or
// its residual (60 values) in a workspace, with the per-Gauss-point
// geometry staged in LDS; then one wave (thread) per node gathers them in
// ascending element order into the node's CSR rows (cell_common.h,
// NPE = 20, incidence entries (e << 5) | a) or residual entries.
#include <cstdint>

#include "cell_common.h"
#include "common.h"
#include "feaxhip.h"
#include "hex8_common.h"
#include "reduce.h"

namespace fx {
namespace {

constexpr int kH20Threads = 64;      // one wave per element
constexpr int kH20Shift = 5;         // incidence entry (e << 5) | a
constexpr int kH20MaxDeg = 96;       // neighbours of a HEX20 node (81 in a structured mesh)

// Corner pairs of the 12 edge nodes (local nodes 8..19) and the three edge
// nodes at each corner.
__constant__ int kH20Edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
__constant__ int kH20EdgesAt[8][3] = {{8, 11, 16}, {8, 9, 17},  {9, 10, 18},  {10, 11, 19},
                                      {12, 15, 16}, {12, 13, 17}, {13, 14, 18}, {14, 15, 19}};
// The 8 nodes of each facet (the HEX8 facet corners, then its edge nodes).
__constant__ int kH20FaceNodes[6][8] = {{0, 1, 3, 2, 8, 9, 10, 11},   {0, 1, 4, 5, 8, 12, 16, 17},
                                        {0, 3, 4, 7, 11, 15, 16, 19}, {1, 2, 5, 6, 9, 13, 17, 18},
                                        {3, 2, 7, 6, 10, 14, 18, 19}, {4, 5, 7, 6, 12, 13, 14, 15}};

// The 20 basis functions and their reference gradients at (x0, x1, x2).
__device__ inline void h20_shape(double x0, double x1, double x2, double (&N)[20], double (&dN)[20][3]) {
  const double s[3] = {2.0 * x0 - 1.0, 2.0 * x1 - 1.0, 2.0 * x2 - 1.0};
  double psi[20], dpsi[20][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int c[3];
    corner(a, c[0], c[1], c[2]);
    const double sg[3] = {2.0 * c[0] - 1.0, 2.0 * c[1] - 1.0, 2.0 * c[2] - 1.0};
    const double f0 = 1.0 + sg[0] * s[0], f1 = 1.0 + sg[1] * s[1], f2 = 1.0 + sg[2] * s[2];
    const double g = sg[0] * s[0] + sg[1] * s[1] + sg[2] * s[2] - 2.0;
    const double P = f0 * f1 * f2;
    psi[a] = 0.125 * P * g;
    // d/dxi = 2 d/ds
    dpsi[a][0] = 0.25 * sg[0] * (f1 * f2 * g + P);
    dpsi[a][1] = 0.25 * sg[1] * (f0 * f2 * g + P);
    dpsi[a][2] = 0.25 * sg[2] * (f0 * f1 * g + P);
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    int p[3], q[3];
    corner(kH20Edge[k][0], p[0], p[1], p[2]);
    corner(kH20Edge[k][1], q[0], q[1], q[2]);
    const int ax = p[0] != q[0] ? 0 : (p[1] != q[1] ? 1 : 2);
    const int i = ax == 0 ? 1 : 0, j = ax == 2 ? 1 : 2;
    const double si = 2.0 * p[i] - 1.0, sj = 2.0 * p[j] - 1.0;
    const double fi = 1.0 + si * s[i], fj = 1.0 + sj * s[j], qk = 1.0 - s[ax] * s[ax];
    psi[8 + k] = 0.25 * qk * fi * fj;
    dpsi[8 + k][ax] = -s[ax] * fi * fj;
    dpsi[8 + k][i] = 0.5 * qk * si * fj;
    dpsi[8 + k][j] = 0.5 * qk * fi * sj;
  }
#pragma unroll
  for (int a = 8; a < 20; ++a) {
    N[a] = 1.5 * psi[a];
#pragma unroll
    for (int d = 0; d < 3; ++d) dN[a][d] = 1.5 * dpsi[a][d];
  }
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    double v = psi[a], g0 = dpsi[a][0], g1 = dpsi[a][1], g2 = dpsi[a][2];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int e = kH20EdgesAt[a][t];
      v -= 0.25 * psi[e];
      g0 -= 0.25 * dpsi[e][0];
      g1 -= 0.25 * dpsi[e][1];
      g2 -= 0.25 * dpsi[e][2];
    }
    N[a] = v;
    dN[a][0] = g0;
    dN[a][1] = g1;
    dN[a][2] = g2;
  }
}

// J_ij = sum_a X_a,i dN_a/dxi_j; returns det J, writes J^-1 (row-major).
__device__ inline double h20_jac_inv(const double (*X)[3], const double (&dN)[20][3], double (&Ji)[3][3]) {
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < 20; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) J[i][j] += X[a][i] * dN[a][j];
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  Ji[0][0] = c00 * id;
  Ji[1][0] = c01 * id;
  Ji[2][0] = c02 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

// Per-element geometry staged in LDS by one wave: coordinates, basis
// values at the 8 Gauss points, physical gradients g[q][a][d], JxW[q].
struct H20Geom {
  double X[20][3];
  double N[8][20];
  double g[8][20][3];
  double w[8];
};

// Fills G for element e (threads 0..19 load coordinates, threads 0..7 one
// Gauss point each); ends with a barrier.
__device__ inline void h20_geometry(const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, const int* __restrict__ cn, H20Geom& G) {
  const int t = threadIdx.x;
  if (t < 20) {
    const int nd = cn[t];
    G.X[t][0] = x[nd];
    G.X[t][1] = y[nd];
    G.X[t][2] = z[nd];
  }
  __syncthreads();
  if (t < 8) {
    double p0, p1, p2;
    gauss_point(t, p0, p1, p2);
    double N[20], dN[20][3], Ji[3][3];
    h20_shape(p0, p1, p2, N, dN);
    const double det = h20_jac_inv(G.X, dN, Ji);
    G.w[t] = det * 0.125;
    for (int a = 0; a < 20; ++a) {
      G.N[t][a] = N[a];
#pragma unroll
      for (int d = 0; d < 3; ++d) G.g[t][a][d] = dN[a][0] * Ji[0][d] + dN[a][1] * Ji[1][d] + dN[a][2] * Ji[2][d];
    }
  }
  __syncthreads();
}

// Material parameter at Gauss point q: sum_a N_a p_a (node variable) or the
// cell value.
__device__ inline double h20_param(const MatParams& m, const double* __restrict__ param, int e,
                                   const int* __restrict__ cn, const H20Geom& G, int q) {
  if (m.loc == FX_PARAM_NODE) {
    double p = 0.0;
    for (int a = 0; a < 20; ++a) p += G.N[q][a] * param[cn[a]];
    return p;
  }
  return param[e];
}

// v_i,d = sum_a f[3 cn[a] + i] g[q][a][d]; rows with mask set read as 0.
__device__ inline void h20_field_grad(const double* __restrict__ f, const uint8_t* __restrict__ mask,
                                      const int* __restrict__ cn, const H20Geom& G, int q, double (&v)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[i][d] = 0.0;
  for (int a = 0; a < 20; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int r = 3 * cn[a] + i;
      const double fa = (mask && mask[r]) ? 0.0 : f[r];
#pragma unroll
      for (int d = 0; d < 3; ++d) v[i][d] += fa * G.g[q][a][d];
    }
  }
}

// One wave per element: the 400 3x3 blocks K_ab of the element matrix at
// ke[3600 e + 9 (20 a + b)], summed over the 8 Gauss points in order:
//   linear / SIMP: E(p) JxW (lam0 ga_i gb_j + mu0 ga_j gb_i + mu0 d_ij ga.gb)
//   Neo-Hookean:   JxW (d_ij mu ga.gb + c1 fa_i fb_j - c2 fb_i fa_j),
//                  fa = F^-T ga, c1 = kappa (2J - 1) J, c2 = -mu + kappa J (J - 1).
template <bool NH>
__global__ __launch_bounds__(kH20Threads) void h20_elem_matrix_kernel(int num_cells, const double* __restrict__ x,
                                                                      const double* __restrict__ y,
                                                                      const double* __restrict__ z,
                                                                      const int* __restrict__ cells, MatParams mat,
                                                                      const double* __restrict__ param,
                                                                      const double* __restrict__ u,
                                                                      double* __restrict__ ke) {
  const int e = blockIdx.x;
  if (e >= num_cells) return;
  __shared__ H20Geom G;
  __shared__ double Q[8][13];  // per Gauss point: E w | w, mu w, c1 w, c2 w, F^-T
  const int* cn = cells + 20 * static_cast<long>(e);
  h20_geometry(x, y, z, cn, G);
  const int t = threadIdx.x;
  if (t < 8) {
    const double p = h20_param(mat, param, e, cn, G, t);
    if constexpr (NH) {
      double ug[3][3], FiT[3][3];
      h20_field_grad(u, nullptr, cn, G, t, ug);
      const double J = inv_transpose_F<double>(ug, FiT);
      const double kap = mat.kappa, w = G.w[t];
      Q[t][0] = w * p;
      Q[t][1] = w * kap * (2.0 * J - 1.0) * J;
      Q[t][2] = w * (-p + kap * J * (J - 1.0));
#pragma unroll
      for (int k = 0; k < 9; ++k) Q[t][3 + k] = FiT[k / 3][k % 3];
    } else {
      Q[t][0] = modulus_of<double>(mat, p) * G.w[t];
    }
  }
  __syncthreads();
  const double nu = mat.nu;
  const double lam0 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu0 = 1.0 / (2.0 * (1.0 + nu));
  for (int blk = t; blk < 400; blk += kH20Threads) {
    const int a = blk / 20, b = blk - 20 * (blk / 20);
    double K[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) K[k] = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double ga[3] = {G.g[q][a][0], G.g[q][a][1], G.g[q][a][2]};
      const double gb[3] = {G.g[q][b][0], G.g[q][b][1], G.g[q][b][2]};
      const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
      if constexpr (NH) {
        double fa[3], fb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          fa[i] = Q[q][3 + 3 * i] * ga[0] + Q[q][4 + 3 * i] * ga[1] + Q[q][5 + 3 * i] * ga[2];
          fb[i] = Q[q][3 + 3 * i] * gb[0] + Q[q][4 + 3 * i] * gb[1] + Q[q][5 + 3 * i] * gb[2];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            K[3 * i + j] += (i == j ? Q[q][0] * gg : 0.0) + Q[q][1] * fa[i] * fb[j] - Q[q][2] * fb[i] * fa[j];
      } else {
        const double Ew = Q[q][0];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            K[3 * i + j] += Ew * (lam0 * ga[i] * gb[j] + mu0 * ga[j] * gb[i] + (i == j ? mu0 * gg : 0.0));
      }
    }
    double* dst = ke + 3600 * static_cast<long>(e) + 9 * blk;
#pragma unroll
    for (int k = 0; k < 9; ++k) dst[k] = K[k];
  }
}

// One wave per element: r_(a,i) = sum_q sum_d sigma_q,id g_q,a,d JxW_q at re[60 e + 3 a + i].
__global__ __launch_bounds__(kH20Threads) void h20_elem_residual_kernel(int num_cells, const double* __restrict__ x,
                                                                        const double* __restrict__ y,
                                                                        const double* __restrict__ z,
                                                                        const int* __restrict__ cells, MatParams mat,
                                                                        const double* __restrict__ param,
                                                                        const double* __restrict__ u,
                                                                        double* __restrict__ re) {
  const int e = blockIdx.x;
  if (e >= num_cells) return;
  __shared__ H20Geom G;
  __shared__ double S[8][9];
  const int* cn = cells + 20 * static_cast<long>(e);
  h20_geometry(x, y, z, cn, G);
  const int t = threadIdx.x;
  if (t < 8) {
    double ug[3][3], sig[3][3];
    h20_field_grad(u, nullptr, cn, G, t, ug);
    material_stress<double>(mat, h20_param(mat, param, e, cn, G, t), ug, sig);
#pragma unroll
    for (int k = 0; k < 9; ++k) S[t][k] = sig[k / 3][k % 3] * G.w[t];
  }
  __syncthreads();
  if (t < 60) {
    const int a = t / 3, i = t - 3 * (t / 3);
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      r += S[q][3 * i] * G.g[q][a][0] + S[q][3 * i + 1] * G.g[q][a][1] + S[q][3 * i + 2] * G.g[q][a][2];
    re[60 * static_cast<long>(e) + t] = r;
  }
}

// Face quadrature of facet lf of element cn at point q: basis values N and
// the Nanson scale |n_ref J^-1| det J / 4.
__device__ inline double h20_face_point(const double* __restrict__ x, const double* __restrict__ y,
                                        const double* __restrict__ z, const int* __restrict__ cn, int lf, int q,
                                        double (&N)[20]) {
  double X[20][3];
  for (int a = 0; a < 20; ++a) {
    const int nd = cn[a];
    X[a][0] = x[nd];
    X[a][1] = y[nd];
    X[a][2] = z[nd];
  }
  double p0, p1, p2, dN[20][3], Ji[3][3], nrm[3];
  face_point(lf, q, p0, p1, p2);
  h20_shape(p0, p1, p2, N, dN);
  const double det = h20_jac_inv(X, dN, Ji);
  face_normal(lf, nrm);
  double m[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) m[i] = nrm[0] * Ji[0][i] + nrm[1] * Ji[1][i] + nrm[2] * Ji[2][i];
  return sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) * det * 0.25;
}

// One thread per node: the element residuals of its elements in ascending
// order, then its loaded faces: r_a += t_f sum_q svar[4 f + q] N_a(p_q) nanson_q.
__global__ __launch_bounds__(256) void h20_node_residual_kernel(
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
    const double* src = re + 60 * static_cast<long>(pk >> kH20Shift) + 3 * (pk & 31);
    r0 += src[0];
    r1 += src[1];
    r2 += src[2];
  }
  if (finc_ptr) {
    for (int k = finc_ptr[nd]; k < finc_ptr[nd + 1]; ++k) {
      const int pk = finc[k];
      const int f = pk >> kH20Shift, a = pk & 31;
      const int e = faces[2 * f], lf = faces[2 * f + 1];
      const int* cn = cells + 20 * static_cast<long>(e);
      double v = 0.0;
      for (int q = 0; q < 4; ++q) {
        double N[20];
        const double ns = h20_face_point(x, y, z, cn, lf, q, N);
        v += svar[4 * f + q] * N[a] * ns;
      }
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
// Per Gauss point of an element, w_q = dR_e/dp . lambda_e there:
//   linear / SIMP: sigma_1(grad u) : grad lambda JxW dE/dp,
//   Neo-Hookean:   (F - F^-T) : grad lambda JxW;
// SELF: lambda = u (the compliance case).  A cell variable gets
// scale * sum_q w_q, a node variable keeps we[8 e + q] for the node pass
// (scale * sum over its elements and Gauss points of N_a(q) w_q).
template <bool SELF>
__global__ __launch_bounds__(kH20Threads) void h20_param_vjp_kernel(
    int num_cells, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int* __restrict__ cells, MatParams mat, const double* __restrict__ param, const double* __restrict__ u,
    const double* __restrict__ lam, const uint8_t* __restrict__ bc_mask, double scale, double* __restrict__ out_cell,
    double* __restrict__ we) {
  const int e = blockIdx.x;
  if (e >= num_cells) return;
  __shared__ H20Geom G;
  __shared__ double W[8];
  const int* cn = cells + 20 * static_cast<long>(e);
  h20_geometry(x, y, z, cn, G);
  const int t = threadIdx.x;
  if (t < 8) {
    double ug[3][3], lg[3][3];
    h20_field_grad(u, nullptr, cn, G, t, ug);
    if constexpr (SELF) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) lg[i][d] = ug[i][d];
    } else {
      h20_field_grad(lam, bc_mask, cn, G, t, lg);
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
      modulus_of<double>(mat, h20_param(mat, param, e, cn, G, t), &dEdp);
      dens *= dEdp;
    }
    W[t] = dens * G.w[t];
  }
  __syncthreads();
  if (t == 0) {
    if (out_cell) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) s += W[q];
      out_cell[e] = scale * s;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) we[8 * static_cast<long>(e) + q] = W[q];
    }
  }
}

__global__ __launch_bounds__(256) void h20_param_vjp_node_kernel(int num_nodes, const int* __restrict__ inc_ptr,
                                                                 const int* __restrict__ inc,
                                                                 const double* __restrict__ we, double scale,
                                                                 double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num_nodes) return;
  double s = 0.0;
  for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
    const int pk = inc[k];
    const int e = pk >> kH20Shift, a = pk & 31;
    for (int q = 0; q < 8; ++q) {
      double p0, p1, p2, N[20], dN[20][3];
      gauss_point(q, p0, p1, p2);
      h20_shape(p0, p1, p2, N, dN);
      s += N[a] * we[8 * static_cast<long>(e) + q];
    }
  }
  out[i] = scale * s;
}

// One thread per (face, Gauss point): grad_svar[4 f + q] = scale * nanson_q *
// sum over the facet's 8 nodes of N_a(p_q) (dir_f . lambda_a).
__global__ __launch_bounds__(256) void h20_traction_vjp_kernel(int num_faces, const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ z,
                                                               const int* __restrict__ cells,
                                                               const int* __restrict__ faces,
                                                               const double* __restrict__ fdir,
                                                               const double* __restrict__ lam,
                                                               const uint8_t* __restrict__ bc_mask, double scale,
                                                               double* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 4 * num_faces) return;
  const int f = t >> 2, q = t & 3;
  const int e = faces[2 * f], lf = faces[2 * f + 1];
  const int* cn = cells + 20 * static_cast<long>(e);
  double N[20];
  const double ns = h20_face_point(x, y, z, cn, lf, q, N);
  double s = 0.0;
  for (int v = 0; v < 8; ++v) {
    const int a = kH20FaceNodes[lf][v];
    double dl = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int r = 3 * cn[a] + i;
      dl += fdir[3 * f + i] * ((bc_mask && bc_mask[r]) ? 0.0 : lam[r]);
    }
    s += N[a] * dl;
  }
  out[t] = scale * s * ns;
}

inline bool h20_law(const fx_material* m) {
  return m->kind == FX_MAT_LINEAR_ELASTIC || m->kind == FX_MAT_SIMP || m->kind == FX_MAT_NEO_HOOKEAN;
}

template <typename T>
int h20_assemble(void* stream, int num_nodes, int num_cells, const double* x, const double* y, const double* z,
                 const int* cells, const int* inc_ptr, const int* inc, const uint8_t* rank, const int* indptr,
                 const int* indices, const fx_material* mat, const double* param, const double* u,
                 const uint8_t* bc_mask, int symmetric, int flags, double jacobi_shift, double* dinv, T* data,
                 void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !rank || !indptr ||
      !indices || !mat || !param || !data)
    return fail(FX_EINVAL, "fx_hex20_assemble: bad arguments");
  if (!h20_law(mat)) return fail(FX_EINVAL, "fx_hex20_assemble: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_hex20_assemble: thermal strain is built for HEX8");
  const bool nh = mat->kind == FX_MAT_NEO_HOOKEAN;
  if (nh && !u) return fail(FX_EINVAL, "fx_hex20_assemble: the Neo-Hookean tangent needs u");
  const size_t need = 3600 * sizeof(double) * static_cast<size_t>(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex20_assemble: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  const int diff = (flags & FX_ASM_DIFF3) != 0;
  const MatParams mp = to_params(mat);
  hipStream_t s = as_stream(stream);
  double* ke = static_cast<double*>(workspace);
  if (nh)
    hipLaunchKernelGGL(h20_elem_matrix_kernel<true>, dim3(num_cells), dim3(kH20Threads), 0, s, num_cells, x, y, z,
                       cells, mp, param, u, ke);
  else
    hipLaunchKernelGGL(h20_elem_matrix_kernel<false>, dim3(num_cells), dim3(kH20Threads), 0, s, num_cells, x, y, z,
                       cells, mp, param, u, ke);
  FX_LAUNCHED("h20_elem_matrix_kernel");
  hipLaunchKernelGGL((gather_rows_kernel<T, double, 20, 3, kH20Shift, kH20MaxDeg>),
                     dim3((num_nodes + kNodesPerBlock - 1) / kNodesPerBlock), dim3(64 * kNodesPerBlock), 0, s,
                     num_nodes, nullptr, inc_ptr, inc, rank, indptr, indices, ke, bc_mask, symmetric, diff,
                     jacobi_shift, dinv, data, nullptr);
  FX_LAUNCHED("gather_rows_kernel (HEX20)");
  return FX_OK;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_hex20_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags) {
  (void)value_bytes;
  (void)flags;
  return num_cells < 0 ? 0 : 3600 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex20_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes) {
  return h20_assemble<double>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr,
                              indices, mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace,
                              ws_bytes);
}

int fx_hex20_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes) {
  return h20_assemble<float>(stream, num_nodes, num_cells, x, y, z, cells, inc_ptr, inc, elem_rank, indptr, indices,
                             mat, param, u, bc_mask, symmetric, flags, jacobi_shift, dinv, data, workspace,
                             ws_bytes);
}

size_t fx_hex20_residual_workspace_bytes(int32_t num_cells) {
  return num_cells < 0 ? 0 : 60 * sizeof(double) * static_cast<size_t>(num_cells);
}

int fx_hex20_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !inc_ptr || !inc || !mat || !param || !u ||
      !res)
    return fail(FX_EINVAL, "fx_hex20_residual_f64: bad arguments");
  if (!h20_law(mat)) return fail(FX_EINVAL, "fx_hex20_residual_f64: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_hex20_residual_f64: thermal strain is built for HEX8");
  if (num_faces > 0 && (!faces || !finc_ptr || !finc || !svar || !fdir))
    return fail(FX_EINVAL, "fx_hex20_residual_f64: face arrays missing");
  const size_t need = fx_hex20_residual_workspace_bytes(num_cells);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex20_residual_f64: workspace %zu < %zu", workspace ? ws_bytes : 0, need);
  hipStream_t s = as_stream(stream);
  const bool has_faces = num_faces > 0;
  double* re = static_cast<double*>(workspace);
  hipLaunchKernelGGL(h20_elem_residual_kernel, dim3(num_cells), dim3(kH20Threads), 0, s, num_cells, x, y, z, cells,
                     to_params(mat), param, u, re);
  hipLaunchKernelGGL(h20_node_residual_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, x, y, z,
                     cells, inc_ptr, inc, re, has_faces ? faces : nullptr, has_faces ? finc_ptr : nullptr,
                     has_faces ? finc : nullptr, svar, fdir, negate, res);
  FX_LAUNCHED("fx_hex20_residual_f64");
  if (num_bc > 0 && bc_rows) {
    if (!bc_vals) return fail(FX_EINVAL, "fx_hex20_residual_f64: bc_vals missing");
    hipLaunchKernelGGL(bc_res_kernel, dim3((num_bc + 255) / 256), dim3(256), 0, s, num_bc, bc_rows, bc_vals, u,
                       negate, res);
    FX_LAUNCHED("bc_res_kernel");
  }
  return FX_OK;
}

size_t fx_hex20_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc) {
  if (num_cells < 0) return 0;
  return param_loc == FX_PARAM_NODE ? 8 * static_cast<size_t>(num_cells) * sizeof(double) : 0;
}

int fx_hex20_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes) {
  if (num_nodes <= 0 || num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !lam || !grad)
    return fail(FX_EINVAL, "fx_hex20_param_vjp_f64: bad arguments");
  if (!h20_law(mat)) return fail(FX_EINVAL, "fx_hex20_param_vjp_f64: unknown material %d", mat->kind);
  if (mat->temperature) return fail(FX_EINVAL, "fx_hex20_param_vjp_f64: thermal strain is built for HEX8");
  if (mat->param_loc != FX_PARAM_NODE && mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_hex20_param_vjp_f64: unknown parameter location %d", mat->param_loc);
  hipStream_t s = as_stream(stream);
  const MatParams mp = to_params(mat);
  if (mat->param_loc == FX_PARAM_CELL) {
    hipLaunchKernelGGL(h20_param_vjp_kernel<false>, dim3(num_cells), dim3(kH20Threads), 0, s, num_cells, x, y, z,
                       cells, mp, param, u, lam, bc_mask, scale, grad, nullptr);
    FX_LAUNCHED("h20_param_vjp_kernel");
    return FX_OK;
  }
  if (!inc_ptr || !inc) return fail(FX_EINVAL, "fx_hex20_param_vjp_f64: node variables need the incidence");
  const size_t need = fx_hex20_param_vjp_workspace_bytes(num_cells, FX_PARAM_NODE);
  if (!workspace || ws_bytes < need)
    return fail(FX_EWORKSPACE, "fx_hex20_param_vjp_f64: workspace %zu < %zu bytes", ws_bytes, need);
  double* we = static_cast<double*>(workspace);
  hipLaunchKernelGGL(h20_param_vjp_kernel<false>, dim3(num_cells), dim3(kH20Threads), 0, s, num_cells, x, y, z,
                     cells, mp, param, u, lam, bc_mask, 1.0, nullptr, we);
  hipLaunchKernelGGL(h20_param_vjp_node_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc, we, scale, grad);
  FX_LAUNCHED("h20_param_vjp_node_kernel");
  return FX_OK;
}

int fx_hex20_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc) {
  if (num_cells <= 0 || !x || !y || !z || !cells || !mat || !param || !u || !dc)
    return fail(FX_EINVAL, "fx_hex20_simp_compliance_sens_f64: bad arguments");
  if (mat->kind != FX_MAT_SIMP && mat->kind != FX_MAT_LINEAR_ELASTIC)
    return fail(FX_EINVAL, "fx_hex20_simp_compliance_sens_f64: needs a linear law (SIMP or linear elastic)");
  if (mat->param_loc != FX_PARAM_CELL)
    return fail(FX_EINVAL, "fx_hex20_simp_compliance_sens_f64: needs a cell variable");
  hipLaunchKernelGGL(h20_param_vjp_kernel<true>, dim3(num_cells), dim3(kH20Threads), 0, as_stream(stream), num_cells,
                     x, y, z, cells, to_params(mat), param, u, nullptr, nullptr, -1.0, dc, nullptr);
  FX_LAUNCHED("h20_param_vjp_kernel");
  return FX_OK;
}

int fx_hex20_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar) {
  if (num_faces < 0 || (num_faces > 0 && (!x || !y || !z || !cells || !faces || !fdir || !lam || !grad_svar)))
    return fail(FX_EINVAL, "fx_hex20_traction_vjp_f64: bad arguments");
  if (num_faces == 0) return FX_OK;
  hipLaunchKernelGGL(h20_traction_vjp_kernel, dim3((4 * num_faces + 255) / 256), dim3(256), 0, as_stream(stream),
                     num_faces, x, y, z, cells, faces, fdir, lam, bc_mask, scale, grad_svar);
  FX_LAUNCHED("h20_traction_vjp_kernel");
  return FX_OK;
}

}  // extern "C"
