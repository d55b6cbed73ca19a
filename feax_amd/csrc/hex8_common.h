// HEX8 element helpers shared by the assembly, residual and adjoint kernels:
// trilinear shape functions on the reference cell [0,1]^3, 2x2x2 Gauss
// points, the isoparametric Jacobian, material parameters at Gauss points
// and the faces of the reference cell.
#pragma once

#include <cstdint>

#include "common.h"

namespace fx {

__device__ __forceinline__ double gauss_pt(int i) {
  // (1 -+ 1/sqrt(3)) / 2: the 2-point Gauss-Legendre rule on [0, 1].
  return i ? 0.78867513459481288225 : 0.21132486540518711775;
}

// Reference-cell corner of feax local node a (basix vertex order re-ordered
// by [0,1,3,2,4,5,7,6], feax/basis.py:65-66).
__host__ __device__ __forceinline__ void corner(int a, int& c0, int& c1, int& c2) {
  c0 = (a == 1 || a == 2 || a == 5 || a == 6);
  c1 = (a == 2 || a == 3 || a == 6 || a == 7);
  c2 = (a >= 4);
}

template <typename T>
__device__ __forceinline__ void shape(int a, T x0, T x1, T x2, T& N, T& d0, T& d1, T& d2) {
  int c0, c1, c2;
  corner(a, c0, c1, c2);
  const T f0 = c0 ? x0 : T(1) - x0, f1 = c1 ? x1 : T(1) - x1, f2 = c2 ? x2 : T(1) - x2;
  const T s0 = c0 ? T(1) : T(-1), s1 = c1 ? T(1) : T(-1), s2 = c2 ? T(1) : T(-1);
  N = f0 * f1 * f2;
  d0 = s0 * f1 * f2;
  d1 = f0 * s1 * f2;
  d2 = f0 * f1 * s2;
}

// Gauss point q of the 2x2x2 rule: coordinate k uses bit (2 - k) of q.
__device__ __forceinline__ void gauss_point(int q, double& x0, double& x1, double& x2) {
  x0 = gauss_pt(q >> 2);
  x1 = gauss_pt((q >> 1) & 1);
  x2 = gauss_pt(q & 1);
}

// J_ij = sum_c X_c,i dN_c/dxi_j; returns det J and writes J^-1 (feax/fe.py:96-129).
template <typename T>
__device__ __forceinline__ T jac_inv(const T (&X)[8][3], T x0, T x1, T x2, T (&Ji)[3][3]) {
  T J[3][3] = {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    T N, d0, d1, d2;
    shape<T>(c, x0, x1, x2, N, d0, d1, d2);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      J[i][0] += X[c][i] * d0;
      J[i][1] += X[c][i] * d1;
      J[i][2] += X[c][i] * d2;
    }
  }
  const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const T det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const T id = T(1) / det;
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

// Physical gradient g_i = sum_j dN/dxi_j (J^-1)_ji.
template <typename T>
__device__ __forceinline__ void phys_grad(const T (&Ji)[3][3], T d0, T d1, T d2, T (&g)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) g[i] = d0 * Ji[0][i] + d1 * Ji[1][i] + d2 * Ji[2][i];
}

struct MatParams {
  int kind, loc;
  double nu, E0, Emin, penal, kappa;
  const double* temp;  // thermal strain: node temperatures, or nullptr
  double alpha, tref;
};

inline MatParams to_params(const fx_material* m) {
  MatParams p;
  p.kind = m->kind;
  p.loc = m->param_loc;
  p.nu = m->nu;
  p.E0 = m->E0;
  p.Emin = m->Emin;
  p.penal = m->penal;
  p.kappa = m->kappa;
  p.temp = m->temperature;
  p.alpha = m->alpha;
  p.tref = m->t_ref;
  return p;
}

// Element coordinates relative to local node 0 (differences taken in FP64,
// so FP32 geometry does not lose digits to large absolute coordinates).
template <typename T>
__device__ __forceinline__ void load_coords(const double* __restrict__ x, const double* __restrict__ y,
                                            const double* __restrict__ z, const int* __restrict__ cn,
                                            T (&X)[8][3]) {
  const int n0 = cn[0];
  const double ox = x[n0], oy = y[n0], oz = z[n0];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int nd = cn[c];
    X[c][0] = static_cast<T>(x[nd] - ox);
    X[c][1] = static_cast<T>(y[nd] - oy);
    X[c][2] = static_cast<T>(z[nd] - oz);
  }
}

// Volume variable at the Gauss point (feax/assembler.py:227-269: node vars
// interpolated with N, cell vars broadcast).
template <typename T>
__device__ __forceinline__ T param_at(const MatParams& m, const double* __restrict__ param, int e,
                                      const int* __restrict__ cn, T x0, T x1, T x2) {
  if (m.loc == FX_PARAM_NODE) {
    T p = T(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      T N, d0, d1, d2;
      shape<T>(c, x0, x1, x2, N, d0, d1, d2);
      p += N * static_cast<T>(param[cn[c]]);
    }
    return p;
  }
  return static_cast<T>(param[e]);
}

// Young's modulus from the parameter: E = p (linear elastic) or the SIMP
// law E = Emin + p^penal (E0 - Emin); dE/dp in *dEdp when given.
template <typename T>
__device__ __forceinline__ T modulus_of(const MatParams& m, T p, T* dEdp = nullptr) {
  if (m.kind == FX_MAT_SIMP) {
    const T pen = static_cast<T>(m.penal), span = static_cast<T>(m.E0 - m.Emin);
    if (dEdp) *dEdp = pen * pow(p, pen - T(1)) * span;
    return static_cast<T>(m.Emin) + pow(p, pen) * span;
  }
  if (dEdp) *dEdp = T(1);
  return p;
}

template <typename T>
__device__ __forceinline__ T modulus_at(const MatParams& m, const double* __restrict__ param, int e,
                                        const int* __restrict__ cn, T x0, T x1, T x2) {
  return modulus_of<T>(m, param_at<T>(m, param, e, cn, x0, x1, x2));
}

// Fused multiply-add in the element's arithmetic type.
__device__ __forceinline__ double fma_c(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_c(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ T xor_sum(T v, int mask) {
  return v + __shfl_xor(v, mask, 64);
}

// F^-T of F = I + ug (ug[i][d] = du_i/dX_d); returns J = det F.
template <typename T>
__device__ __forceinline__ T inv_transpose_F(const T (&ug)[3][3], T (&FiT)[3][3]) {
  T F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) F[i][d] = ug[i][d] + (i == d ? T(1) : T(0));
  const T c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
  const T c01 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
  const T c02 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
  const T J = F[0][0] * c00 + F[0][1] * c01 + F[0][2] * c02;
  const T id = T(1) / J;
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

// Stress S[i][d] (contracted with dN/dx_d in the residual) at a Gauss point
// with volume variable p:
//   linear / SIMP   sigma = lam tr(eps) I + 2 mu eps, E = modulus_of(p)
//                   (examples/benchmark/bench_linear_elasticity.py:327-339);
//   Neo-Hookean     P = mu (F - F^-T) + kappa (J - 1) J F^-T, mu = p
//                   (tests/test_newton_solver.py:65-89 of feax).
template <typename T>
__device__ __forceinline__ void material_stress(const MatParams& m, T p, const T (&ug)[3][3], T (&S)[3][3]) {
  if (m.kind == FX_MAT_NEO_HOOKEAN) {
    T FiT[3][3];
    const T J = inv_transpose_F<T>(ug, FiT);
    const T kj = static_cast<T>(m.kappa) * (J - T(1)) * J;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) S[i][d] = p * (ug[i][d] + (i == d ? T(1) : T(0)) - FiT[i][d]) + kj * FiT[i][d];
    return;
  }
  const T E = modulus_of<T>(m, p);
  const T nu = static_cast<T>(m.nu);
  const T lam0 = nu / ((T(1) + nu) * (T(1) - T(2) * nu)), mu0 = T(1) / (T(2) * (T(1) + nu));
  const T lam = E * lam0, mu = E * mu0;
  const T tr = ug[0][0] + ug[1][1] + ug[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) S[i][d] = mu * (ug[i][d] + ug[d][i]) + (i == d ? lam * tr : T(0));
}

// Thermal strain of the linear / SIMP laws: sigma = C : (eps - alpha dT I)
// shifts the stress by -kth I, kth = E (3 lam0 + 2 mu0) alpha (T_q - t_ref)
// (the thermo-elastic tensor map of feax's examples/basic/
// thermo_mech_coupling.py:95-103).  0 without a temperature field.
template <typename T>
__device__ __forceinline__ T thermal_shift(const MatParams& m, T p, const int* __restrict__ cn, T x0, T x1, T x2,
                                           T* dkth_dE = nullptr) {
  if (!m.temp) return T(0);
  T tq = T(0);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    T N, d0, d1, d2;
    shape<T>(c, x0, x1, x2, N, d0, d1, d2);
    tq += N * static_cast<T>(m.temp[cn[c]]);
  }
  const T nu = static_cast<T>(m.nu);
  const T c3 = T(3) * nu / ((T(1) + nu) * (T(1) - T(2) * nu)) + T(1) / (T(1) + nu);  // 3 lam0 + 2 mu0
  const T per_E = c3 * static_cast<T>(m.alpha) * (tq - static_cast<T>(m.tref));
  if (dkth_dE) *dkth_dE = per_E;
  return modulus_of<T>(m, p) * per_E;
}

// Face quadrature point (lf, q) in the reference cell: the 2x2 Gauss points
// of the unit square mapped onto basix facet lf (feax/basis.py:286-305).
__device__ __forceinline__ void face_point(int lf, int q, double& p0, double& p1, double& p2) {
  const double s = gauss_pt(q >> 1), t = gauss_pt(q & 1);
  switch (lf) {
    case 0: p0 = s; p1 = t; p2 = 0.0; break;
    case 1: p0 = s; p1 = 0.0; p2 = t; break;
    case 2: p0 = 0.0; p1 = s; p2 = t; break;
    case 3: p0 = 1.0; p1 = s; p2 = t; break;
    case 4: p0 = s; p1 = 1.0; p2 = t; break;
    default: p0 = s; p1 = t; p2 = 1.0; break;
  }
}

__device__ __forceinline__ void face_normal(int lf, double (&nrm)[3]) {
  nrm[0] = nrm[1] = nrm[2] = 0.0;
  switch (lf) {
    case 0: nrm[2] = -1.0; break;
    case 1: nrm[1] = -1.0; break;
    case 2: nrm[0] = -1.0; break;
    case 3: nrm[0] = 1.0; break;
    case 4: nrm[1] = 1.0; break;
    default: nrm[2] = 1.0; break;
  }
}

// Nanson scale |n_ref J^-1| det J w at face point (p0, p1, p2) of face lf,
// w = 1/4 (2x2 Gauss on the unit square) (feax/fe.py:175-239).
__device__ __forceinline__ double face_scale(const double (&X)[8][3], int lf, double p0, double p1, double p2) {
  double Ji[3][3];
  const double det = jac_inv<double>(X, p0, p1, p2, Ji);
  double nrm[3];
  face_normal(lf, nrm);
  double m[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) m[i] = nrm[0] * Ji[0][i] + nrm[1] * Ji[1][i] + nrm[2] * Ji[2][i];
  return sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) * det * 0.25;
}

}  // namespace fx
