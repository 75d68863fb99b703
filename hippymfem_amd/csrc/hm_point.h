// Pointwise quantities of the element kernels, FP64: the geometry of the element map at a
// quadrature point, the fields there, and the density's second derivatives in the
// reference-gradient basis.
#pragma once

#include <hip/hip_runtime.h>

#include "hm_common.h"

namespace hm {

// Orders a wavefront's LDS writes before its later reads (the lanes of one wavefront share
// element data in LDS; no block barrier).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Everything one quadrature point of one element needs, in the point type PT: FP64, or FP32
// in the precision forms (whose geometry is made in FP64 and rounded once).
template <typename PT>
struct PointStateT {
  PT w;           // w_q |det J|
  PT Ji[3][3];    // Jinv[k][d] = dxi_k / dx_d
  PT K[3][3];     // sum_d Jinv[k][d] Jinv[l][d], the weights of the reference-gradient pairs
  PT kap, g, g1, g2;
  PT gu[3], gp[3];  // physical gradients of u and p
  PT dot;           // grad u . grad p
  PT uq;            // value of u
};
using PointState = PointStateT<double>;

// The geometry of the element map at a point (of the whole element where the map is affine):
// |det J| in FP64, J^-1 and the pair weights K in PT, each rounded once from FP64.
template <typename PT>
struct GeomT {
  double adet;
  PT Ji[3][3];
  PT K[3][3];
};

// The geometry from the Jacobian of the element map J[d][k] = dx_d / dxi_k: its determinant and
// inverse in FP64, whatever PT is.
template <typename PT>
__device__ inline void geom_from_J(const double J[3][3], GeomT<PT>& G) {
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  // inverse of J (rows k = reference direction, columns d = physical)
  G.Ji[0][0] = (PT)(c00 * id);
  G.Ji[1][0] = (PT)(c01 * id);
  G.Ji[2][0] = (PT)(c02 * id);
  G.Ji[0][1] = (PT)((J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id);
  G.Ji[1][1] = (PT)((J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id);
  G.Ji[2][1] = (PT)((J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id);
  G.Ji[0][2] = (PT)((J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id);
  G.Ji[1][2] = (PT)((J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id);
  G.Ji[2][2] = (PT)((J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id);
  G.adet = fabs(det);
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l)
      G.K[k][l] = G.Ji[k][0] * G.Ji[l][0] + G.Ji[k][1] * G.Ji[l][1] + G.Ji[k][2] * G.Ji[l][2];
}

// The geometry of an axis-aligned box element, J = diag(h0, h1, h2) (the edge lengths): the
// diagonal terms of geom_from_J's cofactor formulas in FP64, the other terms zero.  The element's
// geometry once, as the reference keeps one J^-1 an element for affine groups
// (hippymfem/fem/elementbatch.py:206-223), instead of the trilinear map's multiply-adds and the
// full inverse at every point.
template <typename PT>
__device__ inline void geom_diag(double h0, double h1, double h2, GeomT<PT>& G) {
  const double c0 = h1 * h2, c1 = h0 * h2, c2 = h0 * h1;
  const double det = h0 * c0;
  const double id = 1.0 / det;
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      G.Ji[k][l] = (PT)0;
      G.K[k][l] = (PT)0;
    }
  G.Ji[0][0] = (PT)(c0 * id);
  G.Ji[1][1] = (PT)(c1 * id);
  G.Ji[2][2] = (PT)(c2 * id);
  G.adet = fabs(det);
  for (int k = 0; k < 3; ++k) G.K[k][k] = G.Ji[k][k] * G.Ji[k][k];
}

// The point state from the geometry at the point, the point's quadrature weight and the
// fields' values and reference gradients there, all after the geometry in PT.
template <typename PT>
__device__ inline void point_fields(const hm_problem& P, double wq, const GeomT<PT>& G, PT uq, const PT ru[3],
                                    const PT rp[3], PT mq, PointStateT<PT>& S) {
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      S.Ji[k][l] = G.Ji[k][l];
      S.K[k][l] = G.K[k][l];
    }
  S.w = (PT)(wq * G.adet);
  for (int d = 0; d < 3; ++d) {
    S.gu[d] = S.Ji[0][d] * ru[0] + S.Ji[1][d] * ru[1] + S.Ji[2][d] * ru[2];
    S.gp[d] = S.Ji[0][d] * rp[0] + S.Ji[1][d] * rp[1] + S.Ji[2][d] * rp[2];
  }
  S.dot = S.gu[0] * S.gp[0] + S.gu[1] * S.gp[1] + S.gu[2] * S.gp[2];
  S.kap = exp(mq);
  S.uq = uq;
  if (P.density == HM_DENSITY_NONLINEAR) {
    S.g = (PT)1 + uq * uq;
    S.g1 = (PT)2 * uq;
    S.g2 = (PT)2;
  } else {
    S.g = (PT)1;
    S.g1 = (PT)0;
    S.g2 = (PT)0;
  }
}

// The point state from the Jacobian of the element map at the point.
template <typename PT>
__device__ inline void point_finish(const hm_problem& P, double wq, const double J[3][3], PT uq, const PT ru[3],
                                    const PT rp[3], PT mq, PointStateT<PT>& S) {
  GeomT<PT> G;
  geom_from_J<PT>(J, G);
  point_fields<PT>(P, wq, G, uq, ru, rp, mq, S);
}

// tab_x / tab_m / tab_u may point at LDS copies of the problem's tables (the default: the
// problem's own, in global memory).  PT as point_finish: the fields in PT (dof values and tables
// rounded as read), the geometry in FP64.
template <typename PT>
__device__ inline void point_state(const hm_problem& P, int q, const double* Xe,
                                   const double* ue, const double* me, const double* pe,
                                   PointStateT<PT>& S, bool fields = true,
                                   const double* tab_x = nullptr, const double* tab_m = nullptr,
                                   const double* tab_u = nullptr) {
  const int nq = P.nq;
  // J[d][k] = sum_v X_v[d] dN_v/dxi_k (trilinear map)
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double* gx = (tab_x ? tab_x : P.tab_x) + (size_t)q * 24;
  // loops unrolled by two: fully unrolled, every operand load is in flight at once and the
  // registers they take limit the wavefronts a compute unit holds
#pragma unroll 2
  for (int v = 0; v < 8; ++v)
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) J[d][k] = fma(Xe[v * 3 + d], gx[v * 3 + k], J[d][k]);
  const int nu = P.nd_u, nm = P.nd_m;
  const double* tu = tab_u ? tab_u : P.tab_u;
  const double* Nu = tu + (size_t)q * nu;
  const double* Gu = tu + (size_t)nq * nu + (size_t)q * nu * 3;
  const double* Nm = (tab_m ? tab_m : P.tab_m) + (size_t)q * nm;
  PT uq = 0, mq = 0, ru[3] = {0, 0, 0}, rp[3] = {0, 0, 0};
#pragma unroll 2
  for (int a = 0; a < (fields ? nu : 0); ++a) {
    uq = fma((PT)Nu[a], (PT)ue[a], uq);
    for (int k = 0; k < 3; ++k) {
      ru[k] = fma((PT)Gu[a * 3 + k], (PT)ue[a], ru[k]);
      rp[k] = fma((PT)Gu[a * 3 + k], (PT)pe[a], rp[k]);
    }
  }
  for (int a = 0; a < nm; ++a) mq = fma((PT)Nm[a], (PT)me[a], mq);
  point_finish<PT>(P, P.wq[q], J, uq, ru, rp, mq, S);
}

// The point state of element e at point q with every operand read from global memory as it
// is used (no per-thread staging arrays).
__device__ inline void point_state_stream(const hm_problem& P, int64_t e, int q, bool fields, PointState& S) {
  const int nq = P.nq;
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double* gx = P.tab_x + (size_t)q * 24;
  for (int v = 0; v < 8; ++v) {
    const int32_t iv = P.elem_verts[e * 8 + v];
    const double x0 = P.verts[(size_t)iv * 3], x1 = P.verts[(size_t)iv * 3 + 1], x2 = P.verts[(size_t)iv * 3 + 2];
    for (int k = 0; k < 3; ++k) {
      const double g = gx[v * 3 + k];
      J[0][k] = fma(x0, g, J[0][k]);
      J[1][k] = fma(x1, g, J[1][k]);
      J[2][k] = fma(x2, g, J[2][k]);
    }
  }
  const int nu = P.nd_u, nm = P.nd_m;
  double uq = 0, mq = 0, ru[3] = {0, 0, 0}, rp[3] = {0, 0, 0};
  if (fields) {
    const double* Nu = P.tab_u + (size_t)q * nu;
    const double* Gu = P.tab_u + (size_t)nq * nu + (size_t)q * nu * 3;
    for (int a = 0; a < nu; ++a) {
      const int32_t d = P.edofs_u[e * nu + a];
      const double ua = P.u[d], pa = P.p[d];
      uq = fma(Nu[a], ua, uq);
      for (int k = 0; k < 3; ++k) {
        ru[k] = fma(Gu[a * 3 + k], ua, ru[k]);
        rp[k] = fma(Gu[a * 3 + k], pa, rp[k]);
      }
    }
  }
  const double* Nm = P.tab_m + (size_t)q * nm;
  for (int a = 0; a < nm; ++a) mq = fma(Nm[a], P.m[P.edofs_m[e * nm + a]], mq);
  point_finish(P, P.wq[q], J, uq, ru, rp, mq, S);
}

// The pointwise block of test slot s and trial slot j, in the reference-gradient basis,
// D[r][c] with r, c in {value, d/dxi_0, d/dxi_1, d/dxi_2}, times w.  The closed forms of
// the density r = exp(m) g(u) grad u . grad p - f p (second derivatives per part).
template <typename PT>
__device__ inline void point_block(int s, int j, const PointStateT<PT>& S, PT D[4][4]) {
  PT vv = 0, vg[3] = {0, 0, 0}, gv[3] = {0, 0, 0}, gg = 0;
  const PT kg = S.kap * S.g, kg1 = S.kap * S.g1;
  const int key = s * 3 + j;
  switch (key) {
    case HM_STATE * 3 + HM_STATE:
      vv = S.kap * S.g2 * S.dot;
      for (int d = 0; d < 3; ++d) vg[d] = gv[d] = kg1 * S.gp[d];
      break;
    case HM_STATE * 3 + HM_PARAMETER:
      vv = kg1 * S.dot;
      for (int d = 0; d < 3; ++d) gv[d] = kg * S.gp[d];
      break;
    case HM_PARAMETER * 3 + HM_STATE:
      vv = kg1 * S.dot;
      for (int d = 0; d < 3; ++d) vg[d] = kg * S.gp[d];
      break;
    case HM_STATE * 3 + HM_ADJOINT:
      for (int d = 0; d < 3; ++d) vg[d] = kg1 * S.gu[d];
      gg = kg;
      break;
    case HM_ADJOINT * 3 + HM_STATE:
      for (int d = 0; d < 3; ++d) gv[d] = kg1 * S.gu[d];
      gg = kg;
      break;
    case HM_PARAMETER * 3 + HM_PARAMETER:
      vv = kg * S.dot;
      break;
    case HM_PARAMETER * 3 + HM_ADJOINT:
      for (int d = 0; d < 3; ++d) vg[d] = kg * S.gu[d];
      break;
    case HM_ADJOINT * 3 + HM_PARAMETER:
      for (int d = 0; d < 3; ++d) gv[d] = kg * S.gu[d];
      break;
    default:
      break;
  }
  const PT w = S.w;
  D[0][0] = w * vv;
  for (int k = 0; k < 3; ++k) {
    D[0][1 + k] = w * (S.Ji[k][0] * vg[0] + S.Ji[k][1] * vg[1] + S.Ji[k][2] * vg[2]);
    D[1 + k][0] = w * (S.Ji[k][0] * gv[0] + S.Ji[k][1] * gv[1] + S.Ji[k][2] * gv[2]);
  }
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) D[1 + k][1 + l] = w * gg * S.K[k][l];
}

// Which parts of the (s, j) block can be nonzero: bit 0 the value row, bit 1 the gradient
// rows, bit 2 the value column, bit 3 the gradient columns (the reference skips the
// identically zero pointwise blocks the same way, kernel.py:1606-1646).
__host__ __device__ inline int block_parts(int s, int j, int density) {
  const bool nl = density == HM_DENSITY_NONLINEAR;
  bool vv = false, vg = false, gv = false, gg = false;
  switch (s * 3 + j) {
    case 0: vv = vg = gv = nl; break;                    // (u, u)
    case 1: vv = nl; gv = true; break;                   // (u, m)
    case 3: vv = nl; vg = true; break;                   // (m, u)
    case 2: vg = nl; gg = true; break;                   // (u, p)
    case 6: gv = nl; gg = true; break;                   // (p, u)
    case 4: vv = true; break;                            // (m, m)
    case 5: vg = true; break;                            // (m, p)
    case 7: gv = true; break;                            // (p, m)
    default: break;
  }
  return ((vv || vg) ? 1 : 0) | ((gv || gg) ? 2 : 0) | ((vv || gv) ? 4 : 0) | ((vg || gg) ? 8 : 0);
}

}  // namespace hm
