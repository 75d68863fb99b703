// Element kernels of the assembly path, FP64: geometry of the element map at every
// quadrature point, the fields there, the density's pointwise second derivatives and the
// contraction E = Psi_t^T D Psi_c, plus the element vectors and the element-wise products.
//
// Formulation.  The reference's quadrature route (hippymfem/fem/kernel.py:1431-1485) forms
// D[s][r][c] (value/gradient parts of the density's second derivatives) and contracts it
// with the physical basis Psi = (N, G Jinv).  Here the inverse Jacobian is folded into D
// instead, so the basis is the reference one, identical for every element and held in LDS:
//   D^_vv = w vv,  D^_vg = w Jinv vg,  D^_gv = w Jinv gv,  D^_gg = w c Jinv Jinv^T,
// with w = w_q |det J_q|.  The two are equal in exact arithmetic.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "hm_common.h"
#include "hm_device.h"
#include "hm_point.h"

namespace hm {

// Loads the element's vertex coordinates and its three dof vectors into LDS.
__device__ inline void load_element(const hm_problem& P, int64_t e, double* Xe, double* ue,
                                    double* me, double* pe) {
  for (int i = threadIdx.x; i < 24; i += blockDim.x) {
    const int32_t v = P.elem_verts[e * 8 + i / 3];
    Xe[i] = P.verts[(size_t)v * 3 + i % 3];
  }
  for (int a = threadIdx.x; a < P.nd_u; a += blockDim.x) {
    const int32_t d = P.edofs_u[e * P.nd_u + a];
    ue[a] = P.u[d];
    pe[a] = P.p[d];
  }
  for (int a = threadIdx.x; a < P.nd_m; a += blockDim.x) me[a] = P.m[P.edofs_m[e * P.nd_m + a]];
}

// Psi[q][r][a] in LDS from N (nq, nd) and G (nq, nd, 3).
__device__ inline void load_basis(const double* tab, int nq, int nd, double* Psi) {
  for (int i = threadIdx.x; i < nq * 4 * nd; i += blockDim.x) {
    const int a = i % nd, r = (i / nd) % 4, q = i / (4 * nd);
    Psi[i] = r == 0 ? tab[q * nd + a] : tab[(size_t)nq * nd + ((size_t)q * nd + a) * 3 + (r - 1)];
  }
}

template <typename OutT>
__global__ void __launch_bounds__(256) k_element_matrices(hm_problem P, int s, int j, int64_t e0,
                                                          int64_t e1, int parts, OutT* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const int nd_t = s == HM_PARAMETER ? P.nd_m : P.nd_u;
  const int nd_c = j == HM_PARAMETER ? P.nd_m : P.nd_u;
  const double* tab_t = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  const double* tab_c = j == HM_PARAMETER ? P.tab_m : P.tab_u;
  const bool same = (tab_t == tab_c);
  double* Pt = lds;
  double* Pc = same ? Pt : Pt + nq * 4 * nd_t;
  double* T = Pc + (same ? nq * 4 * nd_t : nq * 4 * nd_c);
  double* D = T + nq * 4 * nd_c;
  double* Xe = D + nq * 16;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  load_basis(tab_t, nq, nd_t, Pt);
  if (!same) load_basis(tab_c, nq, nd_c, Pc);
  const int r0 = (parts & 1) ? 0 : 1, r1 = (parts & 2) ? 4 : 1;
  const int c0 = (parts & 4) ? 0 : 1, c1 = (parts & 8) ? 4 : 1;
  for (int64_t e = e0 + blockIdx.x; e < e1; e += gridDim.x) {
    __syncthreads();
    load_element(P, e, Xe, ue, me, pe);
    __syncthreads();
    // (the FP32 form: fields, kappa, the pairs and the contraction in FP32, the geometry in FP64)
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
      PointStateT<OutT> S;
      point_state<OutT>(P, q, Xe, ue, me, pe, S);
      OutT Dq[4][4];
      point_block(s, j, S, Dq);
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) D[q * 16 + r * 4 + c] = Dq[r][c];
    }
    __syncthreads();
    // T[q][r][b] = sum_c D[q][r][c] Psi_c[q][c][b]
    for (int i = threadIdx.x; i < nq * 4 * nd_c; i += blockDim.x) {
      const int b = i % nd_c, r = (i / nd_c) % 4, q = i / (4 * nd_c);
      OutT acc = 0;
      if (r >= r0 && r < r1)
        for (int c = c0; c < c1; ++c) acc = fma((OutT)D[q * 16 + r * 4 + c], (OutT)Pc[(q * 4 + c) * nd_c + b], acc);
      T[i] = acc;
    }
    __syncthreads();
    // E[a][b] = sum_q sum_r Psi_t[q][r][a] T[q][r][b]
    OutT* Eo = out + (size_t)(e - e0) * nd_t * nd_c;
    for (int i = threadIdx.x; i < nd_t * nd_c; i += blockDim.x) {
      const int a = i / nd_c, b = i % nd_c;
      OutT acc = 0;
      for (int q = 0; q < nq; ++q)
        for (int r = r0; r < r1; ++r) acc = fma((OutT)Pt[(q * 4 + r) * nd_t + a], (OutT)T[(q * 4 + r) * nd_c + b], acc);
      Eo[i] = acc;
    }
  }
}

// The element kernels of the vectors and the element-wise products give each wavefront
// 64 / L elements, L = max(nq, nd_u, nd_m) lanes each.  The lanes of an element load its data
// into the wavefront's own LDS region, act as its quadrature points (the pointwise weights
// into LDS), then as its dofs (the sum over the points), with no block barrier in the loop.
// The basis Psi[q][r][a] of the block's slots is in LDS for the whole block.
struct WaveElement {
  int L, EPW, slot, r;
  bool in_use;
};

__device__ inline WaveElement wave_element(const hm_problem& P) {
  WaveElement w;
  w.L = P.nq;
  if (P.nd_u > w.L) w.L = P.nd_u;
  if (P.nd_m > w.L) w.L = P.nd_m;
  w.EPW = 64 / w.L;
  const int lane = threadIdx.x & 63;
  w.slot = lane / w.L;
  w.r = lane % w.L;
  w.in_use = w.slot < w.EPW;
  return w;
}

// The element's vertex coordinates and dof vectors into the wavefront's region, lane r
// taking vertex r and dofs r.
__device__ inline void load_element_lanes(const hm_problem& P, int64_t e, int r, double* Xe, double* ue,
                                          double* me, double* pe) {
  if (r < 8) {
    const int32_t v = P.elem_verts[e * 8 + r];
    Xe[r * 3] = P.verts[(size_t)v * 3];
    Xe[r * 3 + 1] = P.verts[(size_t)v * 3 + 1];
    Xe[r * 3 + 2] = P.verts[(size_t)v * 3 + 2];
  }
  if (r < P.nd_u) {
    const int32_t d = P.edofs_u[e * P.nd_u + r];
    ue[r] = P.u[d];
    pe[r] = P.p[d];
  }
  if (r < P.nd_m) me[r] = P.m[P.edofs_m[e * P.nd_m + r]];
}

// Element vectors dR_e / d(slot i dofs): sum_q w [val N_a + (Jinv grad) . G_a].
__global__ void __launch_bounds__(256) k_element_vectors(hm_problem P, int s, int64_t e0, int64_t e1,
                                                         double* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const int nd = s == HM_PARAMETER ? P.nd_m : P.nd_u;
  const double* tab = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  double* Ps = lds;
  load_basis(tab, nq, nd, Ps);
  __syncthreads();
  const WaveElement w = wave_element(P);
  const int words = nq * 4 + 24 + 2 * P.nd_u + P.nd_m;
  double* W = Ps + nq * 4 * nd + (size_t)((threadIdx.x >> 6) * w.EPW + (w.in_use ? w.slot : 0)) * words;
  double* Xe = W + nq * 4;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * w.EPW;
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * w.EPW; base < ne;
       base += stride) {
    const int64_t el = base + w.slot;
    const bool act = w.in_use && el < ne;
    if (act) load_element_lanes(P, e0 + el, w.r, Xe, ue, me, pe);
    wave_sync();
    if (act && w.r < nq) {
      const int q = w.r;
      PointState S;
      point_state(P, q, Xe, ue, me, pe, S);
      double val = 0, gr[3] = {0, 0, 0};
      if (s == HM_ADJOINT) {
        val = -P.f;
        for (int d = 0; d < 3; ++d) gr[d] = S.kap * S.g * S.gu[d];
      } else if (s == HM_STATE) {
        val = S.kap * S.g1 * S.dot;
        for (int d = 0; d < 3; ++d) gr[d] = S.kap * S.g * S.gp[d];
      } else {
        val = S.kap * S.g * S.dot;
      }
      W[q * 4] = S.w * val;
      for (int k = 0; k < 3; ++k)
        W[q * 4 + 1 + k] = S.w * (S.Ji[k][0] * gr[0] + S.Ji[k][1] * gr[1] + S.Ji[k][2] * gr[2]);
    }
    wave_sync();
    if (act && w.r < nd) {
      const int a = w.r;
      double acc = 0;
      for (int q = 0; q < nq; ++q)
        for (int r = 0; r < 4; ++r) acc = fma(Ps[(q * 4 + r) * nd + a], W[q * 4 + r], acc);
      out[(size_t)el * nd + a] = acc;
    }
    // the region is rewritten by the next element
    wave_sync();
  }
}

// Block (s, j) applied to the slot-j direction, element by element:
// y_a = sum_q Psi_t[q][:][a] . D[q] (Psi_c[q] d).
__global__ void __launch_bounds__(256) k_element_hvp(hm_problem P, int s, int j, const double* dir,
                                                     int64_t e0, int64_t e1, double* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const int nd_t = s == HM_PARAMETER ? P.nd_m : P.nd_u;
  const int nd_c = j == HM_PARAMETER ? P.nd_m : P.nd_u;
  const double* tab_t = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  const double* tab_c = j == HM_PARAMETER ? P.tab_m : P.tab_u;
  const int32_t* ed_c = j == HM_PARAMETER ? P.edofs_m : P.edofs_u;
  const bool same = tab_t == tab_c;
  double* Pt = lds;
  double* Pc = same ? Pt : Pt + nq * 4 * nd_t;
  load_basis(tab_t, nq, nd_t, Pt);
  if (!same) load_basis(tab_c, nq, nd_c, Pc);
  __syncthreads();
  const WaveElement w = wave_element(P);
  const int words = nq * 4 + nd_c + 24 + 2 * P.nd_u + P.nd_m;
  double* W = Pc + nq * 4 * nd_c + (size_t)((threadIdx.x >> 6) * w.EPW + (w.in_use ? w.slot : 0)) * words;
  double* de = W + nq * 4;
  double* Xe = de + nd_c;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * w.EPW;
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * w.EPW; base < ne;
       base += stride) {
    const int64_t el = base + w.slot;
    const bool act = w.in_use && el < ne;
    if (act) {
      const int64_t e = e0 + el;
      load_element_lanes(P, e, w.r, Xe, ue, me, pe);
      if (w.r < nd_c) de[w.r] = dir[ed_c[e * nd_c + w.r]];
    }
    wave_sync();
    if (act && w.r < nq) {
      const int q = w.r;
      PointState S;
      point_state(P, q, Xe, ue, me, pe, S);
      double Dq[4][4];
      point_block(s, j, S, Dq);
      double z[4] = {0, 0, 0, 0};
      for (int c = 0; c < 4; ++c)
        for (int b = 0; b < nd_c; ++b) z[c] = fma(Pc[(q * 4 + c) * nd_c + b], de[b], z[c]);
      for (int r = 0; r < 4; ++r) {
        double acc = 0;
        for (int c = 0; c < 4; ++c) acc = fma(Dq[r][c], z[c], acc);
        W[q * 4 + r] = acc;
      }
    }
    wave_sync();
    if (act && w.r < nd_t) {
      const int a = w.r;
      double acc = 0;
      for (int q = 0; q < nq; ++q)
        for (int r = 0; r < 4; ++r) acc = fma(Pt[(q * 4 + r) * nd_t + a], W[q * 4 + r], acc);
      out[(size_t)el * nd_t + a] = acc;
    }
    wave_sync();
  }
}

// Element scalars R_e = sum_q w (exp(m) g(u) grad u . grad p - f p): the functional every
// derivative above is taken of.  Lanes act as the element's points (the weighted density
// value into the wavefront's region), then its first lane sums them in point order.
__global__ void __launch_bounds__(256) k_element_scalars(hm_problem P, int64_t e0, int64_t e1, double* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const WaveElement w = wave_element(P);
  const int words = nq + 24 + 2 * P.nd_u + P.nd_m;
  double* W = lds + (size_t)((threadIdx.x >> 6) * w.EPW + (w.in_use ? w.slot : 0)) * words;
  double* Xe = W + nq;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * w.EPW;
  bool bad = false;
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * w.EPW; base < ne;
       base += stride) {
    const int64_t el = base + w.slot;
    const bool act = w.in_use && el < ne;
    if (act) load_element_lanes(P, e0 + el, w.r, Xe, ue, me, pe);
    wave_sync();
    if (act && w.r < nq) {
      const int q = w.r;
      PointState S;
      point_state(P, q, Xe, ue, me, pe, S);
      const double* Nu = P.tab_u + (size_t)q * P.nd_u;
      double pq = 0;
      for (int a = 0; a < P.nd_u; ++a) pq = fma(Nu[a], pe[a], pq);
      W[q] = S.w * (S.kap * S.g * S.dot - P.f * pq);
    }
    wave_sync();
    if (act && w.r == 0) {
      double acc = 0;
      for (int q = 0; q < nq; ++q) acc += W[q];
      out[el] = acc;
      bad |= !isfinite(acc);
    }
    wave_sync();
  }
  if (bad && P.nonfinite) atomicOr(P.nonfinite, 1);
}

// ---------------------------------------------------------------- caller-supplied densities
// A density the caller evaluates itself: the fields at every point go out (hm_point_fields),
// its pointwise derivatives come back in the physical basis (value, d/dx_0..2) and are
// contracted here, with w |det J| and J^-1 folded in as for the closed-form densities:
//   D^_vv = w D_vv,  D^_vk = w sum_d D_vd Jinv[k][d],  D^_kv = w sum_d Jinv[k][d] D_dv,
//   D^_kl = w sum_de Jinv[k][d] D_de Jinv[l][e].

// F (e1-e0, nq, 13): w |det J|, u, grad u, m, grad m, p, grad p (physical gradients).
__global__ void __launch_bounds__(256) k_point_fields(hm_problem P, int64_t e0, int64_t e1, double* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const WaveElement w = wave_element(P);
  const int words = 24 + 2 * P.nd_u + P.nd_m;
  double* Xe = lds + (size_t)((threadIdx.x >> 6) * w.EPW + (w.in_use ? w.slot : 0)) * words;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * w.EPW;
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * w.EPW; base < ne;
       base += stride) {
    const int64_t el = base + w.slot;
    const bool act = w.in_use && el < ne;
    if (act) load_element_lanes(P, e0 + el, w.r, Xe, ue, me, pe);
    wave_sync();
    if (act && w.r < nq) {
      const int q = w.r;
      PointState S;
      point_state(P, q, Xe, ue, me, pe, S);
      const int nu = P.nd_u, nm = P.nd_m;
      const double* Nu = P.tab_u + (size_t)q * nu;
      const double* Nm = P.tab_m + (size_t)q * nm;
      const double* Gm = P.tab_m + (size_t)nq * nm + (size_t)q * nm * 3;
      double pq = 0, mq = 0, rm[3] = {0, 0, 0};
      for (int a = 0; a < nu; ++a) pq = fma(Nu[a], pe[a], pq);
      for (int a = 0; a < nm; ++a) {
        mq = fma(Nm[a], me[a], mq);
        for (int k = 0; k < 3; ++k) rm[k] = fma(Gm[a * 3 + k], me[a], rm[k]);
      }
      double* F = out + ((size_t)el * nq + q) * 13;
      F[0] = S.w;
      F[1] = S.uq;
      for (int d = 0; d < 3; ++d) F[2 + d] = S.gu[d];
      F[5] = mq;
      for (int d = 0; d < 3; ++d) F[6 + d] = S.Ji[0][d] * rm[0] + S.Ji[1][d] * rm[1] + S.Ji[2][d] * rm[2];
      F[9] = pq;
      for (int d = 0; d < 3; ++d) F[10 + d] = S.gp[d];
    }
    wave_sync();
  }
}

// the caller's physical-basis block D (4 x 4) at a point, folded into the reference basis
__device__ inline void fold_pointwise(const double* D, const PointState& S, double Dh[4][4]) {
  double L[4][4];  // L[r][c'] = sum_d Jinv-side of the rows: row r of (Jinv-folded D)
  for (int c = 0; c < 4; ++c) {
    L[0][c] = D[c];
    for (int k = 0; k < 3; ++k)
      L[1 + k][c] = S.Ji[k][0] * D[4 + c] + S.Ji[k][1] * D[8 + c] + S.Ji[k][2] * D[12 + c];
  }
  for (int r = 0; r < 4; ++r) {
    Dh[r][0] = S.w * L[r][0];
    for (int l = 0; l < 3; ++l)
      Dh[r][1 + l] = S.w * (L[r][1] * S.Ji[l][0] + L[r][2] * S.Ji[l][1] + L[r][3] * S.Ji[l][2]);
  }
}

// Element matrices of block (s, j) from the caller's pointwise blocks Din (e1-e0, nq, 4, 4):
// the generic contraction of k_element_matrices.
template <typename OutT>
__global__ void __launch_bounds__(256) k_element_matrices_pointwise(hm_problem P, int s, int j, const double* Din,
                                                                    int64_t e0, int64_t e1, OutT* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const int nd_t = s == HM_PARAMETER ? P.nd_m : P.nd_u;
  const int nd_c = j == HM_PARAMETER ? P.nd_m : P.nd_u;
  const double* tab_t = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  const double* tab_c = j == HM_PARAMETER ? P.tab_m : P.tab_u;
  const bool same = (tab_t == tab_c);
  double* Pt = lds;
  double* Pc = same ? Pt : Pt + nq * 4 * nd_t;
  double* T = Pc + (same ? nq * 4 * nd_t : nq * 4 * nd_c);
  double* D = T + nq * 4 * nd_c;
  double* Xe = D + nq * 16;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  load_basis(tab_t, nq, nd_t, Pt);
  if (!same) load_basis(tab_c, nq, nd_c, Pc);
  for (int64_t e = e0 + blockIdx.x; e < e1; e += gridDim.x) {
    __syncthreads();
    load_element(P, e, Xe, ue, me, pe);
    __syncthreads();
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
      PointState S;
      point_state(P, q, Xe, ue, me, pe, S, false);
      double Dq[4][4];
      fold_pointwise(Din + ((size_t)(e - e0) * nq + q) * 16, S, Dq);
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) D[q * 16 + r * 4 + c] = Dq[r][c];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nq * 4 * nd_c; i += blockDim.x) {
      const int b = i % nd_c, r = (i / nd_c) % 4, q = i / (4 * nd_c);
      double acc = 0;
      for (int c = 0; c < 4; ++c) acc = fma(D[q * 16 + r * 4 + c], Pc[(q * 4 + c) * nd_c + b], acc);
      T[i] = acc;
    }
    __syncthreads();
    OutT* Eo = out + (size_t)(e - e0) * nd_t * nd_c;
    for (int i = threadIdx.x; i < nd_t * nd_c; i += blockDim.x) {
      const int a = i / nd_c, b = i % nd_c;
      double acc = 0;
      for (int q = 0; q < nq; ++q)
        for (int r = 0; r < 4; ++r) acc = fma(Pt[(q * 4 + r) * nd_t + a], T[(q * 4 + r) * nd_c + b], acc);
      Eo[i] = (OutT)acc;
    }
  }
}

// Element vectors of slot s from the caller's pointwise first derivatives Vin (e1-e0, nq, 4):
// sum_q w [V_v N_a + (Jinv V_g) . G_a].
__global__ void __launch_bounds__(256) k_element_vectors_pointwise(hm_problem P, int s, const double* Vin,
                                                                   int64_t e0, int64_t e1, double* out) {
  extern __shared__ double lds[];
  const int nq = P.nq;
  const int nd = s == HM_PARAMETER ? P.nd_m : P.nd_u;
  const double* tab = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  double* Ps = lds;
  load_basis(tab, nq, nd, Ps);
  __syncthreads();
  const WaveElement w = wave_element(P);
  const int words = nq * 4 + 24 + 2 * P.nd_u + P.nd_m;
  double* W = Ps + nq * 4 * nd + (size_t)((threadIdx.x >> 6) * w.EPW + (w.in_use ? w.slot : 0)) * words;
  double* Xe = W + nq * 4;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * w.EPW;
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * w.EPW; base < ne;
       base += stride) {
    const int64_t el = base + w.slot;
    const bool act = w.in_use && el < ne;
    if (act) load_element_lanes(P, e0 + el, w.r, Xe, ue, me, pe);
    wave_sync();
    if (act && w.r < nq) {
      const int q = w.r;
      PointState S;
      point_state(P, q, Xe, ue, me, pe, S, false);
      const double* V = Vin + ((size_t)el * nq + q) * 4;
      W[q * 4] = S.w * V[0];
      for (int k = 0; k < 3; ++k)
        W[q * 4 + 1 + k] = S.w * (S.Ji[k][0] * V[1] + S.Ji[k][1] * V[2] + S.Ji[k][2] * V[3]);
    }
    wave_sync();
    if (act && w.r < nd) {
      const int a = w.r;
      double acc = 0;
      for (int q = 0; q < nq; ++q)
        for (int r = 0; r < 4; ++r) acc = fma(Ps[(q * 4 + r) * nd + a], W[q * 4 + r], acc);
      out[(size_t)el * nd + a] = acc;
    }
    wave_sync();
  }
}

static int check_problem(const hm_problem* P) {
  HM_REQUIRE(P != nullptr, "no problem given");
  HM_REQUIRE(P->ne >= 0 && P->nq > 0 && P->nq <= 64, "quadrature points per element %d outside 1..64", P->nq);
  HM_REQUIRE(P->nd_u > 0 && P->nd_u <= 64 && P->nd_m > 0 && P->nd_m <= 64,
             "dofs per element (%d, %d) outside 1..64", P->nd_u, P->nd_m);
  HM_REQUIRE(P->density == HM_DENSITY_LINEAR || P->density == HM_DENSITY_NONLINEAR,
             "unknown density %d", P->density);
  return HM_OK;
}

// The consistent single-precision block (kernel.py:1503-1571): one thread per element.  K is
// the element matrix in FP32; Rj = E c_j 1 and Ri = E^T c_i 1 are its exact (FP64) products
// with the normalized constants of the two slots, c = 1 / sqrt(nd).  The result, in FP64, is K
// corrected by rank-one terms so that its products with the constants are Rj and Ri:
//   two-sided  K + Dj c_j 1^T + c_i 1 Di^T - c_i^2 c_j (1^T Dj) 1 1^T,
//              Dj = Rj - c_j K 1,  Di = Ri - c_i K^T 1;
//   symmetric  Ks + c (Dj 1^T + 1 Dj^T) - c^2 (c 1^T Dj) 1 1^T,  Ks = (K + K^T) / 2,
//              Dj = Rj - c Ks 1, each value computed once for (a, b) and (b, a).
__global__ void __launch_bounds__(256) k_consistent_block(const float* __restrict__ K,
                                                          const double* __restrict__ Rj,
                                                          const double* __restrict__ Ri, int64_t ne,
                                                          int nd_i, int nd_j, int symmetric,
                                                          double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const float* Ke = K + (size_t)e * nd_i * nd_j;
    double* Oe = out + (size_t)e * nd_i * nd_j;
    const double cj = 1.0 / sqrt((double)nd_j), ci = 1.0 / sqrt((double)nd_i);
    double Dj[HM_MAXND];
    double sum_dj = 0.0;
    if (symmetric) {
      const int n = nd_i;
      for (int a = 0; a < n; ++a) {
        double row = 0.0;
        for (int b = 0; b < n; ++b) row += 0.5 * ((double)Ke[a * n + b] + (double)Ke[b * n + a]);
        Dj[a] = Rj[e * n + a] - cj * row;
        sum_dj += Dj[a];
      }
      const double w = cj * cj * (cj * sum_dj);
      for (int a = 0; a < n; ++a)
        for (int b = a; b < n; ++b) {
          const double v = 0.5 * ((double)Ke[a * n + b] + (double)Ke[b * n + a]) + cj * (Dj[a] + Dj[b]) - w;
          Oe[a * n + b] = v;
          Oe[b * n + a] = v;
        }
    } else {
      for (int a = 0; a < nd_i; ++a) {
        double row = 0.0;
        for (int b = 0; b < nd_j; ++b) row += (double)Ke[a * nd_j + b];
        Dj[a] = Rj[e * nd_i + a] - cj * row;
        sum_dj += Dj[a];
      }
      const double w = ci * ci * cj * sum_dj;
      for (int b = 0; b < nd_j; ++b) {
        double col = 0.0;
        for (int a = 0; a < nd_i; ++a) col += (double)Ke[a * nd_j + b];
        const double Di = Ri[e * nd_j + b] - ci * col;
        for (int a = 0; a < nd_i; ++a) Oe[a * nd_j + b] = (double)Ke[a * nd_j + b] + Dj[a] * cj + ci * Di - w;
      }
    }
  }
}

// Elements a wavefront of the vector and product kernels takes (wave_element on the device).
static int wave_elements(const hm_problem& P) {
  int L = P.nq;
  if (P.nd_u > L) L = P.nd_u;
  if (P.nd_m > L) L = P.nd_m;
  return 64 / L;
}

// Dynamic LDS of a launch within the device's limit (the attribute set above 48 KiB).
static int fit_lds(const void* kernel, size_t bytes) {
  int limit = 64 * 1024, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) limit = v;
  }
  HM_REQUIRE(bytes <= (size_t)limit, "element tables need %zu bytes of LDS, above the %d a block holds", bytes,
             limit);
  if (bytes > 48 * 1024) {
    const hipError_t err = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  return HM_OK;
}

static int grid_for(int64_t n) {
  const int64_t g = n < 1 ? 1 : n;
  return (int)(g < 65536 ? g : 65536);
}

}  // namespace hm

using namespace hm;

extern "C" int hm_element_matrices(const hm_problem* P, int i, int j, int64_t e0, int64_t e1,
                                   int precision, void* out, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(i >= 0 && i < 3 && j >= 0 && j < 3, "slots (%d, %d) outside 0..2", i, j);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements [%lld, %lld) outside [0, %lld)",
             (long long)e0, (long long)e1, (long long)P->ne);
  if (e1 == e0) return HM_OK;
  const int nd_ti = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int nd_cj = j == HM_PARAMETER ? P->nd_m : P->nd_u;
  // route: "sumfact" (default where the problem has tensor structure), "mfma", "valu"
  const char* route = getenv("HM_CONTRACTION");
  const bool valu = route && route[0] == 'v';
  const bool mfma = route && route[0] == 'm';
  // BF16 / FP16 operands: the sum-factorized contraction only
  if (precision == HM_BF16 || precision == HM_FP16)
    return hm_element_matrices_sumfact(P, i, j, e0, e1, precision, out, stream);
  HM_REQUIRE(precision == HM_FP64 || precision == HM_FP32, "unknown precision %d", precision);
  if (!valu && !mfma) {
    const int rc = hm_element_matrices_sumfact(P, i, j, e0, e1, precision, out, stream);
    if (rc != HM_ERR_UNSUPPORTED) return rc;
  }
  // the other routes: a pass over the written matrices sets the non-finite flag
  const size_t nent = (size_t)(e1 - e0) * nd_ti * nd_cj;
  if (nd_ti <= 32 && nd_cj <= 32 && (!valu || precision == HM_FP32)) {
    const int rc = hm_element_matrices_mfma(P, i, j, e0, e1, precision, out, stream);
    if (rc == HM_OK)
      return hm_flag_nonfinite(out, nent, precision == HM_FP64 ? 8 : 4, P->nonfinite, (hipStream_t)stream);
    if (rc != HM_ERR_UNSUPPORTED) return rc;
    // the matrix-core tables do not fit (many points and parts): the generic contraction
  }
  // The generic contraction, in FP64, or in FP32 for the FP32 form (the route of the
  // single-precision forms where the others do not apply).
  const int nq = P->nq;
  const int nd_t = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int nd_c = j == HM_PARAMETER ? P->nd_m : P->nd_u;
  const bool same = (i == HM_PARAMETER) == (j == HM_PARAMETER);
  size_t words = (size_t)nq * 4 * nd_t + (same ? 0 : (size_t)nq * 4 * nd_c) + (size_t)nq * 4 * nd_c +
                 (size_t)nq * 16 + 24 + 2 * P->nd_u + P->nd_m;
  const size_t bytes = words * sizeof(double);
  const int parts = block_parts(i, j, P->density);
  hipStream_t s = (hipStream_t)stream;
  if (precision == HM_FP64) {
    if (int rc = fit_lds((const void*)k_element_matrices<double>, bytes)) return rc;
    hipLaunchKernelGGL(k_element_matrices<double>, dim3(grid_for(e1 - e0)), dim3(256), bytes, s, *P, i,
                       j, e0, e1, parts, (double*)out);
  } else {
    if (int rc = fit_lds((const void*)k_element_matrices<float>, bytes)) return rc;
    hipLaunchKernelGGL(k_element_matrices<float>, dim3(grid_for(e1 - e0)), dim3(256), bytes, s, *P, i,
                       j, e0, e1, parts, (float*)out);
  }
  if (int rc = hip_check("hm_element_matrices")) return rc;
  return hm_flag_nonfinite(out, nent, precision == HM_FP64 ? 8 : 4, P->nonfinite, s);
}

extern "C" int hm_element_matrices_pair(const hm_problem* P, int i1, int j1, int i2, int j2, int64_t e0,
                                        int64_t e1, int precision, void* out1, void* out2, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(i1 >= 0 && i1 < 3 && j1 >= 0 && j1 < 3 && i2 >= 0 && i2 < 3 && j2 >= 0 && j2 < 3,
             "slots (%d, %d), (%d, %d) outside 0..2", i1, j1, i2, j2);
  HM_REQUIRE(i1 * 3 + j1 != i2 * 3 + j2, "block (%d, %d) twice", i1, j1);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements [%lld, %lld) outside [0, %lld)",
             (long long)e0, (long long)e1, (long long)P->ne);
  HM_REQUIRE(out1 && out2 && out1 != out2, "two distinct outputs needed");
  if (e1 == e0) return HM_OK;
  const char* route = getenv("HM_CONTRACTION");
  if (route && (route[0] == 'v' || route[0] == 'm'))
    HM_FAIL(HM_ERR_UNSUPPORTED, "two blocks a pass on the sum-factorized route only");
  return hm_element_matrices_sumfact_pair(P, i1, j1, i2, j2, e0, e1, precision, out1, out2, stream);
}

extern "C" int hm_consistent_block(const float* K, const double* Rj, const double* Ri, int64_t ne, int nd_i,
                                   int nd_j, int symmetric, double* out, void* stream) {
  HM_REQUIRE(ne >= 0 && nd_i > 0 && nd_j > 0 && nd_i <= HM_MAXND && nd_j <= HM_MAXND,
             "bad element shape (%d, %d)", nd_i, nd_j);
  HM_REQUIRE(!symmetric || nd_i == nd_j, "a symmetric block is square");
  HM_REQUIRE(symmetric || Ri != nullptr, "a two-sided correction needs the transposed product");
  HM_REQUIRE(K && Rj && out, "null array");
  if (ne == 0) return HM_OK;
  const int64_t nb = (ne + 255) / 256;
  hipLaunchKernelGGL(k_consistent_block, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0,
                     (hipStream_t)stream, K, Rj, Ri, ne, nd_i, nd_j, symmetric, out);
  return hip_check("hm_consistent_block");
}

extern "C" int hm_element_vectors(const hm_problem* P, int i, int64_t e0, int64_t e1, double* out,
                                  void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(i >= 0 && i < 3, "slot %d outside 0..2", i);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements outside the batch");
  if (e1 == e0) return HM_OK;
  // tensor problems: the sum-factorized point state (the route of the element matrices)
  const char* route = getenv("HM_CONTRACTION");
  if (!(route && (route[0] == 'v' || route[0] == 'm'))) {
    const int rc = hm_element_vectors_sumfact(P, i, e0, e1, out, stream);
    if (rc != HM_ERR_UNSUPPORTED) return rc;
  }
  const int nd = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int epb = 4 * wave_elements(*P);
  const size_t bytes = ((size_t)P->nq * 4 * nd +
                        (size_t)epb * ((size_t)P->nq * 4 + 24 + 2 * P->nd_u + P->nd_m)) * 8;
  if (int rc = fit_lds((const void*)k_element_vectors, bytes)) return rc;
  hipLaunchKernelGGL(k_element_vectors, dim3(grid_for((e1 - e0 + epb - 1) / epb)), dim3(256), bytes,
                     (hipStream_t)stream, *P, i, e0, e1, out);
  return hip_check("hm_element_vectors");
}

extern "C" int hm_element_scalars(const hm_problem* P, int64_t e0, int64_t e1, double* out, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements outside the batch");
  HM_REQUIRE(e1 == e0 || out, "no output given");
  if (e1 == e0) return HM_OK;
  const int epb = 4 * wave_elements(*P);
  const size_t bytes = (size_t)epb * ((size_t)P->nq + 24 + 2 * P->nd_u + P->nd_m) * 8;
  if (int rc = fit_lds((const void*)k_element_scalars, bytes)) return rc;
  hipLaunchKernelGGL(k_element_scalars, dim3(grid_for((e1 - e0 + epb - 1) / epb)), dim3(256), bytes,
                     (hipStream_t)stream, *P, e0, e1, out);
  return hip_check("hm_element_scalars");
}

extern "C" int hm_point_fields(const hm_problem* P, int64_t e0, int64_t e1, double* out, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements outside the batch");
  HM_REQUIRE(e1 == e0 || out, "no output given");
  if (e1 == e0) return HM_OK;
  const int epb = 4 * wave_elements(*P);
  const size_t bytes = (size_t)epb * (24 + 2 * (size_t)P->nd_u + P->nd_m) * 8;
  if (int rc = fit_lds((const void*)k_point_fields, bytes)) return rc;
  hipLaunchKernelGGL(k_point_fields, dim3(grid_for((e1 - e0 + epb - 1) / epb)), dim3(256), bytes,
                     (hipStream_t)stream, *P, e0, e1, out);
  return hip_check("hm_point_fields");
}

extern "C" int hm_element_matrices_pointwise(const hm_problem* P, int i, int j, const double* D, int64_t e0,
                                             int64_t e1, int precision, void* out, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(i >= 0 && i < 3 && j >= 0 && j < 3, "slots (%d, %d) outside 0..2", i, j);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements outside the batch");
  HM_REQUIRE(precision == HM_FP64 || precision == HM_FP32, "precision %d: FP64 or FP32 results", precision);
  HM_REQUIRE(e1 == e0 || (D && out), "no blocks or output given");
  if (e1 == e0) return HM_OK;
  const int nq = P->nq;
  const int nd_t = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int nd_c = j == HM_PARAMETER ? P->nd_m : P->nd_u;
  const bool same = (i == HM_PARAMETER) == (j == HM_PARAMETER);
  const size_t words = (size_t)nq * 4 * nd_t + (same ? 0 : (size_t)nq * 4 * nd_c) + (size_t)nq * 4 * nd_c +
                       (size_t)nq * 16 + 24 + 2 * P->nd_u + P->nd_m;
  const size_t bytes = words * sizeof(double);
  hipStream_t s = (hipStream_t)stream;
  if (precision == HM_FP64) {
    if (int rc = fit_lds((const void*)k_element_matrices_pointwise<double>, bytes)) return rc;
    hipLaunchKernelGGL(k_element_matrices_pointwise<double>, dim3(grid_for(e1 - e0)), dim3(256), bytes, s, *P,
                       i, j, D, e0, e1, (double*)out);
  } else {
    if (int rc = fit_lds((const void*)k_element_matrices_pointwise<float>, bytes)) return rc;
    hipLaunchKernelGGL(k_element_matrices_pointwise<float>, dim3(grid_for(e1 - e0)), dim3(256), bytes, s, *P,
                       i, j, D, e0, e1, (float*)out);
  }
  if (int rc = hip_check("hm_element_matrices_pointwise")) return rc;
  return hm_flag_nonfinite(out, (size_t)(e1 - e0) * nd_t * nd_c, precision == HM_FP64 ? 8 : 4, P->nonfinite, s);
}

extern "C" int hm_element_vectors_pointwise(const hm_problem* P, int i, const double* V, int64_t e0, int64_t e1,
                                            double* out, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(i >= 0 && i < 3, "slot %d outside 0..2", i);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements outside the batch");
  HM_REQUIRE(e1 == e0 || (V && out), "no derivatives or output given");
  if (e1 == e0) return HM_OK;
  const int nd = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int epb = 4 * wave_elements(*P);
  const size_t bytes = ((size_t)P->nq * 4 * nd +
                        (size_t)epb * ((size_t)P->nq * 4 + 24 + 2 * P->nd_u + P->nd_m)) * 8;
  if (int rc = fit_lds((const void*)k_element_vectors_pointwise, bytes)) return rc;
  hipLaunchKernelGGL(k_element_vectors_pointwise, dim3(grid_for((e1 - e0 + epb - 1) / epb)), dim3(256), bytes,
                     (hipStream_t)stream, *P, i, V, e0, e1, out);
  return hip_check("hm_element_vectors_pointwise");
}

extern "C" int hm_element_hvp(const hm_problem* P, int i, int j, const double* dir, int64_t e0,
                              int64_t e1, double* out, void* stream) {
  if (int rc = check_problem(P)) return rc;
  HM_REQUIRE(i >= 0 && i < 3 && j >= 0 && j < 3, "slots (%d, %d) outside 0..2", i, j);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= P->ne, "elements outside the batch");
  if (e1 == e0) return HM_OK;
  const int nd_t = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int nd_c = j == HM_PARAMETER ? P->nd_m : P->nd_u;
  const bool same = (i == HM_PARAMETER) == (j == HM_PARAMETER);
  const int epb = 4 * wave_elements(*P);
  const size_t bytes = ((size_t)P->nq * 4 * (nd_t + (same ? 0 : nd_c)) +
                        (size_t)epb * ((size_t)P->nq * 4 + nd_c + 24 + 2 * P->nd_u + P->nd_m)) * 8;
  if (int rc = fit_lds((const void*)k_element_hvp, bytes)) return rc;
  hipLaunchKernelGGL(k_element_hvp, dim3(grid_for((e1 - e0 + epb - 1) / epb)), dim3(256), bytes,
                     (hipStream_t)stream, *P, i, j, dir, e0, e1, out);
  return hip_check("hm_element_hvp");
}
