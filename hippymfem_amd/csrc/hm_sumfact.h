// Pieces the sum-factorized element kernels share (hm_sumfact.hip, hm_fused.hip): the
// plan of a block's pointwise pairs, the operand rounding of the low-precision forms, the
// sum-factorized point state and the LDS sizes of the tables and of an element.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_point.h"

namespace hm {

constexpr int SF_MAXPAIR = 32;  // 16 (r, c) pairs a block, two blocks a pass

struct SumfactPlan {
  int npair;
  int kmask;                           // bit 4 r + c set for every pair, pairs in the order of k
  int pr[SF_MAXPAIR], pc[SF_MAXPAIR];  // (test part r, trial part c) of each pair
  int tmask[4];                        // pairs of each third-direction type (r == 3, c == 3)
  int offdiag;                         // the pairs of two different reference derivatives
  int need_fields;                     // the pointwise blocks read u or p
  // a second block of the same shape from the same pass (two-block kernels): its slots, its
  // pairs (after the first block's in pr/pc) and the pairs of each block
  int s2, j2, kmask2;
  int bmask[2];
};

// Pairs left out of an element's contraction, their share of every entry being below the
// rounding of the sum: on an axis-aligned box element (J diagonal) the pairs of two different
// reference derivatives, whose weights (Jinv Jinv^T off the diagonal) are round-off there; in
// the low-precision forms, any pair whose pointwise weight stays below SF_DROP of the element's
// largest.
#define SF_DROP 1e-16

// The operand format of the contraction: QP = HM_FP64 or HM_FP32 leave operands as they are;
// HM_BF16 / HM_FP16 round every operand of a product (the scaled pointwise pairs, the 1-D
// table products and the stage results fed to the next stage) to that format, while the
// products and sums stay FP32 (a product of two such operands is exact in FP32).
template <int QP, typename T>
__device__ inline T quant(T x) {
  // the FP32 value is made opaque first: seen through, an FP64 source turned the hardware
  // conversions into a software FP64 -> FP16 rounding of some thirty instructions
  if constexpr (QP == HM_BF16) {
    float f = (float)x;
    asm("" : "+v"(f));
    return __bfloat162float(__float2bfloat16(f));
  } else if constexpr (QP == HM_FP16) {
    float f = (float)x;
    asm("" : "+v"(f));
    return __half2float(__float2half(f));
  } else {
    return x;
  }
}

// The lanes of a wavefront see each other's LDS writes: a wavefront's LDS operations complete
// in order, so a fence against compiler reordering and a wavefront barrier suffice.

// LDS words of one element: its pairs W[npair][nq], vertex coordinates [24], u and p dof
// values [nd_u] each, m dof values [nd_m]; even, so every element starts 16-byte aligned.
__host__ __device__ inline int sf_element_words(int npair, int nd_u, int nd_m, int nq) {
  const int w = npair * nq + 24 + 2 * nd_u + nd_m;
  return (w + 1) & ~1;
}

// The point state of point q (NQ points a direction, q = q1 + NQ (q2 + NQ q3)) from LDS: the
// element's vertex coordinates X, dof values M, and U, P in tensor order (a1 + n1 (a2 + n1 a3));
// the tables stored point-minor (entry (row, q) at row * NQ^3 + q: lanes reading their own
// points hit consecutive words) -- TX[(v * 3 + k)][q], TM[a][q] -- and the state space's 1-D
// basis T1U[2][n1][NQ] (values, then derivatives).
// The fields u, p and their reference gradients are summed one direction at a time: for each
// (a2, a3) the sums over a1 of u and p against the value and the derivative of the 1-D basis
// at q1, then those times the 1-D factors at q2 and q3.
//
// PT: the point type, FP64, or FP32 in the precision forms: the fields, their gradients and
// everything after the geometry are then FP32 (the dof values and tables rounded as read), the
// geometry FP64 (point_finish).
// The fields at point q: u and its reference gradient, p's, m.
template <int NQ, typename PT>
__device__ inline void sf_fields(const hm_problem& P, int q, const double* U, const double* M, const double* Pv,
                                 bool fields, const double* TM, const double* T1U, PT& uq, PT ru[3], PT rp[3],
                                 PT& mq) {
  const int nm = P.nd_m;
  uq = 0;
  mq = 0;
  for (int k = 0; k < 3; ++k) ru[k] = rp[k] = 0;
  if (fields) {
    const int n1 = P.n1_u;
    const int q1 = q % NQ, q2 = (q / NQ) % NQ, q3 = q / (NQ * NQ);
    // the 1-D factors are read from LDS as they are used (few registers: the point phase
    // shares the kernel's register budget with the contraction)
#pragma unroll 1
    for (int a3 = 0; a3 < n1; ++a3) {
      const PT l3 = (PT)T1U[a3 * NQ + q3], d3 = (PT)T1U[(n1 + a3) * NQ + q3];
#pragma unroll 1
      for (int a2 = 0; a2 < n1; ++a2) {
        const PT l2 = (PT)T1U[a2 * NQ + q2], d2 = (PT)T1U[(n1 + a2) * NQ + q2];
        const double* u = U + n1 * (a2 + n1 * a3);
        const double* pv = Pv + n1 * (a2 + n1 * a3);
        PT svu = 0, sdu = 0, svp = 0, sdp = 0;
#pragma unroll
        for (int a1 = 0; a1 < 3; ++a1) {
          if (a1 >= n1) break;
          const PT l1 = (PT)T1U[a1 * NQ + q1], d1 = (PT)T1U[(n1 + a1) * NQ + q1];
          svu = fma(l1, (PT)u[a1], svu);
          sdu = fma(d1, (PT)u[a1], sdu);
          svp = fma(l1, (PT)pv[a1], svp);
          sdp = fma(d1, (PT)pv[a1], sdp);
        }
        const PT f0 = l2 * l3, f1 = d2 * l3, f2 = l2 * d3;
        uq = fma(f0, svu, uq);
        ru[0] = fma(f0, sdu, ru[0]);
        ru[1] = fma(f1, svu, ru[1]);
        ru[2] = fma(f2, svu, ru[2]);
        rp[0] = fma(f0, sdp, rp[0]);
        rp[1] = fma(f1, svp, rp[1]);
        rp[2] = fma(f2, svp, rp[2]);
      }
    }
  }
  for (int a = 0; a < nm; ++a) mq = fma((PT)TM[a * (NQ * NQ * NQ) + q], (PT)M[a], mq);
}

template <int NQ, typename PT = double>
__device__ inline void sf_point_state(const hm_problem& P, int q, const double* X, const double* U,
                                      const double* M, const double* Pv, bool fields, const double* TX,
                                      const double* TM, const double* TW, const double* T1U, PointStateT<PT>& S,
                                      bool aligned = false) {
  GeomT<PT> G;
  if (aligned) {
    // an axis-aligned box element (every vertex coordinate that of the vertex with the same
    // index along its axis; the caller tests it once an element): J the diagonal of its edge
    // lengths
    geom_diag<PT>(X[1 * 3 + 0] - X[0], X[2 * 3 + 1] - X[1], X[4 * 3 + 2] - X[2], G);
  } else {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 2
    for (int v = 0; v < 8; ++v)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double g = TX[(v * 3 + k) * (NQ * NQ * NQ) + q];
        J[0][k] = fma(X[v * 3], g, J[0][k]);
        J[1][k] = fma(X[v * 3 + 1], g, J[1][k]);
        J[2][k] = fma(X[v * 3 + 2], g, J[2][k]);
      }
    geom_from_J<PT>(J, G);
  }
  PT uq, mq, ru[3], rp[3];
  sf_fields<NQ, PT>(P, q, U, M, Pv, fields, TM, T1U, uq, ru, rp, mq);
  point_fields<PT>(P, TW[q], G, uq, ru, rp, mq, S);
}

// LDS words of the tables every element reads (nq points, nq1 a direction): the trilinear
// map's gradients [nq][8][3], the parameter basis [nq][nd_m], the quadrature weights [nq]
// (read from LDS, not memory: a memory read in the point phase would wait for the
// element-matrix stores issued before it, as loads and stores share one counter) and, when
// the fields are read, the state space's 1-D basis [2][n1_u][nq1] (room for n1_u up to 3).
__host__ __device__ inline int sf_table_words(int nd_u, int nd_m, bool fields, int nq1) {
  (void)nd_u;
  const int nq = nq1 * nq1 * nq1;
  const int w = nq * 24 + nq * nd_m + nq + (fields ? 6 * nq1 : 0);
  return (w + 1) & ~1;
}

// The nonzero (r, c) pairs of block (i, j), r, c in {value, d/dxi_1, d/dxi_2, d/dxi_3}: the
// vv, vg, gv, gg parts of the density; appended to pr/pc from np on, their k = 4 r + c bits in
// *kmask.  Returns the new count.
static inline int block_pairs(int i, int j, bool nl, int* pr, int* pc, int np, int* kmask) {
  bool vv = false, vg = false, gv = false, gg = false;
  switch (i * 3 + j) {
    case 0: vv = vg = gv = nl; break;
    case 1: vv = nl; gv = true; break;
    case 3: vv = nl; vg = true; break;
    case 2: vg = nl; gg = true; break;
    case 6: gv = nl; gg = true; break;
    case 4: vv = true; break;
    case 5: vg = true; break;
    case 7: gv = true; break;
    default: break;
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const bool on = (r == 0 && c == 0) ? vv : (r == 0) ? vg : (c == 0) ? gv : gg;
      if (on) {
        pr[np] = r;
        pc[np] = c;
        *kmask |= 1 << (4 * r + c);
        ++np;
      }
    }
  return np;
}

static inline void plan_masks(SumfactPlan& SP) {
  for (int p = 0; p < SP.npair; ++p) {
    SP.tmask[((SP.pr[p] == 3) << 1) | (SP.pc[p] == 3)] |= (int)(1u << p);
    if (SP.pr[p] >= 1 && SP.pc[p] >= 1 && SP.pr[p] != SP.pc[p]) SP.offdiag |= (int)(1u << p);
  }
}

// the problem carries the tensor structure the sum-factorized contraction needs
static inline int check_tensor(const hm_problem* P) {
  if ((P->nq1 != 3 && P->nq1 != 4) || P->nq != P->nq1 * P->nq1 * P->nq1 || !P->tab1_u || !P->tab1_m ||
      !P->lex_u || !P->lex_m)
    HM_FAIL(HM_ERR_UNSUPPORTED, "no tensor structure for the sum-factorized contraction");
  if (P->nd_u > 32 || P->nd_m > 32) HM_FAIL(HM_ERR_UNSUPPORTED, "more than 32 dofs an element");
  if (P->nd_u != P->n1_u * P->n1_u * P->n1_u || P->nd_m != P->n1_m * P->n1_m * P->n1_m)
    HM_FAIL(HM_ERR_UNSUPPORTED, "the element dofs are not the tensor products of the 1-D nodes");
  return HM_OK;
}

}  // namespace hm
