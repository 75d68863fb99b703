// Sum-factorized element contraction for tensor-product hexahedra.
//
// The element matrix of a block is E[a][b] = sum_q sum_(r,c) D^[q][r][c] Psi_r(a, q) Psi_c(b, q),
// r, c in {value, d/dxi_1, d/dxi_2, d/dxi_3}.  On a tensor element every Psi factorizes over
// the three directions, Psi_r(a, q) = X1(a1, q1) X2(a2, q2) X3(a3, q3) with Xd the 1-D basis or
// its derivative (the derivative where r = d), so the sum over the 27 points is taken one
// direction at a time (stage 1 over q1, stage 2 over q2, stage 3 over q3).  Pairs (r, c) that
// share their third-direction factor are added before stage 3.  For the Jacobian of the
// diffusion density this is about 17.5k fused multiply-adds an element instead of 59k for the
// dense contraction; the arithmetic is the reference's quadrature route (kernel.py:1431-1485)
// with its sums reordered.
//
// One wavefront contracts one element; the pointwise blocks D^ of the block's nonzero
// (r, c) pairs are computed for the 8 elements of a block at once (216 of 256 threads) into
// LDS, in FP64.  CT is the arithmetic of the stages (FP64, or FP32 for the precision form).
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_device.h"
#include "hm_point.h"

namespace hm {

constexpr int SF_MAXPAIR = 16;
constexpr int SF_EPB = 8;  // elements per block iteration (two per wavefront)

struct SumfactPlan {
  int npair;
  int pr[SF_MAXPAIR], pc[SF_MAXPAIR];  // (test part r, trial part c) of each pair
  int ngroup;                          // groups of pairs sharing the third-direction factor
  int gstart[5], gpair[SF_MAXPAIR];    // pairs of group g: gpair[gstart[g] .. gstart[g+1])
  int need_fields;                     // the pointwise blocks read u or p
};

template <typename CT, int NT, int NC>
__global__ void __launch_bounds__(256) k_element_sumfact(hm_problem P, int s, int j, int64_t e0,
                                                         int64_t e1, SumfactPlan SP, void* out) {
  constexpr int NQ = 3;
  constexpr int TC = NT * NC;  // 1-D (test node, trial node) pairs
  extern __shared__ double lds[];
  const int nd_t = NT * NT * NT, nd_c = NC * NC * NC;
  const int npair = SP.npair;
  // LDS: 1-D tables (test/trial, value/derivative), local dof maps, then the pointwise
  // blocks of the 8 elements (FP64), then per wave S1 and S2 (CT).
  double* Bt = lds;                       // [2][NT][NQ]
  double* Bc = Bt + 2 * NT * NQ;          // [2][NC][NQ]
  int* lext = (int*)(Bc + 2 * NC * NQ);   // [nd_t]
  int* lexc = lext + 32;                  // [nd_c]
  double* W = (double*)(lexc + 32);       // [SF_EPB][npair][27]
  double* elem = W + SF_EPB * npair * 27; // [SF_EPB][24 + 2 nd_u + nd_m]
  const int ew = 24 + 2 * P.nd_u + P.nd_m;
  CT* sbase = (CT*)(elem + SF_EPB * ew);
  const int s1_words = npair * TC * 9;
  const int s2_words = SP.ngroup * TC * TC * NQ;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  CT* S1 = sbase + (size_t)wave * (s1_words + s2_words);
  CT* S2 = S1 + s1_words;

  const double* t1t = s == HM_PARAMETER ? P.tab1_m : P.tab1_u;
  const double* t1c = j == HM_PARAMETER ? P.tab1_m : P.tab1_u;
  const int32_t* lt = s == HM_PARAMETER ? P.lex_m : P.lex_u;
  const int32_t* lc = j == HM_PARAMETER ? P.lex_m : P.lex_u;
  for (int i = threadIdx.x; i < 2 * NT * NQ; i += blockDim.x) Bt[i] = t1t[i];
  for (int i = threadIdx.x; i < 2 * NC * NQ; i += blockDim.x) Bc[i] = t1c[i];
  for (int i = threadIdx.x; i < nd_t; i += blockDim.x) lext[i] = lt[i];
  for (int i = threadIdx.x; i < nd_c; i += blockDim.x) lexc[i] = lc[i];

  for (int64_t base = e0 + (int64_t)blockIdx.x * SF_EPB; base < e1; base += (int64_t)gridDim.x * SF_EPB) {
    const int nel = (int)((e1 - base) < SF_EPB ? (e1 - base) : SF_EPB);
    __syncthreads();
    // element data of the 8 elements
    for (int i = threadIdx.x; i < SF_EPB * ew; i += blockDim.x) {
      const int k = i / ew, w = i % ew;
      if (k >= nel) continue;
      const int64_t e = base + k;
      double v;
      if (w < 24) {
        v = P.verts[(size_t)P.elem_verts[e * 8 + w / 3] * 3 + w % 3];
      } else if (w < 24 + 2 * P.nd_u) {
        const int a = (w - 24) % P.nd_u;
        if (!SP.need_fields) {
          v = 0.0;
        } else {
          const int32_t d = P.edofs_u[e * P.nd_u + a];
          v = (w - 24) < P.nd_u ? P.u[d] : P.p[d];
        }
      } else {
        v = P.m[P.edofs_m[e * P.nd_m + (w - 24 - 2 * P.nd_u)]];
      }
      elem[i] = v;
    }
    __syncthreads();
    // pointwise blocks: thread (k, q)
    if (threadIdx.x < SF_EPB * 27) {
      const int k = threadIdx.x / 27, q = threadIdx.x % 27;
      if (k < nel) {
        const double* Xe = elem + k * ew;
        PointState st;
        point_state(P, q, Xe, Xe + 24, Xe + 24 + 2 * P.nd_u, Xe + 24 + P.nd_u, st, SP.need_fields != 0);
        double Dq[4][4];
        point_block(s, j, st, Dq);
        for (int p = 0; p < npair; ++p) W[(k * npair + p) * 27 + q] = Dq[SP.pr[p]][SP.pc[p]];
      }
    }
    __syncthreads();
    for (int half = 0; half < 2; ++half) {
      const int k = wave * 2 + half;
      if (k >= nel) break;
      const double* Wk = W + k * npair * 27;
      // stage 1: S1[p][a1 b1][q2 + 3 q3] = sum_q1 X1(a1, q1) Y1(b1, q1) W_p(q1, q2, q3)
      for (int task = lane; task < npair * 9 * NT; task += 64) {
        const int p = task / (9 * NT), rem = task % (9 * NT), q23 = rem / NT, a1 = rem % NT;
        const double* Xa = Bt + ((SP.pr[p] == 1) * NT + a1) * NQ;
        const double w0 = Wk[p * 27 + 3 * q23], w1 = Wk[p * 27 + 3 * q23 + 1], w2 = Wk[p * 27 + 3 * q23 + 2];
        const double x0 = Xa[0] * w0, x1 = Xa[1] * w1, x2 = Xa[2] * w2;
#pragma unroll
        for (int b1 = 0; b1 < NC; ++b1) {
          const double* Yb = Bc + ((SP.pc[p] == 1) * NC + b1) * NQ;
          S1[(p * TC + a1 * NC + b1) * 9 + q23] = (CT)(x0 * Yb[0] + x1 * Yb[1] + x2 * Yb[2]);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // stage 2: S2[g][a1b1][a2 b2][q3] = sum_{p in g} sum_q2 X2(a2,q2) Y2(b2,q2) S1[p][a1b1][q2 + 3 q3]
      for (int task = lane; task < SP.ngroup * NQ * TC * NT; task += 64) {
        const int g = task / (NQ * TC * NT), rem = task % (NQ * TC * NT);
        const int q3 = rem / (TC * NT), rem2 = rem % (TC * NT), ab1 = rem2 / NT, a2 = rem2 % NT;
        CT acc[NC];
#pragma unroll
        for (int b2 = 0; b2 < NC; ++b2) acc[b2] = 0;
        for (int gi = SP.gstart[g]; gi < SP.gstart[g + 1]; ++gi) {
          const int p = SP.gpair[gi];
          const double* Xa = Bt + ((SP.pr[p] == 2) * NT + a2) * NQ;
          const CT* s1 = S1 + (p * TC + ab1) * 9 + 3 * q3;
          const CT z0 = (CT)Xa[0] * s1[0], z1 = (CT)Xa[1] * s1[1], z2 = (CT)Xa[2] * s1[2];
#pragma unroll
          for (int b2 = 0; b2 < NC; ++b2) {
            const double* Yb = Bc + ((SP.pc[p] == 2) * NC + b2) * NQ;
            acc[b2] += z0 * (CT)Yb[0] + z1 * (CT)Yb[1] + z2 * (CT)Yb[2];
          }
        }
#pragma unroll
        for (int b2 = 0; b2 < NC; ++b2) S2[((g * TC + ab1) * TC + a2 * NC + b2) * NQ + q3] = acc[b2];
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // stage 3: E[a][b] = sum_g sum_q3 X3_g(a3, q3) Y3_g(b3, q3) S2[g][a1b1][a2b2][q3]
      CT* Eo = (CT*)out + (size_t)(base + k - e0) * nd_t * nd_c;
      for (int task = lane; task < TC * TC * NT; task += 64) {
        const int ab1 = task / (TC * NT), rem = task % (TC * NT), ab2 = rem / NT, a3 = rem % NT;
        CT acc[NC];
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) acc[b3] = 0;
        for (int g = 0; g < SP.ngroup; ++g) {
          const int p = SP.gpair[SP.gstart[g]];  // every pair of a group has these two factors
          const double* Xa = Bt + ((SP.pr[p] == 3) * NT + a3) * NQ;
          const CT* s2 = S2 + ((g * TC + ab1) * TC + ab2) * NQ;
          const CT z0 = (CT)Xa[0] * s2[0], z1 = (CT)Xa[1] * s2[1], z2 = (CT)Xa[2] * s2[2];
#pragma unroll
          for (int b3 = 0; b3 < NC; ++b3) {
            const double* Yb = Bc + ((SP.pc[p] == 3) * NC + b3) * NQ;
            acc[b3] += z0 * (CT)Yb[0] + z1 * (CT)Yb[1] + z2 * (CT)Yb[2];
          }
        }
        const int a1 = ab1 / NC, b1 = ab1 % NC, a2 = ab2 / NC, b2 = ab2 % NC;
        const int a = a1 + NT * (a2 + NT * a3);
        const int row = lext[a];
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) {
          const int b = b1 + NC * (b2 + NC * b3);
          Eo[row * nd_c + lexc[b]] = acc[b3];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
}

template <typename CT, int NT, int NC>
static int launch_sumfact(const hm_problem& P, int i, int j, int64_t e0, int64_t e1, const SumfactPlan& SP,
                          void* out, hipStream_t st) {
  constexpr int TC = NT * NC;
  const int ew = 24 + 2 * P.nd_u + P.nd_m;
  const size_t head = (2 * NT * 3 + 2 * NC * 3) * 8 + 64 * 4;
  const size_t wbytes = (size_t)SF_EPB * SP.npair * 27 * 8 + (size_t)SF_EPB * ew * 8;
  const size_t per_wave = ((size_t)SP.npair * TC * 9 + (size_t)SP.ngroup * TC * TC * 3) * sizeof(CT);
  const size_t bytes = head + wbytes + 4 * per_wave;
  int limit = 64 * 1024, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) limit = v;
  }
  if (bytes > (size_t)limit) HM_FAIL(HM_ERR_UNSUPPORTED, "the sum-factorized contraction needs %zu bytes of LDS", bytes);
  if (bytes > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute((const void*)k_element_sumfact<CT, NT, NC>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  const int64_t groups = (e1 - e0 + SF_EPB - 1) / SF_EPB;
  const int grid = (int)(groups < 8192 ? groups : 8192);
  hipLaunchKernelGGL((k_element_sumfact<CT, NT, NC>), dim3(grid), dim3(256), bytes, st, P, i, j, e0, e1, SP, out);
  return hip_check("k_element_sumfact");
}

}  // namespace hm

using namespace hm;

// The sum-factorized contraction when the problem carries tensor structure with 3 points per
// direction and 1-D orders 1 or 2; returns HM_ERR_UNSUPPORTED otherwise (the caller then takes
// another route).
int hm_element_matrices_sumfact(const hm_problem* P, int i, int j, int64_t e0, int64_t e1, int precision,
                                void* out, void* stream) {
  if (P->nq1 != 3 || P->nq != 27 || !P->tab1_u || !P->tab1_m || !P->lex_u || !P->lex_m)
    HM_FAIL(HM_ERR_UNSUPPORTED, "no tensor structure for the sum-factorized contraction");
  const int nt = i == HM_PARAMETER ? P->n1_m : P->n1_u;
  const int nc = j == HM_PARAMETER ? P->n1_m : P->n1_u;
  if (nt < 2 || nt > 3 || nc < 2 || nc > 3) HM_FAIL(HM_ERR_UNSUPPORTED, "1-D orders outside 1..2");
  const int parts = block_parts(i, j, P->density);
  SumfactPlan SP = {};
  // the nonzero (r, c) pairs of the block, grouped by their third-direction factor type
  const bool row_v = parts & 1, row_g = parts & 2, col_v = parts & 4, col_g = parts & 8;
  // which (r, c) pairs are nonzero: vv, vg, gv, gg parts of the density
  const bool nl = P->density == HM_DENSITY_NONLINEAR;
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
  (void)row_v, (void)row_g, (void)col_v, (void)col_g;
  int np = 0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const bool on = (r == 0 && c == 0) ? vv : (r == 0) ? vg : (c == 0) ? gv : gg;
      if (on) {
        SP.pr[np] = r;
        SP.pc[np] = c;
        ++np;
      }
    }
  SP.npair = np;
  if (np == 0) {
    const size_t n = (size_t)(e1 - e0) * (nt * nt * nt) * (nc * nc * nc);
    return hm_fill_zero(out, n, precision == HM_FP64 ? 8 : 4, (hipStream_t)stream);
  }
  int g = 0, at = 0;
  for (int type = 0; type < 4; ++type) {  // (r == 3, c == 3)
    const int start = at;
    for (int p = 0; p < np; ++p)
      if (((SP.pr[p] == 3) << 1 | (SP.pc[p] == 3)) == type) SP.gpair[at++] = p;
    if (at > start) {
      SP.gstart[g] = start;
      ++g;
    }
  }
  SP.gstart[g] = at;
  SP.ngroup = g;
  SP.need_fields = !(i * 3 + j == 6 && !nl);  // the linear Jacobian reads m and the geometry only
  hipStream_t st = (hipStream_t)stream;
  if (precision == HM_FP64) {
    if (nt == 3 && nc == 3) return launch_sumfact<double, 3, 3>(*P, i, j, e0, e1, SP, out, st);
    if (nt == 3 && nc == 2) return launch_sumfact<double, 3, 2>(*P, i, j, e0, e1, SP, out, st);
    if (nt == 2 && nc == 3) return launch_sumfact<double, 2, 3>(*P, i, j, e0, e1, SP, out, st);
    return launch_sumfact<double, 2, 2>(*P, i, j, e0, e1, SP, out, st);
  }
  if (precision != HM_FP32) HM_FAIL(HM_ERR_UNSUPPORTED, "precision %d", precision);
  if (nt == 3 && nc == 3) return launch_sumfact<float, 3, 3>(*P, i, j, e0, e1, SP, out, st);
  if (nt == 3 && nc == 2) return launch_sumfact<float, 3, 2>(*P, i, j, e0, e1, SP, out, st);
  if (nt == 2 && nc == 3) return launch_sumfact<float, 2, 3>(*P, i, j, e0, e1, SP, out, st);
  return launch_sumfact<float, 2, 2>(*P, i, j, e0, e1, SP, out, st);
}
