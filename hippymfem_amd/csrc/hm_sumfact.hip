// Sum-factorized element contraction for tensor-product hexahedra, in two kernels.
//
// The element matrix of a block is E[a][b] = sum_q sum_(r,c) D^[q][r][c] Psi_r(a, q) Psi_c(b, q),
// r, c in {value, d/dxi_1, d/dxi_2, d/dxi_3}.  On a tensor element every Psi factorizes over
// the three directions, Psi_r(a, q) = X1(a1, q1) X2(a2, q2) X3(a3, q3) with Xd the 1-D basis or
// its derivative (the derivative where r = d), so the sum over the 27 points is taken one
// direction at a time (stage 1 over q1, stage 2 over q2, stage 3 over q3).  Pairs (r, c) that
// share their third-direction factor are added before stage 3.  This is the reference's
// quadrature route (kernel.py:1431-1485) with its sums reordered; for the Jacobian of the
// diffusion density it takes 17.5k multiply-adds an element, or 7k where the element is
// axis-aligned, against 59k for the dense contraction.
//
// Kernel 1 (k_point_pairs): one thread per (element, point) computes, in FP64, the geometry of
// the element map, the fields and the block's pointwise pairs W[e][p][q] (D^ of pair p at point
// q) into a workspace.  Kernel 2 (k_sumfact_contract): one wavefront per element reads its
// pairs, leaves out the pairs that are round-off on this element, runs stage 1 into LDS and
// stages 2 and 3 in registers, and writes the element matrix.  Elements go in chunks whose workspace
// (128 MiB) stays in the last-level cache between the two kernels.
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_device.h"
#include "hm_point.h"

namespace hm {

constexpr int SF_MAXPAIR = 16;

struct SumfactPlan {
  int npair;
  int pr[SF_MAXPAIR], pc[SF_MAXPAIR];  // (test part r, trial part c) of each pair
  int tmask[4];                        // pairs of each third-direction type (r == 3, c == 3)
  int need_fields;                     // the pointwise blocks read u or p
};

// Kernel 1: thread (element e, point q) -> W[(e - e0) * npair + p][q].
__global__ void __launch_bounds__(256) k_point_pairs(hm_problem P, int s, int j, int64_t e0, int64_t e1,
                                                     SumfactPlan SP, double* __restrict__ W) {
  // 27 points (the sum-factorized route's rule); a chunk holds fewer than 2^31 points
  const int n = (int)((e1 - e0) * 27);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int el = t / 27, q = t - 27 * (t / 27);
    const int64_t e = e0 + el;
    PointState st;
    point_state_stream(P, e, q, SP.need_fields != 0, st);
    double Dq[4][4];
    point_block(s, j, st, Dq);
    double* w = W + (size_t)el * SP.npair * 27 + q;
    for (int p = 0; p < SP.npair; ++p) {
      const int rr = SP.pr[p], cc = SP.pc[p];
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) v = (r == rr && c == cc) ? Dq[r][c] : v;
      w[p * 27] = v;
    }
  }
}

// The rows of a 1-D table for derivative flag d, from LDS.
template <typename CT, int N>
__device__ inline void tab_rows(const double* tab, int d, CT out[N][3]) {
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int q = 0; q < 3; ++q) out[n][q] = (CT)tab[(d * N + n) * 3 + q];
}

// acc[a][b] += sum_q X[a][q] Y[b][q] z[q]
template <typename CT, int NT, int NC>
__device__ inline void sf_accum(const CT X[NT][3], const CT Y[NC][3], const CT z[3], CT acc[NT][NC]) {
#pragma unroll
  for (int a = 0; a < NT; ++a) {
    const CT x0 = X[a][0] * z[0], x1 = X[a][1] * z[1], x2 = X[a][2] * z[2];
#pragma unroll
    for (int b = 0; b < NC; ++b) acc[a][b] += x0 * Y[b][0] + x1 * Y[b][1] + x2 * Y[b][2];
  }
}

// A pair whose pointwise weight stays below this fraction of the element's largest is left out
// of that element's contraction: its share of every entry is below the rounding of the sum.
// On axis-aligned elements that drops the six off-diagonal gradient pairs of Jinv Jinv^T,
// which are round-off there.
#define SF_DROP 1e-16

constexpr int SF_WPB = 4;  // wavefronts (elements in flight) per block

template <typename CT, int NT, int NC>
__global__ void __launch_bounds__(256) k_sumfact_contract(const double* __restrict__ W, int64_t ne_chunk,
                                                          int s, int j, const hm_problem P, SumfactPlan SP,
                                                          CT* __restrict__ out) {
  constexpr int NQ = 3;
  constexpr int TC = NT * NC;  // 1-D (test node, trial node) pairs
  constexpr int nd_t = NT * NT * NT, nd_c = NC * NC * NC;
  extern __shared__ double lds[];
  const int npair = SP.npair;
  // LDS: local dof maps, the plan, per-wave pair lists, the 1-D tables, then per wave S1 and
  // the staged element matrix.
  int* lext = (int*)lds;                  // [nd_t]
  int* lexc = lext + 32;                  // [nd_c]
  int* spr = lexc + 32;                   // [SF_MAXPAIR]
  int* spc = spr + SF_MAXPAIR;
  int* wap_all = spc + SF_MAXPAIR;        // [SF_WPB][SF_MAXPAIR]
  double* tabt = (double*)(wap_all + SF_WPB * SF_MAXPAIR);  // [2][NT][3]
  double* tabc = tabt + 6 * NT;                             // [2][NC][3]
  unsigned long long* mxp_all = (unsigned long long*)(tabc + 6 * NC);  // [SF_WPB][SF_MAXPAIR]
  const int nw = npair * 27;
  double* Wl_all = (double*)(mxp_all + SF_WPB * SF_MAXPAIR);             // [SF_WPB][nw]
  CT* sbase = (CT*)(Wl_all + SF_WPB * nw);
  const int s1_words = npair * TC * 9;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  CT* S1 = sbase + (size_t)wave * s1_words;
  int* wap = wap_all + wave * SF_MAXPAIR;
  unsigned long long* mxp = mxp_all + wave * SF_MAXPAIR;
  double* Wl = Wl_all + wave * nw;
  {
    const double* t1t = s == HM_PARAMETER ? P.tab1_m : P.tab1_u;
    const double* t1c = j == HM_PARAMETER ? P.tab1_m : P.tab1_u;
    for (int i = threadIdx.x; i < 6 * NT; i += blockDim.x) tabt[i] = t1t[i];
    for (int i = threadIdx.x; i < 6 * NC; i += blockDim.x) tabc[i] = t1c[i];
    const int32_t* lt = s == HM_PARAMETER ? P.lex_m : P.lex_u;
    const int32_t* lc = j == HM_PARAMETER ? P.lex_m : P.lex_u;
    for (int i = threadIdx.x; i < nd_t; i += blockDim.x) lext[i] = lt[i];
    for (int i = threadIdx.x; i < nd_c; i += blockDim.x) lexc[i] = lc[i];
    if (threadIdx.x < SF_MAXPAIR) {
      spr[threadIdx.x] = SP.pr[threadIdx.x];
      spc[threadIdx.x] = SP.pc[threadIdx.x];
    }
  }
  __syncthreads();
  // Each wavefront walks its elements with the next element's pairs in flight: they are
  // loaded into registers (at most 7 a lane) while the current element is contracted.
  constexpr int PRE = (SF_MAXPAIR * 27 + 63) / 64;
  const int64_t stride = (int64_t)gridDim.x * SF_WPB;
  int64_t el = (int64_t)blockIdx.x * SF_WPB + wave;
  double pre[PRE];
#pragma unroll
  for (int k = 0; k < PRE; ++k) {
    const int idx = lane + 64 * k;
    pre[k] = (el < ne_chunk && idx < nw) ? W[el * nw + idx] : 0.0;
  }
  for (; el < ne_chunk; el += stride) {
    // the element's pairs into LDS, each pair's largest magnitude on the way (the bits of a
    // non-negative double order as the double does)
    if (lane < SF_MAXPAIR) mxp[lane] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
      const int idx = lane + 64 * k;
      if (idx < nw) {
        Wl[idx] = pre[k];
        atomicMax(&mxp[idx / 27], (unsigned long long)__double_as_longlong(fabs(pre[k])));
      }
    }
    // the next element's pairs, in flight from here on
    {
      const int64_t nx = el + stride;
#pragma unroll
      for (int k = 0; k < PRE; ++k) {
        const int idx = lane + 64 * k;
        pre[k] = (nx < ne_chunk && idx < nw) ? W[nx * nw + idx] : 0.0;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const double* Wk = Wl;
    // the pairs this element needs
    const double mx = lane < npair ? __longlong_as_double((long long)mxp[lane]) : 0.0;
    double top = mx;
    for (int off = 8; off > 0; off >>= 1) top = fmax(top, __shfl_xor(top, off, 16));
    top = __shfl(top, 0, 64);
    const uint32_t live = (uint32_t)__ballot(lane < npair && mx > SF_DROP * top);
    uint32_t mt[4];
    int ct[4], ot[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      mt[t] = live & (uint32_t)SP.tmask[t];
      ct[t] = __popc(mt[t]);
    }
    ot[0] = 0;
    ot[1] = ct[0];
    ot[2] = ct[0] + ct[1];
    ot[3] = ct[0] + ct[1] + ct[2];
    const int na = ot[3] + ct[3];
    if (lane < npair && ((live >> lane) & 1)) {
      const int t = ((spr[lane] == 3) << 1) | (spc[lane] == 3);
      const uint32_t m = t == 0 ? mt[0] : t == 1 ? mt[1] : t == 2 ? mt[2] : mt[3];
      const int o = t == 0 ? ot[0] : t == 1 ? ot[1] : t == 2 ? ot[2] : ot[3];
      wap[o + __popc(m & ((1u << lane) - 1u))] = lane;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // stage 1: task (active pair, q2 + 3 q3): S1[ip][a1 b1][q23]
    for (int task = lane; task < na * 9; task += 64) {
      const int ip = task / 9, q23 = task % 9;
      const int p = wap[ip];
      const double* wp = Wk + p * 27 + 3 * q23;
      const CT z[3] = {(CT)wp[0], (CT)wp[1], (CT)wp[2]};
      CT Xs[NT][3], Ys[NC][3];
      tab_rows<CT, NT>(tabt, spr[p] == 1, Xs);
      tab_rows<CT, NC>(tabc, spc[p] == 1, Ys);
      CT* o = S1 + ip * TC * 9 + q23;
#pragma unroll
      for (int a1 = 0; a1 < NT; ++a1) {
        const CT x0 = Xs[a1][0] * z[0], x1 = Xs[a1][1] * z[1], x2 = Xs[a1][2] * z[2];
#pragma unroll
        for (int b1 = 0; b1 < NC; ++b1) o[(a1 * NC + b1) * 9] = x0 * Ys[b1][0] + x1 * Ys[b1][1] + x2 * Ys[b1][2];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // stages 2 and 3: task (a1b1, a2b2); per type t of the third-direction factors,
    //   S2_t[q3] = sum_{p of type t} sum_q2 X2_p(a2, q2) Y2_p(b2, q2) S1_p[a1b1][q2 + 3 q3]
    // in registers, then E[a][b] += sum_q3 X3_t(a3, q3) Y3_t(b3, q3) S2_t[q3]; the pair and
    // type loops are the same for every lane
    CT* Eo = out + (size_t)el * nd_t * nd_c;
    for (int task = lane; task < TC * TC; task += 64) {
      const int ab1 = task / TC, ab2 = task % TC;
      const int a2 = ab2 / NC, b2 = ab2 % NC;
      CT acc[NT][NC];
#pragma unroll
      for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NC; ++b) acc[a][b] = 0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (ct[t] == 0) continue;
        CT s2[3] = {0, 0, 0};
        for (int ip = ot[t]; ip < ot[t] + ct[t]; ++ip) {
          const int p = wap[ip];
          const double* xr = tabt + ((spr[p] == 2) * NT + a2) * 3;
          const double* yr = tabc + ((spc[p] == 2) * NC + b2) * 3;
          const CT f0 = (CT)xr[0] * (CT)yr[0], f1 = (CT)xr[1] * (CT)yr[1], f2 = (CT)xr[2] * (CT)yr[2];
          const CT* s1 = S1 + (ip * TC + ab1) * 9;
#pragma unroll
          for (int q3 = 0; q3 < 3; ++q3) s2[q3] += f0 * s1[3 * q3] + f1 * s1[3 * q3 + 1] + f2 * s1[3 * q3 + 2];
        }
        CT Xs[NT][3], Ys[NC][3];
        tab_rows<CT, NT>(tabt, t >> 1, Xs);
        tab_rows<CT, NC>(tabc, t & 1, Ys);
        sf_accum<CT, NT, NC>(Xs, Ys, s2, acc);
      }
      const int a1 = ab1 / NC, b1 = ab1 % NC;
#pragma unroll
      for (int a3 = 0; a3 < NT; ++a3) {
        const int row = lext[a1 + NT * (a2 + NT * a3)];
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) Eo[row * nd_c + lexc[b1 + NC * (b2 + NC * b3)]] = acc[a3][b3];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Kernel 2, lane-mapped: lane (element, a1, b1, a2), NT*NC*NT lanes an element, no LDS.  The
// lane runs stage 1 for its (a1, b1) and stage 2 for its a2 in registers, pair after pair,
// and accumulates stage 3 into its NC*NT*NC entries (b2, a3, b3) of the element matrix.
template <typename CT, int NT, int NC>
__global__ void __launch_bounds__(256) k_sumfact_lanes(const double* __restrict__ W, int64_t ne_chunk,
                                                       int s, int j, const hm_problem P, SumfactPlan SP,
                                                       CT* __restrict__ out) {
  constexpr int LPE = NT * NC * NT;      // lanes an element
  constexpr int EPW = 64 / LPE;          // elements a wavefront
  constexpr int nd_t = NT * NT * NT, nd_c = NC * NC * NC;
  __shared__ double tabs[6 * (NT + NC)];
  const double* t1t = s == HM_PARAMETER ? P.tab1_m : P.tab1_u;
  const double* t1c = j == HM_PARAMETER ? P.tab1_m : P.tab1_u;
  for (int i = threadIdx.x; i < 6 * NT; i += blockDim.x) tabs[i] = t1t[i];
  for (int i = threadIdx.x; i < 6 * NC; i += blockDim.x) tabs[6 * NT + i] = t1c[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / LPE, r = lane % LPE;
  const int a1 = r % NT, b1 = (r / NT) % NC, a2 = r / (NT * NC);
  // the 1-D tables B[d][n][q] at uniform LDS addresses (broadcast reads)
  const double* Bt = tabs;
  const double* Bc = tabs + 6 * NT;
  const int32_t* lt = s == HM_PARAMETER ? P.lex_m : P.lex_u;
  const int32_t* lc = j == HM_PARAMETER ? P.lex_m : P.lex_u;
  const int npair = SP.npair;
  const bool in_use = slot < EPW;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * EPW;
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * EPW; base < ne_chunk; base += stride) {
    // the table reads stay inside the loop (hoisted, the products would take the registers)
    asm volatile("" ::: "memory");
    const int64_t e = base + slot;
    const bool act = in_use && e < ne_chunk;
    const double* We = W + (act ? e : 0) * npair * 27;
    // the pairs this element needs: lane r measures pairs r and r + LPE against the element's
    // largest pointwise weight
    double m0 = 0.0, m1 = 0.0;
    if (act && r < npair)
      for (int q = 0; q < 27; ++q) m0 = fmax(m0, fabs(We[r * 27 + q]));
    if (act && LPE + r < npair)
      for (int q = 0; q < 27; ++q) m1 = fmax(m1, fabs(We[(LPE + r) * 27 + q]));
    double top = 0.0;
    for (int p = 0; p < npair; ++p) {
      const double v0 = __shfl(m0, slot * LPE + p % LPE, 64), v1 = __shfl(m1, slot * LPE + p % LPE, 64);
      top = fmax(top, p < LPE ? v0 : v1);
    }
    const uint64_t bal0 = __ballot(act && r < npair && m0 > SF_DROP * top);
    const uint64_t bal1 = __ballot(act && LPE + r < npair && m1 > SF_DROP * top);
    const uint64_t lpe_mask = (1ull << LPE) - 1ull;
    const int sh = in_use ? slot * LPE : 0;
    const uint64_t both = ((bal0 >> sh) & lpe_mask) | (((bal1 >> sh) & lpe_mask) << LPE);
    const uint32_t live = (uint32_t)both & ((1u << npair) - 1u);
    CT acc[NC][NT][NC];
#pragma unroll
    for (int b2 = 0; b2 < NC; ++b2)
#pragma unroll
      for (int a3 = 0; a3 < NT; ++a3)
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) acc[b2][a3][b3] = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t tm = live & (uint32_t)SP.tmask[t];
      if (!act || tm == 0) continue;
      CT s2[NC][3];
#pragma unroll
      for (int b2 = 0; b2 < NC; ++b2) s2[b2][0] = s2[b2][1] = s2[b2][2] = 0;
      for (int p = 0; p < npair; ++p) {
        if (!((tm >> p) & 1)) continue;
        const int d1x = SP.pr[p] == 1, d1y = SP.pc[p] == 1, d2x = SP.pr[p] == 2, d2y = SP.pc[p] == 2;
        const double* w = We + p * 27;
        // stage 1: s1[q2 + 3 q3] = sum_q1 X(a1,q1) Y(b1,q1) W(q1, q2, q3)
        CT f1[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) f1[q] = (CT)(Bt[((d1x ? NT : 0) + a1) * 3 + q] * Bc[((d1y ? NC : 0) + b1) * 3 + q]);
        CT s1[9];
#pragma unroll
        for (int q23 = 0; q23 < 9; ++q23)
          s1[q23] = f1[0] * (CT)w[3 * q23] + f1[1] * (CT)w[3 * q23 + 1] + f1[2] * (CT)w[3 * q23 + 2];
        // stage 2: s2[b2][q3] += sum_q2 X(a2,q2) Y(b2,q2) s1[q2 + 3 q3]
        CT x2[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) x2[q] = (CT)Bt[((d2x ? NT : 0) + a2) * 3 + q];
#pragma unroll
        for (int b2 = 0; b2 < NC; ++b2) {
          CT f2[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) f2[q] = x2[q] * (CT)Bc[((d2y ? NC : 0) + b2) * 3 + q];
#pragma unroll
          for (int q3 = 0; q3 < 3; ++q3)
            s2[b2][q3] += f2[0] * s1[3 * q3] + f2[1] * s1[3 * q3 + 1] + f2[2] * s1[3 * q3 + 2];
        }
      }
      // stage 3 for this type of third-direction factors
#pragma unroll
      for (int a3 = 0; a3 < NT; ++a3)
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) {
          CT f3[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) f3[q] = (CT)(Bt[((t >> 1) * NT + a3) * 3 + q] * Bc[((t & 1) * NC + b3) * 3 + q]);
#pragma unroll
          for (int b2 = 0; b2 < NC; ++b2)
            acc[b2][a3][b3] += f3[0] * s2[b2][0] + f3[1] * s2[b2][1] + f3[2] * s2[b2][2];
        }
    }
    if (act) {
      CT* Eo = out + (size_t)e * nd_t * nd_c;
#pragma unroll
      for (int a3 = 0; a3 < NT; ++a3) {
        CT* Er = Eo + lt[a1 + NT * (a2 + NT * a3)] * nd_c;
#pragma unroll
        for (int b2 = 0; b2 < NC; ++b2)
#pragma unroll
          for (int b3 = 0; b3 < NC; ++b3) Er[lc[b1 + NC * (b2 + NC * b3)]] = acc[b2][a3][b3];
      }
    }
  }
}

// Per-device workspace of the pointwise pairs, allocated on first use and kept for the life of
// the process (128 MiB: a chunk of elements whose pairs stay in the last-level cache).
constexpr size_t SF_WORKSPACE = size_t(128) << 20;

static double* workspace() {
  static double* ws[64] = {nullptr};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  if (!ws[dev]) {
    void* p = nullptr;
    if (hipMalloc(&p, SF_WORKSPACE) != hipSuccess) return nullptr;
    ws[dev] = (double*)p;
  }
  return ws[dev];
}

template <typename CT, int NT, int NC>
static int launch_sumfact(const hm_problem& P, int i, int j, int64_t e0, int64_t e1, const SumfactPlan& SP,
                          void* out, hipStream_t st) {
  constexpr int TC = NT * NC;
  const size_t head = (64 + 2 * SF_MAXPAIR + SF_WPB * SF_MAXPAIR) * 4 + 6 * (NT + NC) * 8;
  const size_t per_wave = (size_t)SP.npair * TC * 9 * sizeof(CT) + (size_t)SP.npair * 27 * 8 + SF_MAXPAIR * 8;
  const size_t bytes = head + SF_WPB * per_wave;
  int limit = 64 * 1024, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) limit = v;
  }
  if (bytes > (size_t)limit)
    HM_FAIL(HM_ERR_UNSUPPORTED, "the sum-factorized contraction needs %zu bytes of LDS", bytes);
  if (bytes > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute((const void*)k_sumfact_contract<CT, NT, NC>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  const char* lv = getenv("HM_SF_LANES");
  const bool lanes = lv && lv[0] == '1';
  double* W = workspace();
  if (!W) HM_FAIL(HM_ERR_HIP, "the %zu-byte workspace of the contraction could not be allocated", SF_WORKSPACE);
  const int64_t chunk = (int64_t)(SF_WORKSPACE / ((size_t)SP.npair * P.nq * sizeof(double)));
  const size_t nd = (size_t)(NT * NT * NT) * (NC * NC * NC);
  for (int64_t a = e0; a < e1; a += chunk) {
    const int64_t b = a + chunk < e1 ? a + chunk : e1;
    const int64_t nthreads = (b - a) * P.nq;
    const int64_t nb = (nthreads + 255) / 256;
    hipLaunchKernelGGL(k_point_pairs, dim3((unsigned)(nb < 1048576 ? nb : 1048576)), dim3(256), 0, st, P, i, j,
                       a, b, SP, W);
    if (lanes) {
      constexpr int EPB = (64 / (NT * NC * NT)) * 4;  // elements a block of 4 wavefronts
      const int64_t groups = (b - a + EPB - 1) / EPB;
      const int grid = (int)(groups < 8192 ? groups : 8192);
      hipLaunchKernelGGL((k_sumfact_lanes<CT, NT, NC>), dim3(grid), dim3(256), 0, st, W, b - a, i, j, P, SP,
                         (CT*)out + (size_t)(a - e0) * nd);
    } else {
      // persistent: a few blocks per compute unit, each wavefront walking many elements
      const int64_t groups = (b - a + SF_WPB - 1) / SF_WPB;
      const int grid = (int)(groups < 1024 ? groups : 1024);
      hipLaunchKernelGGL((k_sumfact_contract<CT, NT, NC>), dim3(grid), dim3(256), bytes, st, W, b - a, i, j, P,
                         SP, (CT*)out + (size_t)(a - e0) * nd);
    }
  }
  return hip_check("k_sumfact_contract");
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
  if (P->nd_u > 32 || P->nd_m > 32) HM_FAIL(HM_ERR_UNSUPPORTED, "more than 32 dofs an element");
  const int nt = i == HM_PARAMETER ? P->n1_m : P->n1_u;
  const int nc = j == HM_PARAMETER ? P->n1_m : P->n1_u;
  if (nt < 2 || nt > 3 || nc < 2 || nc > 3) HM_FAIL(HM_ERR_UNSUPPORTED, "1-D orders outside 1..2");
  SumfactPlan SP = {};
  // the nonzero (r, c) pairs of the block: the vv, vg, gv, gg parts of the density
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
  for (int p = 0; p < np; ++p) SP.tmask[((SP.pr[p] == 3) << 1) | (SP.pc[p] == 3)] |= 1 << p;
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
