// Sum-factorized element matrices for tensor-product hexahedra, one kernel.
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
// A wavefront holds 64 / (NT NC max(NT, NC)) elements (at most 3 with 4 points a direction).  The lanes of an element first stage its vertex
// coordinates and dof values in LDS, then act as its quadrature points and write the block's
// pointwise pairs W[p][q] (D^ of pair p at point q, FP64) to LDS, then act as its contraction
// tasks (a1, b1, x2), x2 = b2 or a2: each runs stages 1 and 2 in registers and stage 3 into its
// entries (a2 or b2, a3, b3) of the element matrix, which it writes to memory.
// The pointwise pairs never leave the compute unit.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hm_common.h"
#include "hm_device.h"
#include "hm_point.h"
#include "hm_sumfact.h"

namespace hm {

constexpr int SF_WPB = 4;  // wavefronts a block

// Elements a wavefront: as many as its 64 lanes hold contraction tasks (NT NC max(NT, NC) an
// element), at most 3 with 4 points a direction, whose pointwise pairs take 64 LDS words each
// (8 P1 x P1 elements a wavefront would need more LDS than a compute unit has)
template <int NT, int NC, int NQ>
constexpr int sf_epw() {
  constexpr int lpe = NT * NC * (NC >= NT ? NC : NT);
  return (NQ > 3 && 64 / lpe > 3) ? 3 : 64 / lpe;
}

template <typename CT, int NT, int NC, int NQ>
constexpr int sf_waves() {
  return NQ > 3 ? 2 : (sizeof(CT) == 4 && NT == 3 && NC == 3) ? 4 : 3;
}

// NB = 2: the kernel writes two blocks of one shape (test and trial orders) from one pass over
// the elements: geometry, fields and the pointwise pairs of both blocks once, then the two
// contractions one after the other (the first block's matrix is stored at once, the second's
// as the single block's is).
template <typename CT, int NT, int NC, int QP, int NB, int NQ>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sf_waves<CT, NT, NC, NQ>(), sf_waves<CT, NT, NC, NQ>()))) k_sumfact(const hm_problem P, int s, int j, int64_t e0, int64_t e1,
                                                 SumfactPlan SP, const double* __restrict__ t1t,
                                                 const double* __restrict__ t1c, CT* __restrict__ out,
                                                 CT* __restrict__ out2) {
  // Lane task (a1, b1, x2): x2 = b2 with the a2 loop in registers where NC >= NT (for a fixed
  // entry index the lanes of one a1 then write NC^2 adjacent columns), x2 = a2 with the b2 loop
  // otherwise (fewer registers for the wide-by-narrow blocks).
  constexpr bool BMAP = NC >= NT;
  constexpr int N2 = BMAP ? NT : NC;          // the stage-2 index held in registers
  constexpr int LPE = NT * NC * (BMAP ? NC : NT);  // lanes an element
  constexpr int EPW = sf_epw<NT, NC, NQ>();  // elements a wavefront
  constexpr int NPF = (27 + LPE - 1) / LPE;  // dofs a lane fetches
  constexpr int nd_t = NT * NT * NT, nd_c = NC * NC * NC;
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ;  // points a plane, points
  // the products of the 1-D tables, PT[dt][dc][a][b][q] = Xt_dt(a, q) Xc_dc(b, q), d the
  // derivative flags: every factor of stages 1, 2 and 3
  __shared__ double PT[4 * NT * NC * NQ];
  __shared__ int32_t LX[nd_t + nd_c];  // local dof of each tensor index, test then trial
  __shared__ int32_t UT[32];            // tensor index of each local state dof
  extern __shared__ double lds[];
  const int npair = SP.npair, nu = P.nd_u, nm = P.nd_m;
  const bool fields = SP.need_fields != 0;
  double* TX = lds;
  double* TM = TX + NQ3 * 24;
  double* TW = TM + NQ3 * nm;
  double* T1U = TW + NQ3;
  {
    const int32_t* lxt = s == HM_PARAMETER ? P.lex_m : P.lex_u;
    const int32_t* lxc = j == HM_PARAMETER ? P.lex_m : P.lex_u;
    for (int i = threadIdx.x; i < nd_t; i += blockDim.x) LX[i] = lxt[i];
    for (int i = threadIdx.x; i < nd_c; i += blockDim.x) LX[nd_t + i] = lxc[i];
    for (int i = threadIdx.x; i < 4 * NT * NC * NQ; i += blockDim.x) {
      const int q = i % NQ, b = (i / NQ) % NC, a = (i / (NQ * NC)) % NT, dd = i / (NQ * NC * NT);
      PT[i] = t1t[((dd >> 1) * NT + a) * NQ + q] * t1c[((dd & 1) * NC + b) * NQ + q];
    }
    // the tables point-minor: entry (row, q) at row * NQ3 + q
    for (int i = threadIdx.x; i < NQ3 * 24; i += blockDim.x) TX[(i % 24) * NQ3 + i / 24] = P.tab_x[i];
    for (int i = threadIdx.x; i < NQ3 * nm; i += blockDim.x) TM[(i % nm) * NQ3 + i / nm] = P.tab_m[i];
    for (int i = threadIdx.x; i < NQ3; i += blockDim.x) TW[i] = P.wq[i];
    if (fields) {
      for (int i = threadIdx.x; i < 2 * P.n1_u * NQ; i += blockDim.x) T1U[i] = P.tab1_u[i];
      for (int i = threadIdx.x; i < nu; i += blockDim.x) UT[P.lex_u[i]] = i;
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  // The lane's fixed quantities (its element slot, task and LDS addresses).  The Q2 x Q2 forms
  // derive them again every iteration from a value the compiler cannot see through: kept live
  // across the loop they took registers that the FP64 form then spilled to scratch memory and
  // reloaded every iteration through the vector memory path the element-matrix stores need,
  // and that the single-precision forms need to fit four wavefronts a SIMD.
  constexpr bool REDERIVE = NT == 3 && NC == 3;
  int lane, slot, r, a1, b1, x2;
  bool in_use;
  double *Wl, *Xl, *Ul, *Pl, *Ml;
  auto derive = [&](int lo) {
    lane = lo & 63;
    slot = lane / LPE;
    r = lane % LPE;
    in_use = slot < EPW;
    a1 = r % NT;
    b1 = (r / NT) % NC;
    x2 = r / (NT * NC);
    Wl = lds + sf_table_words(nu, nm, fields, NQ) +
         (size_t)(wave * EPW + (in_use ? slot : 0)) * sf_element_words(npair, nu, nm, NQ3);
    Xl = Wl + npair * NQ3;
    Ul = Xl + 24;
    Pl = Ul + nu;
    Ml = Pl + nu;
  };
  derive((int)threadIdx.x);
  // the (a2, b2) of register index i
  auto a2_of = [&](int i) { return BMAP ? i : x2; };
  auto b2_of = [&](int i) { return BMAP ? x2 : i; };
  const int32_t* lt = LX;
  const int32_t* lc = LX + nd_t;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * SF_WPB * EPW;
  // The next element's data travels in two steps, so that no wait falls on the stores of an
  // element matrix (a wait for a load also waits for the memory operations issued before it):
  // its ids (vertex, dofs) are loaded early in an iteration, its vertex coordinates and dof
  // values after the contraction and before the stores, and they are put in LDS at the top of
  // the next iteration.  Lane r holds vertex r and dofs r, r + LPE, ...
  int32_t ix = 0, iu[NPF], im[NPF];
  double fx[3] = {0.0, 0.0, 0.0}, fu[NPF], fp[NPF], fm[NPF];
  auto fetch_ids = [&](int64_t el) {
    const bool ok = in_use && el < ne;
    const int64_t e = e0 + (ok ? el : 0);
    if (ok && r < 8) ix = P.elem_verts[e * 8 + r];
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int a = r + k * LPE;
      if (ok && fields && a < nu) iu[k] = P.edofs_u[e * nu + a];
      if (ok && a < nm) im[k] = P.edofs_m[e * nm + a];
    }
  };
  auto fetch_values = [&](int64_t el) {
    const bool ok = in_use && el < ne;
    // one wait for the ids on every path (vmcnt(0) only: expcnt and lgkmcnt at their maxima),
    // so that the coordinate and dof-value loads below leave together: waits placed by the
    // compiler at each use of an id also waited for the loads issued just before them (after
    // a store, loads and stores share the counter and it waits for all)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (ok && r < 8) {
      fx[0] = P.verts[(int64_t)ix * 3];
      fx[1] = P.verts[(int64_t)ix * 3 + 1];
      fx[2] = P.verts[(int64_t)ix * 3 + 2];
    }
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int a = r + k * LPE;
      if (ok && fields && a < nu) {
        fu[k] = P.u[iu[k]];
        fp[k] = P.p[iu[k]];
      }
      if (ok && a < nm) fm[k] = P.m[im[k]];
    }
  };
  // the element matrix of the previous iteration (-1: none), written once the next element's
  // values have arrived, so the stores never hold up a wait
  constexpr bool LOWP = QP == HM_BF16 || QP == HM_FP16;
  static_assert(NB == 1 || NB == 2, "one or two blocks a pass");
  static_assert(NB == 1 || !LOWP, "two blocks a pass in FP64 or FP32 only");
  CT acc[N2][NT][NC];
  int64_t prev_el = -1;
  double prev_scale = 1.0;
  bool bad = false;  // a value stored was not finite
  auto store_acc = [&](CT* dst, int64_t e_, double scale) {
    CT* Eo = dst + (size_t)e_ * nd_t * nd_c;
#pragma unroll
    for (int a3 = 0; a3 < NT; ++a3)
#pragma unroll
      for (int i = 0; i < N2; ++i) {
        CT* Er = Eo + lt[a1 + NT * (a2_of(i) + NT * a3)] * nd_c;
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) {
          const CT v = LOWP ? (CT)(acc[i][a3][b3] * scale) : acc[i][a3][b3];
          Er[lc[b1 + NC * (b2_of(i) + NC * b3)]] = v;
          bad |= !__builtin_isfinite(v);
        }
      }
  };
  auto store_prev = [&]() {
    if (prev_el < 0) return;
    store_acc(NB == 2 ? out2 : out, prev_el, prev_scale);
  };
  int64_t base = ((int64_t)blockIdx.x * SF_WPB + wave) * EPW;
  fetch_ids(base + slot);
  fetch_values(base + slot);
  for (; base < ne; base += stride) {
    // the table reads stay inside the loop (hoisted, the products would take the registers)
    asm volatile("" ::: "memory");
    if constexpr (REDERIVE) {
      int lo = (int)threadIdx.x;
      asm volatile("" : "+v"(lo));
      derive(lo);
    }
    const int64_t el = base + slot;
    const bool act = in_use && el < ne;
    // 1. the element's vertex coordinates and dof values into LDS; the next element's ids
    if (act) {
      if (r < 8) {
        Xl[r * 3] = fx[0];
        Xl[r * 3 + 1] = fx[1];
        Xl[r * 3 + 2] = fx[2];
      }
#pragma unroll
      for (int k = 0; k < NPF; ++k) {
        const int a = r + k * LPE;
        if (fields && a < nu) {
          Ul[UT[a]] = fu[k];
          Pl[UT[a]] = fp[k];
        }
        if (a < nm) Ml[a] = fm[k];
      }
    }
    // Is the element an axis-aligned box (every vertex coordinate that of the vertex with
    // the same index along that axis)?  Then J is diagonal and the pairs of two different
    // derivatives are round-off: they are left out of its contraction.
    const uint64_t lpe_mask = (1ull << LPE) - 1ull;
    const int sh = in_use ? slot * LPE : 0;
    bool aligned;
    {
      const double rx = __shfl(fx[0], sh + (r & 1), 64), ry = __shfl(fx[1], sh + (r & 2), 64),
                   rz = __shfl(fx[2], sh + (r & 4), 64);
      const uint64_t bad = __ballot(act && r < 8 && !(fx[0] == rx && fx[1] == ry && fx[2] == rz));
      aligned = ((bad >> sh) & lpe_mask) == 0;
    }
    store_prev();
    fetch_ids(el + stride);
    wave_sync();
    // 2. lane = quadrature point: the pointwise pairs into LDS, in the order of k = 4 r + c
    // (the precision forms make the fields, kappa and the pairs in FP32, the geometry in FP64)
    auto point_task = [&](int q, double* We, const double* Xe, const double* Ue, const double* Me,
                          const double* Pe, bool al) {
      using PT = typename std::conditional<QP == HM_FP64, double, float>::type;
      PointStateT<PT> S;
      sf_point_state<NQ, PT>(P, q, Xe, Ue, Me, Pe, fields, TX, TM, TW, T1U, S, al);
      PT D[4][4];
      point_block(s, j, S, D);
      int p = 0;
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if ((SP.kmask >> k) & 1) {
          const double v = D[k >> 2][k & 3];
          We[p * NQ3 + q] = v;
          ++p;
        }
      if constexpr (NB == 2) {
        point_block(SP.s2, SP.j2, S, D);
#pragma unroll
        for (int k = 0; k < 16; ++k)
          if ((SP.kmask2 >> k) & 1) {
            We[p * NQ3 + q] = D[k >> 2][k & 3];
            ++p;
          }
      }
    };
    if constexpr (NQ > 3) {
      // more points than an element's lanes: every lane of the wavefront takes the next
      // (element, point) of its elements, so no round leaves lanes idle but the last
      const int words = sf_element_words(npair, nu, nm, NQ3);
      double* W0 = lds + sf_table_words(nu, nm, fields, NQ) + (size_t)wave * EPW * words;
      // the alignment of each element of the wavefront (its first lane's flag)
      const uint64_t amask = __ballot(aligned && r == 0);
      for (int t = lane; t < EPW * NQ3; t += 64) {
        const int ee = t / NQ3, q = t - ee * NQ3;
        if (base + ee < ne) {
          double* We = W0 + (size_t)ee * words;
          const double* Xe = We + npair * NQ3;
          point_task(q, We, Xe, Xe + 24, Xe + 24 + 2 * nu, Xe + 24 + nu, (amask >> (ee * LPE)) & 1);
        }
      }
    } else if (act) {
      for (int q = r; q < NQ3; q += LPE) point_task(q, Wl, Xl, Ul, Ml, Pl, aligned);
    }
    wave_sync();
    // 3. the pairs this element needs.  The low-precision forms also measure every pair
    // against the element's largest pointwise weight, their scale, and leave out the pairs
    // below SF_DROP of it.
    const uint32_t every = npair >= 32 ? ~0u : (1u << npair) - 1u;
    uint32_t live = every & (aligned ? ~(uint32_t)SP.offdiag : ~0u);
    double top = 1.0;
    if constexpr (LOWP) {
      double m0 = 0.0, m1 = 0.0;
      if (act && r < npair)
        for (int q = 0; q < NQ3; ++q) m0 = fmax(m0, fabs(Wl[r * NQ3 + q]));
      if (act && LPE + r < npair)
        for (int q = 0; q < NQ3; ++q) m1 = fmax(m1, fabs(Wl[(LPE + r) * NQ3 + q]));
      top = 0.0;
      for (int p = 0; p < npair; ++p) {
        const double v0 = __shfl(m0, slot * LPE + p % LPE, 64), v1 = __shfl(m1, slot * LPE + p % LPE, 64);
        top = fmax(top, p < LPE ? v0 : v1);
      }
      const uint64_t bal0 = __ballot(act && r < npair && m0 > SF_DROP * top);
      const uint64_t bal1 = __ballot(act && LPE + r < npair && m1 > SF_DROP * top);
      const uint64_t both = ((bal0 >> sh) & lpe_mask) | (((bal1 >> sh) & lpe_mask) << LPE);
      live = (uint32_t)both & every;
    }
    // 4. lane = contraction task (a1, b1, x2).  The low-precision forms scale the element's
    // pairs by its largest pointwise weight (operands within [-1, 1]) and the result back.
    const double wscale = (LOWP && top > 0.0) ? 1.0 / top : 1.0;
#pragma unroll 1
    for (int blk = 0; blk < NB; ++blk) {
    const uint32_t bm = NB == 2 ? (uint32_t)SP.bmask[blk] : ~0u;
#pragma unroll
    for (int i = 0; i < N2; ++i)
#pragma unroll
      for (int a3 = 0; a3 < NT; ++a3)
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) acc[i][a3][b3] = 0;
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
      const uint32_t tm = live & (uint32_t)SP.tmask[t] & bm;
      if (!act || tm == 0) continue;
      CT s2[N2][NQ];
#pragma unroll
      for (int i = 0; i < N2; ++i)
#pragma unroll
        for (int q = 0; q < NQ; ++q) s2[i][q] = 0;
      for (int p = 0; p < npair; ++p) {
        if (!((tm >> p) & 1)) continue;
        const int d1 = ((SP.pr[p] == 1) << 1) | (SP.pc[p] == 1);
        const int d2 = ((SP.pr[p] == 2) << 1) | (SP.pc[p] == 2);
        // every LDS operand of the pair in flight at once (one wait, not one a use)
        const double* wv = Wl + p * NQ3;
        const double* f1 = PT + ((d1 * NT + a1) * NC + b1) * NQ;
        double gd[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) gd[q] = f1[q];
        double hd[N2][NQ];
#pragma unroll
        for (int i = 0; i < N2; ++i) {
          const double* f2 = PT + ((d2 * NT + a2_of(i)) * NC + b2_of(i)) * NQ;
#pragma unroll
          for (int q = 0; q < NQ; ++q) hd[i][q] = f2[q];
        }
        __builtin_amdgcn_sched_barrier(0);
        CT g[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) g[q] = quant<QP>((CT)gd[q]);
        // one q3 plane at a time:
        //   stage 1: s1[q2] = sum_q1 X(a1, q1) Y(b1, q1) W(q1, q2, q3)
        //   stage 2: s2[i][q3] += sum_q2 X(a2, q2) Y(b2, q2) s1[q2], (a2, b2) of i
#pragma unroll
        for (int q3 = 0; q3 < NQ; ++q3) {
          CT s1[NQ];
#pragma unroll
          for (int q2 = 0; q2 < NQ; ++q2) {
            const double* wr = wv + NQ * q2 + NQ2 * q3;
            if constexpr (LOWP && NQ == 3) {
              const CT w0 = quant<QP>((CT)(wr[0] * wscale)), w1 = quant<QP>((CT)(wr[1] * wscale)),
                       w2 = quant<QP>((CT)(wr[2] * wscale));
              s1[q2] = quant<QP>(fma(g[2], w2, fma(g[1], w1, g[0] * w0)));
            } else if constexpr (LOWP) {
              CT w[NQ];
#pragma unroll
              for (int q1 = 0; q1 < NQ; ++q1) w[q1] = quant<QP>((CT)(wr[q1] * wscale));
              CT a = g[0] * w[0];
#pragma unroll
              for (int q1 = 1; q1 < NQ; ++q1) a = fma(g[q1], w[q1], a);
              s1[q2] = quant<QP>(a);
            } else if constexpr (NQ == 3) {
              s1[q2] = fma(g[2], (CT)wr[2], fma(g[1], (CT)wr[1], g[0] * (CT)wr[0]));
            } else {
              CT a = g[0] * (CT)wr[0];
#pragma unroll
              for (int q1 = 1; q1 < NQ; ++q1) a = fma(g[q1], (CT)wr[q1], a);
              s1[q2] = a;
            }
          }
          if constexpr (NQ == 3) {
#pragma unroll
            for (int i = 0; i < N2; ++i)
              s2[i][q3] = fma(quant<QP>((CT)hd[i][2]), s1[2],
                              fma(quant<QP>((CT)hd[i][1]), s1[1], fma(quant<QP>((CT)hd[i][0]), s1[0], s2[i][q3])));
          } else {
#pragma unroll
            for (int i = 0; i < N2; ++i) {
              CT a = s2[i][q3];
#pragma unroll
              for (int q2 = 0; q2 < NQ; ++q2) a = fma(quant<QP>((CT)hd[i][q2]), s1[q2], a);
              s2[i][q3] = a;
            }
          }
        }
      }
      // stage 3 for this type of third-direction factors, the 1-D rows read as wavefront-uniform
      // values from memory (scalar loads):
      //   E[i][a3][b3] += sum_q3 Xt(a3, q3) (Yc(b3, q3) s2[i][q3])
      const double* bt = t1t + (t >> 1) * NT * NQ;
      const double* bc = t1c + (t & 1) * NC * NQ;
      if constexpr (LOWP) {
        // the stage-3 operands: s2 and the table products, each rounded
#pragma unroll
        for (int i = 0; i < N2; ++i)
#pragma unroll
          for (int q = 0; q < NQ; ++q) s2[i][q] = quant<QP>(s2[i][q]);
#pragma unroll
        for (int a3 = 0; a3 < NT; ++a3)
#pragma unroll
          for (int b3 = 0; b3 < NC; ++b3) {
            const double* f3 = PT + ((t * NT + a3) * NC + b3) * NQ;
            if constexpr (NQ == 3) {
              const CT k0 = quant<QP>((CT)f3[0]), k1 = quant<QP>((CT)f3[1]), k2 = quant<QP>((CT)f3[2]);
#pragma unroll
              for (int i = 0; i < N2; ++i)
                acc[i][a3][b3] = fma(k2, s2[i][2], fma(k1, s2[i][1], fma(k0, s2[i][0], acc[i][a3][b3])));
            } else {
              CT k[NQ];
#pragma unroll
              for (int q = 0; q < NQ; ++q) k[q] = quant<QP>((CT)f3[q]);
#pragma unroll
              for (int i = 0; i < N2; ++i) {
                CT a = acc[i][a3][b3];
#pragma unroll
                for (int q = 0; q < NQ; ++q) a = fma(k[q], s2[i][q], a);
                acc[i][a3][b3] = a;
              }
            }
          }
        continue;
      }
#pragma unroll
      for (int b3 = 0; b3 < NC; ++b3) {
        CT u[N2][NQ];
#pragma unroll
        for (int i = 0; i < N2; ++i)
#pragma unroll
          for (int q = 0; q < NQ; ++q) u[i][q] = (CT)bc[b3 * NQ + q] * s2[i][q];
#pragma unroll
        for (int a3 = 0; a3 < NT; ++a3) {
          if constexpr (NQ == 3) {
            const CT k0 = (CT)bt[a3 * 3], k1 = (CT)bt[a3 * 3 + 1], k2 = (CT)bt[a3 * 3 + 2];
#pragma unroll
            for (int i = 0; i < N2; ++i)
              acc[i][a3][b3] = fma(k2, u[i][2], fma(k1, u[i][1], fma(k0, u[i][0], acc[i][a3][b3])));
          } else {
            CT k[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) k[q] = (CT)bt[a3 * NQ + q];
#pragma unroll
            for (int i = 0; i < N2; ++i) {
              CT a = acc[i][a3][b3];
#pragma unroll
              for (int q = 0; q < NQ; ++q) a = fma(k[q], u[i][q], a);
              acc[i][a3][b3] = a;
            }
          }
        }
      }
    }
    // the first of two blocks leaves at once (its accumulators are needed for the second)
    if constexpr (NB == 2)
      if (blk == 0 && act) store_acc(out, el, 1.0);
    }
    // the next element's values in flight; this element's matrix leaves at the top of the
    // next iteration, after those values are waited for
    fetch_values(el + stride);
    prev_el = act ? el : -1;
    prev_scale = top;
    // the LDS of this element is rewritten by the next
    wave_sync();
  }
  store_prev();
  if (bad && P.nonfinite) atomicOr(P.nonfinite, 1);
}

template <typename CT, int NT, int NC, int QP, int NB, int NQ>
static int launch_sumfact(const hm_problem& P, int i, int j, int64_t e0, int64_t e1, const SumfactPlan& SP,
                          void* out, void* out2, hipStream_t st) {
  constexpr int EPW = sf_epw<NT, NC, NQ>();
  const size_t bytes = ((size_t)sf_table_words(P.nd_u, P.nd_m, SP.need_fields != 0, NQ) +
                        (size_t)SF_WPB * EPW * sf_element_words(SP.npair, P.nd_u, P.nd_m, NQ * NQ * NQ)) *
                       sizeof(double);
  int limit = 64 * 1024, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) limit = v;
  }
  if (bytes + 4 * NT * NC * NQ * sizeof(double) > (size_t)limit)
    HM_FAIL(HM_ERR_UNSUPPORTED, "the sum-factorized contraction needs %zu bytes of LDS", bytes);
  if (bytes > 48 * 1024) {
    hipError_t err = hipFuncSetAttribute((const void*)k_sumfact<CT, NT, NC, QP, NB, NQ>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  // persistent: four times as many blocks as the device holds at once, each walking many
  // elements (the tables are loaded once a block and the next element's data is in flight
  // meanwhile); the later blocks fill the slots the first ones free, which balances the end
  // of the launch (measured at 64^3: up to 3 % faster than one block a slot)
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sumfact<CT, NT, NC, QP, NB, NQ>, 256, bytes) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  const int64_t groups = (e1 - e0 + SF_WPB * EPW - 1) / (SF_WPB * EPW);
  const int64_t full = (int64_t)per_cu * cus * 4;
  const int grid = (int)(groups < full ? groups : full);
  hipLaunchKernelGGL((k_sumfact<CT, NT, NC, QP, NB, NQ>), dim3(grid), dim3(256), bytes, st, P, i, j, e0, e1, SP,
                     (const double*)(i == HM_PARAMETER ? P.tab1_m : P.tab1_u),
                     (const double*)(j == HM_PARAMETER ? P.tab1_m : P.tab1_u), (CT*)out, (CT*)out2);
  return hip_check("k_sumfact");
}

template <typename CT, int QP, int NB = 1>
static int dispatch_orders(const hm_problem& P, int i, int j, int64_t e0, int64_t e1, int nt, int nc,
                           const SumfactPlan& SP, void* out, hipStream_t st, void* out2 = nullptr) {
  // 3 points a direction (the 3 x 3 x 3 rule) or 4 (4 x 4 x 4, degree 6 or 7)
  if (P.nq1 == 4) {
    if (nt == 3 && nc == 3) return launch_sumfact<CT, 3, 3, QP, NB, 4>(P, i, j, e0, e1, SP, out, out2, st);
    if (nt == 3 && nc == 2) return launch_sumfact<CT, 3, 2, QP, NB, 4>(P, i, j, e0, e1, SP, out, out2, st);
    if (nt == 2 && nc == 3) return launch_sumfact<CT, 2, 3, QP, NB, 4>(P, i, j, e0, e1, SP, out, out2, st);
    return launch_sumfact<CT, 2, 2, QP, NB, 4>(P, i, j, e0, e1, SP, out, out2, st);
  }
  if (nt == 3 && nc == 3) return launch_sumfact<CT, 3, 3, QP, NB, 3>(P, i, j, e0, e1, SP, out, out2, st);
  if (nt == 3 && nc == 2) return launch_sumfact<CT, 3, 2, QP, NB, 3>(P, i, j, e0, e1, SP, out, out2, st);
  if (nt == 2 && nc == 3) return launch_sumfact<CT, 2, 3, QP, NB, 3>(P, i, j, e0, e1, SP, out, out2, st);
  return launch_sumfact<CT, 2, 2, QP, NB, 3>(P, i, j, e0, e1, SP, out, out2, st);
}

// Element vectors dR_e / d(slot s dofs) = sum_q w [val N_a + (Jinv grad) . G_a] on tensor
// problems: 64 / NQ^3 elements a wavefront (NQ^3 lanes each: two for the 3 x 3 x 3 rule, one
// for 4 x 4 x 4), the lanes first the element's points (the
// geometry and the fields by the sum-factorized point state, from the LDS tables of k_sumfact)
// writing the pointwise weights into the wavefront's own LDS region, then the element's dofs
// (the sum over the points against the basis, in LDS for the block).
template <int NQ>
__global__ void __launch_bounds__(256) k_sumfact_vectors(const hm_problem P, int s, int64_t e0, int64_t e1,
                                                         double* __restrict__ out) {
  constexpr int NQ3 = NQ * NQ * NQ, EPW = 64 / NQ3;
  extern __shared__ double lds[];
  __shared__ int32_t UT[32];
  const int nu = P.nd_u, nm = P.nd_m;
  const int nd = s == HM_PARAMETER ? nm : nu;
  const double* tab = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  double* TX = lds;
  double* TM = TX + NQ3 * 24;
  double* TW = TM + NQ3 * nm;
  double* T1U = TW + NQ3;
  double* Ps = T1U + 6 * NQ;  // Ps[q][r][a]: r = 0 the value, 1..3 the reference derivatives
  for (int i = threadIdx.x; i < NQ3 * 24; i += blockDim.x) TX[(i % 24) * NQ3 + i / 24] = P.tab_x[i];
  for (int i = threadIdx.x; i < NQ3 * nm; i += blockDim.x) TM[(i % nm) * NQ3 + i / nm] = P.tab_m[i];
  for (int i = threadIdx.x; i < NQ3; i += blockDim.x) TW[i] = P.wq[i];
  for (int i = threadIdx.x; i < 2 * P.n1_u * NQ; i += blockDim.x) T1U[i] = P.tab1_u[i];
  for (int i = threadIdx.x; i < nu; i += blockDim.x) UT[P.lex_u[i]] = i;
  for (int i = threadIdx.x; i < NQ3 * 4 * nd; i += blockDim.x) {
    const int a = i % nd, r = (i / nd) % 4, q = i / (4 * nd);
    Ps[i] = r == 0 ? tab[q * nd + a] : tab[NQ3 * nd + (q * nd + a) * 3 + (r - 1)];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = lane / NQ3, r = lane % NQ3;
  const bool in_use = slot < EPW;
  const int words = NQ3 * 4 + 24 + 2 * nu + nm;
  double* W = Ps + NQ3 * 4 * nd + (size_t)(wave * EPW + (in_use ? slot : 0)) * words;
  double* X = W + NQ3 * 4;
  double* U = X + 24;
  double* Pv = U + nu;
  double* M = Pv + nu;
  const int64_t ne = e1 - e0;
  const int64_t stride = (int64_t)gridDim.x * 4 * EPW;
  for (int64_t base = ((int64_t)blockIdx.x * 4 + wave) * EPW; base < ne; base += stride) {
    const int64_t el = base + slot;
    const bool act = in_use && el < ne;
    if (act) {
      const int64_t e = e0 + el;
      if (r < 8) {
        const int32_t v = P.elem_verts[e * 8 + r];
        X[r * 3] = P.verts[(size_t)v * 3];
        X[r * 3 + 1] = P.verts[(size_t)v * 3 + 1];
        X[r * 3 + 2] = P.verts[(size_t)v * 3 + 2];
      }
      if (r < nu) {
        const int32_t d = P.edofs_u[e * nu + r];
        U[UT[r]] = P.u[d];
        Pv[UT[r]] = P.p[d];
      }
      if (r < nm) M[r] = P.m[P.edofs_m[e * nm + r]];
    }
    wave_sync();
    if (act) {
      const int q = r;
      PointState S;
      sf_point_state<NQ>(P, q, X, U, M, Pv, true, TX, TM, TW, T1U, S);
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
    if (act && r < nd) {
      double acc = 0;
      for (int q = 0; q < NQ3; ++q)
        for (int k = 0; k < 4; ++k) acc = fma(Ps[(q * 4 + k) * nd + r], W[q * 4 + k], acc);
      out[(size_t)el * nd + r] = acc;
    }
    // the region is rewritten by the next element
    wave_sync();
  }
}

}  // namespace hm

using namespace hm;

// The sum-factorized contraction when the problem carries tensor structure with 3 points per
// direction and 1-D orders 1 or 2; returns HM_ERR_UNSUPPORTED otherwise (the caller then takes
// another route).
int hm_element_matrices_sumfact(const hm_problem* P, int i, int j, int64_t e0, int64_t e1, int precision,
                                void* out, void* stream) {
  if (int rc = check_tensor(P)) return rc;
  const int nt = i == HM_PARAMETER ? P->n1_m : P->n1_u;
  const int nc = j == HM_PARAMETER ? P->n1_m : P->n1_u;
  if (nt < 2 || nt > 3 || nc < 2 || nc > 3) HM_FAIL(HM_ERR_UNSUPPORTED, "1-D orders outside 1..2");
  SumfactPlan SP = {};
  const bool nl = P->density == HM_DENSITY_NONLINEAR;
  SP.npair = block_pairs(i, j, nl, SP.pr, SP.pc, 0, &SP.kmask);
  if (SP.npair == 0) {
    const size_t n = (size_t)(e1 - e0) * (nt * nt * nt) * (nc * nc * nc);
    return hm_fill_zero(out, n, precision == HM_FP64 ? 8 : 4, (hipStream_t)stream);
  }
  plan_masks(SP);
  SP.need_fields = !(i * 3 + j == 6 && !nl);  // the linear Jacobian reads m and the geometry only
  if (SP.need_fields && (P->n1_u < 2 || P->n1_u > 3))
    HM_FAIL(HM_ERR_UNSUPPORTED, "state fields of 1-D order outside 1..2 for the sum-factorized contraction");
  hipStream_t st = (hipStream_t)stream;
  switch (precision) {
    case HM_FP64: return dispatch_orders<double, HM_FP64>(*P, i, j, e0, e1, nt, nc, SP, out, st);
    case HM_FP32: return dispatch_orders<float, HM_FP32>(*P, i, j, e0, e1, nt, nc, SP, out, st);
    case HM_BF16: return dispatch_orders<float, HM_BF16>(*P, i, j, e0, e1, nt, nc, SP, out, st);
    case HM_FP16: return dispatch_orders<float, HM_FP16>(*P, i, j, e0, e1, nt, nc, SP, out, st);
    default: HM_FAIL(HM_ERR_UNSUPPORTED, "precision %d", precision);
  }
}

// Two blocks of one shape from one pass (FP64 or FP32); HM_ERR_UNSUPPORTED where the pair
// does not qualify (different shapes, a block without pairs, or no tensor structure): the
// caller then assembles the two blocks one by one.
int hm_element_matrices_sumfact_pair(const hm_problem* P, int i1, int j1, int i2, int j2, int64_t e0, int64_t e1,
                                     int precision, void* out1, void* out2, void* stream) {
  if (int rc = check_tensor(P)) return rc;
  const int nt = i1 == HM_PARAMETER ? P->n1_m : P->n1_u;
  const int nc = j1 == HM_PARAMETER ? P->n1_m : P->n1_u;
  if (nt != (i2 == HM_PARAMETER ? P->n1_m : P->n1_u) || nc != (j2 == HM_PARAMETER ? P->n1_m : P->n1_u))
    HM_FAIL(HM_ERR_UNSUPPORTED, "blocks (%d, %d) and (%d, %d) differ in shape", i1, j1, i2, j2);
  if (nt < 2 || nt > 3 || nc < 2 || nc > 3) HM_FAIL(HM_ERR_UNSUPPORTED, "1-D orders outside 1..2");
  if (precision != HM_FP64 && precision != HM_FP32)
    HM_FAIL(HM_ERR_UNSUPPORTED, "two blocks a pass in FP64 or FP32 only");
  SumfactPlan SP = {};
  const bool nl = P->density == HM_DENSITY_NONLINEAR;
  int pr2[SF_MAXPAIR], pc2[SF_MAXPAIR];
  const int n1 = block_pairs(i1, j1, nl, SP.pr, SP.pc, 0, &SP.kmask);
  const int n2 = block_pairs(i2, j2, nl, pr2, pc2, 0, &SP.kmask2);
  if (n1 == 0 || n2 == 0 || n1 + n2 > SF_MAXPAIR)
    HM_FAIL(HM_ERR_UNSUPPORTED, "blocks (%d, %d) and (%d, %d) do not share a pass", i1, j1, i2, j2);
  for (int p = 0; p < n2; ++p) {
    SP.pr[n1 + p] = pr2[p];
    SP.pc[n1 + p] = pc2[p];
  }
  SP.npair = n1 + n2;
  SP.bmask[0] = (int)((1u << n1) - 1u);
  SP.bmask[1] = (int)(((1u << n2) - 1u) << n1);
  SP.s2 = i2;
  SP.j2 = j2;
  plan_masks(SP);
  SP.need_fields = !(i1 * 3 + j1 == 6 && !nl) || !(i2 * 3 + j2 == 6 && !nl);
  if (SP.need_fields && (P->n1_u < 2 || P->n1_u > 3))
    HM_FAIL(HM_ERR_UNSUPPORTED, "state fields of 1-D order outside 1..2 for the sum-factorized contraction");
  hipStream_t st = (hipStream_t)stream;
  if (precision == HM_FP64)
    return dispatch_orders<double, HM_FP64, 2>(*P, i1, j1, e0, e1, nt, nc, SP, out1, st, out2);
  return dispatch_orders<float, HM_FP32, 2>(*P, i1, j1, e0, e1, nt, nc, SP, out1, st, out2);
}

// Element vectors on tensor problems (k_sumfact_vectors); HM_ERR_UNSUPPORTED otherwise.
int hm_element_vectors_sumfact(const hm_problem* P, int i, int64_t e0, int64_t e1, double* out, void* stream) {
  if (int rc = check_tensor(P)) return rc;
  if (P->n1_u < 2 || P->n1_u > 3) HM_FAIL(HM_ERR_UNSUPPORTED, "state fields of 1-D order outside 1..2");
  const int nd = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int nq1 = P->nq1, nq = P->nq, epb = 4 * (64 / nq);
  const size_t bytes = ((size_t)nq * 24 + (size_t)nq * P->nd_m + nq + 6 * nq1 + (size_t)nq * 4 * nd +
                        (size_t)epb * (nq * 4 + 24 + 2 * P->nd_u + P->nd_m)) * sizeof(double);
  const void* kern = nq1 == 4 ? (const void*)k_sumfact_vectors<4> : (const void*)k_sumfact_vectors<3>;
  int limit = 64 * 1024, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) limit = v;
  }
  if (bytes > (size_t)limit) HM_FAIL(HM_ERR_UNSUPPORTED, "the element vectors need %zu bytes of LDS", bytes);
  if (bytes > 48 * 1024) {
    const hipError_t err = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  const int64_t groups = (e1 - e0 + epb - 1) / epb;
  const unsigned grid = (unsigned)(groups < 65536 ? (groups < 1 ? 1 : groups) : 65536);
  if (nq1 == 4)
    hipLaunchKernelGGL(k_sumfact_vectors<4>, dim3(grid), dim3(256), bytes, (hipStream_t)stream, *P, i, e0, e1, out);
  else
    hipLaunchKernelGGL(k_sumfact_vectors<3>, dim3(grid), dim3(256), bytes, (hipStream_t)stream, *P, i, e0, e1, out);
  return hip_check("k_sumfact_vectors");
}
