// Fused assembly of a block on a box mesh: the element kernel and the scatter into CSR values
// in one kernel, so the element matrices never go to memory.
//
// Owner computes.  The n_x x n_y x n_z hexahedra (element e = ex + n_x (ey + n_y ez)) are cut into
// columns of bx x by elements (tiles), each swept from bottom to top by one block.  A tile owns the
// CSR rows of the test-lattice points (p_t the test order) with x in [p_t x0, p_t x1) (the last tile
// also x = p_t n_x), likewise y, every z.  Each layer of the sweep takes the tile's elements and a
// one-element halo on its lower x and y sides (the elements whose upper faces hold owned rows):
// (bx + 1)(by + 1) elements a layer.  A block of threads
//   1. puts the layer's element data (vertex coordinates, dof values) in LDS (loaded a layer ahead),
//   2. acts as the (element, point) pairs of the layer: geometry, fields and the block's pointwise
//      pairs into LDS (the same code as the element-matrix kernel, hm_sumfact.hip),
//   3. acts as the contraction tasks (element, a1, b1, a2) of the owned rows only: stages 1-3 of the
//      sum factorization in registers (b2, a3, b3 in registers), the same operations in the same
//      order as the element-matrix kernel, so every entry has its bits,
//   4. adds the entries into LDS row buffers of the lattice plane the rows lie in, in sub-steps
//      that order the adds of every CSR value by ascending element: a3 = 0, 1, 2 (the plane of
//      z = p_t ez holds the previous layer's a3 = p_t adds already), and within one a3 the rows of
//      local y >= 1 before y = 0 and of local x >= 1 before x = 0.  Two elements of one sub-step
//      never share a row (rows of one parity in x come from one local x), so the adds of a sub-step
//      are plain read-add-writes without atomics, and the sub-steps are ordered by block barriers.
//      Every CSR value is so the sum of its entries in ascending element order, from 0: in FP64
//      the order and the bits of np.bincount (hippymfem/fem/pattern.py:788) and of the two-pass
//      scatter of the same element matrices; the FP32 form sums its FP32 entries in FP32 in the
//      same order (the two-pass scatter sums them in FP64),
//   5. writes a plane out once no later layer adds to it (after a3 = 0 and a3 = 1; the plane of
//      a3 = p_t is carried into the next layer) as whole CSR rows, and clears its buffer.
// A row buffer holds the row's columns as a dense stencil (the trial-lattice box the row couples
// to), so an entry's place follows from its local indices; a per-entry map of the plan (sigma, in
// tile order, one byte an entry) gives the stencil place of every CSR position, with the dropped
// columns of the elimination flagged.  Tiles may be cut into chunks of layers, each recomputing
// the layer below it (its a3 = p_t rows only) as a halo, to balance the blocks over the device.
// The element matrices, the pointwise pairs and the stencil rows never leave the compute unit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "hm_common.h"
#include "hm_device.h"
#include "hm_point.h"
#include "hm_sumfact.h"

namespace hm {

constexpr int FU_MAXL = 17;     // lattice points of a tile a direction (p_t bx + 1)
constexpr int FU_MAXM = 64;     // element slots of a layer
constexpr int FU_MAXT = 512;    // contraction tasks of a layer
constexpr int FU_ITEMS = 2;     // element-data loads a thread a layer
constexpr int FU_SIGV = 2;      // 16-byte stencil-map loads a thread a layer
constexpr int FU_SLOTS = 2;     // planes of staged tables
constexpr int FU_MAXZ = 128;    // layers of a chunk
constexpr int FU_MAXS2D = 2048;  // stencil places of a tile's lattice plane a unit of height
constexpr int FU_EB = 8;         // plane entries a thread writes out in one batch
constexpr int FU_MAXPL = 16;     // pairs of a third-direction type (a block has at most 16)

// The box, the tile grid and the plan arrays (kernel argument).
struct FusedGeom {
  int nx, ny, nz;
  int bx, by, ntx, nty;
  int lz_chunk, nchunk;
  const int2* hdr;              // every row's {CSR offset, row mode}, in tile order
  const int64_t* hdr_off;       // (nplane, tiles): first row of each (plane, tile)
  const uint8_t* sigma;         // stencil place of every CSR entry, tile order; bit 7: dropped
  const int64_t* sig_off;       // (nplane, tiles): 16-byte aligned start of each (plane, tile)
  double diag;
  // the LDS slot of every pointwise pair (pairs of equal values share one: the symmetric
  // reference-gradient pairs (k, l), (l, k), and value/gradient pairs of a block of one slot
  // with itself), the slots, and the pairs that write their slot
  int ws[SF_MAXPAIR];
  int nslot;
  uint32_t wfirst;
  // the pairs of every third-direction type tt (test factor d/dxi_3 or not, trial likewise) in
  // ascending order, for an element of the general form ([0]) and an axis-aligned one ([1]):
  // d1 | d2 << 2 | slot << 4, with d1, d2 the derivative flags of the first two directions
  int pl[2][4][SF_MAXPAIR];
  int pn[2][4];
};

// Global memory (address space 1): loads through pointers read from LDS are global loads, not
// flat ones (a flat load also counts as an LDS operation, so every later LDS wait would wait
// for it).
#define GLOBAL __attribute__((address_space(1)))

// The type of the plane buffers' sums: FP64 for FP64 values (the bits of np.bincount); the
// FP32 form sums in FP32, in the same ascending element order, so that its buffers take half
// the LDS and two blocks share a compute unit.
template <typename CT> struct FuSum { using T = double; };
template <> struct FuSum<float> { using T = float; };

// Sizes of a tile's LDS, for the launch and the host's choice of tile.
struct FusedSizes {
  int m_max, t_max, rows_max, sx_max, sy_max, dw, pw;
  size_t table_words, wa_words, elem_words, plane0_words, plane1_words, sig_bytes, bytes;
};

// The stencil width a direction of a lattice point l of order pt (n elements), trial order pc.
__host__ __device__ inline int fu_width(int l, int pt, int n, int pc) {
  const bool shared = (l % pt == 0) && l > 0 && l < pt * n;
  return pc * (shared ? 2 : 1) + 1;
}

template <typename CT, int NT, int NC, int NQ>
__host__ inline FusedSizes fu_sizes(int bx, int by, int nslot, int nd_u, int nd_m, bool fields, bool aligned,
                                    int nz_max) {
  constexpr int pt = NT - 1, pc = NC - 1, NQ3 = NQ * NQ * NQ;
  FusedSizes z{};
  z.m_max = (bx + 1) * (by + 1);
  z.t_max = NC * (bx * by * NT * NT + (bx + by) * NT + 1);
  z.rows_max = (pt * bx + 1) * (pt * by + 1);
  z.sx_max = bx * (2 * pc + 1) + (pt - 1) * bx * (pc + 1) + (pc + 1);
  z.sy_max = by * (2 * pc + 1) + (pt - 1) * by * (pc + 1) + (pc + 1);
  // an element slot's data (words of FP64): 2 mod 4 words, so that the slots' regions start on
  // 16 different LDS banks
  z.dw = ((24 + nd_m + (fields ? 2 * nd_u : 0) + 1) & ~1) | 2;
  // an element slot's pairs in the contraction's type (words of FP64, even)
  z.pw = (int)((((size_t)nslot * NQ3 * sizeof(CT) + 15) & ~(size_t)15) / sizeof(double));
  // the tables; the trilinear map's gradients only where an element may not be axis-aligned
  z.table_words = (size_t)sf_table_words(nd_u, nd_m, fields, NQ) - (aligned ? 24 * NQ * NQ * NQ : 0) +
                  (((4 * NT * NC + (sizeof(CT) < 8 ? 2 * NT + 2 * NC : 0)) * NQ * sizeof(CT) + 15) & ~(size_t)15) /
                      sizeof(double);
  const size_t s2d = (size_t)z.sx_max * z.sy_max;
  z.sig_bytes = ((s2d * (2 * pc + 1) + 15) & ~(size_t)15);
  // the pairs of every element slot, and in the same words, once the contraction is over, the
  // staged stencil maps and row headers of FU_SLOTS planes; then every slot's data
  const size_t stage_words = ((FU_SLOTS * z.sig_bytes + FU_SLOTS * (size_t)z.rows_max * sizeof(int2) + 15) & ~(size_t)15) /
                             sizeof(double);
  z.wa_words = std::max((size_t)z.m_max * z.pw, stage_words);
  z.elem_words = z.wa_words + (size_t)z.m_max * z.dw;
  // plane buffers in words of their sum type, 16 bytes whole, so that the staged maps after
  // them start 16-byte aligned
  using BT = typename FuSum<CT>::T;
  constexpr size_t A16 = 16 / sizeof(BT);
  z.plane0_words = (s2d * (2 * pc + 1) + A16 - 1) & ~(A16 - 1);
  z.plane1_words = (s2d * (pt == 2 ? pc + 1 : 2 * pc + 1) + A16 - 1) & ~(A16 - 1);
  z.bytes = (z.table_words + z.elem_words) * sizeof(double) + (z.plane0_words + z.plane1_words) * sizeof(BT) +
            (size_t)(pt * nz_max + 1) * 8;
  return z;
}

// The sub-step of the adds of a task's rows (local x a1, y a2 within its element): rows of local
// y >= 1 before those of y = 0, and within each, local x >= 1 before x = 0.
__host__ __device__ inline int fu_group(int a1, int a2) { return (a2 >= 1 ? 0 : 2) + (a1 >= 1 ? 0 : 1); }

// BS threads a block, NB blocks a compute unit (the register budget: NB BS / 256 waves a SIMD).
template <typename CT, int NT, int NC, int QP, int NQ, typename OT, int BS, int NB, bool AL>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, BS), amdgpu_waves_per_eu((NB * BS + 255) / 256)))
k_fused(const hm_problem P, int s, int j, SumfactPlan SP, FusedGeom G, FusedSizes Z, const double* __restrict__ t1t,
        const double* __restrict__ t1c, OT* __restrict__ data) {
  constexpr int pt = NT - 1, pc = NC - 1;
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ;
  using BT = typename FuSum<CT>::T;
  extern __shared__ double lds[];
  __shared__ int32_t UT[32];
  __shared__ int32_t TASK[FU_MAXT];
  __shared__ int32_t B2D[FU_MAXL * FU_MAXL];
  __shared__ int32_t WXs[FU_MAXL], WYs[FU_MAXL];
  __shared__ int32_t EFLAG[FU_MAXM];
  __shared__ int32_t TCOUNT[4];            // tasks of each sub-step group
  __shared__ uint8_t ROWOF[FU_MAXS2D];  // the row of every stencil place of a unit-height plane
  // the element-data arrays by item kind (vertex, parameter dof, state dof): values (verts, m,
  // u), p, and the element tables (elem_verts, edofs_m, edofs_u).  Read from LDS where a lane
  // picks its item's array: held in scalar registers through the layer loop they spill, and a
  // lane's choice between them costs a reload of each from the spill lanes.
  __shared__ uint64_t PTAB[7];
  __shared__ int32_t EOFF[FU_MAXM];  // every element slot's element within its layer, ex + n_x ey
  // (general form) the pair lists of FusedGeom, read per lane: an element's list depends on its
  // alignment, and a per-lane read of the kernel arguments is a global load and a wait for every
  // load in flight
  __shared__ int32_t PLS[2 * 4 * FU_MAXPL];
  __shared__ int32_t PNS[2 * 4];

  const int nu = P.nd_u, nm = P.nd_m;
  const bool fields = SP.need_fields != 0;
  // the thread's index, made opaque again at the start of every phase: per-thread quantities
  // derived from it are recomputed there rather than kept live across the contraction, whose
  // registers they would otherwise spill to scratch memory (a global-latency reload inside
  // every barrier-bound step)
  int tid = threadIdx.x, lane = tid & 63;
  auto rederive = [&]() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    tid = t;
    lane = t & 63;
  };
  // ------------------------------------------------------------------ the tile
  const int ntile = G.ntx * G.nty;
  const int tile = blockIdx.x % ntile, chunk = blockIdx.x / ntile;
  const int X = tile % G.ntx, Y = tile / G.ntx;
  const int x0 = X * G.bx, x1 = min(x0 + G.bx, G.nx);
  const int y0 = Y * G.by, y1 = min(y0 + G.by, G.ny);
  const int z0 = chunk * G.lz_chunk, z1 = min(z0 + G.lz_chunk, G.nz);
  const int xs = x0 > 0 ? x0 - 1 : 0, ys = y0 > 0 ? y0 - 1 : 0, zs = z0 > 0 ? z0 - 1 : 0;
  const int cx = x1 - xs, cy = y1 - ys, M = cx * cy;
  const int lx_lo = pt * x0, lx_hi = pt * x1 + (x1 == G.nx ? 1 : 0);
  const int ly_lo = pt * y0, ly_hi = pt * y1 + (y1 == G.ny ? 1 : 0);
  const int lz_lo = pt * z0, lz_hi = pt * z1 + (z1 == G.nz ? 1 : 0);
  const int cxl = lx_hi - lx_lo, cyl = ly_hi - ly_lo, nrows = cxl * cyl;
  const int64_t plane_rows = (int64_t)(pt * G.nx + 1) * (pt * G.ny + 1);

  // ------------------------------------------------------------------ LDS
  // the tables: the parameter basis, the weights, the state's 1-D basis (fields), (general form)
  // the trilinear map's gradients, then in the contraction's type the 1-D table products PT and
  // (FP32) the test and trial 1-D tables, read as they are used: no conversion in the contraction
  double* TM = lds;
  double* TW = TM + NQ3 * nm;
  double* T1U = TW + NQ3;
  double* TX = lds + sf_table_words(nu, nm, fields, NQ) - NQ3 * 24;
  CT* PT = (CT*)(AL ? TX : TX + NQ3 * 24);
  CT* T1T = PT + 4 * NT * NC * NQ;
  CT* T1C = T1T + 2 * NT * NQ;
  // The element slots' pairs share their words with the staged tables of the planes a layer
  // writes out: the pairs are read by the contraction only, and every wavefront's contraction is
  // over at the first sub-step barrier of the adds; the tables are staged at the last sub-step of
  // phase A and read by the write-outs, all before the next layer's top barrier, after which its
  // point phase writes the pairs again.
  double* WA = lds + Z.table_words;                  // pairs of the element slots
  double* DA = WA + Z.wa_words;                      // data of the element slots
  BT* BUF0 = (BT*)(WA + Z.elem_words);               // plane buffers
  BT* BUF1 = BUF0 + Z.plane0_words;
  uint8_t* SIG = (uint8_t*)WA;                       // staged stencil maps, FU_SLOTS planes
  int2* HDR = (int2*)(SIG + FU_SLOTS * Z.sig_bytes); // staged row headers, FU_SLOTS planes
  int64_t* SOFF = (int64_t*)(BUF1 + Z.plane1_words); // the chunk's planes' map starts
  auto Wreg = [&](int m) { return (CT*)(WA + (size_t)m * Z.pw); };
  auto Dreg = [&](int m) { return DA + (size_t)m * Z.dw; };

  {
    for (int i = tid; i < 4 * NT * NC * NQ; i += BS) {
      const int q = i % NQ, b = (i / NQ) % NC, a = (i / (NQ * NC)) % NT, dd = i / (NQ * NC * NT);
      PT[i] = (CT)(t1t[((dd >> 1) * NT + a) * NQ + q] * t1c[((dd & 1) * NC + b) * NQ + q]);
    }
    if constexpr (sizeof(CT) < 8) {
      for (int i = tid; i < 2 * NT * NQ; i += BS) T1T[i] = (CT)t1t[i];
      for (int i = tid; i < 2 * NC * NQ; i += BS) T1C[i] = (CT)t1c[i];
    }
    if constexpr (!AL)
      for (int i = tid; i < NQ3 * 24; i += BS) TX[(i % 24) * NQ3 + i / 24] = P.tab_x[i];
    for (int i = tid; i < NQ3 * nm; i += BS) TM[(i % nm) * NQ3 + i / nm] = P.tab_m[i];
    for (int i = tid; i < NQ3; i += BS) TW[i] = P.wq[i];
    if (tid == 0) {
      PTAB[0] = (uint64_t)P.verts;
      PTAB[1] = (uint64_t)P.m;
      PTAB[2] = (uint64_t)P.u;
      PTAB[3] = (uint64_t)P.p;
      PTAB[4] = (uint64_t)P.elem_verts;
      PTAB[5] = (uint64_t)P.edofs_m;
      PTAB[6] = (uint64_t)P.edofs_u;
    }
    for (int m = tid; m < M; m += BS) EOFF[m] = (xs + m % cx) + G.nx * (ys + m / cx);
    if constexpr (!AL) {
      for (int i = tid; i < 2 * 4 * FU_MAXPL; i += BS)
        PLS[i] = G.pl[i / (4 * FU_MAXPL)][(i / FU_MAXPL) % 4][i % FU_MAXPL];
      for (int i = tid; i < 2 * 4; i += BS) PNS[i] = G.pn[i / 4][i % 4];
    }
    if (fields) {
      for (int i = tid; i < 2 * P.n1_u * NQ; i += BS) T1U[i] = P.tab1_u[i];
      for (int i = tid; i < nu; i += BS) UT[P.lex_u[i]] = i;
    }
    for (size_t i = tid; i < Z.plane0_words + Z.plane1_words; i += BS) BUF0[i] = 0;
    for (int i = tid; i < lz_hi - lz_lo; i += BS) SOFF[i] = G.sig_off[(int64_t)(lz_lo + i) * ntile + tile];
    // stencil widths of the tile's lattice points, and every row's place in a plane buffer
    for (int i = tid; i < cxl; i += BS) WXs[i] = fu_width(lx_lo + i, pt, G.nx, pc);
    for (int i = tid; i < cyl; i += BS) WYs[i] = fu_width(ly_lo + i, pt, G.ny, pc);
    // contraction tasks: (a1, a2) of the owned rows of every element slot, every b1, counted by
    // the sub-step of their adds (thread g counts group g)
    if (tid < 4) {
      int n = 0;
      for (int m = 0; m < M; ++m) {
        const int ex = xs + m % cx, ey = ys + m / cx;
        const int a1lo = max(lx_lo - pt * ex, 0), a1hi = min(lx_hi - pt * ex, NT);
        const int a2lo = max(ly_lo - pt * ey, 0), a2hi = min(ly_hi - pt * ey, NT);
        for (int a2 = a2lo; a2 < a2hi; ++a2)
          for (int a1 = a1lo; a1 < a1hi; ++a1) n += fu_group(a1, a2) == tid ? NC : 0;
      }
      TCOUNT[tid] = n;
    }
  }
  const int64_t hdr0 = G.hdr_off[tile];  // the tile's first row in every plane (plane 0's)
  __syncthreads();
  if (tid == 0) {
    int sx = 0, px[FU_MAXL];
    for (int i = 0; i < cxl; ++i) {
      px[i] = sx;
      sx += WXs[i];
    }
    int py = 0;
    for (int iy = 0; iy < cyl; ++iy) {
      for (int ix = 0; ix < cxl; ++ix) B2D[iy * cxl + ix] = sx * py + WYs[iy] * px[ix];
      py += WYs[iy];
    }
  }
  __syncthreads();
  // the tasks in the order of the sub-steps (thread g writes group g), in each group by element
  // slot, a2, a1, b1: the lanes of a wavefront mostly share one sub-step, and a wavefront passes
  // over the others as a whole
  if (tid < 4) {
    int k = 0;
    for (int g = 0; g < tid; ++g) k += TCOUNT[g];
    for (int m = 0; m < M; ++m) {
      const int ex = xs + m % cx, ey = ys + m / cx;
      const int a1lo = max(lx_lo - pt * ex, 0), a1hi = min(lx_hi - pt * ex, NT);
      const int a2lo = max(ly_lo - pt * ey, 0), a2hi = min(ly_hi - pt * ey, NT);
      for (int a2 = a2lo; a2 < a2hi; ++a2)
        for (int a1 = a1lo; a1 < a1hi; ++a1)
          if (fu_group(a1, a2) == tid)
            for (int b1 = 0; b1 < NC; ++b1) TASK[k++] = m | (a1 << 8) | (b1 << 12) | (a2 << 16);
    }
  }
  for (int rho = tid; rho < nrows; rho += BS) {
    const int b = B2D[rho], w = WXs[rho % cxl] * WYs[rho / cxl];
    for (int q = 0; q < w; ++q) ROWOF[b + q] = (uint8_t)rho;
  }
  __syncthreads();
  const int ntask = TCOUNT[0] + TCOUNT[1] + TCOUNT[2] + TCOUNT[3];
  const int s2d_row = B2D[nrows - 1] + WXs[cxl - 1] * WYs[cyl - 1];  // stencil entries / wz

  // ------------------------------------------------------------------ element data
  // items of a layer: per element slot 8 vertices, nd_m parameter dofs, and (fields) nd_u
  // state/adjoint dofs; thread tid takes items tid, tid + BS, ...  A layer's data is loaded
  // into registers during the layer before (before its contraction, so that the loads are in
  // flight during the contraction and are not queued behind its stores) and put in LDS after the
  // contraction (aligned form) or after the first adds (general form), before the layer's planes
  // are written out (a wait for the loads there waits for no store).  An item's array is chosen per lane from the pointer table in LDS
  // (PTAB), with no branch: loads of divergent branches into one register would wait for each
  // other.
  const int NI = 8 + nm + (fields ? nu : 0);
  int32_t nid[FU_ITEMS];
  double nv[FU_ITEMS][3];
  const int64_t lstride = (int64_t)G.nx * G.ny;  // elements a layer
  // an item's element slot it / NI by a rounded FP32 product with 1 / NI, fewer operations than an
  // integer division: exact for it < 2^13 and NI <= 128 (the product is within 2^-9 of
  // (it + 1/2) / NI, even with 1 / NI a unit in the last place off, and (it + 1/2) / NI is at
  // least 1 / (2 NI) >= 2^-8 from an integer)
  const float rNI = 1.0f / (float)NI;
  auto item_slot = [&](int it) { return (int)__fmul_rn((float)it + 0.5f, rNI); };
  auto item_elem = [&](int m, int ez) -> int64_t { return (int64_t)EOFF[m] + lstride * ez; };
  auto load_ids = [&](int ez) {
#pragma unroll
    for (int k = 0; k < FU_ITEMS; ++k) {
      const int it = tid + k * BS;
      if (it >= M * NI || ez >= z1) continue;
      const int m = item_slot(it);
      const int64_t e = item_elem(m, ez);
      const int w = it - m * NI;
      const int kind = w < 8 ? 0 : w < 8 + nm ? 1 : 2;
      const int nd = kind == 0 ? 8 : kind == 1 ? nm : nu;
      const int sub = w - (kind == 0 ? 0 : kind == 1 ? 8 : 8 + nm);
      nid[k] = ((const GLOBAL int32_t*)PTAB[4 + kind])[e * nd + sub];
    }
  };
  auto load_values = [&](int ez) {
#pragma unroll
    for (int k = 0; k < FU_ITEMS; ++k) {
      const int it = tid + k * BS;
      if (it >= M * NI || ez >= z1) continue;
      const int w = it - item_slot(it) * NI;
      const int64_t d = nid[k];
      // three loads an item, whatever its kind: a vertex's coordinates, a parameter dof's value
      // (three times), a state dof's u and p
      const int kind = w < 8 ? 0 : w < 8 + nm ? 1 : 2;
      const GLOBAL double* s0 = (const GLOBAL double*)PTAB[kind] + d * (kind == 0 ? 3 : 1);
      const GLOBAL double* s1 = kind == 0 ? s0 + 1 : kind == 1 ? s0 : (const GLOBAL double*)PTAB[3] + d;
      const GLOBAL double* s2 = kind == 0 ? s0 + 2 : s0;
      nv[k][0] = *s0;
      nv[k][1] = *s1;
      nv[k][2] = *s2;
    }
  };
  // the element ids loaded a layer ahead waited for here, before the layer's stores: a wait for
  // them after those stores (loads and stores complete in order) would wait for every store
  auto settle_ids = [&]() {
#pragma unroll
    for (int k = 0; k < FU_ITEMS; ++k) asm volatile("" ::"v"(nid[k]));
  };
  auto store_values = [&](int ez) {
#pragma unroll
    for (int k = 0; k < FU_ITEMS; ++k) {
      const int it = tid + k * BS;
      if (it >= M * NI || ez >= z1) continue;
      const int m = item_slot(it), w = it - m * NI;
      double* D = Dreg(m);
      if (w < 8) {
        D[w * 3] = nv[k][0];
        D[w * 3 + 1] = nv[k][1];
        D[w * 3 + 2] = nv[k][2];
      } else if (w < 8 + nm) {
        D[24 + (w - 8)] = nv[k][0];
      } else {
        const int a = UT[w - 8 - nm];
        D[24 + nm + a] = nv[k][0];
        D[24 + nm + nu + a] = nv[k][1];
      }
    }
  };

  // ------------------------------------------------------------------ planes written out
  // the planes a layer completes: z = p_t ez + a3 for a3 < p_t that the chunk owns (staged
  // slots 0, 1), and at the top of the box the plane z = p_t n_z after the last layer (staged
  // in slot 0 once the first plane of that layer is out)
  auto retire_plane = [&](int ez, int k) -> int {
    int n = 0;
    for (int a3 = 0; a3 < pt; ++a3) {
      const int lz = pt * ez + a3;
      if (lz >= lz_lo && lz < lz_hi) {
        if (n == k) return lz;
        ++n;
      }
    }
    return -1;
  };
  auto wz_of = [&](int lz) { return fu_width(lz, pt, G.nz, pc); };
  // row headers and stencil maps of a layer's planes, in registers between loads and LDS stores
  int2 hv[FU_SLOTS];
  uint4 sv[FU_SIGV];
  auto load_planes = [&](int ez) {
    int lzs[FU_SLOTS], chunks[FU_SLOTS];
#pragma unroll
    for (int k = 0; k < FU_SLOTS; ++k) {
      lzs[k] = ez < z1 ? retire_plane(ez, k) : -1;
      chunks[k] = lzs[k] >= 0 ? (s2d_row * wz_of(lzs[k]) + 15) >> 4 : 0;
      if (lzs[k] >= 0 && tid < nrows) hv[k] = G.hdr[hdr0 + (int64_t)lzs[k] * plane_rows + tid];
    }
#pragma unroll
    for (int v = 0; v < FU_SIGV; ++v) {
      const int c = tid + v * BS;
      sv[v] = make_uint4(0, 0, 0, 0);
      if (c < chunks[0])
        sv[v] = ((const uint4*)(G.sigma + SOFF[lzs[0] - lz_lo]))[c];
      else if (c - chunks[0] < chunks[1])
        sv[v] = ((const uint4*)(G.sigma + SOFF[lzs[1] - lz_lo]))[c - chunks[0]];
    }
  };
  auto store_planes = [&](int ez) {
    int lzs[FU_SLOTS], chunks[FU_SLOTS];
#pragma unroll
    for (int k = 0; k < FU_SLOTS; ++k) {
      lzs[k] = ez < z1 ? retire_plane(ez, k) : -1;
      chunks[k] = lzs[k] >= 0 ? (s2d_row * wz_of(lzs[k]) + 15) >> 4 : 0;
      if (lzs[k] >= 0 && tid < nrows) HDR[k * Z.rows_max + tid] = hv[k];
    }
#pragma unroll
    for (int v = 0; v < FU_SIGV; ++v) {
      const int c = tid + v * BS;
      if (c < chunks[0])
        ((uint4*)SIG)[c] = sv[v];
      else if (c - chunks[0] < chunks[1])
        ((uint4*)(SIG + Z.sig_bytes))[c - chunks[0]] = sv[v];
    }
  };
  // the top plane's tables into slot 0, directly (once a chunk, at the top of the box)
  auto stage_top = [&](int lz) {
    if (tid < nrows) HDR[tid] = G.hdr[hdr0 + (int64_t)lz * plane_rows + tid];
    const int chunks = (s2d_row * wz_of(lz) + 15) >> 4;
    const uint4* src = (const uint4*)(G.sigma + SOFF[lz - lz_lo]);
    for (int c = tid; c < chunks; c += BS) ((uint4*)SIG)[c] = src[c];
  };
  // every row of a staged plane as whole CSR rows, its buffer cleared: thread tid takes the
  // plane's places tid, tid + BS, ... (in row order, so the stores of a row are contiguous),
  // every read of the batch issued before its first use.  A place's row: ROWOF[k / wz] (a row
  // of the plane holds w_x w_y wz places from B2D wz on); k / wz by a multiply and a shift
  // (exact for k < 2^16 / wz).  The index products are of 24-bit operands (full-rate multiplies).
  auto write_plane = [&](int lz, int k) {
    rederive();
    const int wz = wz_of(lz);
    const int S = s2d_row * wz;
    const uint32_t mul = (65536u + (uint32_t)wz - 1u) / (uint32_t)wz;
    BT* buf = (lz & 1) ? BUF1 : BUF0;
    const BT diagv = (BT)G.diag;
    const uint8_t* sg = SIG + k * Z.sig_bytes;
    const int2* hdr = HDR + k * Z.rows_max;
    for (int base = 0; base < S; base += FU_EB * BS) {
      int kk[FU_EB], rb[FU_EB], rho[FU_EB];
      unsigned c[FU_EB];
      int2 hd[FU_EB];
      BT v[FU_EB];
      // every read unconditional (a place past the plane reads the last one, and is not
      // written), so that each of the three rounds of reads is issued whole before its first use
#pragma unroll
      for (int e = 0; e < FU_EB; ++e) {
        kk[e] = base + tid + e * BS;
        const int kc = min(kk[e], S - 1);
        rho[e] = ROWOF[__umul24((uint32_t)kc, mul) >> 16];
        c[e] = sg[kc];
      }
#pragma unroll
      for (int e = 0; e < FU_EB; ++e) {
        hd[e] = hdr[rho[e]];
        rb[e] = (int)__umul24((uint32_t)B2D[rho[e]], (uint32_t)wz);
      }
#pragma unroll
      for (int e = 0; e < FU_EB; ++e) v[e] = buf[rb[e] + (c[e] & 0x7F)];
#pragma unroll
      for (int e = 0; e < FU_EB; ++e)
        if (kk[e] < S) {
          buf[rb[e] + (c[e] & 0x7F)] = 0;
          const int t = kk[e] - rb[e];
          // (in the sums' type: the FP32 form converts nothing here)
          BT o = (c[e] & 0x80) ? (BT)0 : v[e];
          if (hd[e].y != 0) o = (hd[e].y == 2 && t == 0) ? diagv : (BT)0;
          data[(int64_t)hd[e].x + t] = (OT)o;
        }
    }
  };

  // ------------------------------------------------------------------ the sweep
  bool bad = false;
  CT acc[NC][NT][NC];
  load_ids(zs);
  load_values(zs);
  store_values(zs);
  load_ids(zs + 1);
  settle_ids();
  for (int ez = zs; ez < z1; ++ez) {
    rederive();
    // 1. this layer's element data, put in LDS by the layer before
    __syncthreads();
    // each element's alignment (every vertex coordinate that of the vertex with the same index
    // along its axis: J diagonal, the pairs of two different derivatives left out of its
    // contraction, as the element kernel does; lane v of the eight lanes of element m checks
    // vertex v, a ballot joins them).  On a mesh of axis-aligned elements (AL) every element is.
    if constexpr (!AL) {
      const int m = tid >> 3, v = tid & 7;
      bool ok = true;
      if (m < M) {
        const double* Xe = Dreg(m);
        ok = (Xe[v * 3] == Xe[(v & 1) * 3]) & (Xe[v * 3 + 1] == Xe[(v & 2) * 3 + 1]) &
             (Xe[v * 3 + 2] == Xe[(v & 4) * 3 + 2]);
      }
      const uint64_t all = __ballot(ok);
      if (m < M && v == 0) EFLAG[m] = ((all >> (lane & 56)) & 0xFFu) == 0xFFu ? 1 : 0;
      __syncthreads();
    }
    // 2. (element, point) pairs: geometry, fields, pointwise pairs (one LDS slot for the pairs
    // of equal values)
    for (int t = tid; t < M * NQ3; t += BS) {
      const int m = t / NQ3, q = t - m * NQ3;
      const double* D = Dreg(m);
      CT* We = Wreg(m);
      // (the FP32 form makes the fields, kappa and the pairs in FP32, the geometry in FP64)
      PointStateT<CT> S;
      sf_point_state<NQ, CT>(P, q, D, D + 24 + nm, D + 24, D + 24 + nm + nu, fields, TX, TM, TW, T1U, S,
                             AL || EFLAG[m] != 0);
      CT Db[4][4];
      point_block(s, j, S, Db);
      int p = 0;
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if ((SP.kmask >> k) & 1) {
          if (G.wfirst & (1u << p)) We[G.ws[p] * NQ3 + q] = (CT)Db[k >> 2][k & 3];
          ++p;
        }
    }
    __syncthreads();
    // 3. the contraction of this thread's task, the next layer's element data in flight
    rederive();
    load_values(ez + 1);
    load_ids(ez + 2);
    bool has = tid < ntask;
    int m = 0, a1 = 0, b1 = 0, a2 = 0;
    if (has) {
      const int tk = TASK[tid];
      m = tk & 0xFF;
      a1 = (tk >> 8) & 0xF;
      b1 = (tk >> 12) & 0xF;
      a2 = (tk >> 16) & 0xF;
    }
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int a3 = 0; a3 < NT; ++a3)
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) acc[i][a3][b3] = 0;
    if (has) {
      // the element's pairs of each third-direction type, from the plan's lists (the aligned
      // element's leave out the pairs of two different reference derivatives)
      const int kind = (AL || EFLAG[m]) ? 1 : 0;
      const CT* Wl = Wreg(m);
      // one third-direction type t (FP64 aligned form: the four unrolled, which lets the
      // compiler read every list length ahead; elsewhere that costs more than it gains)
      auto type_step = [&](int t) {
        const int np = AL ? G.pn[1][t] : PNS[kind * 4 + t];
        if (np == 0) return;
        CT s2[NC][NQ];
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
          for (int q = 0; q < NQ; ++q) s2[i][q] = 0;
        for (int k = 0; k < np; ++k) {
          const int pw = AL ? G.pl[1][t][k] : PLS[(kind * 4 + t) * FU_MAXPL + k];
          const int d1 = pw & 3, d2 = (pw >> 2) & 3;
          const CT* wv = Wl + (pw >> 4) * NQ3;
          const CT* f1 = PT + ((d1 * NT + a1) * NC + b1) * NQ;
          CT gd[NQ];
#pragma unroll
          for (int q = 0; q < NQ; ++q) gd[q] = f1[q];
          CT hd[NC][NQ];
#pragma unroll
          for (int i = 0; i < NC; ++i) {
            const CT* f2 = PT + ((d2 * NT + a2) * NC + i) * NQ;
#pragma unroll
            for (int q = 0; q < NQ; ++q) hd[i][q] = f2[q];
          }
          __builtin_amdgcn_sched_barrier(0);
          CT g[NQ];
#pragma unroll
          for (int q = 0; q < NQ; ++q) g[q] = gd[q];
#pragma unroll
          for (int q3 = 0; q3 < NQ; ++q3) {
            // the slice's NQ^2 pair values read together: one LDS round trip a slice
            CT wq[NQ2];
#pragma unroll
            for (int r = 0; r < NQ2; ++r) wq[r] = wv[NQ2 * q3 + r];
            __builtin_amdgcn_sched_barrier(0);
            CT s1[NQ];
#pragma unroll
            for (int q2 = 0; q2 < NQ; ++q2) {
              const CT* wr = wq + NQ * q2;
              if constexpr (NQ == 3) {
                s1[q2] = fma(g[2], wr[2], fma(g[1], wr[1], g[0] * wr[0]));
              } else {
                CT a = g[0] * wr[0];
#pragma unroll
                for (int q1 = 1; q1 < NQ; ++q1) a = fma(g[q1], wr[q1], a);
                s1[q2] = a;
              }
            }
            if constexpr (NQ == 3) {
#pragma unroll
              for (int i = 0; i < NC; ++i)
                s2[i][q3] = fma(hd[i][2], s1[2], fma(hd[i][1], s1[1], fma(hd[i][0], s1[0], s2[i][q3])));
            } else {
#pragma unroll
              for (int i = 0; i < NC; ++i) {
                CT a = s2[i][q3];
#pragma unroll
                for (int q2 = 0; q2 < NQ; ++q2) a = fma(hd[i][q2], s1[q2], a);
                s2[i][q3] = a;
              }
            }
          }
        }
        // (FP64: the kernel's arguments, read into scalar registers; FP32: the LDS copies)
        const CT* bt = (std::is_same<CT, double>::value ? (const CT*)t1t : T1T) + (t >> 1) * NT * NQ;
        const CT* bc = (std::is_same<CT, double>::value ? (const CT*)t1c : T1C) + (t & 1) * NC * NQ;
#pragma unroll
        for (int b3 = 0; b3 < NC; ++b3) {
          CT u[NC][NQ];
#pragma unroll
          for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int q = 0; q < NQ; ++q) u[i][q] = bc[b3 * NQ + q] * s2[i][q];
#pragma unroll
          for (int a3 = 0; a3 < NT; ++a3) {
            if constexpr (NQ == 3) {
              const CT k0 = bt[a3 * 3], k1 = bt[a3 * 3 + 1], k2 = bt[a3 * 3 + 2];
#pragma unroll
              for (int i = 0; i < NC; ++i)
                acc[i][a3][b3] = fma(k2, u[i][2], fma(k1, u[i][1], fma(k0, u[i][0], acc[i][a3][b3])));
            } else {
              CT k[NQ];
#pragma unroll
              for (int q = 0; q < NQ; ++q) k[q] = bt[a3 * NQ + q];
#pragma unroll
              for (int i = 0; i < NC; ++i) {
                CT a = acc[i][a3][b3];
#pragma unroll
                for (int q = 0; q < NQ; ++q) a = fma(k[q], u[i][q], a);
                acc[i][a3][b3] = a;
              }
            }
          }
        }
      };
      if constexpr (AL && sizeof(CT) == 8) {
#pragma unroll
        for (int t = 0; t < 4; ++t) type_step(t);
      } else {
#pragma unroll 1
        for (int t = 0; t < 4; ++t) type_step(t);
      }
    }
    rederive();
    // the next layer's element data into LDS (the point phase, the only reader of the region, is
    // over, and the loads had the contraction to arrive).  The general form keeps part of the
    // loaded values in scratch memory over the contraction, and stores them after the first adds.
    if constexpr (AL) {
      store_values(ez + 1);
      settle_ids();
    }
    // the tables of the planes this layer completes, in flight during the adds of phase A
    load_planes(ez);
    // the task again, from LDS (nothing of it kept live through the contraction)
    has = tid < ntask;
    {
      const int tk = has ? TASK[tid] : 0;
      m = tk & 0xFF;
      a1 = (tk >> 8) & 0xF;
      b1 = (tk >> 12) & 0xF;
      a2 = (tk >> 16) & 0xF;
    }
    // 4./5. the adds in ascending element order, and the planes written out.  Phase A takes the
    // rows of a3 = 0 and 1 (two planes in two buffers), phase B those of a3 = 2 (p_t = 2; into
    // the buffer of a3 = 0, cleared when that plane was written out).  Within a phase the four
    // sub-steps of local (y, x) order the adds of every place by element; the places of one
    // lane are distinct, so a lane reads all of them before it writes any.
    const int ex = xs + m % cx, ey = ys + m / cx;
    const int ix = pt * ex + a1 - lx_lo, iy = pt * ey + a2 - ly_lo;
    const int grp = fu_group(a1, a2);
    const int ox = pc * ((a1 == 0 && ex > 0) ? 1 : 0) + b1;
    const int oy0 = pc * ((a2 == 0 && ey > 0) ? 1 : 0);
    const int wx = has ? WXs[ix] : 0, wy = has ? WYs[iy] : 0;
    const int b2d = has ? B2D[iy * cxl + ix] : 0;
    int rk = 0;            // staged planes written so far
    bool staged = AL;      // the next layer's element data in LDS
    constexpr int NA = NT < 2 ? NT : 2;  // rows of phase A: a3 < NA
#pragma unroll
    for (int ph = 0; ph < 2; ++ph) {
      const int a3lo = ph == 0 ? 0 : NA, a3hi = ph == 0 ? NA : NT;
      if (a3lo >= a3hi) continue;
      bool any = false;  // the phase holds a plane the chunk owns (uniform over the block)
      for (int a3 = a3lo; a3 < a3hi; ++a3) {
        const int lz = pt * ez + a3;
        any = any || (lz >= lz_lo && lz < lz_hi);
      }
      if (!any) continue;
#pragma unroll 1
      for (int g = 0; g < 4; ++g) {
        if (has && grp == g) {
#pragma unroll
          for (int a3 = a3lo; a3 < a3hi; ++a3) {
            const int lz = pt * ez + a3;
            if (lz < lz_lo || lz >= lz_hi) continue;  // a row of the chunk below (halo)
            const int wz = wz_of(lz);
            BT* row = ((lz & 1) ? BUF1 : BUF0) + (size_t)b2d * wz;
            const int oz0 = pc * ((a3 == 0 && ez > 0) ? 1 : 0);
            BT old[NC][NC];
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
              for (int b3 = 0; b3 < NC; ++b3) old[i][b3] = row[ox + wx * ((oy0 + i) + wy * (oz0 + b3))];
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
              for (int b3 = 0; b3 < NC; ++b3) {
                const CT v = acc[i][a3][b3];
                bad |= !__builtin_isfinite(v);
                row[ox + wx * ((oy0 + i) + wy * (oz0 + b3))] = old[i][b3] + (BT)v;
              }
          }
        }
        if (ph == 0 && g == 3) store_planes(ez);  // ordered before the write-out by the barrier
        __syncthreads();
      }
      // the next layer's element data into LDS (its loads had the adds to arrive; this layer's
      // point phase, the only reader of the region, is over), before this phase's stores
      if (!staged) {
        store_values(ez + 1);
        settle_ids();
        staged = true;
      }
      // the planes of this phase no later layer adds to
      for (int a3 = a3lo; a3 < a3hi; ++a3) {
        const int lz = pt * ez + a3;
        if (lz < lz_lo || lz >= lz_hi) continue;
        if (a3 < pt) {
          write_plane(lz, rk++);
          // two planes of a phase lie in different buffers and slots: one barrier after both
          if (a3 + 1 >= a3hi || lz + 1 >= lz_hi || a3 + 1 >= pt) __syncthreads();
        } else if (ez == G.nz - 1 && z1 == G.nz) {
          // the top of the box: its tables into slot 0, then written out
          rederive();
          stage_top(lz);
          __syncthreads();
          write_plane(lz, 0);
          __syncthreads();
        }
      }
    }
    if (!staged) {
      store_values(ez + 1);
      settle_ids();
    }
  }
  if (bad && P.nonfinite) atomicOr(P.nonfinite, 1);
}

// Whether a tile of sizes z (bx x by elements, test 1-D order nt - 1) runs in blocks of bs threads:
// a task a thread, eight lanes an element slot (the alignment test), a row a thread (the plane
// tables), and the element data and stencil maps of a layer in FU_ITEMS, FU_SIGV loads a thread.
static bool fu_fits(const hm_problem& P, const SumfactPlan& SP, const FusedSizes& z, int bs, int bx, int by, int nt) {
  const int items = (z.m_max * (8 + P.nd_m + (SP.need_fields ? P.nd_u : 0)) + bs - 1) / bs;
  const int sigv = (int)((FU_SLOTS * z.sig_bytes / 16 + bs - 1) / bs);
  return 8 * z.m_max <= bs && z.t_max <= bs && z.t_max <= FU_MAXT && z.m_max <= FU_MAXM && items <= FU_ITEMS &&
         sigv <= FU_SIGV && z.rows_max <= bs && z.rows_max <= 256 && (size_t)z.sx_max * z.sy_max <= FU_MAXS2D &&
         z.sig_bytes <= 8192 && (nt - 1) * bx + 1 <= FU_MAXL && (nt - 1) * by + 1 <= FU_MAXL;
}

template <typename CT, int NT, int NC, int NQ>
static FusedSizes fu_pick(const hm_problem& P, const SumfactPlan& SP, int nslot, bool aligned, int bs, size_t limit,
                          int* bx, int* by) {
  static const int cand[][2] = {{4, 4}, {4, 3}, {3, 3}, {3, 2}, {2, 2}, {2, 1}, {1, 1}};
  FusedSizes z{};
  for (const auto& c : cand) {
    z = fu_sizes<CT, NT, NC, NQ>(c[0], c[1], nslot, P.nd_u, P.nd_m, SP.need_fields != 0, aligned, FU_MAXZ);
    if (z.bytes <= limit && fu_fits(P, SP, z, bs, c[0], c[1], NT)) {
      *bx = c[0];
      *by = c[1];
      return z;
    }
  }
  *bx = *by = 0;
  return z;
}

// the static LDS of k_fused (task table, row places, widths, flags), left out of its dynamic LDS
// (with 256 bytes for the alignment of the arrays)
constexpr size_t FU_STATIC_LDS =
    sizeof(int32_t) * (32 + FU_MAXT + FU_MAXL * FU_MAXL + 2 * FU_MAXL + 2 * FU_MAXM + 4 + 8 * FU_MAXPL + 8) + FU_MAXS2D +
    7 * sizeof(uint64_t) + 256;

// The dynamic LDS a block may take when `blocks` blocks share a compute unit.
static size_t lds_limit(int blocks = 1) {
  int dev = 0, v = 0, w = 0;
  size_t lim = 64 * 1024;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
    lim = (size_t)v;
  if (blocks > 1) {
    if (hipDeviceGetAttribute(&w, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || w <= 0)
      return 0;
    lim = std::min(lim, ((size_t)w / blocks) & ~(size_t)1023);
  }
  return lim > FU_STATIC_LDS ? lim - FU_STATIC_LDS : 0;
}

// The LDS slots of the pointwise pairs: a pair whose values are those of an earlier one takes
// its slot (the reference-gradient pairs (k, l) and (l, k), whose weights w g J^-1 J^-T are
// symmetric to the bit, and every (r, c), (c, r) of a block of a slot with itself).  On a mesh of
// axis-aligned elements the pairs of two different reference derivatives are left out of every
// element's contraction and take no slot (-1).
struct FuSlots {
  int ws[SF_MAXPAIR];
  int nslot;
  uint32_t wfirst;
};

static FuSlots fu_slots(const SumfactPlan& SP, int i, int j, bool aligned) {
  FuSlots W{};
  for (int p = 0; p < SP.npair; ++p) {
    const int r = SP.pr[p], c = SP.pc[p];
    if (aligned && ((SP.offdiag >> p) & 1)) {
      W.ws[p] = -1;
      continue;
    }
    int q = -1;
    if (r > c && ((r >= 1 && c >= 1) || i == j))
      for (int t = 0; t < p; ++t)
        if (SP.pr[t] == c && SP.pc[t] == r && W.ws[t] >= 0) q = t;
    if (q >= 0) {
      W.ws[p] = W.ws[q];
    } else {
      W.ws[p] = W.nslot++;
      W.wfirst |= 1u << p;
    }
  }
  return W;
}

// The pairs and fields of block (i, j), as the element-matrix kernel plans them.
static int fu_plan(const hm_problem* P, int i, int j, SumfactPlan& SP, int* nt, int* nc) {
  if (int rc = check_tensor(P)) return rc;
  *nt = i == HM_PARAMETER ? P->n1_m : P->n1_u;
  *nc = j == HM_PARAMETER ? P->n1_m : P->n1_u;
  if (*nt < 2 || *nt > 3 || *nc < 2 || *nc > 3) HM_FAIL(HM_ERR_UNSUPPORTED, "1-D orders outside 1..2");
  if (*nt < *nc) HM_FAIL(HM_ERR_UNSUPPORTED, "fused assembly takes test orders of at least the trial order");
  SP = SumfactPlan{};
  const bool nl = P->density == HM_DENSITY_NONLINEAR;
  SP.npair = block_pairs(i, j, nl, SP.pr, SP.pc, 0, &SP.kmask);
  if (SP.npair == 0) HM_FAIL(HM_ERR_UNSUPPORTED, "block (%d, %d) is zero for this density", i, j);
  plan_masks(SP);
  SP.need_fields = !(i * 3 + j == 6 && !nl);
  if (SP.need_fields && (P->n1_u < 2 || P->n1_u > 3))
    HM_FAIL(HM_ERR_UNSUPPORTED, "state fields of 1-D order outside 1..2");
  return HM_OK;
}

// The blocks of (i, j) that share a compute unit on tile bx x by, in the form of the mesh's kind:
// one of 768 threads (512 for the 4-point rule); the FP32 form of the 3-point rule two of 512
// where the tile's LDS fits half of the unit's (their barriers and load waits overlap).
template <typename CT, int NT, int NC, int NQ>
static int fu_blocks(const hm_problem& P, const SumfactPlan& SP, int i, int j, bool aligned, int bx, int by) {
  if (!(std::is_same<CT, float>::value && NQ == 3)) return 1;
  const FusedSizes Z = fu_sizes<CT, NT, NC, NQ>(bx, by, fu_slots(SP, i, j, aligned).nslot, P.nd_u, P.nd_m,
                                                SP.need_fields != 0, aligned, FU_MAXZ);
  return Z.bytes <= lds_limit(2) && fu_fits(P, SP, Z, 512, bx, by, NT) ? 2 : 1;
}

// The tile of block (i, j): the largest candidate that fits the general form (every element
// tested for its alignment), so that a plan made for it serves a mesh of either kind; and the
// blocks a compute unit on it in the form of flags (nb).
template <typename CT, int NT, int NC, int NQ>
static int fu_tile(const hm_problem* P, const SumfactPlan& SP, int i, int j, int flags, int* bx, int* by, int* nb) {
  constexpr int BS = NQ == 3 ? 768 : 512;
  fu_pick<CT, NT, NC, NQ>(*P, SP, fu_slots(SP, i, j, false).nslot, false, BS, lds_limit(), bx, by);
  if (*bx == 0) HM_FAIL(HM_ERR_UNSUPPORTED, "no tile of the fused assembly fits the LDS");
  *nb = fu_blocks<CT, NT, NC, NQ>(*P, SP, i, j, (flags & HM_FUSED_ALIGNED) != 0, *bx, *by);
  return HM_OK;
}

template <typename CT>
static int fu_tile_orders(const hm_problem* P, const SumfactPlan& SP, int i, int j, int nt, int nc, int flags,
                          int* bx, int* by, int* nb) {
  if (P->nq1 == 4) {
    if (nt == 3 && nc == 3) return fu_tile<CT, 3, 3, 4>(P, SP, i, j, flags, bx, by, nb);
    if (nt == 3) return fu_tile<CT, 3, 2, 4>(P, SP, i, j, flags, bx, by, nb);
    return fu_tile<CT, 2, 2, 4>(P, SP, i, j, flags, bx, by, nb);
  }
  if (nt == 3 && nc == 3) return fu_tile<CT, 3, 3, 3>(P, SP, i, j, flags, bx, by, nb);
  if (nt == 3) return fu_tile<CT, 3, 2, 3>(P, SP, i, j, flags, bx, by, nb);
  return fu_tile<CT, 2, 2, 3>(P, SP, i, j, flags, bx, by, nb);
}

static int fu_query(const hm_problem* P, int i, int j, int precision, int flags, int* bx, int* by, int* nb) {
  if (precision != HM_FP64 && precision != HM_FP32)
    HM_FAIL(HM_ERR_UNSUPPORTED, "fused assembly in FP64 or FP32 only");
  SumfactPlan SP;
  int nt = 0, nc = 0;
  if (int rc = fu_plan(P, i, j, SP, &nt, &nc)) return rc;
  if (precision == HM_FP64) return fu_tile_orders<double>(P, SP, i, j, nt, nc, flags, bx, by, nb);
  return fu_tile_orders<float>(P, SP, i, j, nt, nc, flags, bx, by, nb);
}

template <typename CT, int NT, int NC, int QP, int NQ, typename OT, int BS, int NB, bool AL>
static int fu_run(const hm_problem& P, int i, int j, const SumfactPlan& SP, const FuSlots& W, const FusedSizes& Z,
                  const hm_fused_plan& F, void* data, hipStream_t st) {
  const int bx = F.bx, by = F.by;
  auto kern = k_fused<CT, NT, NC, QP, NQ, OT, BS, NB, AL>;
  if (Z.bytes > 48 * 1024) {
    const hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)Z.bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  FusedGeom G{};
  G.nx = F.nx;
  G.ny = F.ny;
  G.nz = F.nz;
  G.bx = bx;
  G.by = by;
  G.ntx = (F.nx + bx - 1) / bx;
  G.nty = (F.ny + by - 1) / by;
  G.hdr = (const int2*)F.hdr;
  G.hdr_off = F.hdr_off;
  G.sigma = F.sigma;
  G.sig_off = F.sig_off;
  G.diag = F.diag;
  for (int p = 0; p < SF_MAXPAIR; ++p) G.ws[p] = W.ws[p];
  G.nslot = W.nslot;
  G.wfirst = W.wfirst;
  // the pairs of each third-direction type, ascending, for a general and an aligned element
  for (int kind = 0; kind < 2; ++kind)
    for (int t = 0; t < 4; ++t) {
      int n = 0;
      G.pn[kind][t] = 0;
      if (AL && kind == 0) continue;  // every element aligned: the general lists are not read
      for (int p = 0; p < SP.npair; ++p) {
        if (!((SP.tmask[t] >> p) & 1) || (kind == 1 && ((SP.offdiag >> p) & 1))) continue;
        const int d1 = ((SP.pr[p] == 1) << 1) | (SP.pc[p] == 1);
        const int d2 = ((SP.pr[p] == 2) << 1) | (SP.pc[p] == 2);
        if (n < FU_MAXPL) G.pl[kind][t][n] = d1 | (d2 << 2) | (W.ws[p] << 4);
        ++n;
      }
      G.pn[kind][t] = n;
      if (n > FU_MAXPL) HM_FAIL(HM_ERR_UNSUPPORTED, "more than %d pairs of one third-direction type", FU_MAXPL);
    }
  // layers a chunk: the fewest rounds of blocks over the compute units (NB blocks on each)
  // times the layers a block sweeps (one more for the halo layer of a cut column)
  int cus = 256, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  if (cus < 1) cus = 256;
  cus *= NB;
  const int64_t ntile = (int64_t)G.ntx * G.nty;
  int best = 1;
  int64_t best_cost = INT64_MAX;
  for (int c = 1; c <= F.nz; ++c) {
    const int len = (F.nz + c - 1) / c;
    if (c > 1 && len < 4) break;
    if (len > FU_MAXZ) continue;
    const int nch = (F.nz + len - 1) / len;
    const int64_t rounds = (ntile * nch + cus - 1) / cus;
    const int64_t cost = rounds * (len + (nch > 1 ? 1 : 0));
    if (cost < best_cost) {
      best_cost = cost;
      best = len;
    }
  }
  G.lz_chunk = best;
  G.nchunk = (F.nz + best - 1) / best;
  const int64_t grid = ntile * G.nchunk;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BS), Z.bytes, st, P, i, j, SP, G, Z,
                     (const double*)(i == HM_PARAMETER ? P.tab1_m : P.tab1_u),
                     (const double*)(j == HM_PARAMETER ? P.tab1_m : P.tab1_u), (OT*)data);
  return hip_check("k_fused");
}

template <typename CT, int NT, int NC, int QP, int NQ, typename OT, bool AL>
static int fu_launch(const hm_problem& P, int i, int j, const SumfactPlan& SP, const hm_fused_plan& F, void* data,
                     hipStream_t st) {
  constexpr int BS = NQ == 3 ? 768 : 512;
  const FuSlots W = fu_slots(SP, i, j, AL);
  int bx = 0, by = 0;
  fu_pick<CT, NT, NC, NQ>(P, SP, fu_slots(SP, i, j, false).nslot, false, BS, lds_limit(), &bx, &by);
  if (bx != F.bx || by != F.by)
    HM_FAIL(HM_ERR_ARG, "the plan is laid out for tiles of %d x %d, this block takes %d x %d", F.bx, F.by, bx, by);
  const FusedSizes Z =
      fu_sizes<CT, NT, NC, NQ>(bx, by, W.nslot, P.nd_u, P.nd_m, SP.need_fields != 0, AL, FU_MAXZ);
  if constexpr (std::is_same<CT, float>::value && NQ == 3) {
    if (fu_blocks<CT, NT, NC, NQ>(P, SP, i, j, AL, bx, by) == 2)
      return fu_run<CT, NT, NC, QP, NQ, OT, 512, 2, AL>(P, i, j, SP, W, Z, F, data, st);
  }
  return fu_run<CT, NT, NC, QP, NQ, OT, BS, 1, AL>(P, i, j, SP, W, Z, F, data, st);
}

template <typename CT, int QP, typename OT, bool AL>
static int fu_orders(const hm_problem& P, int i, int j, int nt, int nc, const SumfactPlan& SP,
                     const hm_fused_plan& F, void* data, hipStream_t st) {
  if (P.nq1 == 4) {
    if (nt == 3 && nc == 3) return fu_launch<CT, 3, 3, QP, 4, OT, AL>(P, i, j, SP, F, data, st);
    if (nt == 3 && nc == 2) return fu_launch<CT, 3, 2, QP, 4, OT, AL>(P, i, j, SP, F, data, st);
    return fu_launch<CT, 2, 2, QP, 4, OT, AL>(P, i, j, SP, F, data, st);
  }
  if (nt == 3 && nc == 3) return fu_launch<CT, 3, 3, QP, 3, OT, AL>(P, i, j, SP, F, data, st);
  if (nt == 3 && nc == 2) return fu_launch<CT, 3, 2, QP, 3, OT, AL>(P, i, j, SP, F, data, st);
  return fu_launch<CT, 2, 2, QP, 3, OT, AL>(P, i, j, SP, F, data, st);
}

// the mesh's kind: every element axis-aligned (plan flag), or any
template <typename CT, int QP, typename OT>
static int fu_form(const hm_problem& P, int i, int j, int nt, int nc, const SumfactPlan& SP, const hm_fused_plan& F,
                   void* data, hipStream_t st) {
  if (F.flags & HM_FUSED_ALIGNED) return fu_orders<CT, QP, OT, true>(P, i, j, nt, nc, SP, F, data, st);
  return fu_orders<CT, QP, OT, false>(P, i, j, nt, nc, SP, F, data, st);
}

}  // namespace hm

using namespace hm;

extern "C" int hm_fused_tile(const hm_problem* P, int i, int j, int precision, int32_t* bx, int32_t* by) {
  HM_REQUIRE(P && bx && by, "bad fused tile arguments");
  int x = 0, y = 0, nb = 1;
  const int rc = fu_query(P, i, j, precision, 0, &x, &y, &nb);
  *bx = x;
  *by = y;
  return rc;
}

extern "C" int hm_fused_blocks(const hm_problem* P, int i, int j, int precision, int32_t flags, int32_t* blocks) {
  HM_REQUIRE(P && blocks, "bad fused blocks arguments");
  int x = 0, y = 0, nb = 1;
  const int rc = fu_query(P, i, j, precision, flags, &x, &y, &nb);
  *blocks = nb;
  return rc;
}

extern "C" int hm_assemble_fused(const hm_problem* P, int i, int j, const hm_fused_plan* F, int precision,
                                 void* data, int out_dtype, void* stream) {
  HM_REQUIRE(P && F && data, "bad fused assembly arguments");
  HM_REQUIRE(F->nx >= 1 && F->ny >= 1 && F->nz >= 1 && (int64_t)F->nx * F->ny * F->nz == P->ne,
             "a box of %d x %d x %d elements against %lld", F->nx, F->ny, F->nz, (long long)P->ne);
  HM_REQUIRE(F->hdr && F->hdr_off && F->sigma && F->sig_off, "the fused plan is incomplete");
  HM_REQUIRE((int64_t)F->nx * F->ny < ((int64_t)1 << 31), "a layer of %d x %d elements", F->nx, F->ny);
  SumfactPlan SP;
  int nt = 0, nc = 0;
  if (int rc = fu_plan(P, i, j, SP, &nt, &nc)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (precision == HM_FP64 && out_dtype == HM_FP64)
    return fu_form<double, HM_FP64, double>(*P, i, j, nt, nc, SP, *F, data, st);
  if (precision == HM_FP32 && out_dtype == HM_FP32)
    return fu_form<float, HM_FP32, float>(*P, i, j, nt, nc, SP, *F, data, st);
  HM_FAIL(HM_ERR_UNSUPPORTED, "fused assembly of precision %d into values of type %d", precision, out_dtype);
}
