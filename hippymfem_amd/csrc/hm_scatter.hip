// The scatter of element arrays into CSR values and dof vectors: a deterministic
// segmented reduction with no atomics on values (in memory or in LDS).
//
// Every CSR row owns its sum.  Its entries come from its element rows (runs t = e*nd_t + a,
// ascending, from the pattern build), each a contiguous run of nd_c entries of E, and the
// place of every entry within the row (its code, stored in run order so that a row's codes
// lie together).  A row is reduced by one group of lanes
// (half a wavefront when nd_c <= 32) into a row buffer in LDS, run after run in ascending
// order, so every CSR value is the sum of its entries in ascending entry order: the order of
// np.bincount, the reference's host scatter (hippymfem/fem/pattern.py:788).  The row buffer
// then goes out with contiguous stores.  Eliminated rows are written as zero rows (with the
// diagonal of the policy at their first place, which is the diagonal on a square block) and
// never read E; eliminated columns carry the code "all ones" and are dropped
// (pattern.py:693-713, elimination.py:53-72).
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_device.h"

namespace hm {

// LDS of the row buffers of one block (the dynamic LDS a launch takes without asking for more)
constexpr size_t SCATTER_LDS = 64 * 1024;

template <typename CodeT>
struct Drop;
template <>
struct Drop<uint8_t> {
  static constexpr unsigned v = 0xFFu;
};
template <>
struct Drop<uint16_t> {
  static constexpr unsigned v = 0xFFFFu;
};

// Data-parallel moves between the lanes of a row of 16 (and from one row into the next):
// a double moved as its two halves.  Lanes of rows left out of ROWS get 0.
template <int CTRL, int ROWS = 0xF>
__device__ inline double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// The sum of d over the LPR lanes of a group (groups aligned to LPR lanes), in a fixed order
// on the vector units: pairs within quads, then the mirror pairs of 8 and of 16 lanes, then
// the rows of 16 broadcast into the next.  Valid in the group's last lane.
template <int LPR>
__device__ inline double group_sum(double d) {
  d += dpp_f64<0xB1>(d);  // quad_perm [1, 0, 3, 2]
  d += dpp_f64<0x4E>(d);  // quad_perm [2, 3, 0, 1]
  if constexpr (LPR >= 8) d += dpp_f64<0x141>(d);        // row_half_mirror
  if constexpr (LPR >= 16) d += dpp_f64<0x140>(d);       // row_mirror
  if constexpr (LPR >= 32) d += dpp_f64<0x142, 0xA>(d);  // row_bcast:15 into rows 1 and 3
  if constexpr (LPR >= 64) d += dpp_f64<0x143, 0xC>(d);  // row_bcast:31 into rows 2 and 3
  return d;
}

// LPR lanes reduce one row (LPR a power of two up to 64, at least the element rows' entries: lane
// b takes entry b of every run); a block of 256 threads holds 256/LPR row buffers.  The runs
// of a row are taken eight at a time: their ids are read by one load and spread with a lane
// broadcast, all eight runs' entries and codes are loaded before any is added (eight loads
// in flight per lane), then added run after run in ascending order.
template <typename CodeT, typename ET, typename OT, int LPR>
__global__ void __launch_bounds__(256) k_scatter_rows(hm_scatter_plan S, const ET* __restrict__ E,
                                                      int64_t t0, int64_t t1, OT* __restrict__ data,
                                                      int accumulate, double* __restrict__ defect) {
  extern __shared__ double rowbuf[];
  constexpr int groups = 256 / LPR;
  constexpr int BATCH = 8;
  const int lane = threadIdx.x % LPR;
  const int grp = threadIdx.x / LPR;
  double* buf = rowbuf + (size_t)grp * S.max_row;
  const CodeT* codes = (const CodeT*)S.codes;
  const int nd_c = S.nd_c;
  constexpr unsigned DROP = Drop<CodeT>::v;
  // with row headers (LPR >= 16), a row's tables and first eight run ids arrive in one load
  const bool rh = LPR >= 16 && S.rowhead != nullptr;
  for (int64_t r = (int64_t)blockIdx.x * groups + grp; r < S.nrow; r += (int64_t)gridDim.x * groups) {
    int32_t lo, len, p0, p1, head = 0;
    int mode;
    if (rh) {
      head = lane < HM_ROWHEAD ? S.rowhead[r * HM_ROWHEAD + lane] : 0;
      const int32_t packed = __shfl(head, 1, LPR);
      lo = __shfl(head, 0, LPR);
      len = packed & 0xFFFF;
      mode = (packed >> 16) & 0xFF;
      p0 = __shfl(head, 2, LPR);
      p1 = p0 + ((packed >> 24) & 0xFF);
      head = __shfl(head, 3 + (lane & 7), LPR);  // lane k < 8: the row's run k
    } else {
      lo = S.indptr[r];
      len = S.indptr[r + 1] - lo;
      mode = S.rowmode ? S.rowmode[r] : 0;
      p0 = S.rptr[r];
      p1 = S.rptr[r + 1];
    }
    if (mode != 0) {
      for (int i = lane; i < len; i += LPR) data[lo + i] = (OT)((mode == 2 && i == 0) ? S.diag : 0.0);
      if (defect && lane == 0) defect[r] = 0.0;
      continue;
    }
    for (int i = lane; i < len; i += LPR) buf[i] = accumulate ? (double)data[lo + i] : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int32_t p = p0; p < p1; p += BATCH) {
      const int m = p1 - p < BATCH ? p1 - p : BATCH;
      const int32_t mine = (rh && p == p0) ? head : (lane < m ? S.runs[p + lane] : -1);
      // the runs' entries (in their own type) and codes in flight before the first add
      ET v[BATCH];
      unsigned c[BATCH];
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int64_t t = __shfl(mine, k, LPR);
        c[k] = DROP;
        v[k] = (ET)0;
        if (k < m && lane < nd_c && t >= t0 && t < t1) {
          c[k] = codes[(int64_t)(p + k) * nd_c + lane];
          v[k] = E[(t - t0) * nd_c + lane];
        }
      }
      // one read-add-write a run, run after run: the entries of one run have distinct places
      // (no lane of the row reads a place another lane writes in the same run) and a
      // wavefront's LDS operations complete in issue order, so every sum takes its entries in
      // ascending run order; no atomics
#pragma unroll
      for (int k = 0; k < BATCH; ++k)
        if (c[k] != DROP) buf[c[k]] = buf[c[k]] + (double)v[k];
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (defect) {
      // what the rounding to the values' type takes from the row's sum: each lane its
      // places in order, then a fixed tree over the row's lanes
      double d = 0.0;
      for (int i = lane; i < len; i += LPR) {
        const OT v = (OT)buf[i];
        data[lo + i] = v;
        d += buf[i] - (double)v;
      }
      d = group_sum<LPR>(d);
      if (lane == LPR - 1) defect[r] = d;
    } else {
      for (int i = lane; i < len; i += LPR) data[lo + i] = (OT)buf[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// One thread per dof: the element-vector entries of its runs, in ascending order.
__global__ void __launch_bounds__(256) k_scatter_vector(int64_t n, const int32_t* __restrict__ rptr,
                                                        const int32_t* __restrict__ runs,
                                                        const double* __restrict__ V, int64_t t0,
                                                        int64_t t1, const int8_t* __restrict__ zero_rows,
                                                        double* __restrict__ out, int accumulate) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    if (zero_rows && zero_rows[r]) {
      out[r] = 0.0;
      continue;
    }
    double acc = accumulate ? out[r] : 0.0;
    for (int32_t p = rptr[r]; p < rptr[r + 1]; ++p) {
      const int64_t t = runs[p];
      if (t >= t0 && t < t1) acc += V[t - t0];
    }
    out[r] = acc;
  }
}

// y = A x: 16 lanes per row, a fixed butterfly for the row sum (deterministic).
template <typename VT>
__global__ void __launch_bounds__(256) k_csr_matvec(int64_t nrow, const int32_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices,
                                                    const VT* __restrict__ data,
                                                    const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x & 15;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; r < nrow;
       r += ((int64_t)gridDim.x * blockDim.x) >> 4) {
    double acc = 0.0;
    for (int32_t k = indptr[r] + lane; k < indptr[r + 1]; k += 16) acc = fma((double)data[k], x[indices[k]], acc);
    for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 16);
    if (lane == 0) y[r] = acc;
  }
}

template <typename CodeT, typename ET, typename OT, int LPR>
static void launch_lpr(const hm_scatter_plan& S, const void* E, int64_t t0, int64_t t1, void* data,
                       int accumulate, double* defect, hipStream_t st) {
  constexpr int groups = 256 / LPR;
  const int64_t nb = (S.nrow + groups - 1) / groups;
  const int grid = (int)(nb < 131072 ? (nb < 1 ? 1 : nb) : 131072);
  hipLaunchKernelGGL((k_scatter_rows<CodeT, ET, OT, LPR>), dim3(grid), dim3(256),
                     (size_t)groups * S.max_row * sizeof(double), st, S, (const ET*)E, t0, t1, (OT*)data,
                     accumulate, defect);
}

// Lanes a row: the least power of two that holds an element row (measured on MI355X at 64^3:
// the reduction is latency-bound, so more rows in flight win over wider rows).
template <typename CodeT, typename ET, typename OT>
static void launch_rows(const hm_scatter_plan& S, const void* E, int64_t t0, int64_t t1, void* data,
                        int accumulate, double* defect, hipStream_t st) {
  if (S.nd_c <= 8)
    launch_lpr<CodeT, ET, OT, 8>(S, E, t0, t1, data, accumulate, defect, st);
  else if (S.nd_c <= 16)
    launch_lpr<CodeT, ET, OT, 16>(S, E, t0, t1, data, accumulate, defect, st);
  else if (S.nd_c <= 32)
    launch_lpr<CodeT, ET, OT, 32>(S, E, t0, t1, data, accumulate, defect, st);
  else
    launch_lpr<CodeT, ET, OT, 64>(S, E, t0, t1, data, accumulate, defect, st);
}

template <typename CodeT>
static int dispatch_types(const hm_scatter_plan& S, const void* E, int e_dtype, int64_t t0, int64_t t1,
                          void* data, int out_dtype, int accumulate, double* defect, hipStream_t st) {
  if (e_dtype == HM_FP64 && out_dtype == HM_FP64)
    launch_rows<CodeT, double, double>(S, E, t0, t1, data, accumulate, defect, st);
  else if (e_dtype == HM_FP32 && out_dtype == HM_FP32)
    launch_rows<CodeT, float, float>(S, E, t0, t1, data, accumulate, defect, st);
  else if (e_dtype == HM_FP32 && out_dtype == HM_FP64)
    launch_rows<CodeT, float, double>(S, E, t0, t1, data, accumulate, defect, st);
  else if (e_dtype == HM_FP64 && out_dtype == HM_FP32)
    launch_rows<CodeT, double, float>(S, E, t0, t1, data, accumulate, defect, st);
  else
    HM_FAIL(HM_ERR_ARG, "element dtype %d / value dtype %d not supported", e_dtype, out_dtype);
  return HM_OK;
}

}  // namespace hm

using namespace hm;

namespace {

static int scatter_matrix(const hm_scatter_plan* S, const void* E, int e_dtype, int64_t e0, int64_t e1, void* data,
                          int out_dtype, int accumulate, double* defect, void* stream) {
  HM_REQUIRE(S != nullptr, "no scatter plan given");
  HM_REQUIRE(S->nrow >= 0 && S->nd_t > 0 && S->nd_c > 0, "bad scatter plan shape");
  if (S->nd_c > HM_MAXND)
    HM_FAIL(HM_ERR_UNSUPPORTED, "element rows of %d entries; the scatter takes at most %d", S->nd_c, HM_MAXND);
  HM_REQUIRE(e0 >= 0 && e0 <= e1 && e1 <= S->ne, "elements [%lld, %lld) outside [0, %lld)",
             (long long)e0, (long long)e1, (long long)S->ne);
  HM_REQUIRE(S->code_bytes == 1 || S->code_bytes == 2, "code width %d is neither 1 nor 2", S->code_bytes);
  HM_REQUIRE(S->max_row >= 0, "negative longest row %d", S->max_row);
  {
    // a block holds 256 / LPR row buffers of max_row fp64 values, LPR as launch_rows picks it
    const int lpr = S->nd_c <= 8 ? 8 : S->nd_c <= 16 ? 16 : S->nd_c <= 32 ? 32 : 64;
    const size_t need = (size_t)(256 / lpr) * (size_t)S->max_row * sizeof(double);
    if (need > SCATTER_LDS)
      HM_FAIL(HM_ERR_UNSUPPORTED,
              "a row of %d entries needs %zu bytes of row buffers for %d rows a block; the scatter "
              "takes at most %zu",
              S->max_row, need, 256 / lpr, (size_t)SCATTER_LDS);
  }
  if (S->nrow == 0) return HM_OK;
  const int64_t t0 = e0 * S->nd_t, t1 = e1 * S->nd_t;
  hipStream_t st = (hipStream_t)stream;
  int rc = S->code_bytes == 1
               ? dispatch_types<uint8_t>(*S, E, e_dtype, t0, t1, data, out_dtype, accumulate, defect, st)
               : dispatch_types<uint16_t>(*S, E, e_dtype, t0, t1, data, out_dtype, accumulate, defect, st);
  if (rc) return rc;
  return hip_check("hm_scatter_matrix");
}

}  // namespace

extern "C" int hm_scatter_matrix(const hm_scatter_plan* S, const void* E, int e_dtype, int64_t e0,
                                 int64_t e1, void* data, int out_dtype, int accumulate, void* stream) {
  return scatter_matrix(S, E, e_dtype, e0, e1, data, out_dtype, accumulate, nullptr, stream);
}

extern "C" int hm_scatter_matrix_defect(const hm_scatter_plan* S, const void* E, int e_dtype, int64_t e0,
                                        int64_t e1, void* data, int out_dtype, double* row_defect,
                                        void* stream) {
  HM_REQUIRE(row_defect != nullptr, "no row defect array given");
  return scatter_matrix(S, E, e_dtype, e0, e1, data, out_dtype, 0, row_defect, stream);
}

extern "C" int hm_scatter_vector(int64_t n, int nd, const int32_t* rptr, const int32_t* runs,
                                 const double* V, int64_t e0, int64_t e1, const int8_t* zero_rows,
                                 double* out, int accumulate, void* stream) {
  HM_REQUIRE(n >= 0 && nd > 0 && e0 >= 0 && e0 <= e1, "bad vector scatter arguments");
  if (n == 0) return HM_OK;
  const int64_t nb = (n + 255) / 256;
  const int grid = (int)(nb < 65536 ? nb : 65536);
  hipLaunchKernelGGL(k_scatter_vector, dim3(grid), dim3(256), 0, (hipStream_t)stream, n, rptr, runs, V,
                     e0 * nd, e1 * nd, zero_rows, out, accumulate);
  return hip_check("hm_scatter_vector");
}

extern "C" int hm_csr_matvec(int64_t nrow, const int32_t* indptr, const int32_t* indices,
                             const void* data, int val_dtype, const double* x, double* y, void* stream) {
  HM_REQUIRE(nrow >= 0, "negative row count");
  if (nrow == 0) return HM_OK;
  const int64_t nb = (nrow * 16 + 255) / 256;
  const int grid = (int)(nb < 131072 ? nb : 131072);
  hipStream_t st = (hipStream_t)stream;
  if (val_dtype == HM_FP64)
    hipLaunchKernelGGL(k_csr_matvec<double>, dim3(grid), dim3(256), 0, st, nrow, indptr, indices,
                       (const double*)data, x, y);
  else if (val_dtype == HM_FP32)
    hipLaunchKernelGGL(k_csr_matvec<float>, dim3(grid), dim3(256), 0, st, nrow, indptr, indices,
                       (const float*)data, x, y);
  else
    HM_FAIL(HM_ERR_ARG, "value dtype %d not supported", val_dtype);
  return hip_check("hm_csr_matvec");
}
