// The pattern build on the device: the arrays of hm_pattern_build and hm_pattern_place, bit
// for bit, from device dof tables.
//
// The steps are the reference's sort-free build (hippymfem/fem/patternbuild.py:209-253,
// hippymfem/fem/_patternkernels.py:110-183), laid out for the GPU:
//   1. every element row t = e nd_t + a counts into its row dof; an exclusive scan gives rptr;
//   2. every element row takes a place in its row's run list (an integer atomic cursor, in any
//      order), then each row sorts its few run ids ascending, which gives the runs of the
//      counting pass exactly;
//   3. a wavefront a row reads the row's entries (its runs' columns, in walk order), keyed by
//      code = 0 for the diagonal of a square pattern and column + 1 otherwise; an entry's
//      place is the number of distinct codes below its own; the row's length is the number
//      of distinct codes, and an element row with two entries of one code is flagged;
//   4. a scan of the lengths gives indptr; every entry writes its column at its place.
// Integer atomics only: the counts are order-free and every list is sorted before use, so
// the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_device.h"

namespace hm {

constexpr int PB_SCAN = 1024;     // rows a block of the scan
constexpr int PB_MAXENT = 2048;   // entries of one row in the order step (LDS)

// int32 scalars of the workspace
enum { PB_BAD_ROW, PB_BAD_COL, PB_LONGEST, PB_WIDEST, PB_REPEAT, PB_TOOLONG, PB_NFLAG = 8 };

struct PbWork {
  int64_t* tot;     // block totals, then block offsets
  int64_t* total;   // 2: the totals of the two scans
  int32_t* flags;   // PB_NFLAG
  int32_t* cnt;     // nrow + 1: element rows a row, then entries a row
  int32_t* cursor;  // nrow + 1
};

static int64_t scan_blocks(int64_t nrow) { return (nrow + PB_SCAN - 1) / PB_SCAN + 1; }

static int64_t work_bytes(int64_t nrow) {
  return 8 * (scan_blocks(nrow) + 2) + 4 * (PB_NFLAG + 2 * (nrow + 1));
}

static PbWork carve(void* w, int64_t nrow) {
  PbWork p;
  p.tot = (int64_t*)w;
  p.total = p.tot + scan_blocks(nrow);
  p.flags = (int32_t*)(p.total + 2);
  p.cnt = p.flags + PB_NFLAG;
  p.cursor = p.cnt + (nrow + 1);
  return p;
}

__global__ void k_pb_count(const int32_t* __restrict__ tdofs, int64_t nrun, int64_t nrow, int32_t* cnt,
                           int32_t* flags) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrun; t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = tdofs[t];
    if (r < 0 || r >= nrow)
      atomicOr(&flags[PB_BAD_ROW], 1);
    else
      atomicAdd(&cnt[r], 1);
  }
}

__global__ void k_pb_check_cols(const int32_t* __restrict__ cdofs, int64_t n, int64_t ncol, int32_t* flags) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    if (cdofs[k] < 0 || cdofs[k] >= ncol) atomicOr(&flags[PB_BAD_COL], 1);
}

// Step one of the exclusive scan of in[0..n) into out[0..n]: offsets within each block, the
// block's total in tot, the largest input in *mx.  A block's total stays below 2^31: an input
// is a count of element rows or of distinct entries of one row, below 2^16 unless the row is
// refused anyway, and a block holds 2^10 of them.
__global__ void __launch_bounds__(PB_SCAN) k_pb_scan_blocks(const int32_t* __restrict__ in, int64_t n,
                                                           int32_t* __restrict__ out, int64_t* tot, int32_t* mx) {
  __shared__ int32_t s[PB_SCAN];
  const int64_t i = (int64_t)blockIdx.x * PB_SCAN + threadIdx.x;
  const int32_t v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < PB_SCAN; off <<= 1) {
    const int32_t add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == PB_SCAN - 1) tot[blockIdx.x] = s[threadIdx.x];
  if (v > 0) atomicMax(mx, v);
}

// Step two: the block totals into block offsets (one block walks them), and the grand total.
__global__ void __launch_bounds__(PB_SCAN) k_pb_scan_totals(int64_t* tot, int64_t nb, int64_t* total) {
  __shared__ int64_t s[PB_SCAN];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += PB_SCAN) {
    const int64_t i = b0 + threadIdx.x;
    const int64_t v = i < nb ? tot[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < PB_SCAN; off <<= 1) {
      const int64_t add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nb) tot[i] = carry + s[threadIdx.x] - v;
    carry += s[PB_SCAN - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Step three: block offsets added, out[n] = the total.  Offsets beyond int32 wrap here; the
// host refuses such a result from the total.
__global__ void __launch_bounds__(PB_SCAN) k_pb_scan_add(int32_t* out, int64_t n, const int64_t* tot,
                                                        const int64_t* total) {
  const int64_t i = (int64_t)blockIdx.x * PB_SCAN + threadIdx.x;
  if (i < n) out[i] = (int32_t)(out[i] + tot[blockIdx.x]);
  if (i == 0) out[n] = (int32_t)*total;
}

// Every element row takes a place in its row's list.
__global__ void k_pb_fill(const int32_t* __restrict__ tdofs, int64_t nrun, int64_t nrow,
                          const int32_t* __restrict__ rptr, int32_t* cursor, int32_t* __restrict__ runs) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrun; t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = tdofs[t];
    if (r < 0 || r >= nrow) continue;
    const int32_t k = atomicAdd(&cursor[r], 1);
    runs[rptr[r] + k] = (int32_t)t;
  }
}

// Each row's run ids ascending (a row holds a few: insertion sort).  A row of more than
// PB_MAXENT entries is refused by the build (k_pb_order flags it) and left unsorted: one thread
// would otherwise sort a pathological row of tens of thousands of runs, for minutes.
__global__ void k_pb_sort_runs(const int32_t* __restrict__ rptr, int64_t nrow, int nd_c, int32_t* __restrict__ runs) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrow; r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t lo = rptr[r], hi = rptr[r + 1];
    if ((int64_t)(hi - lo) * nd_c > PB_MAXENT) continue;
    for (int32_t i = lo + 1; i < hi; ++i) {
      const int32_t v = runs[i];
      int32_t j = i - 1;
      while (j >= lo && runs[j] > v) {
        runs[j + 1] = runs[j];
        --j;
      }
      runs[j + 1] = v;
    }
  }
}

// One wavefront a row: the place of each of its entries and its length.
__global__ void __launch_bounds__(64) k_pb_order(const int32_t* __restrict__ cdofs, int64_t nrow, int nd_t,
                                                 int nd_c, int diag_first, const int32_t* __restrict__ rptr,
                                                 const int32_t* __restrict__ runs, uint16_t* __restrict__ pos,
                                                 int32_t* __restrict__ rowlen, int32_t* flags) {
  __shared__ uint32_t code[PB_MAXENT];
  __shared__ uint8_t first[PB_MAXENT];
  const int lane = threadIdx.x;
  for (int64_t r = blockIdx.x; r < nrow; r += gridDim.x) {
    const int32_t p0 = rptr[r];
    const int64_t m64 = (int64_t)(rptr[r + 1] - p0) * nd_c;
    if (m64 > PB_MAXENT) {
      if (lane == 0) {
        atomicOr(&flags[PB_TOOLONG], 1);
        rowlen[r] = 0;
      }
      continue;
    }
    const int m = (int)m64;
    // the row's entries in walk order: run after run, columns in element order
    for (int q = lane; q < m; q += 64) {
      const int64_t t = runs[p0 + q / nd_c];
      const int32_t c = cdofs[(t / nd_t) * nd_c + q % nd_c];
      code[q] = (diag_first && c == r) ? 0u : (uint32_t)c + 1u;
    }
    __syncthreads();
    // first occurrence of each code; two entries of one code in one element row are flagged
    for (int q = lane; q < m; q += 64) {
      const uint32_t cq = code[q];
      const int k0 = (q / nd_c) * nd_c;
      int f = 1;
      for (int k = 0; k < k0; ++k)
        if (code[k] == cq) {
          f = 0;
          break;
        }
      for (int k = k0; k < q; ++k)
        if (code[k] == cq) {
          f = 0;
          atomicOr(&flags[PB_REPEAT], 1);
          break;
        }
      first[q] = (uint8_t)f;
    }
    __syncthreads();
    // place = distinct codes below the entry's own; length = distinct codes
    int cnt = 0;
    for (int q = lane; q < m; q += 64) {
      const uint32_t cq = code[q];
      int below = 0;
      for (int k = 0; k < m; ++k) below += (first[k] && code[k] < cq) ? 1 : 0;
      const int64_t t = runs[p0 + q / nd_c];
      pos[t * nd_c + q % nd_c] = (uint16_t)below;
      cnt += first[q];
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (lane == 0) rowlen[r] = cnt;
    __syncthreads();
  }
}

__global__ void k_pb_place(const int32_t* __restrict__ cdofs, int64_t nrun, int64_t nrow, int nd_t, int nd_c,
                           const int32_t* __restrict__ tdofs, const int32_t* __restrict__ indptr,
                           const uint16_t* __restrict__ pos, int32_t* __restrict__ indices) {
  const int64_t n = nrun * nd_c;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = k / nd_c;
    const int32_t r = tdofs[t];
    if (r < 0 || r >= nrow) continue;
    const int32_t at = indptr[r] + pos[k];
    if (at < indptr[r + 1]) indices[at] = cdofs[(t / nd_t) * nd_c + k % nd_c];
  }
}

static unsigned grid_of(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : (g < 65536 ? g : 65536));
}

static void scan(const int32_t* in, int64_t n, int32_t* out, int64_t* tot, int64_t* total, int32_t* mx,
                 hipStream_t st) {
  const int64_t nb = (n + PB_SCAN - 1) / PB_SCAN;
  const unsigned g = (unsigned)(nb < 1 ? 1 : nb);
  hipLaunchKernelGGL(k_pb_scan_blocks, dim3(g), dim3(PB_SCAN), 0, st, in, n, out, tot, mx);
  hipLaunchKernelGGL(k_pb_scan_totals, dim3(1), dim3(PB_SCAN), 0, st, tot, (int64_t)g, total);
  hipLaunchKernelGGL(k_pb_scan_add, dim3(g), dim3(PB_SCAN), 0, st, out, n, tot, total);
}

}  // namespace hm

using namespace hm;

extern "C" int64_t hm_pattern_work_bytes(int64_t nrow) { return nrow < 0 ? 0 : work_bytes(nrow); }

extern "C" int hm_pattern_build_device(const int32_t* tdofs, const int32_t* cdofs, int64_t ne, int nd_t, int nd_c,
                                       int64_t nrow, int64_t ncol, int diag_first, int32_t* rptr, int32_t* runs,
                                       uint16_t* pos, int32_t* indptr, void* work, int64_t* info, void* stream) {
  HM_REQUIRE(ne >= 0 && nd_t > 0 && nd_c > 0 && nrow >= 0 && ncol >= 0,
             "bad pattern shape (ne=%lld, nd_t=%d, nd_c=%d, nrow=%lld, ncol=%lld)", (long long)ne, nd_t, nd_c,
             (long long)nrow, (long long)ncol);
  HM_REQUIRE(nrow < (int64_t(1) << 31) && ncol < (int64_t(1) << 31),
             "%lld rows and %lld columns do not fit 32-bit local indices", (long long)nrow, (long long)ncol);
  HM_REQUIRE(ne * nd_t < (int64_t(1) << 31), "%lld element rows do not fit int32 runs", (long long)(ne * nd_t));
  HM_REQUIRE(rptr && indptr && work && info && (ne == 0 || (tdofs && cdofs && runs && pos)), "null array");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nrun = ne * nd_t;
  PbWork w = carve(work, nrow);
  if (int rc = hm_fill_zero(work, (size_t)(work_bytes(nrow) / 4), 4, st)) return rc;
  if (nrun > 0) {
    hipLaunchKernelGGL(k_pb_count, dim3(grid_of(nrun, 256)), dim3(256), 0, st, tdofs, nrun, nrow, w.cnt, w.flags);
    hipLaunchKernelGGL(k_pb_check_cols, dim3(grid_of(ne * nd_c, 256)), dim3(256), 0, st, cdofs, ne * nd_c, ncol,
                       w.flags);
  }
  scan(w.cnt, nrow, rptr, w.tot, &w.total[0], &w.flags[PB_LONGEST], st);
  if (nrun > 0) {
    hipLaunchKernelGGL(k_pb_fill, dim3(grid_of(nrun, 256)), dim3(256), 0, st, tdofs, nrun, nrow, rptr, w.cursor,
                       runs);
    hipLaunchKernelGGL(k_pb_sort_runs, dim3(grid_of(nrow, 256)), dim3(256), 0, st, rptr, nrow, nd_c, runs);
    const unsigned gr = (unsigned)(nrow < 1 ? 1 : (nrow < 262144 ? nrow : 262144));
    hipLaunchKernelGGL(k_pb_order, dim3(gr), dim3(64), 0, st, cdofs, nrow, nd_t, nd_c, diag_first, rptr, runs,
                       pos, w.cnt, w.flags);
  }
  scan(w.cnt, nrow, indptr, w.tot, &w.total[1], &w.flags[PB_WIDEST], st);
  if (int rc = hip_check("hm_pattern_build_device")) return rc;
  // the totals and flags back to the host (this call synchronizes the stream)
  struct {
    int64_t total[2];
    int32_t flags[PB_NFLAG];
  } h;
  hipError_t e = hipMemcpyAsync(&h, w.total, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) HM_FAIL(HM_ERR_HIP, "hm_pattern_build_device: %s", hipGetErrorString(e));
  if (h.flags[PB_BAD_ROW]) HM_FAIL(HM_ERR_ARG, "a row dof is outside [0, %lld)", (long long)nrow);
  if (h.flags[PB_BAD_COL]) HM_FAIL(HM_ERR_ARG, "a column dof is outside [0, %lld)", (long long)ncol);
  if ((int64_t)h.flags[PB_LONGEST] * nd_c >= (int64_t(1) << 16))
    HM_FAIL(HM_ERR_ROW_LONG, "a row holds %lld entries, more than the packed per-row key allows (65535)",
            (long long)h.flags[PB_LONGEST] * nd_c);
  if (h.flags[PB_TOOLONG])
    HM_FAIL(HM_ERR_UNSUPPORTED, "a row holds more than %d entries, the device build's limit; build on the host",
            PB_MAXENT);
  if (h.total[1] >= (int64_t(1) << 31))
    HM_FAIL(HM_ERR_ARG, "the sparsity pattern has %lld entries, beyond int32", (long long)h.total[1]);
  info[0] = h.total[1];           // nnz
  info[1] = h.flags[PB_WIDEST];   // longest row
  info[2] = h.flags[PB_REPEAT];   // an element row repeats a column
  return HM_OK;
}

extern "C" int hm_pattern_place_device(const int32_t* tdofs, const int32_t* cdofs, int64_t ne, int nd_t, int nd_c,
                                       int64_t nrow, const int32_t* indptr, const uint16_t* pos, int32_t* indices,
                                       void* stream) {
  HM_REQUIRE(ne >= 0 && nd_t > 0 && nd_c > 0 && nrow >= 0, "bad pattern shape");
  const int64_t nrun = ne * nd_t;
  if (nrun == 0) return HM_OK;
  HM_REQUIRE(tdofs && cdofs && indptr && pos && indices, "null array");
  hipLaunchKernelGGL(k_pb_place, dim3(grid_of(nrun * nd_c, 256)), dim3(256), 0, (hipStream_t)stream, cdofs, nrun,
                     nrow, nd_t, nd_c, tdofs, indptr, pos, indices);
  return hip_check("hm_pattern_place_device");
}
