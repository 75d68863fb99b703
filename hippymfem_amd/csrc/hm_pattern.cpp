// Host pattern build: the CSR graph of the union of element couplings and the place of
// every element-matrix entry in it, without a global sort.
//
// The algorithm is the reference's sort-free build (hippymfem/fem/patternbuild.py:209-253,
// hippymfem/fem/_patternkernels.py:110-183): element rows are grouped by their row dof with
// one counting pass; each CSR row then orders its entries by the packed key
// (code << 16) | q, code 0 for the diagonal of a square pattern and column + 1 otherwise,
// q the entry's place in the row walk, so the row comes out diagonal first, columns
// ascending, and every entry learns its place from one scan.  The arrays are therefore the
// ones the reference's sort route gives (pattern.py:548-663).  Rows are split over host
// threads; each keeps its key buffer in a small per-thread vector.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "hm_common.h"

namespace {

constexpr int kQBits = 16;
constexpr int64_t kQMask = (int64_t(1) << kQBits) - 1;

int thread_count(int nthreads) {
  if (nthreads > 0) return nthreads;
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? int(hc) : 1;
}

// Calls fn(lo, hi) on [0, n) split into blocks over `threads` threads.
template <class F>
void parallel_rows(int64_t n, int threads, F fn) {
  if (threads <= 1 || n < 4096) {
    fn(int64_t(0), n);
    return;
  }
  const int64_t block = 4096;
  const int64_t nblk = (n + block - 1) / block;
  std::vector<std::thread> pool;
  std::atomic<int64_t> next{0};
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&] {
      for (;;) {
        int64_t b = next.fetch_add(1);
        if (b >= nblk) break;
        fn(b * block, std::min(n, (b + 1) * block));
      }
    });
  }
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" int hm_pattern_build(const int32_t* tdofs, const int32_t* cdofs, int64_t ne,
                                int nd_t, int nd_c, int64_t nrow, int64_t ncol,
                                int diag_first, int nthreads, int64_t* rptr, int32_t* runs,
                                uint16_t* pos, int64_t* indptr, int64_t* nnz,
                                int64_t* max_row) {
  HM_REQUIRE(ne >= 0 && nd_t > 0 && nd_c > 0 && nrow >= 0 && ncol >= 0,
             "bad pattern shape (ne=%lld, nd_t=%d, nd_c=%d, nrow=%lld, ncol=%lld)",
             (long long)ne, nd_t, nd_c, (long long)nrow, (long long)ncol);
  HM_REQUIRE(nrow < (int64_t(1) << 31) && ncol < (int64_t(1) << 31),
             "%lld rows and %lld columns do not fit 32-bit local indices", (long long)nrow,
             (long long)ncol);
  HM_REQUIRE(ne * nd_t < (int64_t(1) << 31), "%lld element rows do not fit int32 runs",
             (long long)(ne * nd_t));
  const int threads = thread_count(nthreads);
  const int64_t nrun = ne * nd_t;
  // group_table: rows of every element row, counted then placed in ascending t
  std::memset(rptr, 0, sizeof(int64_t) * size_t(nrow + 1));
  for (int64_t t = 0; t < nrun; ++t) {
    int32_t r = tdofs[t];
    if (r < 0 || r >= nrow) HM_FAIL(HM_ERR_ARG, "row dof %d out of [0, %lld)", r, (long long)nrow);
    rptr[r + 1] += 1;
  }
  int64_t longest = 0;
  for (int64_t r = 0; r < nrow; ++r) {
    longest = std::max(longest, rptr[r + 1]);
    rptr[r + 1] += rptr[r];
  }
  if (longest * nd_c >= (int64_t(1) << kQBits))
    HM_FAIL(HM_ERR_ROW_LONG,
            "a row holds %lld entries, more than the packed per-row key allows (%lld)",
            (long long)(longest * nd_c), (long long)kQMask);
  {
    std::vector<int64_t> fill(rptr, rptr + nrow);
    for (int64_t t = 0; t < nrun; ++t) runs[fill[tdofs[t]]++] = int32_t(t);
  }
  for (int64_t k = 0; k < ne * nd_c; ++k)
    if (cdofs[k] < 0 || cdofs[k] >= ncol)
      HM_FAIL(HM_ERR_ARG, "column dof %d out of [0, %lld)", cdofs[k], (long long)ncol);

  // order_table: each row's entries by (code << 16) | q, their places, the row's count
  std::vector<int64_t> ucount(size_t(nrow), 0);
  parallel_rows(nrow, threads, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> key;
    std::vector<int64_t> ent;
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t m = (rptr[r + 1] - rptr[r]) * nd_c;
      if (m == 0) continue;
      key.resize(size_t(m));
      ent.resize(size_t(m));
      int64_t q = 0;
      for (int64_t p = rptr[r]; p < rptr[r + 1]; ++p) {
        const int64_t t = runs[p];
        const int64_t e = t / nd_t;
        const int64_t k0 = t * nd_c;
        for (int b = 0; b < nd_c; ++b) {
          const int64_t c = cdofs[e * nd_c + b];
          const int64_t code = (diag_first && c == r) ? 0 : c + 1;
          key[size_t(q)] = (code << kQBits) | q;
          ent[size_t(q)] = k0 + b;
          ++q;
        }
      }
      std::sort(key.begin(), key.begin() + m);
      int64_t at = -1, prev = -1;
      for (int64_t s = 0; s < m; ++s) {
        const int64_t code = key[size_t(s)] >> kQBits;
        if (code != prev) {
          ++at;
          prev = code;
        }
        pos[ent[size_t(key[size_t(s)] & kQMask)]] = uint16_t(at);
      }
      ucount[size_t(r)] = at + 1;
    }
  });
  indptr[0] = 0;
  int64_t widest = 0;
  for (int64_t r = 0; r < nrow; ++r) {
    indptr[r + 1] = indptr[r] + ucount[size_t(r)];
    widest = std::max(widest, ucount[size_t(r)]);
  }
  *nnz = indptr[nrow];
  *max_row = widest;
  return HM_OK;
}

extern "C" int hm_pattern_place(const int32_t* cdofs, int64_t ne, int nd_t, int nd_c,
                                int64_t nrow, const int64_t* rptr, const int32_t* runs,
                                const int64_t* indptr, const uint16_t* pos, int nthreads,
                                int32_t* indices, int32_t* slot) {
  HM_REQUIRE(ne >= 0 && nd_t > 0 && nd_c > 0 && nrow >= 0, "bad pattern shape");
  HM_REQUIRE(indptr[nrow] < (int64_t(1) << 31),
             "the pattern has %lld entries, beyond 32-bit CSR indices",
             (long long)indptr[nrow]);
  const int threads = thread_count(nthreads);
  // place_table: every entry writes its column at its place and takes its row's offset
  parallel_rows(nrow, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t base = indptr[r];
      for (int64_t p = rptr[r]; p < rptr[r + 1]; ++p) {
        const int64_t t = runs[p];
        const int64_t e = t / nd_t;
        const int64_t k0 = t * nd_c;
        for (int b = 0; b < nd_c; ++b) {
          const int64_t k = k0 + b;
          indices[base + pos[k]] = cdofs[e * nd_c + b];
          if (slot) slot[k] = int32_t(base + pos[k]);
        }
      }
    }
  });
  return HM_OK;
}

// Whether an element row puts two of its entries in one place (a repeated column): the
// device scatter adds the entries of one element row at once and needs them distinct.
extern "C" int hm_pattern_repeats(int64_t ne, int nd_t, int nd_c, const uint16_t* pos,
                                  int nthreads, int32_t* repeats) {
  const int threads = thread_count(nthreads);
  std::atomic<int32_t> found{0};
  parallel_rows(ne * nd_t, threads, [&](int64_t lo, int64_t hi) {
    std::vector<uint16_t> v(static_cast<size_t>(nd_c));
    for (int64_t t = lo; t < hi && !found.load(std::memory_order_relaxed); ++t) {
      std::memcpy(v.data(), pos + t * nd_c, sizeof(uint16_t) * size_t(nd_c));
      std::sort(v.begin(), v.end());
      if (std::adjacent_find(v.begin(), v.end()) != v.end()) found.store(1);
    }
  });
  *repeats = found.load();
  return HM_OK;
}
