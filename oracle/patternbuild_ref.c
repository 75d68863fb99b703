/* C restatement of the reference's sort-free pattern build -- TEST INFRASTRUCTURE ONLY.
 *
 * Follows hippymfem/fem/_patternkernels.py:110-183 (group_table, order_table,
 * place_table) and hippymfem/fem/patternbuild.py:209-253 (build_tables), serially:
 *   - element rows t = e*ndt + a are grouped by their row dof (a counting pass);
 *   - each row orders its entries by the packed key (code << 16) | q, where q is the
 *     entry's place in the row walk and code is 0 for the diagonal of a square pattern,
 *     else column + 1 (so rows lead with their diagonal, then ascend);
 *   - one scan gives the distinct columns and each entry's place; the places are then
 *     offset by the row start.
 * Output: indptr (nrow+1, int64), indices (nnz, int32), slot (nent, int64).
 * Two phases: hm_ref_count fills indptr and returns nnz (or a negative error),
 * hm_ref_place fills indices and slot.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define QBITS 16
#define QMASK ((1LL << QBITS) - 1)

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* rcount (nrow+1) and runs (ne*ndt) as group_table builds them (_patternkernels.py:110-127) */
static void group_table(const int64_t* tdofs, int64_t ne, int ndt, int64_t nrow,
                        int64_t* rcount, int64_t* runs) {
    memset(rcount, 0, sizeof(int64_t) * (size_t)(nrow + 1));
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < ndt; ++a) rcount[tdofs[e * ndt + a] + 1] += 1;
    for (int64_t r = 0; r < nrow; ++r) rcount[r + 1] += rcount[r];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrow + 1));
    memcpy(fill, rcount, sizeof(int64_t) * (size_t)(nrow + 1));
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < ndt; ++a) {
            int64_t r = tdofs[e * ndt + a];
            runs[fill[r]++] = e * ndt + a;
        }
    free(fill);
}

/* Returns nnz, or -1 when a row holds 2**16 entries or more (patternbuild.py:233-236). */
int64_t hm_ref_build(const int64_t* tdofs, const int64_t* cdofs, int64_t ne, int ndt,
                     int ndr, int64_t nrow, int diag_first, int64_t* indptr,
                     int32_t* indices /* nullable: count only */, int64_t* slot) {
    int64_t* rcount = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrow + 1));
    int64_t* runs = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ne * ndt + 1));
    group_table(tdofs, ne, ndt, nrow, rcount, runs);
    int64_t maxlen = 1;
    for (int64_t r = 0; r < nrow; ++r) {
        int64_t m = (rcount[r + 1] - rcount[r]) * ndr;
        if (m > maxlen) maxlen = m;
    }
    if (maxlen >= (1LL << QBITS)) {
        free(rcount);
        free(runs);
        return -1;
    }
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)maxlen);
    int64_t* ent = (int64_t*)malloc(sizeof(int64_t) * (size_t)maxlen);
    /* order_table (_patternkernels.py:129-161) */
    indptr[0] = 0;
    for (int64_t r = 0; r < nrow; ++r) {
        int64_t m = (rcount[r + 1] - rcount[r]) * ndr;
        int64_t q = 0;
        for (int64_t p = rcount[r]; p < rcount[r + 1]; ++p) {
            int64_t t = runs[p], e = t / ndt, k0 = t * ndr;
            for (int b = 0; b < ndr; ++b) {
                int64_t c = cdofs[e * ndr + b];
                int64_t code = (diag_first && c == r) ? 0 : c + 1;
                key[q] = (code << QBITS) | q;
                ent[q] = k0 + b;
                ++q;
            }
        }
        qsort(key, (size_t)m, sizeof(int64_t), cmp_i64);
        int64_t pos = -1, prev = -1;
        for (int64_t s = 0; s < m; ++s) {
            int64_t code = key[s] >> QBITS;
            if (code != prev) {
                ++pos;
                prev = code;
            }
            slot[ent[key[s] & QMASK]] = pos;
        }
        indptr[r + 1] = indptr[r] + (pos + 1);
    }
    /* place_table (_patternkernels.py:163-183) */
    for (int64_t r = 0; r < nrow; ++r) {
        int64_t base = indptr[r];
        for (int64_t p = rcount[r]; p < rcount[r + 1]; ++p) {
            int64_t t = runs[p], e = t / ndt, k0 = t * ndr;
            for (int b = 0; b < ndr; ++b) {
                int64_t k = k0 + b;
                if (indices) indices[base + slot[k]] = (int32_t)cdofs[e * ndr + b];
                slot[k] += base;
            }
        }
    }
    free(key);
    free(ent);
    free(rcount);
    free(runs);
    return indptr[nrow];
}

/* np.bincount(slot, weights, minlength=nnz) in entry order (pattern.py:788). */
void hm_ref_bincount(const int64_t* slot, const double* w, int64_t n, double* out,
                     int64_t nnz) {
    memset(out, 0, sizeof(double) * (size_t)nnz);
    for (int64_t k = 0; k < n; ++k) out[slot[k]] += w[k];
}
