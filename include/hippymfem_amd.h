/* hippymfem_amd: C ABI of the MI355X (gfx950) finite-element assembly path.
 *
 * Plain pointers and sizes only.  Every device pointer is owned by the caller (torch
 * tensors in the Python host layer); the library allocates nothing on the device.  Each
 * device call is ordered on the stream it is given and returns 0 on success or a negative
 * error code, whose text hm_last_error() gives.  Host calls (the pattern build) are
 * synchronous and use the host's threads.
 *
 * Conventions (those of the reference):
 *   slots STATE = 0 (u, H1 order p), PARAMETER = 1 (m, H1 order 1), ADJOINT = 2 (p);
 *   block (i, j) = d2R / d(slot i dofs) d(slot j dofs): rows test slot i, columns trial
 *   slot j (hippymfem/fem/kernel.py:1707-1709);
 *   element-matrix entry k = (e * nd_i + a) * nd_j + b (hippymfem/fem/_patternkernels.py:103-105);
 *   CSR rows of a square block lead with their diagonal, the other columns ascend
 *   (hippymfem/fem/pattern.py:583-587), int32 indptr/indices as hypre's diag block.
 */
#ifndef HIPPYMFEM_AMD_H
#define HIPPYMFEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { HM_STATE = 0, HM_PARAMETER = 1, HM_ADJOINT = 2 };
enum { HM_DENSITY_LINEAR = 0, HM_DENSITY_NONLINEAR = 1 };
/* arithmetic of an element contraction: FP64; FP32; BF16 or FP16 products with FP32
 * accumulation and per-element scaling.  Geometry (J, det J, J^-1) is FP64 in all. */
enum { HM_FP64 = 0, HM_FP32 = 1, HM_BF16 = 2, HM_FP16 = 3 };

/* Errors. */
enum {
  HM_OK = 0,
  HM_ERR_ARG = -1,       /* a bad argument (ValueError on the Python side)      */
  HM_ERR_HIP = -2,       /* a HIP runtime error                                 */
  HM_ERR_ROW_LONG = -3,  /* a row holds 2**16 entries or more                   */
  HM_ERR_DUP = -4,       /* an element row repeats a column                     */
  HM_ERR_NODIAG = -5,    /* an eliminated row has no diagonal entry            */
  HM_ERR_UNSUPPORTED = -6
};

/* The text of the last error of this thread. */
const char* hm_last_error(void);
/* Library version, for the loader's check. */
int hm_version(void);

/* ------------------------------------------------------------------ the problem
 * Everything the element kernels read, all device pointers.
 *   verts      (nvert, 3) fp64       vertex coordinates
 *   elem_verts (ne, 8)   int32       element vertices, local v = vx + 2 vy + 4 vz (trilinear map)
 *   edofs_u    (ne, nd_u) int32      state/adjoint dofs of each element, in the table order
 *   edofs_m    (ne, nd_m) int32      parameter dofs of each element
 *   tab_u      N_u (nq, nd_u) then G_u (nq, nd_u, 3)  fp64 reference basis of the state space
 *   tab_m      N_m (nq, nd_m) then G_m (nq, nd_m, 3)  fp64 reference basis of the parameter space
 *   tab_x      G_x (nq, 8, 3)  fp64 gradients of the trilinear vertex basis
 *   wq         (nq,) fp64 quadrature weights
 *   u, m, p    dof vectors of the three slots (fp64)
 * Replaces the element data of hippymfem/fem/elementbatch.py:159-223 (ElementGroup) and
 * :413-507 (SpaceTables), which the reference keeps per element group. */
typedef struct hm_problem {
  int64_t ne;
  int32_t nq, nd_u, nd_m, density;
  double f;
  const double* verts;
  const int32_t* elem_verts;
  const int32_t* edofs_u;
  const int32_t* edofs_m;
  const double* tab_u;
  const double* tab_m;
  const double* tab_x;
  const double* wq;
  const double* u;
  const double* m;
  const double* p;
  /* Tensor-product structure, for the sum-factorized contraction (nq1 = 0: none).
   *   nq1          Gauss points per direction (nq = nq1^3, point (i, j, k) at i + nq1 (j + nq1 k));
   *                the sum-factorized kernels take 3 or 4 (degree 4 or 5, and 6 or 7), other
   *                rules run on the generic kernels
   *   n1_u, n1_m   1-D nodes per direction of the two spaces (order + 1)
   *   tab1_u/_m    1-D basis at the 1-D points: [2][n1][nq1], values then derivatives
   *   lex_u/_m     local dof of the tensor index a1 + n1 (a2 + n1 a3), (n1^3,) int32 */
  int32_t nq1, n1_u, n1_m, reserved;
  const double* tab1_u;
  const double* tab1_m;
  const int32_t* lex_u;
  const int32_t* lex_m;
  /* NULL, or one int32 the element-matrix entries set to 1 when one of the values they write
   * is not finite (the caller clears it; it is never cleared here).  The reference's callers
   * check every element matrix so (require_finite_arrays,
   * hippymfem/modeling/PDEVariationalProblem.py:1844-1866); here the kernels check the values
   * as they store them and no second pass reads the matrices. */
  int32_t* nonfinite;
} hm_problem;

/* Element matrices of block (i, j) for elements [e0, e1): out (e1-e0, nd_i, nd_j), fp64
 * for HM_FP64, fp32 for the other precisions.  Replaces GroupKernel.hess_block /
 * _hess_slot_quadrature_fn (hippymfem/fem/kernel.py:1704-1720, :1431-1485). */
int hm_element_matrices(const hm_problem* P, int i, int j, int64_t e0, int64_t e1,
                        int precision, void* out, void* stream);

/* Element matrices of blocks (i1, j1) and (i2, j2) of one shape from one pass over the
 * elements: the geometry, the fields and the pointwise pairs of both blocks at the points once,
 * then the two contractions; out1 and out2 as hm_element_matrices writes them, to the same bits.
 * HM_FP64 or HM_FP32.  Returns HM_ERR_UNSUPPORTED, having written nothing, where the blocks do
 * not share a pass (different shapes, a block that is zero for the density, no tensor
 * structure): the caller then makes two single calls.
 * Replaces GroupKernel.element_matrices_many (hippymfem/fem/kernel.py:2048-2068) for two blocks. */
int hm_element_matrices_pair(const hm_problem* P, int i1, int j1, int i2, int j2, int64_t e0,
                             int64_t e1, int precision, void* out1, void* out2, void* stream);

/* The consistent single-precision block: K (ne, nd_i, nd_j) fp32 element matrices,
 * Rj (ne, nd_i) and Ri (ne, nd_j) their exact fp64 products with the normalized constants of
 * slot j (1/sqrt(nd_j) at every dof) and, transposed, of slot i.  out (ne, nd_i, nd_j) fp64 is
 * K plus the rank-one corrections that make its products with those constants Rj and Ri;
 * symmetric = 1 symmetrizes K first, takes Rj for both sides (Ri may be NULL) and gives a
 * result symmetric to the last bit.  Replaces GroupKernel.consistent_block
 * (hippymfem/fem/kernel.py:1503-1571) for scalar spaces (one constant mode a slot). */
int hm_consistent_block(const float* K, const double* Rj, const double* Ri, int64_t ne, int nd_i,
                        int nd_j, int symmetric, double* out, void* stream);

/* Element vectors dR_e/d(slot i) for [e0, e1): out (e1-e0, nd_i) fp64.
 * Replaces GroupKernel.grad_block (hippymfem/fem/kernel.py:1343-1355). */
int hm_element_vectors(const hm_problem* P, int i, int64_t e0, int64_t e1, double* out,
                       void* stream);

/* Element scalars R_e = sum_q w_q |det J| (exp(m) g(u) grad u . grad p - f p) for [e0, e1):
 * out (e1-e0) fp64, each the sum over its points in point order; sets *nonfinite where a value
 * is not finite.  Replaces GroupKernel.residual_value and QuadratureKernel.element_values
 * (hippymfem/fem/kernel.py:1859-1865, :2129-2136) on the functional of
 * GroupKernel._functional (kernel.py:977-1009). */
int hm_element_scalars(const hm_problem* P, int64_t e0, int64_t e1, double* out, void* stream);

/* Caller-supplied densities (QuadratureKernel(density, ...), hippymfem/fem/kernel.py:1917-1937,
 * whose pointwise derivatives the reference forms by automatic differentiation in
 * _point_blocks, kernel.py:1581-1646): the caller differentiates its density at the points
 * and the library contracts the result.
 * hm_point_fields: F (e1-e0, nq, 13) fp64 = w |det J|, u, grad u, m, grad m, p, grad p at
 *   every point (physical gradients).
 * hm_element_matrices_pointwise: block (i, j) element matrices from D (e1-e0, nq, 4, 4) fp64,
 *   d2r / d(slot i value, grad) d(slot j value, grad) in the physical basis; out as
 *   hm_element_matrices (FP64 or FP32 results); sets *nonfinite.
 * hm_element_vectors_pointwise: slot i element vectors from V (e1-e0, nq, 4) fp64,
 *   dr / d(slot i value, grad); out (e1-e0, nd_i) fp64. */
int hm_point_fields(const hm_problem* P, int64_t e0, int64_t e1, double* out, void* stream);
int hm_element_matrices_pointwise(const hm_problem* P, int i, int j, const double* D, int64_t e0,
                                  int64_t e1, int precision, void* out, void* stream);
int hm_element_vectors_pointwise(const hm_problem* P, int i, const double* V, int64_t e0,
                                 int64_t e1, double* out, void* stream);

/* Block (i, j) applied element by element to the slot-j dof vector dir:
 * out (e1-e0, nd_i) fp64.  Replaces GroupKernel.hvp_block (hippymfem/fem/kernel.py:1756-1782). */
int hm_element_hvp(const hm_problem* P, int i, int j, const double* dir, int64_t e0,
                   int64_t e1, double* out, void* stream);

/* ---------------------------------------------------------------- the pattern (host)
 * The sort-free pattern build of hippymfem/fem/patternbuild.py:209-253 (build_tables) and
 * hippymfem/fem/_patternkernels.py:110-183, on the host's threads, in two calls.
 *
 * hm_pattern_build: tdofs (ne, nd_t), cdofs (ne, nd_c) int32 host arrays.  Writes
 *   rptr (nrow+1) int64 and runs (ne*nd_t) int32: the element rows t = e*nd_t + a of
 *   every CSR row, ascending (group_table); pos (ne*nd_t*nd_c) uint16: each entry's place
 *   in its row (order_table); indptr (nrow+1) int64; *nnz; *max_row (longest row).
 *   diag_first: rows lead with their diagonal (a square pattern).  nthreads <= 0: all.
 * hm_pattern_place: indices (nnz) int32 and, if slot is not NULL, slot (ne*nd_t*nd_c)
 *   int32 = indptr[row] + pos (place_table). */
int hm_pattern_build(const int32_t* tdofs, const int32_t* cdofs, int64_t ne, int nd_t,
                     int nd_c, int64_t nrow, int64_t ncol, int diag_first, int nthreads,
                     int64_t* rptr, int32_t* runs, uint16_t* pos, int64_t* indptr,
                     int64_t* nnz, int64_t* max_row);
int hm_pattern_place(const int32_t* cdofs, int64_t ne, int nd_t, int nd_c, int64_t nrow,
                     const int64_t* rptr, const int32_t* runs, const int64_t* indptr,
                     const uint16_t* pos, int nthreads, int32_t* indices, int32_t* slot);
/* *repeats = 1 when an element row puts two entries in one place (a repeated column),
 * which the device scatter does not take (it adds one element row's entries at once). */
int hm_pattern_repeats(int64_t ne, int nd_t, int nd_c, const uint16_t* pos, int nthreads,
                       int32_t* repeats);

/* ---------------------------------------------------------------- the pattern (device)
 * The same build from device dof tables, on the given stream, with the same arrays bit for
 * bit (indptr and rptr as int32, hypre's diag-block index type): the device counterpart of
 * build_tables (hippymfem/fem/patternbuild.py:209-253) and of the counting, ordering and
 * placing kernels (hippymfem/fem/_patternkernels.py:110-183).  Integer atomics only; every
 * list is sorted before use, so the result does not depend on scheduling.
 *
 * hm_pattern_work_bytes(nrow): size of the caller's device workspace for hm_pattern_build_device.
 * hm_pattern_build_device: tdofs (ne, nd_t), cdofs (ne, nd_c) int32 device arrays.  Writes rptr
 *   (nrow+1), runs (ne*nd_t), pos (ne*nd_t*nd_c) uint16 and indptr (nrow+1); info (host, 3
 *   int64) = {nnz, longest row, 1 if an element row repeats a column}.  Synchronizes the
 *   stream to read the totals.  Errors as hm_pattern_build; rows of more than 2048 entries
 *   give HM_ERR_UNSUPPORTED (build those on the host).
 * hm_pattern_place_device: indices (nnz) int32 from the build's indptr and pos. */
int64_t hm_pattern_work_bytes(int64_t nrow);
int hm_pattern_build_device(const int32_t* tdofs, const int32_t* cdofs, int64_t ne, int nd_t,
                            int nd_c, int64_t nrow, int64_t ncol, int diag_first, int32_t* rptr,
                            int32_t* runs, uint16_t* pos, int32_t* indptr, void* work,
                            int64_t* info, void* stream);
int hm_pattern_place_device(const int32_t* tdofs, const int32_t* cdofs, int64_t ne, int nd_t,
                            int nd_c, int64_t nrow, const int32_t* indptr, const uint16_t* pos,
                            int32_t* indices, void* stream);

/* ------------------------------------------------------------------ the scatter
 * Device plan of one block's scatter: every CSR row sums the entries of its element rows
 * in ascending entry order, which is np.bincount's order (hippymfem/fem/pattern.py:788),
 * with no atomics.  Folded elimination: rowmode[r] = 0 assembles the row, 1 writes a zero
 * row, 2 a zero row with `diag` at its first (diagonal) place; an entry whose code is
 * all ones (0xFF / 0xFFFF) is dropped (an eliminated column).
 *   indptr (nrow+1) int32, rptr (nrow+1) int32, runs (ne*nd_t) int32,
 *   codes (ne*nd_t*nd_c) uint8 (code_bytes 1) or uint16 (2): the place within its row of
 *   entry b of the element row at run position p (runs[p]) is codes[p*nd_c + b], so that a
 *   row's codes lie together; rowmode (nrow) int8 or NULL. */
typedef struct hm_scatter_plan {
  int64_t nrow, nnz, ne;
  int32_t nd_t, nd_c, code_bytes, max_row;
  const int32_t* indptr;
  const int32_t* rptr;
  const int32_t* runs;
  const void* codes;
  const int8_t* rowmode;
  double diag;
  /* Optional (NULL: not used), for element rows of more than 8 entries: the same row tables
   * packed HM_ROWHEAD int32 a row, so that a row's arrive in one load: indptr[r];
   * (indptr[r+1] - indptr[r]) | rowmode[r] << 16 | (rptr[r+1] - rptr[r]) << 24; rptr[r];
   * runs[rptr[r] + k] for k < 8 (-1 past the row's runs); one unused.  A row of more than 8
   * runs reads the others from runs. */
  const int32_t* rowhead;
} hm_scatter_plan;
enum { HM_ROWHEAD = 12 };

/* data[nnz] (+)= the entries of the element matrices of elements [e0, e1) (E holds
 * exactly those, (e1-e0, nd_t, nd_c)).  e_dtype / out_dtype: HM_FP64 or HM_FP32.
 * accumulate = 0 overwrites data (a whole batch), 1 adds to it (a chunk of a fused
 * scatter, in order).  Replaces ScatterPattern.data / fused_add (hippymfem/fem/pattern.py:
 * 749-791, :864-886) and the folded elimination (pattern.py:693-713; elimination.py:53-72). */
int hm_scatter_matrix(const hm_scatter_plan* S, const void* E, int e_dtype, int64_t e0,
                      int64_t e1, void* data, int out_dtype, int accumulate, void* stream);

/* hm_scatter_matrix of a whole batch (data overwritten) that also writes row_defect[nrow]
 * fp64: what the rounding of each row's FP64 sums to the values' type takes from the row's
 * sum, sum_j (a_ij - fl(a_ij)) (0 for FP64 values and for eliminated rows), so that the
 * values plus diag(row_defect) have the row sums of the FP64 sums.  Replaces the row defect
 * of a single-precision matrix (hippymfem/algorithms/singlesolve.py:682-700, the
 * SINGLE_ROW_SUMS default of hippymfem/fem/tdofassemble.py:107-125), which the reference
 * finds from the rounded residues; here it is summed in FP64 beside the values. */
int hm_scatter_matrix_defect(const hm_scatter_plan* S, const void* E, int e_dtype, int64_t e0,
                             int64_t e1, void* data, int out_dtype, double* row_defect,
                             void* stream);

/* out[n] (+)= the element vectors V (e1-e0, nd) of [e0, e1) by dof, in entry order, then
 * out[r] = 0 where zero_rows[r] != 0 (zero_rows may be NULL).  rptr/runs as the plan's.
 * Replaces VectorPattern.scatter (hippymfem/fem/pattern.py:1186-1209) and zero_essential
 * (hippymfem/fem/bcs.py:80-102). */
int hm_scatter_vector(int64_t n, int nd, const int32_t* rptr, const int32_t* runs,
                      const double* V, int64_t e0, int64_t e1, const int8_t* zero_rows,
                      double* out, int accumulate, void* stream);

/* y = A x for a CSR block (int32 indptr/indices, fp64 or fp32 values by val_dtype, fp64
 * vectors).  Replaces HypreParMatrix::Mult on one rank (PDEVariationalProblem.apply_ij,
 * hippymfem/modeling/PDEVariationalProblem.py:1629-1670). */
int hm_csr_matvec(int64_t nrow, const int32_t* indptr, const int32_t* indices,
                  const void* data, int val_dtype, const double* x, double* y, void* stream);

/* ------------------------------------------------------------------ fused assembly
 * Element kernel and scatter in one kernel on a box mesh of nx x ny x nz hexahedra (element
 * e = ex + nx (ey + ny ez), the problem's elem_verts and dof tables as for the other calls), so
 * that no element matrix goes to memory.  Columns of bx x by elements are swept bottom to top,
 * each block owning the CSR rows of its lattice points and recomputing a one-element halo on its
 * lower sides; every CSR value is the sum of its entries in ascending element order, the order of
 * np.bincount (hippymfem/fem/pattern.py:788), with the element-matrix kernel's bits for every
 * entry.  FP64: the values are those of hm_element_matrices followed by hm_scatter_matrix, bit for
 * bit.  FP32: the FP32 entries are summed in FP32 in that order (hm_scatter_matrix sums them in
 * FP64), so that the sums take half the LDS and two blocks may share a compute unit.
 * The plan holds, in tile order (planes of the test lattice, then tiles x fastest, then the
 * tile's lattice points x fastest), every row's CSR offset and row mode (0 assemble, 1 zero row,
 * 2 zero row with diag first) and every CSR entry's place in the row's dense stencil (the box of
 * trial-lattice points the row couples to, x fastest; bit 7 set: a column dropped by the
 * elimination, written as 0).
 * Replaces the assembly of a block from its element matrices (PDEVariationalProblem._block /
 * _jacobian / setLinearizationPoint, hippymfem/modeling/PDEVariationalProblem.py:476-659,
 * :1372-1482: QuadratureKernel.element_matrices then assemble_matrix with the folded
 * elimination, hippymfem/fem/assemble.py:34-69, pattern.py:693-791).
 *
 * hm_fused_tile: the tile (bx, by) the kernel takes for block (i, j) of this problem at this
 *   precision (HM_FP64 or HM_FP32): the layout a plan must be made for.  HM_ERR_UNSUPPORTED where
 *   the problem has no tensor structure or the block is zero for the density.
 * hm_assemble_fused: data (nnz) fp64 for HM_FP64, fp32 for HM_FP32, every value written. */
typedef struct hm_fused_plan {
  int32_t nx, ny, nz;
  int32_t bx, by;
  int32_t flags;           /* HM_FUSED_ALIGNED: every element an axis-aligned box (J diagonal) */
  const int32_t* hdr;      /* (nrow, 2): CSR offset, row mode; tile order                    */
  const int64_t* hdr_off;  /* (planes * tiles): first row of each (plane, tile) in hdr       */
  const uint8_t* sigma;    /* stencil place of every CSR entry, tile order, bit 7: dropped   */
  const int64_t* sig_off;  /* (planes * tiles): start of each (plane, tile), 16-byte aligned */
  double diag;             /* the diagonal of rows of mode 2                                 */
} hm_fused_plan;

/* hm_fused_plan.flags: HM_FUSED_ALIGNED when every element of the mesh is an axis-aligned box
 * (each vertex coordinate equals that of the vertex with the same index along its axis, the test
 * the element kernels make element by element): no element then needs the pointwise pairs of two
 * different reference derivatives, and the kernel leaves out the per-element test.  A plan made
 * for a mesh stays with that mesh's coordinates. */
#define HM_FUSED_ALIGNED 1

int hm_fused_tile(const hm_problem* P, int i, int j, int precision, int32_t* bx, int32_t* by);

/* hm_fused_blocks: no counterpart in the reference (the route choice of the fused assembly).
 * The blocks of the fused kernel that share a compute unit for block (i, j) at this precision on
 * a plan with these flags: 1, or 2 where the FP32 form's tile fits half the unit's LDS (its plane
 * sums then FP32).  Errors as hm_fused_tile. */
int hm_fused_blocks(const hm_problem* P, int i, int j, int precision, int32_t flags, int32_t* blocks);
int hm_assemble_fused(const hm_problem* P, int i, int j, const hm_fused_plan* F, int precision,
                      void* data, int out_dtype, void* stream);

/* ------------------------------------------------------------------ serving inputs
 * No counterpart in the reference: the seeded inputs of the serving form (assemble-and-apply
 * requests), drawn on the device from a seed held in device memory so that a captured HIP
 * graph of a whole request draws new inputs at every replay.  Deterministic: the same seed
 * gives the same bits in any launch shape.
 *
 * hm_serving_normal: out (n) fp64 N(0, 1) values of (*seed, salt).
 * hm_serving_field: the benchmarks' parameter field on the n1^3 vertex lattice (x fastest):
 *   N(0, 1) noise of (*seed, 0), `passes` 7-point averages over the neighbours that exist,
 *   then mean 0 and population standard deviation 1 (divisor n, as numpy's std).  work (2 n1^3 + 2) fp64; out (n1^3) fp64
 *   in lattice order, or out[lattice_to_dof[i]] when lattice_to_dof is not NULL. */
int hm_serving_normal(const int64_t* seed, int64_t salt, int64_t n, double* out, void* stream);
int hm_serving_field(const int64_t* seed, int n1, int passes, double* work,
                     const int64_t* lattice_to_dof, double* out, void* stream);

/* hm_serving_field_noalloc: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc): hm_serving_field with its workspace in a static device array of the library
 * (2 HM_SERVING_FIELD_NOALLOC_MAX_N1^3 + 2 doubles, about 4.4 MB, held on every device the
 * library runs on until the process ends) instead of the caller's.  Calls no hipMalloc or
 * hipFree; n1 above HM_SERVING_FIELD_NOALLOC_MAX_N1 gives HM_ERR_ARG.  Every call on one
 * device uses that device's array: calls that overlap on one device (two streams, two
 * threads) race and must be ordered by the caller.  Nothing in the array needs to start at
 * zero (every call writes what it reads). */
#define HM_SERVING_FIELD_NOALLOC_MAX_N1 65
int hm_serving_field_noalloc(const int64_t* seed, int n1, int passes,
                             const int64_t* lattice_to_dof, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
