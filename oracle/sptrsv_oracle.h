/*
 * sptrsv_oracle.h -- CPU restatement of the reference's sync-free SpTRSV path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links, calls or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout, harunsasmaz/Benchmark-SpTRSV).  Parity pins for each
 * function (what checks it besides retyping the formula) are listed next to it.
 */
#ifndef SPTRSV_ORACLE_H
#define SPTRSV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_SUBSTITUTION_FORWARD 0  /* SpTRSV_cuda/common.h:25 */
#define ORACLE_SUBSTITUTION_BACKWARD 1 /* SpTRSV_cuda/common.h:26 */

/* In-place exclusive prefix sum (SpTRSV_cuda/utils.h:374-391).
 * Pinned by: numpy.cumsum in tests/test_oracle.py and the _ref build. */
void oracle_exclusive_scan(int *a, int len);

/* Stable CSR->CSC transpose (SpTRSV_cuda/tranpose.h:6-43).  Also transposes
 * CSC->CSR when called with the CSC arrays in the CSR slots, exactly as
 * findlevel_csc does (SpTRSV_cuda/findlevel.h:94-96).
 * Pinned by: scipy.sparse tocsc(), brute force, the _ref build. */
void oracle_transpose_f64(int m, int n, int nnz, const int *rowPtr,
                          const int *colIdx, const double *val, int *outIdx,
                          int *outPtr, double *outVal);
void oracle_transpose_f32(int m, int n, int nnz, const int *rowPtr,
                          const int *colIdx, const float *val, int *outIdx,
                          int *outPtr, float *outVal);

/* In-degree analyser: row histogram of the CSC row indices, diagonal included
 * (SpTRSV_cuda/sptrsv_syncfree_serialref.h:6-23).
 * Pinned by: closed form diff(csrRowPtr), the _ref build. */
void oracle_analyser(const int *cscRowIdx, int m, int nnz, int *inDegree);

/* Serial sync-free executor (SpTRSV_cuda/sptrsv_syncfree_serialref.h:25-108),
 * FP64 and FP32.  x and b are row-major [n][rhs] (SpTRSV_cuda/main.cu:332-336).
 * Returns 0, or -2 when the in-degree wait of :51-55 / :80-84 could never be
 * satisfied (the reference would spin forever there).
 * Pinned by: x=1 Laplacian known answer, integer unit-lower known answer,
 * scipy spsolve_triangular fixtures, dense numpy solve, the _ref build
 * (bit-exact: same operation order). */
int oracle_executor_f64(const int *cscColPtr, const int *cscRowIdx,
                        const double *cscVal, const int *inDegree, int m, int n,
                        int substitution, int rhs, const double *b, double *x);
int oracle_executor_f32(const int *cscColPtr, const int *cscRowIdx,
                        const float *cscVal, const int *inDegree, int m, int n,
                        int substitution, int rhs, const float *b, float *x);

/* b = L * x with the CSC SpMV the driver uses to build the right-hand side
 * (SpTRSV_cuda/main.cu:347-358).  b is overwritten (the driver zeroes it at
 * :343-344 first).  Pinned by: scipy.sparse matmul. */
void oracle_spmv_csc(const int *cscColPtr, const int *cscRowIdx,
                     const double *cscVal, int m, int n, int rhs,
                     const double *x, double *b);

/* Relative L1 error sum|x - x_ref| / sum|x_ref| (sptrsv_syncfree_cuda.h:555-565;
 * when sum|x_ref| == 0 the plain sum is returned, as at :565). */
double oracle_rel_err_l1(const double *x, const double *x_ref, long long len);

/* Level sets by Kahn BFS from rows whose in-degree is 1 (findlevel.h:6-69),
 * plus the statistics of findlevel_csc (findlevel.h:71-147), including its
 * else-if in the min/max update (:127-133).  Returns 0, -1 if m != n, -2 if the
 * graph is not a DAG over all rows.  Pinned by: closed forms nx+ny-1 (2D),
 * nx+ny+nz-2 (3D), n (chain). */
int oracle_findlevel_csc(const int *cscColPtr, const int *cscRowIdx, int m,
                         int n, int nnz, int *nlevel, int *pmin, int *pavg,
                         int *pmax);

/* Column partitioners (SpTRSV_cuda/utils.h:7-17 and :19-67).  For the nnz
 * partitioner counts[0] and displs[0] must be zero on entry, as the callers
 * arrange (naive_unified.h:130-131); parts the greedy split never reaches are
 * left untouched (utils.h:55-66).  Pinned by: brute-force enumeration of all
 * contiguous splits on tiny inputs in tests/test_oracle.py. */
void oracle_equal_col_partition(int nparts, int cols, int *counts, int *displs);
void oracle_equal_nnz_partition(int nparts, int ncols, int nnz,
                                const int *colPtr, int *counts, int *displs);

/* Mixed-precision iterative refinement, the precision form added on top of the
 * reference (it has no counterpart there): x0 = solve32(b); repeat
 * r = b - L x (FP64, CSC), d = solve32(r), x += d, for `iters` steps.
 * Pinned by: convergence to oracle_executor_f64 (rel err <= 1e-10). */
int oracle_refine_f32(const int *cscColPtr, const int *cscRowIdx,
                      const double *cscVal, int n, int substitution, int rhs,
                      const double *b, double *x, int iters);

#ifdef __cplusplus
}
#endif
#endif
