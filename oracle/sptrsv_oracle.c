/*
 * sptrsv_oracle.c -- CPU restatement of the reference's sync-free SpTRSV path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (see the header
 * for who may load it).  Built with -ffp-contract=off so that every a*b+c is
 * rounded twice, as the reference's host code is on x86-64 without FMA.
 */
#include "sptrsv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* SpTRSV_cuda/utils.h:374-391: out[0] = 0, out[i] = in[0] + ... + in[i-1]. */
void oracle_exclusive_scan(int *a, int len) {
  if (len <= 1) {
    if (len == 1) a[0] = 0;
    return;
  }
  int running = 0;
  for (int i = 0; i < len; i++) {
    int v = a[i];
    a[i] = running;
    running += v;
  }
}

/* SpTRSV_cuda/tranpose.h:6-43.  Histogram of the minor index (:17-21), scan
 * (:24), then a scatter that walks the major index in ascending order with a
 * per-bucket cursor (:26-40), so each output bucket keeps the input order. */
#define ORACLE_TRANSPOSE_BODY(VT)                                              \
  memset(outPtr, 0, sizeof(int) * (size_t)(n + 1));                            \
  for (int j = 0; j < nnz; j++) outPtr[colIdx[j]]++;                           \
  oracle_exclusive_scan(outPtr, n + 1);                                        \
  int *cursor = (int *)malloc(sizeof(int) * (size_t)(n + 1));                  \
  memcpy(cursor, outPtr, sizeof(int) * (size_t)(n + 1));                       \
  for (int row = 0; row < m; row++) {                                          \
    for (int j = rowPtr[row]; j < rowPtr[row + 1]; j++) {                      \
      int dst = cursor[colIdx[j]]++;                                           \
      outIdx[dst] = row;                                                       \
      outVal[dst] = val[j];                                                    \
    }                                                                          \
  }                                                                            \
  free(cursor);

void oracle_transpose_f64(int m, int n, int nnz, const int *rowPtr,
                          const int *colIdx, const double *val, int *outIdx,
                          int *outPtr, double *outVal) {
  ORACLE_TRANSPOSE_BODY(double)
}

void oracle_transpose_f32(int m, int n, int nnz, const int *rowPtr,
                          const int *colIdx, const float *val, int *outIdx,
                          int *outPtr, float *outVal) {
  ORACLE_TRANSPOSE_BODY(float)
}

/* SpTRSV_cuda/sptrsv_syncfree_serialref.h:6-23 */
void oracle_analyser(const int *cscRowIdx, int m, int nnz, int *inDegree) {
  memset(inDegree, 0, sizeof(int) * (size_t)m);
  for (int j = 0; j < nnz; j++) inDegree[cscRowIdx[j]] += 1;
}

/* SpTRSV_cuda/sptrsv_syncfree_serialref.h:25-108.  Columns are visited in
 * solve order (ascending for forward, descending for backward).  A column is
 * solved once all of its row's off-diagonal contributions have arrived
 * (arrived == inDegree - 1, :48-55); x = (b - left_sum) / diag with a real
 * division (:59, :88); then x * L(r, col) is added into left_sum[r] for every
 * off-diagonal entry of the column, in stored order (:63-70, :92-100).  The
 * diagonal is the first stored entry of the column for forward and the last for
 * backward. */
#define ORACLE_EXECUTOR_BODY(VT)                                               \
  int *arrived = (int *)calloc((size_t)m, sizeof(int));                        \
  VT *left_sum = (VT *)calloc((size_t)m * (size_t)rhs, sizeof(VT));            \
  int status = 0;                                                              \
  const int backward = substitution == ORACLE_SUBSTITUTION_BACKWARD;           \
  for (int step = 0; step < n; step++) {                                       \
    const int i = backward ? n - 1 - step : step;                              \
    if (inDegree[i] - 1 != arrived[i]) {                                       \
      status = -2;                                                             \
      break;                                                                   \
    }                                                                          \
    const int first = cscColPtr[i], last = cscColPtr[i + 1];                   \
    const VT diag = backward ? cscVal[last - 1] : cscVal[first];               \
    for (int k = 0; k < rhs; k++) {                                            \
      const size_t p = (size_t)i * rhs + k;                                    \
      x[p] = (b[p] - left_sum[p]) / diag;                                      \
    }                                                                          \
    const int lo = backward ? first : first + 1;                               \
    const int hi = backward ? last - 1 : last;                                 \
    for (int j = lo; j < hi; j++) {                                            \
      const int r = cscRowIdx[j];                                              \
      for (int k = 0; k < rhs; k++) {                                          \
        const VT prod = x[(size_t)i * rhs + k] * cscVal[j];                    \
        left_sum[(size_t)r * rhs + k] += prod;                                 \
      }                                                                        \
      arrived[r] += 1;                                                         \
    }                                                                          \
  }                                                                            \
  free(arrived);                                                               \
  free(left_sum);                                                              \
  return status;

int oracle_executor_f64(const int *cscColPtr, const int *cscRowIdx,
                        const double *cscVal, const int *inDegree, int m, int n,
                        int substitution, int rhs, const double *b, double *x) {
  ORACLE_EXECUTOR_BODY(double)
}

int oracle_executor_f32(const int *cscColPtr, const int *cscRowIdx,
                        const float *cscVal, const int *inDegree, int m, int n,
                        int substitution, int rhs, const float *b, float *x) {
  ORACLE_EXECUTOR_BODY(float)
}

/* SpTRSV_cuda/main.cu:347-358: column sweep, b[row] += L(row, col) * x[col]. */
void oracle_spmv_csc(const int *cscColPtr, const int *cscRowIdx,
                     const double *cscVal, int m, int n, int rhs,
                     const double *x, double *b) {
  memset(b, 0, sizeof(double) * (size_t)m * (size_t)rhs);
  for (int c = 0; c < n; c++) {
    for (int j = cscColPtr[c]; j < cscColPtr[c + 1]; j++) {
      const int r = cscRowIdx[j];
      for (int k = 0; k < rhs; k++) {
        const double prod = cscVal[j] * x[(size_t)c * rhs + k];
        b[(size_t)r * rhs + k] += prod;
      }
    }
  }
}

/* SpTRSV_cuda/sptrsv_syncfree_cuda.h:555-565 */
double oracle_rel_err_l1(const double *x, const double *x_ref, long long len) {
  double ref = 0.0, res = 0.0;
  for (long long i = 0; i < len; i++) {
    ref += fabs(x_ref[i]);
    res += fabs(x[i] - x_ref[i]);
  }
  return ref == 0.0 ? res : res / ref;
}

/* SpTRSV_cuda/findlevel.h:6-69 (Kahn BFS) and :71-147 (statistics). */
int oracle_findlevel_csc(const int *cscColPtr, const int *cscRowIdx, int m,
                         int n, int nnz, int *nlevel, int *pmin, int *pavg,
                         int *pmax) {
  if (m != n) return -1;
  int *indeg = (int *)malloc(sizeof(int) * (size_t)m);
  oracle_analyser(cscRowIdx, m, nnz, indeg);
  int *levelPtr = (int *)malloc(sizeof(int) * (size_t)(m + 2));
  int *levelItem = (int *)malloc(sizeof(int) * (size_t)(m + 1));
  int filled = 0;
  levelPtr[0] = 0;
  for (int i = 0; i < m; i++)
    if (indeg[i] == 1) levelItem[filled++] = i;
  levelPtr[1] = filled;
  int lv = 1;
  int status = 0;
  while (levelPtr[lv] != m) {
    for (int t = levelPtr[lv - 1]; t < levelPtr[lv]; t++) {
      const int node = levelItem[t];
      for (int j = cscColPtr[node]; j < cscColPtr[node + 1]; j++) {
        const int v = cscRowIdx[j];
        indeg[v]--;
        if (indeg[v] == 1) levelItem[filled++] = v;
      }
    }
    lv++;
    levelPtr[lv] = filled;
    if (levelPtr[lv] == levelPtr[lv - 1] && filled != m) {
      status = -2; /* the reference loops forever here */
      break;
    }
  }
  if (status == 0) {
    int mn = levelPtr[1] - levelPtr[0];
    int mx = mn;
    for (int l = 1; l < lv; l++) {
      const int w = levelPtr[l + 1] - levelPtr[l];
      if (w > mx)
        mx = w;
      else if (w < mn)
        mn = w;
    }
    *nlevel = lv;
    *pmin = mn;
    *pmax = mx;
    *pavg = m / lv;
  }
  free(indeg);
  free(levelPtr);
  free(levelItem);
  return status;
}

/* SpTRSV_cuda/utils.h:7-17 */
void oracle_equal_col_partition(int nparts, int cols, int *counts, int *displs) {
  const int each = (int)round((double)cols / nparts);
  for (int p = 0; p + 1 < nparts; p++) {
    counts[p] = each;
    displs[p] = p * each;
  }
  counts[nparts - 1] = cols - (nparts - 1) * each;
  displs[nparts - 1] = (nparts - 1) * each;
}

/* Number of contiguous parts the greedy sweep of utils.h:36-44 opens beyond
 * the first when no part may exceed `cap` nonzeros. */
static int nnz_parts_needed(const int *len, int ncols, int cap) {
  int sum = 0, need = 0;
  for (int c = 0; c < ncols; c++) {
    if (sum + len[c] > cap) {
      sum = len[c];
      need++;
    } else {
      sum += len[c];
    }
  }
  return need;
}

/* SpTRSV_cuda/utils.h:19-67: binary search for the smallest cap with fewer
 * than `nparts` extra parts (:29-52), then the greedy split at that cap
 * (:55-66). */
void oracle_equal_nnz_partition(int nparts, int ncols, int nnz,
                                const int *colPtr, int *counts, int *displs) {
  int *len = (int *)malloc(sizeof(int) * (size_t)(ncols > 0 ? ncols : 1));
  for (int c = 0; c < ncols; c++) len[c] = colPtr[c + 1] - colPtr[c];
  int lo = 0, hi = nnz;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (nnz_parts_needed(len, ncols, mid) < nparts)
      hi = mid;
    else
      lo = mid + 1;
  }
  int sum = 0, part = 0;
  for (int c = 0; c < ncols; c++) {
    if (sum + len[c] > lo) {
      sum = len[c];
      part++;
      counts[part] = 1;
      displs[part] = c;
    } else {
      sum += len[c];
      counts[part]++;
    }
  }
  free(len);
}

/* Residual r = b - L x in FP64, column sweep. */
static void residual_csc(const int *colPtr, const int *rowIdx,
                         const double *val, int n, int rhs, const double *b,
                         const double *x, double *r) {
  memcpy(r, b, sizeof(double) * (size_t)n * (size_t)rhs);
  for (int c = 0; c < n; c++)
    for (int j = colPtr[c]; j < colPtr[c + 1]; j++)
      for (int k = 0; k < rhs; k++) {
        const double prod = val[j] * x[(size_t)c * rhs + k];
        r[(size_t)rowIdx[j] * rhs + k] -= prod;
      }
}

int oracle_refine_f32(const int *cscColPtr, const int *cscRowIdx,
                      const double *cscVal, int n, int substitution, int rhs,
                      const double *b, double *x, int iters) {
  const int nnz = cscColPtr[n];
  const size_t len = (size_t)n * (size_t)rhs;
  float *val32 = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
  float *rhs32 = (float *)malloc(sizeof(float) * (len > 0 ? len : 1));
  float *d32 = (float *)malloc(sizeof(float) * (len > 0 ? len : 1));
  double *r = (double *)malloc(sizeof(double) * (len > 0 ? len : 1));
  int *indeg = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < nnz; j++) val32[j] = (float)cscVal[j];
  oracle_analyser(cscRowIdx, n, nnz, indeg);
  int status = 0;
  for (size_t i = 0; i < len; i++) x[i] = 0.0;
  for (int it = 0; it <= iters && status == 0; it++) {
    residual_csc(cscColPtr, cscRowIdx, cscVal, n, rhs, b, x, r);
    for (size_t i = 0; i < len; i++) rhs32[i] = (float)r[i];
    status = oracle_executor_f32(cscColPtr, cscRowIdx, val32, indeg, n, n,
                                 substitution, rhs, rhs32, d32);
    for (size_t i = 0; i < len; i++) x[i] += (double)d32[i];
  }
  free(val32);
  free(rhs32);
  free(d32);
  free(r);
  free(indeg);
  return status;
}
