/*
 * ref_shim.cpp -- exports the reference's own host routines, compiled from the
 * reference checkout where it lies, as plain C symbols for the tests.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/Makefile).  No reference source is
 * copied here: the recipe puts the reference directories on the include path
 * and this file only forwards to the functions those headers define:
 *   matrix_transposition      SpTRSV_cuda/tranpose.h:6-43 (identical OpenCL copy)
 *   exclusive_scan            SpTRSV_opencl_amd/utils.h (cusparse-free copy of
 *                             SpTRSV_cuda/utils.h:374-391)
 *   sptrsv_syncfree_analyser  sptrsv_syncfree_serialref.h:6-23
 *   sptrsv_syncfree_executor  sptrsv_syncfree_serialref.h:25-108
 *   findlevel_csc             SpTRSV_cuda/findlevel.h:71-147
 * VALUE_TYPE is set by the recipe (double or float), one library per type.
 */
#include "common.h"
#include "utils.h"
#include "tranpose.h"
#include "findlevel.h"
#include "sptrsv_syncfree_serialref.h"

extern "C" {

void ref_exclusive_scan(int *a, int len) { exclusive_scan<int>(a, len); }

void ref_transpose(int m, int n, int nnz, const int *rowPtr, const int *colIdx,
                   const VALUE_TYPE *val, int *outIdx, int *outPtr,
                   VALUE_TYPE *outVal) {
  matrix_transposition(m, n, nnz, rowPtr, colIdx, val, outIdx, outPtr, outVal);
}

int ref_analyser(const int *cscRowIdx, int m, int n, int nnz, int *inDegree) {
  return sptrsv_syncfree_analyser(cscRowIdx, m, n, nnz, inDegree);
}

int ref_executor(const int *cscColPtr, const int *cscRowIdx,
                 const VALUE_TYPE *cscVal, const int *inDegree, int m, int n,
                 int substitution, int rhs, const VALUE_TYPE *b, VALUE_TYPE *x) {
  return sptrsv_syncfree_executor(cscColPtr, cscRowIdx, cscVal, inDegree, m, n,
                                  substitution, rhs, b, x);
}

int ref_findlevel_csc(const int *cscColPtr, const int *cscRowIdx,
                      const VALUE_TYPE *cscVal, int m, int n, int nnz,
                      int *nlevel, int *pmin, int *pavg, int *pmax) {
  return findlevel_csc(cscColPtr, cscRowIdx, cscVal, m, n, nnz, nlevel, pmin,
                       pavg, pmax);
}

}  // extern "C"
