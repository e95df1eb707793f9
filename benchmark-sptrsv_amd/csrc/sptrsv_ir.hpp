// sptrsv_ir.hpp -- kernels of the mixed-precision form: FP32 sync-free solves
// with the residual r = b - L x and the update x += d kept in FP64.
//
// The residual is a row-parallel CSR SpMV (each row sums its own entries in
// stored order, so it is deterministic, unlike the CSC scatter of
// SpTRSV_cuda/main.cu:347-358).  One kernel forms x_new = x_old + d, the FP64
// residual of x_new and its FP32 copy for the next correction solve.
#pragma once

#include "sptrsv_device.hpp"

namespace sptrsv {

__global__ void k_to_f32(const double *__restrict__ in, float *__restrict__ out,
                         size_t len) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// One thread per (row, rhs) pair, rhs fastest.  x_old may be null (= 0).
__global__ void k_residual_update(const int *__restrict__ rowPtr,
                                  const int *__restrict__ colIdx,
                                  const double *__restrict__ val, int n,
                                  int rhs, const double *__restrict__ b,
                                  const double *__restrict__ x_old,
                                  const float *__restrict__ d,
                                  double *__restrict__ x_new,
                                  float *__restrict__ r32) {
  const size_t total = (size_t)n * rhs;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(t / rhs);
    const int k = (int)(t - (size_t)row * rhs);
    double s = 0.0;
    for (int j = rowPtr[row]; j < rowPtr[row + 1]; j++) {
      const size_t q = (size_t)colIdx[j] * rhs + k;
      const double xj = (x_old ? x_old[q] : 0.0) + (double)d[q];
      s = fma(val[j], xj, s);
    }
    x_new[t] = (x_old ? x_old[t] : 0.0) + (double)d[t];
    r32[t] = (float)(b[t] - s);
  }
}

__global__ void k_update(const double *__restrict__ x_old,
                         const float *__restrict__ d, double *__restrict__ x,
                         size_t len) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = (x_old ? x_old[i] : 0.0) + (double)d[i];
}

}  // namespace sptrsv
