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
// A row's entries are taken four at a time with their loads issued together
// (a loop of one entry at a time waited out each gather in turn: 3.0 ms for
// 128 right-hand sides on the 128^3 factor); the sum keeps the stored order.
__global__ void k_residual_update(const int *__restrict__ rowPtr,
                                  const int *__restrict__ colIdx,
                                  const double *__restrict__ val, int n,
                                  int rhs, const double *__restrict__ b,
                                  const double *__restrict__ x_old,
                                  const float *__restrict__ d,
                                  double *__restrict__ x_new,
                                  float *__restrict__ r32) {
  constexpr int U = 4;
  const size_t total = (size_t)n * rhs;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(t / rhs);
    const int k = (int)(t - (size_t)row * rhs);
    const double bt = b[t];
    const double xt = (x_old ? x_old[t] : 0.0) + (double)d[t];
    const int j0 = rowPtr[row], j1 = rowPtr[row + 1];
    double s = 0.0;
    for (int j = j0; j < j1; j += U) {
      int c[U];
      double v[U], xj[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool h = j + u < j1;
        c[u] = h ? colIdx[j + u] : 0;
        v[u] = h ? val[j + u] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const size_t q = (size_t)c[u] * rhs + k;
        xj[u] = j + u < j1 ? (x_old ? x_old[q] : 0.0) + (double)d[q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (j + u < j1) s = fma(v[u], xj[u], s);
    }
    x_new[t] = xt;
    r32[t] = (float)(bt - s);
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
