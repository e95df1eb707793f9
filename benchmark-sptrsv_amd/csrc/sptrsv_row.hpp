// sptrsv_row.hpp -- row engine: pull-based sync-free solve over the CSR of
// the triangle, for factors whose dependencies are wide and shallow (random
// power-law rows) rather than long and local.
//
// The reference executor (SpTRSV_cuda/sptrsv_syncfree_cuda.h:16-128) pushes:
// a solved column scatters L(r,c)*x_c into left_sum[r] with atomics and bumps
// in_degree[r], and column r spins until in_degree[r] says every contribution
// is in.  A row that many columns feed (a power-law hub with hundreds of
// entries) then serialises hundreds of atomics on one address.  Pulling
// instead: the lanes of row r read x_c for each of the row's entries and form
// the dot product themselves, so the only shared writes are the solutions.
// Each x_c is its own arrival flag: x is filled with a signalling-NaN
// sentinel before the launch and every solution is written once with an
// agent-scope exchange; a reader polls x_c with an RMW that leaves it
// unchanged (fetch_or 0) until it is not the sentinel.  No counters, no
// fences, no left_sum, and the sum of a row has a fixed order (per lane in
// entry order, then a fixed butterfly), so the result is deterministic.
//
// Work assignment is static: G lanes per (row, right-hand side) item, items
// in solve order, group g of the grid takes items g, g + ngroups, ... .  An
// item only depends on items of lower index, every group takes its items in
// increasing order and every group is resident (the grid is the resident
// capacity), so the lowest unsolved item always has its inputs and is being
// worked on: the solve always completes.
#pragma once

#include "sptrsv_device.hpp"
#include "sptrsv_lane.hpp"  // SlotBits: the sentinel bit patterns

namespace sptrsv {

template <typename T>
struct RowArgs {
  const int *__restrict__ rowPtr;  // [n+1] CSR of the triangle
  const int *__restrict__ colIdx;  // [nnz] columns ascending in each row
  const T *__restrict__ val;       // [nnz]
  const T *__restrict__ b;         // [n][rhs]
  T *x;                            // [n][rhs], the sentinel until solved
  int n, rhs;
  int *err;
  long long wait_cap;
};

constexpr int kRowU = 4;  // entries per lane per poll round

// Read a location other workgroups write with an exchange, through an RMW
// that leaves it unchanged (performed where every XCD sees the latest value).
__device__ __forceinline__ double rmw_read_val(double *p) {
  const unsigned long long v = __hip_atomic_fetch_or(
      reinterpret_cast<unsigned long long *>(p), 0ull, __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_bit_cast(double, v);
}
__device__ __forceinline__ float rmw_read_val(float *p) {
  const unsigned v = __hip_atomic_fetch_or(reinterpret_cast<unsigned *>(p), 0u,
                                           __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_bit_cast(float, v);
}

// G lanes per item (G a power of two, at most 64).  Forward: the factor is
// lower triangular, the diagonal is the last entry of its row; backward:
// upper triangular, the diagonal is the first entry.  The groups of a
// wavefront advance independently (an item may depend on an item of another
// group of the same wavefront): each lane loads kRowU of its entries, polls
// them until all are in, accumulates, and moves on; a group whose lanes have
// all run out of entries reduces and publishes at once.
template <typename T, bool BWD, int G>
__global__ void __launch_bounds__(1024) k_exec_row(RowArgs<T> a) {
  const int lane = __lane_id();
  const int gl = lane & (G - 1);
  const unsigned long long gmask =
      (G == kWave ? ~0ull : ((1ull << G) - 1)) << (lane & ~(G - 1));
  const int waves_per_block = blockDim.x / kWave;
  const int gwave = blockIdx.x * waves_per_block + threadIdx.x / kWave;
  const long long ngroups = (long long)gridDim.x * waves_per_block * (kWave / G);
  const long long gidx = (long long)gwave * (kWave / G) + lane / G;
  const long long items = (long long)a.n * a.rhs;
  StallWatch watch;
  int level = 0;
  bool stopped = false;
  // wave-uniform rounds: in round m the groups take items m * ngroups + gidx
  for (long long base = 0; base < items && !stopped; base += ngroups) {
    const long long t = base + gidx;
    const bool act = t < items;
    int r = 0, q = 0, j = 0, hi = 0;
    T d = T(1), bv = T(0);
    if (act) {
      const int p = (int)(t / a.rhs);
      q = (int)(t - (long long)p * a.rhs);
      r = BWD ? a.n - 1 - p : p;
      const int s0 = a.rowPtr[r], s1 = a.rowPtr[r + 1];
      d = a.val[BWD ? s0 : s1 - 1];
      bv = a.b[(size_t)r * a.rhs + q];
      j = (BWD ? s0 + 1 : s0) + gl;
      hi = BWD ? s1 : s1 - 1;
    }
    bool fin = !act || j >= hi;  // this lane has no entries left
    bool pub = !act;             // the group's item is published
    bool chunk = false;          // entries loaded, not yet accumulated
    unsigned pend = 0;
    int col[kRowU];
    T v[kRowU], xv[kRowU];
    T acc = T(0);
    while (__ballot(!pub) != 0) {
      if (!fin && !chunk) {
#pragma unroll
        for (int u = 0; u < kRowU; u++) {
          const int jj = j + u * G;
          const bool has = jj < hi;
          col[u] = has ? a.colIdx[jj] : 0;
          v[u] = has ? a.val[jj] : T(0);
          xv[u] = T(0);
          if (has) pend |= 1u << u;
        }
        chunk = true;
      }
      unsigned got = 0;
#pragma unroll
      for (int u = 0; u < kRowU; u++)
        if (pend & (1u << u)) {
          const T y = rmw_read_val(&a.x[(size_t)col[u] * a.rhs + q]);
          if (!slot_empty(y)) {
            xv[u] = y;
            got |= 1u << u;
          }
        }
      pend &= ~got;
      if (chunk && pend == 0) {
#pragma unroll
        for (int u = 0; u < kRowU; u++) acc = __builtin_fma(v[u], xv[u], acc);
        chunk = false;
        j += G * kRowU;
        fin = j >= hi;
      }
      // groups whose lanes all ran out of entries reduce and publish
      const bool gdone = !pub && (__ballot(fin) & gmask) == gmask;
      const bool any_done = __ballot(gdone) != 0;
      if (any_done) {
        T s = acc;  // fixed-order butterfly over the group's lanes
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (gdone) {
          if (gl == 0) {
            const T rr = bv - s;
            const T inv = T(1) / d;
            T xr = rr * inv;
            xr = __builtin_fma(__builtin_fma(-xr, d, rr), inv, xr);
            xchg_agent(&a.x[(size_t)r * a.rhs + q], xr);
          }
          pub = true;
        }
      }
      // a wavefront per row: a short fixed pause (its inputs are usually one
      // publish away; 8% faster on the 10^6 power-law factor); narrower
      // groups, many rows per wavefront: exponential back-off (a fixed pause
      // made the 700^2 grid 30% slower)
      const bool any = any_done || __ballot(got != 0) != 0;
      if constexpr (G == kWave) {
        if (!any) __builtin_amdgcn_s_sleep(1);
      } else {
        backoff(level, any);
      }
      if (watch.step(any, a.err, a.wait_cap)) {
        stopped = true;
        break;
      }
    }
  }
}

// Number of off-diagonal entries at least `far` columns from the diagonal
// (one atomic per workgroup).  The default configuration picks the row
// engine when most entries are that far: windows of consecutive columns
// would then route most entries through HBM slots.
__global__ void k_count_far(const int *__restrict__ colPtr,
                            const int *__restrict__ rowIdx, int n, int far,
                            unsigned long long *out) {
  __shared__ unsigned long long s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  unsigned long long c = 0;
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < n;
       col += gridDim.x * blockDim.x)
    for (int j = colPtr[col]; j < colPtr[col + 1]; j++) {
      const int d = rowIdx[j] - col;
      c += (d >= far || d <= -far) ? 1 : 0;
    }
  atomicAdd(&s_cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, s_cnt);
}

}  // namespace sptrsv
