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
// Work is claimed through a ticket, by workgroups, in the level order of the
// rows (the analysis sorts the solve positions by DAG level; a ticket striped
// over eight counters, as the level sweep claims, let the stripes drift apart
// and made the power-law factor 14% slower): a claim is
// `rounds` rounds of one item per group of the workgroup (item = (row,
// right-hand-side tile), tiles fastest), consecutive items going to
// different groups.  Items that follow each other in level order are mostly
// independent, so a workgroup's rounds rarely wait on one another, and the
// claims are few (one atomic per rounds * groups items).  A claimed item
// reads only items of lower levels, claimed earlier by running workgroups;
// every wavefront finishes its items of a claim in order and a workgroup
// claims again only when all its wavefronts are done, so the lowest
// unfinished item always has its inputs and is being worked on: the solve
// completes with any grid size, whether or not every workgroup is resident
// (the reference's dynamic column ticket, sptrsv_syncfree_cuda.h:128-131,
// plays the same role in row order).
#pragma once

#include "sptrsv_device.hpp"
#include "sptrsv_lane.hpp"  // SlotBits: the sentinel bit patterns

namespace sptrsv {

template <typename T>
struct RowArgs {
  const int *__restrict__ rowPtr;  // [n+1] CSR of the triangle
  const int4 *__restrict__ rows;   // [n] in level order: {row, first entry,
                                   // end of its entries, -}
  const int *__restrict__ colIdx;  // [nnz] columns ascending in each row
  const T *__restrict__ val;       // [nnz]
  const T *__restrict__ b;         // [n][rhs]
  T *x;                            // [n][rhs], the sentinel until solved
  int n, rhs, ntiles;              // items: n * ntiles
  int rounds;                      // rounds per workgroup claim
  int *ticket;                     // [2]: claims, finished workgroups
  int *err;
  long long wait_cap;
  IrFuse ir;                       // refinement fusions (FP32 solves)
};

constexpr int kRowU = 4;      // entries per lane per poll round

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

// Waves per SIMD the compiler must leave room for.  Lanes over right-hand
// sides with two per lane: 8 (at most 64 VGPRs, two spilled, so that two
// 1024-thread workgroups fit a CU; unbounded they took 70 VGPRs, one
// workgroup per CU, and 128 rhs on the 128^3 factor 3.49 -> 3.07 ms).  The
// others are left alone: 32 lanes per row over its entries take 65 VGPRs,
// and bounded to 64 (two workgroups per CU) the 27-point stencil ran 12%
// slower; one right-hand side per lane fits 64 already, four take ~110.
constexpr int row_min_waves(int G, int RL, int RPL) {
  return RL > 1 && RPL == 2 ? 8 : 1;
}

// G lanes per item: RL of them over the tile's right-hand sides (RPL values
// each, tile width RL * RPL), NS = G / RL over the row's entries.  RL = 1 is
// the per-(row, right-hand side) pull of a single solution; RL = G reads whole
// rows of x ([n][rhs] row-major, so a tile is one contiguous segment) and
// passes over the factor once for the whole tile.
//
// Forward: the factor is lower triangular, the diagonal is the last entry of
// its row; backward: upper triangular, the diagonal is the first entry.  The
// groups of a wavefront advance independently (an item may depend on an item
// of another group of the same wavefront): each lane loads kRowU of its
// entries, polls them until all are in, accumulates, and moves on; a group
// whose lanes have all run out of entries reduces and publishes at once.
// Polls and publishes: with one value per item (RL = 1) an RMW that leaves
// the value unchanged and an exchange; with whole rows of x, agent-scope
// atomic loads and stores (coalesced, and a right-hand-side tile of a row is
// hundreds of values: RMWs would serialise on the atomic units).
template <typename T, bool BWD, int G, int RL, int RPL>
__global__ void __launch_bounds__(1024, row_min_waves(G, RL, RPL)) k_exec_row(RowArgs<T> a) {
  constexpr int NS = G / RL;  // lanes over entries
  constexpr bool kBulk = RL > 1;
  constexpr int kPer = kWave / G;
  const int lane = __lane_id();
  const int gl = lane & (G - 1);
  const int lr = gl % RL;   // right-hand-side lane
  const int ls = gl / RL;   // entry lane
  const unsigned long long gmask =
      (G == kWave ? ~0ull : ((1ull << (G % kWave)) - 1)) << (lane & ~(G - 1));
  const long long items = (long long)a.n * a.ntiles;
  const int rhs = a.rhs;
  const int gpb = (blockDim.x / kWave) * kPer;  // groups per workgroup
  const int chunk = a.rounds * gpb;             // items per claim
  const int grp = (threadIdx.x / kWave) * kPer + lane / G;
  __shared__ long long s_claim;
  StallWatch watch;
  int level = 0;
  bool stopped = false;
  if (threadIdx.x == 0) s_claim = atomicAdd(a.ticket, chunk);
  __syncthreads();
  for (;;) {
    const long long base = s_claim;
    __syncthreads();  // every wavefront has read the claim
    if (base >= items) break;
    // the next claim is in flight while this one is worked on (after an
    // error every wavefront's stall watch stops it; the remaining claims
    // then pass through without work)
    if (threadIdx.x == 0) s_claim = stopped ? items : (long long)atomicAdd(a.ticket, chunk);
    for (int rd = 0; rd < a.rounds && !stopped; rd++) {
      const long long t = base + (long long)rd * gpb + grp;
      const bool act = t < items;
      int r = 0, tile = 0, j = 0, hi = 0;
      T d = T(1);
      T bv[RPL];
      bool qv[RPL];
      // refinement fusions (FP32 solves with lanes over right-hand sides; a
      // wavefront per row over its entries ran 11% slower on the power-law
      // factor with them compiled in, used or not).  The last correction's
      // iterate is loaded with b, so the FP64 update adds no wait to the
      // publish.
      constexpr bool kUpd = std::is_same<T, float>::value && RL > 1;
      double xo[kUpd ? RPL : 1];
#pragma unroll
      for (int q = 0; q < RPL; q++) {
        bv[q] = T(0);
        qv[q] = false;
        if constexpr (kUpd) xo[q] = 0.0;
      }
      if (act) {
        const int idx = (int)(t / a.ntiles);
        tile = (int)(t - (long long)idx * a.ntiles);
        const int4 ri = a.rows[idx];
        r = ri.x;
        const int s0 = ri.y, s1 = ri.z;
        d = a.val[BWD ? s0 : s1 - 1];
#pragma unroll
        for (int q = 0; q < RPL; q++) {
          const int k = tile * (RL * RPL) + q * RL + lr;
          qv[q] = k < rhs;
          if (ls == 0 && qv[q]) {
            if constexpr (kUpd) {
              bv[q] = load_b(a.b, a.ir, (size_t)r * rhs + k);
              if (a.ir.xout && a.ir.xold) xo[q] = a.ir.xold[(size_t)r * rhs + k];
            } else {
              bv[q] = a.b[(size_t)r * rhs + k];
            }
          }
        }
        j = (BWD ? s0 + 1 : s0) + ls;
        hi = BWD ? s1 : s1 - 1;
      }
      bool fin = !act || j >= hi;  // this lane has no entries left
      bool pub = !act;             // the group's item is published
      bool chunk = false;          // entries loaded, not yet accumulated
      unsigned pend = 0;           // bit u * RPL + q: value still awaited
      int col[kRowU];
      T v[kRowU], xv[kRowU][RPL];
      T acc[RPL];
#pragma unroll
      for (int q = 0; q < RPL; q++) acc[q] = T(0);
      while (__ballot(!pub) != 0) {
        if (!fin && !chunk) {
#pragma unroll
          for (int u = 0; u < kRowU; u++) {
            const int jj = j + u * NS;
            const bool has = jj < hi;
            col[u] = has ? a.colIdx[jj] : 0;
            v[u] = has ? a.val[jj] : T(0);
#pragma unroll
            for (int q = 0; q < RPL; q++) {
              xv[u][q] = T(0);
              if (has && qv[q]) pend |= 1u << (u * RPL + q);
            }
          }
          chunk = true;
        }
        unsigned got = 0;
#pragma unroll
        for (int u = 0; u < kRowU; u++)
#pragma unroll
          for (int q = 0; q < RPL; q++)
            if (pend & (1u << (u * RPL + q))) {
              T *px = &a.x[(size_t)col[u] * rhs + tile * (RL * RPL) + q * RL + lr];
              const T y = kBulk ? ld_agent(px) : rmw_read_val(px);
              if (!slot_empty(y)) {
                xv[u][q] = y;
                got |= 1u << (u * RPL + q);
              }
            }
        pend &= ~got;
        if (chunk && pend == 0) {
#pragma unroll
          for (int u = 0; u < kRowU; u++)
#pragma unroll
            for (int q = 0; q < RPL; q++) acc[q] = fma_t(v[u], xv[u][q], acc[q]);
          chunk = false;
          j += NS * kRowU;
          fin = j >= hi;
        }
        // groups whose lanes all ran out of entries reduce and publish
        const bool gdone = !pub && (__ballot(fin) & gmask) == gmask;
        const bool any_done = __ballot(gdone) != 0;
        if (any_done) {
          // fixed-order butterfly over the entry lanes (every lane takes
          // part; only finished groups use the sums, the others keep
          // accumulating into acc)
          T sum[RPL];
#pragma unroll
          for (int q = 0; q < RPL; q++) {
            T s = acc[q];
#pragma unroll
            for (int off = G / 2; off >= RL; off >>= 1) s += __shfl_xor(s, off);
            sum[q] = s;
          }
          if (gdone) {
            if (ls == 0) {
              const T inv = T(1) / d;
#pragma unroll
              for (int q = 0; q < RPL; q++)
                if (qv[q]) {
                  const T rr = bv[q] - sum[q];
                  T xr = rr * inv;
                  xr = fma_t(fma_t(-xr, d, rr), inv, xr);
                  const size_t e = (size_t)r * rhs + tile * (RL * RPL) + q * RL + lr;
                  T *px = &a.x[e];
                  if (kBulk)
                    st_agent(px, xr);
                  else
                    xchg_agent(px, xr);
                  if constexpr (kUpd)
                    if (a.ir.xout) a.ir.xout[e] = xo[q] + (double)xr;
                }
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
    __syncthreads();  // the claim is done; s_claim holds the next one
  }
  if (threadIdx.x == 0) {
    const int done = atomicAdd(a.ticket + 1, 1);
    if (done == (int)gridDim.x - 1) {
      xchg_agent(a.ticket, 0);
      xchg_agent(a.ticket + 1, 0);
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

// Largest of v[0..n) into *out (*out starts at 0).  One atomicMax per
// workgroup: atomics on one address serialise (~11 ns each), so a launch of
// 256 workgroups pays ~3 us where one atomic per wavefront paid ~95 us.
__global__ void k_max_int(const int *__restrict__ v, int n,
                          unsigned long long *out) {
  __shared__ int s_max;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    m = max(m, v[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if (__lane_id() == 0) atomicMax(&s_max, m);
  __syncthreads();
  if (threadIdx.x == 0 && s_max > 0) atomicMax(out, (unsigned long long)s_max);
}

}  // namespace sptrsv
