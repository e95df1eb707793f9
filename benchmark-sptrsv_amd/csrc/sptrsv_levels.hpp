// sptrsv_levels.hpp -- level analysis of the dependency DAG on the device.
//
// The level of a row is the length of the longest dependency path that ends
// at it: 0 for a row whose only entry is the diagonal, otherwise one more
// than the largest level among the rows it reads.  That is the level set of
// findlevel_csc (SpTRSV_cuda/findlevel.h:6-69: a Kahn sweep from the rows of
// in-degree 1, each row entering the level after its last predecessor's), and
// it is itself a sync-free triangular solve in the (max, +) semiring, so it
// is computed the way the row engine solves: a pull over the CSR where each
// level is written once with an exchange and is its own arrival flag (-1
// while unset).
//
// The same kernel gives the levels of a coarser DAG whose nodes ("units") are
// runs of consecutive solve positions -- the lane engine's batches: a unit's
// level is 0 when no row in it reads a column of another unit, else one more
// than the largest level among the units it reads.
//
// Units are claimed through a striped ticket in solve order, by wavefronts
// (claim_chunk: eight counters, so the claim atomics do not queue on one
// address): a claim is `rounds` rounds of 64 / G units (G lanes per unit).  A claimed unit reads
// only lower units, claimed earlier by running wavefronts, and a wavefront
// finishes its units in claim order, so the lowest unfinished unit always
// has its inputs and the sweep completes with any grid size.
#pragma once

#include "sptrsv_device.hpp"

namespace sptrsv {

struct UnitLevelArgs {
  const int *__restrict__ rowPtr;  // [n+1] CSR of the factor
  const int *__restrict__ colIdx;  // [nnz]
  const int *__restrict__ ufirst;  // [nunits+1] first position of each unit,
                                   // or null: every position is a unit
  const int *__restrict__ uid;     // [n] unit of each solve position, or null
  int *lev;                        // [nunits], -1 until set
  int n, nunits, backward;
  int runs;                        // units are chain runs: level of the last row
  int rounds;                      // rounds of 64 / G units per claim
  int *ticket;                     // [2]: claims, finished workgroups
  int *err;
  long long wait_cap;
};

constexpr int kLevelU = 4;  // entries per lane per poll round

template <int G>
__global__ void __launch_bounds__(256) k_unit_levels(UnitLevelArgs a) {
  constexpr int kPer = kWave / G;  // units per wavefront round
  const int lane = __lane_id();
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const int chunk = a.rounds * kPer;  // units per claim
  const unsigned long long gmask =
      (G == kWave ? ~0ull : ((1ull << (G % kWave)) - 1)) << (lane & ~(G - 1));
  const int n = a.n;
  // a wavefront's claims come from its stripe of the ticket
  const int wpb = blockDim.x / kWave;
  const int nstripes = min(kStripes, (int)gridDim.x * wpb);
  const int stripe = (int)((blockIdx.x * wpb + threadIdx.x / kWave) % nstripes);
  StallWatch watch;
  int level = 0;
  bool stopped = false;
  for (;;) {
    // no claim ahead: units in solve order form chains, and a unit held
    // while its wavefront works on earlier ones would stall its dependents
    long long claim = 0;
    if (lane == 0)
      claim = stopped ? a.nunits : claim_chunk(a.ticket, stripe, nstripes) * chunk;
    const long long base = __shfl(claim, 0);
    if (base >= a.nunits) break;
    for (int rd = 0; rd < a.rounds && !stopped; rd++) {
      const int u = (int)base + rd * kPer + grp;
      const bool act = u < a.nunits;
      int f = 0, j = 0, hi = 0;
      if (act) {
        f = a.ufirst ? a.ufirst[u] : u;
        const int l = a.ufirst ? a.ufirst[u + 1] : u + 1;
        const int r0 = a.backward ? n - l : f, r1 = a.backward ? n - f : l;
        j = a.rowPtr[r0] + gl;
        hi = a.rowPtr[r1];
      }
      bool fin = !act || j >= hi;
      bool pub = !act;
      bool chunk_in = false;
      unsigned pend = 0;
      int dep[kLevelU];
      int m = 0;  // largest level read + 1
      while (__ballot(!pub) != 0) {
        if (!fin && !chunk_in) {
#pragma unroll
          for (int q = 0; q < kLevelU; q++) {
            const int jj = j + q * G;
            dep[q] = -1;
            if (jj < hi) {
              const int c = a.colIdx[jj];
              const int pc = a.backward ? n - 1 - c : c;
              if (pc < f) {  // a column of an earlier unit
                dep[q] = a.uid ? a.uid[pc] : pc;
                pend |= 1u << q;
              }
            }
          }
          chunk_in = true;
        }
        unsigned got = 0;
#pragma unroll
        for (int q = 0; q < kLevelU; q++)
          if (pend & (1u << q)) {
            const int lv = rmw_read(&a.lev[dep[q]]);
            if (lv >= 0) {
              m = max(m, lv + 1);
              got |= 1u << q;
            }
          }
        pend &= ~got;
        if (chunk_in && pend == 0) {
          chunk_in = false;
          j += G * kLevelU;
          fin = j >= hi;
        }
        const bool gdone = !pub && (__ballot(fin) & gmask) == gmask;
        const bool any_done = __ballot(gdone) != 0;
        if (any_done) {
          int mm = m;
#pragma unroll
          for (int off = G / 2; off > 0; off >>= 1) mm = max(mm, __shfl_xor(mm, off));
          if (gdone) {
            // a run unit's level is its last row's
            const int ul = a.runs && a.ufirst ? a.ufirst[u + 1] - f - 1 : 0;
            if (gl == 0) xchg_agent(&a.lev[u], mm + ul);
            pub = true;
          }
        }
        const bool any = any_done || __ballot(got != 0) != 0;
        backoff(level, any);
        if (watch.step(any, a.err, a.wait_cap)) {
          stopped = true;
          break;
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int done = atomicAdd(a.ticket + 1, 1);
    if (done == (int)gridDim.x - 1) {
      reset_stripes(a.ticket);
      xchg_agent(a.ticket + 1, 0);
    }
  }
}

// Runs of a chain, for the row-level sweep: position p continues the run of
// p - 1 when the row at p reads exactly one other column, the one at p - 1,
// and that column is read by no other row.  The rows inside a run are then
// seen by no row outside it, and the run's rows have consecutive levels, so
// the sweep treats a run as one unit (its level: the level of its last row,
// len - 1 more than its first row's) and k_level_expand restores the rows'
// levels.  start[p] = 1 where a unit starts.
__global__ void k_level_runs(const int *__restrict__ colPtr,
                             const int *__restrict__ rowIdx,
                             const int *__restrict__ inDeg, int n, int backward,
                             int *start) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += gridDim.x * blockDim.x) {
    int v = 1;
    if (p > 0) {
      const int c = backward ? n - p : p - 1;     // column at position p - 1
      const int want = backward ? n - 1 - p : p;  // row at position p
      const int f = colPtr[c], l = colPtr[c + 1];
      const int j = backward ? f : f + 1;         // its one off-diagonal entry
      if (l - f == 2 && rowIdx[j] == want && inDeg[want] == 2) v = 0;
    }
    start[p] = v;
  }
}

// Unit of each position and first position of each unit, from the
// inclusive scan of the start flags (ufirst[nunits] = n set by the caller).
__global__ void k_level_units(const int *__restrict__ start,
                              const int *__restrict__ incl, int n, int *uid,
                              int *ufirst) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += gridDim.x * blockDim.x) {
    const int u = incl[p] - 1;
    uid[p] = u;
    if (start[p]) ufirst[u] = p;
  }
}

// Row levels from unit levels: a unit's level is its last row's.
__global__ void k_level_expand(const int *__restrict__ ulev,
                               const int *__restrict__ uid,
                               const int *__restrict__ ufirst, int n, int *lev) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += gridDim.x * blockDim.x) {
    const int u = uid[p];
    lev[p] = ulev[u] - (ufirst[u + 1] - 1 - p);
  }
}

// Level histogram, privatised per workgroup: a workgroup takes a chunk of
// kLvlChunk consecutive positions, whose levels lie close together in solve
// order (a 128^3 grid's chunk spans ~160 levels), counts them in LDS bins
// relative to the chunk's smallest level and touches each global bin once.
// Levels past the LDS bins fall back to global atomics.  One atomic per
// position on a few hundred global bins serialised the 2M positions of a
// 128^3 grid (0.39 ms).
constexpr int kLvlChunk = 4096;
constexpr int kLvlBins = 2048;

__device__ __forceinline__ int chunk_min_level(const int *lev, int i0, int i1,
                                               int *s_red) {
  int mn = 0x7fffffff;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) mn = min(mn, lev[i]);
  for (int off = kWave / 2; off > 0; off >>= 1) mn = min(mn, __shfl_xor(mn, off));
  if (threadIdx.x == 0) *s_red = 0x7fffffff;
  __syncthreads();
  if (__lane_id() == 0) atomicMin(s_red, mn);
  __syncthreads();
  return *s_red;
}

// Histogram of the levels and the largest level (cnt zeroed, *maxlev = -1).
__global__ void __launch_bounds__(256) k_level_histogram(const int *__restrict__ lev,
                                                         int len, int *cnt,
                                                         int *maxlev) {
  __shared__ int s_bin[kLvlBins];
  __shared__ int s_red;
  const int i0 = blockIdx.x * kLvlChunk, i1 = min(len, i0 + kLvlChunk);
  for (int d = threadIdx.x; d < kLvlBins; d += blockDim.x) s_bin[d] = 0;
  const int lo = chunk_min_level(lev, i0, i1, &s_red);  // synchronises
  int mx = -1;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int v = lev[i];
    mx = max(mx, v);
    if (v - lo < kLvlBins)
      atomicAdd(&s_bin[v - lo], 1);
    else
      atomicAdd(&cnt[v], 1);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
  if (__lane_id() == 0) atomicMax(maxlev, mx);
  __syncthreads();
  for (int d = threadIdx.x; d < kLvlBins; d += blockDim.x)
    if (s_bin[d] > 0) atomicAdd(&cnt[lo + d], s_bin[d]);
}

// Level order: off[] holds each level's first place (exclusive scan of the
// widths) and is advanced as places are taken.  One atomic per position,
// unlike the histogram: places taken in about position order keep the rows
// of a level in ascending order, which the row engine's claims read
// faster (per-workgroup reservations, which hand the chunks' places out in
// whatever order the workgroups reach them, made the 10^6 power-law solve
// 7% slower).
__global__ void k_level_scatter(const int *__restrict__ lev, int len, int *off,
                                int *order) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += gridDim.x * blockDim.x)
    order[atomicAdd(&off[lev[i]], 1)] = i;
}

// Row table of the level order: {row, its first entry, the end of its
// entries} for each place, so a claimed item needs one load to find its row.
__global__ void k_level_rows(const int *__restrict__ order,
                             const int *__restrict__ rowPtr, int n, int backward,
                             int4 *rows) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x) {
    const int p = order[k];
    const int r = backward ? n - 1 - p : p;
    rows[k] = make_int4(r, rowPtr[r], rowPtr[r + 1], 0);
  }
}

}  // namespace sptrsv
