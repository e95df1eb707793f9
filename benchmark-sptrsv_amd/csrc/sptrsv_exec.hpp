// sptrsv_exec.hpp -- the sync-free executor for gfx950.
//
// Algorithm (Liu et al. sync-free SpTRSV as in SpTRSV_cuda/sptrsv_syncfree_cuda.h
// :21-107 and :290-407; serial semantics sptrsv_syncfree_serialref.h:25-108):
// a column is solved once every off-diagonal entry of its row has delivered
// its contribution; its x is then (b - left_sum) / diag and x * L(r, col) is
// scattered into left_sum[r] of every later row r, with an arrival counter
// bumped per entry.
//
// gfx950 shape:
//  * Work is claimed in windows of consecutive columns (solve order) through a
//    global ticket, so every claimed window only waits on windows that are
//    already owned by running workgroups (forward progress with any grid size,
//    the role of d_id_extractor at sptrsv_syncfree_cuda.h:128-131).
//  * A window's partial sums and arrival counts live in LDS (the 16-slot shared
//    window of :41-42 / unified_and_shared.h:48-49, generalised to `window`
//    columns); contributions to rows of later windows go to global left_sum /
//    counters with agent-scope atomics.  Which of the two a row waits for is
//    precomputed per window partition (need.x local, need.y global), so a
//    column whose inputs are all local never polls HBM.
//  * G lanes of a 64-wide wavefront cooperate on a column (G = 64 is one
//    wavefront per column as in the reference); RL of them run over right-hand
//    sides, G/RL over the column's nonzeros; each lane holds rpl right-hand
//    sides.  The 64/G column groups of a wavefront make progress independently
//    inside one polling loop, so no lane ever spins on a result its own
//    wavefront has yet to produce.
//  * Global counters are polled with an RMW that leaves them unchanged, and
//    global partial sums are taken with exchange(sum, 0): RMWs are performed
//    at the coherence point, whereas a plain sc1 load may be served by a line
//    resident in this XCD's L2.  Counters and sums are returned to zero by
//    their consumer, so a solve needs no memset and leaves a clean state.
#pragma once

#include "sptrsv_device.hpp"

namespace sptrsv {

template <typename T>
struct ExecArgs {
  const int *__restrict__ colPtr;
  const int *__restrict__ rowIdx;
  const T *__restrict__ val;
  const int2 *__restrict__ need;  // [n] by column: {local, global} arrivals
  const int *__restrict__ winPtr; // [nwin + 1] window bounds, solve order
  T *gsum;                        // [ntiles][n][R] global partial sums
  int *gcnt;                      // [ntiles][n] global arrival counters
  const T *__restrict__ b;        // [n][rhs]
  T *__restrict__ x;              // [n][rhs]
  int n, rhs, nwin, ntiles, maxw, rpl;
  int *ticket;                    // [2]: next ticket, finished workgroups
  int *err;                       // set to 1 on a spin-wait timeout
  long long wait_cap;              // longest single wait, wall-clock ticks
};

constexpr int kPrefetch = 4;  // off-diagonal entries per lane loaded early

template <typename T, int G, int RL, int MAXRPL, bool BWD>
__device__ __forceinline__ void exec_window(const ExecArgs<T> &a, T *s_sum,
                                            int *s_cnt, int wlo, int wlen,
                                            int tile) {
  constexpr int NS = G / RL;  // lanes over nonzeros
  const int R = RL * a.rpl;   // right-hand sides per tile
  const int n = a.n;
  const int tid = threadIdx.x;
  const int ng = blockDim.x / G;
  const int gl = tid & (G - 1);
  const int lr = gl % RL;
  const int ls = gl / RL;
  const int leader = __lane_id() & ~(G - 1);
  const size_t tile_rows = (size_t)tile * n;

  int k = tid / G;
  bool active = k < wlen;
  int c = 0, lo = 0, hi = 0;
  int2 nd = make_int2(0, 0);
  T diag = T(1);
  T bv[MAXRPL];
  int pf_r[kPrefetch];
  T pf_v[kPrefetch];

  auto load_column = [&](int kk) {
    const int p = wlo + kk;
    c = BWD ? n - 1 - p : p;
    const int f = a.colPtr[c], l = a.colPtr[c + 1];
    diag = a.val[BWD ? l - 1 : f];
    lo = BWD ? f : f + 1;
    hi = BWD ? l - 1 : l;
    nd = a.need[c];
#pragma unroll
    for (int q = 0; q < MAXRPL; q++) {
      const int ridx = tile * R + q * RL + lr;
      bv[q] = (q < a.rpl && ridx < a.rhs) ? a.b[(size_t)c * a.rhs + ridx]
                                          : T(0);
    }
#pragma unroll
    for (int i = 0; i < kPrefetch; i++) {
      const int j = lo + ls + i * NS;
      pf_r[i] = j < hi ? a.rowIdx[j] : 0;
      pf_v[i] = j < hi ? a.val[j] : T(0);
    }
  };
  if (active) load_column(k);

  StallWatch watch;
  int level = 0;
  for (;;) {
    if (__ballot(active) == 0) break;
    bool ready = false;
    if (active) {
      ready = ld_wg(&s_cnt[k]) == nd.x;
      if (ready && nd.y != 0) ready = rmw_read(&a.gcnt[tile_rows + c]) == nd.y;
      if (G > 1) ready = __shfl(ready ? 1 : 0, leader) != 0;
    }
    if (ready) {
      const bool any_global_in = __ballot(nd.y != 0) != 0;
      if (any_global_in)
        fence_acq_agent();
      else
        fence_acq_wg();
      T xv[MAXRPL];
#pragma unroll
      for (int q = 0; q < MAXRPL; q++) {
        xv[q] = T(0);
        if (q < a.rpl) {
          const int col_in_tile = q * RL + lr;
          const int ridx = tile * R + col_in_tile;
          T left = s_sum[k * R + col_in_tile];
          if (nd.y != 0) {
            // slice 0 reads and clears the global sum; other slices get it
            // by shuffle, so the clear can never overtake a read
            T g = T(0);
            if (ls == 0)  // read and clear in one RMW at the coherence point
              g = xchg_take(&a.gsum[(tile_rows + c) * R + col_in_tile], T(0));
            if (NS > 1) g = __shfl(g, leader + lr);
            left = g + left;
          }
          xv[q] = (bv[q] - left) / diag;
          if (ls == 0 && ridx < a.rhs) a.x[(size_t)c * a.rhs + ridx] = xv[q];
        }
      }
      if (gl == 0 && nd.y != 0) xchg_agent(&a.gcnt[tile_rows + c], 0);

      // partial sums: LDS for rows in this window, HBM for later windows
      bool remote = false;
      auto add_entry = [&](int r, T v) {
        const int pr = BWD ? n - 1 - r : r;
        const int lk = pr - wlo;
        if (lk < wlen) {
#pragma unroll
          for (int q = 0; q < MAXRPL; q++)
            if (q < a.rpl && tile * R + q * RL + lr < a.rhs)
              add_wg(&s_sum[lk * R + q * RL + lr], xv[q] * v);
        } else {
          remote = true;
#pragma unroll
          for (int q = 0; q < MAXRPL; q++)
            if (q < a.rpl && tile * R + q * RL + lr < a.rhs)
              add_agent(&a.gsum[(tile_rows + r) * R + q * RL + lr], xv[q] * v);
        }
      };
#pragma unroll
      for (int i = 0; i < kPrefetch; i++)
        if (lo + ls + i * NS < hi) add_entry(pf_r[i], pf_v[i]);
      for (int j = lo + ls + kPrefetch * NS; j < hi; j += NS)
        add_entry(a.rowIdx[j], a.val[j]);

      // publish: sums before counters
      if (__ballot(remote) != 0)
        fence_rel_agent();
      else
        fence_rel_wg();
      if (lr == 0) {
        auto count_entry = [&](int r) {
          const int pr = BWD ? n - 1 - r : r;
          const int lk = pr - wlo;
          if (lk < wlen)
            add_wg(&s_cnt[lk], 1);
          else
            add_agent(&a.gcnt[tile_rows + r], 1);
        };
#pragma unroll
        for (int i = 0; i < kPrefetch; i++)
          if (lo + ls + i * NS < hi) count_entry(pf_r[i]);
        for (int j = lo + ls + kPrefetch * NS; j < hi; j += NS)
          count_entry(a.rowIdx[j]);
      }

      k += ng;
      active = k < wlen;
      if (active) load_column(k);
    }
    // back off only while this wavefront waits on HBM counters
    const bool any = __ballot(ready) != 0;
    const bool waits_global = __ballot(active && !ready && nd.y != 0) != 0;
    backoff(level, any || !waits_global);
    if (watch.step(any, a.err, a.wait_cap)) active = false;
  }
}

template <typename T, int G, int RL, int MAXRPL, bool BWD>
__global__ void __launch_bounds__(1024) k_exec(ExecArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = RL * a.rpl;
  T *s_sum = reinterpret_cast<T *>(smem_raw);
  int *s_cnt = reinterpret_cast<int *>(s_sum + (size_t)a.maxw * R);
  __shared__ int s_ticket;
  const int total = a.nwin * a.ntiles;

  for (;;) {
    if (threadIdx.x == 0) {
      s_ticket = ld_agent(a.err) != 0 ? total : atomicAdd(a.ticket, 1);
    }
    __syncthreads();
    const int t = s_ticket;
    if (t >= total) break;
    const int win = t / a.ntiles;
    const int tile = t - win * a.ntiles;
    const int wlo = a.winPtr[win];
    const int wlen = a.winPtr[win + 1] - wlo;
    for (int i = threadIdx.x; i < wlen * R; i += blockDim.x) s_sum[i] = T(0);
    for (int i = threadIdx.x; i < wlen; i += blockDim.x) s_cnt[i] = 0;
    __syncthreads();
    exec_window<T, G, RL, MAXRPL, BWD>(a, s_sum, s_cnt, wlo, wlen, tile);
    __syncthreads();
  }
  // The last workgroup out rewinds the ticket for the next solve.
  if (threadIdx.x == 0) {
    const int done = atomicAdd(a.ticket + 1, 1);
    if (done == (int)gridDim.x - 1) {
      xchg_agent(a.ticket, 0);
      xchg_agent(a.ticket + 1, 0);
    }
  }
}

}  // namespace sptrsv
