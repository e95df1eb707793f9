// sptrsv_lane.hpp -- the lane engine: one lane per column, LDS windows, and
// value-carrying slots for dependencies between windows.
//
// Same solve semantics as the wavefront-group executor (sptrsv_exec.hpp) and
// the serial executor it restates (sptrsv_syncfree_serialref.h:25-108), with
// a memory protocol shaped for gfx950's counters:
//
//  * A workgroup owns a window of consecutive columns (solve order), claimed
//    through the global ticket in increasing order.  Its LDS holds one partial
//    value per column, initialised with b, from which every contribution
//    L(r, c) * x_c is subtracted, plus one arrival counter per column.
//  * Compute wavefronts poll and update LDS only; their loop never waits on a
//    vector-memory access (gfx9's vmcnt counts stores too, so one outstanding
//    store or load would otherwise stall every poll of the wavefront).
//    Column metadata and the first kPF off-diagonal entries of 64 columns are
//    loaded into registers once per batch.
//  * A contribution to a row of a later window is written into that entry's
//    own slot in HBM.  The slot holds a NaN sentinel while empty, so the
//    written value is its own arrival flag: one relaxed exchange by the
//    producer (an RMW, so it lands where every XCD sees it; its result is
//    unused, so nothing waits for it), one relaxed load by the consumer, no
//    fence and no counter.  Slots are grouped by row (rptr) and summed in slot
//    order.
//  * Wavefront 0 of each workgroup is the receiver: it takes the slots of the
//    window's rows with exchange(slot, EMPTY) -- an RMW, so the poll reads the
//    coherence point rather than a possibly stale line of this XCD's L2, and
//    the slot is reset in the same operation -- then folds complete sums into
//    LDS.  Every solve leaves the slots empty for the next one.
#pragma once

#include "sptrsv_device.hpp"

namespace sptrsv {

template <typename T>
struct SlotBits;
template <>
struct SlotBits<double> {
  using U = unsigned long long;
  static constexpr U kEmpty = 0x7ff4deadbeef5afeull;  // signalling-NaN payload
};
template <>
struct SlotBits<float> {
  using U = unsigned int;
  static constexpr U kEmpty = 0x7fa5afe5u;
};

template <typename T>
__device__ __forceinline__ bool slot_empty(T v) {
  return __builtin_bit_cast(typename SlotBits<T>::U, v) == SlotBits<T>::kEmpty;
}
template <typename T>
__device__ __forceinline__ T slot_empty_value() {
  return __builtin_bit_cast(T, SlotBits<T>::kEmpty);
}

template <typename T>
struct LaneArgs {
  const int *__restrict__ colPtr;
  const T *__restrict__ val;
  const int *__restrict__ etgt;   // [nnz]: local -(lk+1) or slot index
  const int2 *__restrict__ need;  // [n] by row: {local, remote} arrivals
  const int *__restrict__ rptr;   // [n+1] slot range of each row
  const int *__restrict__ winPtr; // [nwin + 1] window bounds, solve order
  const int4 *__restrict__ rlist; // columns with remote inputs, window order:
                                  // {local index, first slot, slot count, -}
  const int *__restrict__ roff;   // [nwin + 1] each window's rlist range
  T *slots;                       // [ntiles][nslots]
  const T *__restrict__ b;        // [n][rhs]
  T *__restrict__ x;              // [n][rhs]
  int n, rhs, nwin, ntiles, maxw, nslots;
  int nrecv;                      // receiver wavefronts per workgroup
  int *ticket;
  int *err;
  long long wait_cap;  // longest single wait, wall-clock ticks
};

constexpr int kPF = 4;  // off-diagonal entries per column held in registers

// Wave-uniform broadcast of lane i's value (v_readlane; i is uniform).
__device__ __forceinline__ int bcast(int v, int i) {
  return __builtin_amdgcn_readlane(v, i);
}
__device__ __forceinline__ double bcast(double v, int i) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)u, i);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), i);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float bcast(float v, int i) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), i));
}

// Ready lanes at which a wavefront stops sweeping one lane at a time and
// processes them all in parallel instead.
constexpr int kParallelReady = 16;
// Lanes a sweep solves between two publications of their results.
constexpr int kPublish = 16;

// Compute wavefront cw handles batches of 64 consecutive columns (lane =
// column) cw, cw + ncw, ...  Per refresh it reads every lane's LDS arrival
// count and partial value once.  If many lanes are ready (independent
// columns) they are solved in parallel and publish through LDS.  Otherwise
// the wavefront sweeps the ready lanes in ascending order: the lowest ready
// lane's x and its first kPF entries are broadcast with v_readlane, and every
// lane that is the target of such an entry inside this batch adds it to
// registers at once, so a dependency chain inside the batch costs a few
// scalar/vector instructions per hop instead of an LDS round trip.  The swept
// lanes publish their remaining entries (LDS for the window, slots beyond it)
// together at the end of the sweep.  Every entry is delivered exactly once:
// by register when it is a sweep-time in-batch entry, otherwise by LDS/slot,
// and a lane is ready when LDS arrivals + register arrivals reach its total.
template <typename T, bool BWD>
__device__ __forceinline__ void lane_compute(const LaneArgs<T> &a, T *s_val,
                                            int *s_cnt, int wlo, int wlen,
                                            int tile, int cw, int ncw) {
  const int lane = __lane_id();
  const int n = a.n;
  T *slots = a.slots + (size_t)tile * a.nslots;
  for (int base = cw * kWave; base < wlen; base += ncw * kWave) {
    const int k = base + lane;
    bool active = k < wlen;
    int lo = 0, hi = 0, total = 0, ne = 0;
    T diag = T(1), inv = T(1);
    int tg[kPF];
    T vv[kPF];
    unsigned inb = 0;  // bit e: entry e targets a column of this batch
    int tl[kPF];       // ... and that column's lane
    if (active) {
      const int p = wlo + k;
      const int c = BWD ? n - 1 - p : p;
      const int f = a.colPtr[c], l = a.colPtr[c + 1];
      diag = a.val[BWD ? l - 1 : f];
      inv = T(1) / diag;  // off the critical path
      lo = BWD ? f : f + 1;
      hi = BWD ? l - 1 : l;
      ne = min(hi - lo, kPF);
      const int2 nd = a.need[c];
      total = nd.x + nd.y;
#pragma unroll
      for (int i = 0; i < kPF; i++) {
        const int j = lo + i;
        tg[i] = j < hi ? a.etgt[j] : 0;
        vv[i] = j < hi ? a.val[j] : T(0);
      }
    }
#pragma unroll
    for (int i = 0; i < kPF; i++) {
      // targets lie after the column, so in-batch means below base + 64
      tl[i] = -tg[i] - 1 - base;
      if (active && i < ne && tg[i] < 0 && tl[i] < kWave) inb |= 1u << i;
    }
    T racc = T(0);  // in-batch contributions received by register
    int rcnt = 0;   // ... and their count
    StallWatch watch;
    int level = 0;
    while (__ballot(active) != 0) {
      const int cnt = active ? ld_wg(&s_cnt[k]) : 0;
      bool rdy = active && cnt + rcnt == total;
      unsigned long long m = __ballot(rdy);
      const bool any = m != 0;
      // nothing ready: short back-off, so idle wavefronts leave LDS bandwidth
      // to the ones that are solving
      backoff_short(level, any);
      if (!any) {
        if (watch.step(false, a.err, a.wait_cap)) active = false;
        continue;
      }
      fence_acq_wg();
      const T val = active ? s_val[k] : T(0);
      bool solved = false;    // solved, not yet published
      bool swept = false;     // ... by the sweep (in-batch entries already sent)
      T xv = T(0);
      // publish the solved lanes: LDS sums, then LDS counters and slots
      auto publish = [&]() {
        const unsigned viaLds = swept ? ~inb : ~0u;
        if (solved) {
          s_val[k] = xv;  // the column is final: its LDS value now holds x
#pragma unroll
          for (int e = 0; e < kPF; e++)
            if (e < ne && ((viaLds >> e) & 1u) && tg[e] < 0)
              add_wg(&s_val[-tg[e] - 1], -(xv * vv[e]));
          for (int j = lo + kPF; j < hi; j++) {
            const int t = a.etgt[j];
            if (t < 0) add_wg(&s_val[-t - 1], -(xv * a.val[j]));
          }
        }
        fence_rel_wg();
        if (solved) {
#pragma unroll
          for (int e = 0; e < kPF; e++) {
            if (e < ne && ((viaLds >> e) & 1u)) {
              if (tg[e] < 0)
                add_wg(&s_cnt[-tg[e] - 1], 1);
              else
                xchg_agent(&slots[tg[e]], xv * vv[e]);
            }
          }
          for (int j = lo + kPF; j < hi; j++) {
            const int t = a.etgt[j];
            if (t < 0)
              add_wg(&s_cnt[-t - 1], 1);
            else
              xchg_agent(&slots[t], xv * a.val[j]);
          }
        }
        solved = false;
      };
      if (__popcll(m) >= kParallelReady) {
        if (rdy) {
          const T q = val - racc;
          xv = q * inv;
          xv = __builtin_fma(__builtin_fma(-xv, diag, q), inv, xv);
          solved = true;
          active = false;
        }
        publish();
      } else {
        int since = 0;
        while (m != 0) {
          const int i = __builtin_ctzll(m);
          // every lane evaluates its quotient; lane i's is final
          const T q = val - racc;
          T xx = q * inv;
          xx = __builtin_fma(__builtin_fma(-xx, diag, q), inv, xx);
          const bool me = lane == i;
          xv = me ? xx : xv;
          solved = solved || me;
          swept = swept || me;
          active = active && !me;
          const T xi = bcast(xx, i);
          const unsigned bi = (unsigned)bcast((int)inb, i);
#pragma unroll
          for (int e = 0; e < kPF; e++) {
            if ((bi >> e) & 1u) {  // uniform
              const int t = bcast(tl[e], i);
              const T ve = bcast(vv[e], i);
              const bool hit = lane == t;  // the target is never solved yet
              racc = hit ? __builtin_fma(xi, ve, racc) : racc;
              rcnt += hit ? 1 : 0;
            }
          }
          rdy = active && cnt + rcnt == total;
          m = __ballot(rdy);
          // hand results to other batches and windows every kPublish lanes
          if (++since == kPublish && m != 0) {
            publish();
            since = 0;
          }
        }
        if (__ballot(solved) != 0) publish();
      }
      if (watch.step(any, a.err, a.wait_cap)) active = false;
    }
  }
}

constexpr int kTake = 8;  // slots of one row taken per group

// Receivers: the first nrecv wavefronts walk the window's receive list (only
// the columns that have inputs from earlier windows, precomputed per window
// partition), lane l of receiver r taking entries r*64 + l, stride nrecv*64.
// A column's slots are taken in groups of up to kTake with
// exchange(slot, EMPTY), which both polls and resets a slot.  Only the first
// slot of a group is polled until it arrives; the rest of the group is taken
// after it, and the group is summed in slot order once complete, so the
// remote sum has a fixed order.  A receiver wavefront that took nothing in an
// iteration backs off.
template <typename T>
__device__ __forceinline__ void lane_receive(const LaneArgs<T> &a, T *s_val,
                                            int *s_cnt, int win, int tile,
                                            int rw) {
  T *slots = a.slots + (size_t)tile * a.nslots;
  const T empty = slot_empty_value<T>();
  const int e_end = a.roff[win + 1];
  const int stride = a.nrecv * kWave;
  int e = a.roff[win] + rw * kWave + __lane_id();
  int k = 0, s = 0, r1 = 0, m = 0;
  unsigned got = 0;
  T part[kTake];
  T sum = T(0);
  int4 next = e < e_end ? a.rlist[e] : make_int4(0, 0, 0, 0);
  auto take_next = [&]() {
    k = next.x;
    s = next.y;
    m = next.z;
    r1 = s + m;
    got = 0;
    sum = T(0);
    if (e + stride < e_end) next = a.rlist[e + stride];  // prefetch
  };
  bool active = e < e_end;
  if (active) take_next();
  StallWatch watch;
  int level = 0;
  while (__ballot(active) != 0) {
    bool done = false;
    bool progressed = false;
    if (active) {
      const int g = min(kTake, r1 - s);
      const unsigned before = got;
#pragma unroll
      for (int q = 0; q < kTake; q++) {
        // slot 0 of the group first; the others once slot 0 is in
        if (q < g && !(got & (1u << q)) && (q == 0 || (got & 1u))) {
          const T v = xchg_take(&slots[s + q], empty);
          if (!slot_empty(v)) {
            part[q] = v;
            got |= 1u << q;
          }
        }
      }
      progressed = got != before;
      if (got == (1u << g) - 1) {
#pragma unroll
        for (int q = 0; q < kTake; q++)
          if (q < g) sum += part[q];
        s += g;
        got = 0;
        if (s == r1) {
          add_wg(&s_val[k], -sum);
          fence_rel_wg();
          add_wg(&s_cnt[k], m);
          done = true;
        }
      }
    }
    if (done) {
      e += stride;
      active = e < e_end;
      if (active) take_next();
    }
    const bool any = __ballot(progressed) != 0;
    backoff(level, any);
    if (watch.step(any, a.err, a.wait_cap)) active = false;
  }
}

template <typename T, bool BWD>
__global__ void __launch_bounds__(1024) k_exec_lane(LaneArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *s_val = reinterpret_cast<T *>(smem_raw);
  int *s_cnt = reinterpret_cast<int *>(s_val + a.maxw);
  __shared__ int s_ticket;
  const int total = a.nwin * a.ntiles;
  const int wave = threadIdx.x / kWave;
  const int ncw = blockDim.x / kWave - a.nrecv;
  const int n = a.n;

  for (;;) {
    if (threadIdx.x == 0)
      s_ticket = ld_agent(a.err) != 0 ? total : atomicAdd(a.ticket, 1);
    __syncthreads();
    const int t = s_ticket;
    if (t >= total) break;
    const int win = t / a.ntiles;
    const int tile = t - win * a.ntiles;
    const int wlo = a.winPtr[win];
    const int wlen = a.winPtr[win + 1] - wlo;
    for (int k = threadIdx.x; k < wlen; k += blockDim.x) {
      const int p = wlo + k;
      const int c = BWD ? n - 1 - p : p;
      s_val[k] = a.b[(size_t)c * a.rhs + tile];
      s_cnt[k] = 0;
    }
    __syncthreads();
    if (wave < a.nrecv)
      lane_receive<T>(a, s_val, s_cnt, win, tile, wave);
    else
      lane_compute<T, BWD>(a, s_val, s_cnt, wlo, wlen, tile, wave - a.nrecv,
                           ncw);
    __syncthreads();
    for (int k = threadIdx.x; k < wlen; k += blockDim.x) {
      const int p = wlo + k;
      const int c = BWD ? n - 1 - p : p;
      a.x[(size_t)c * a.rhs + tile] = s_val[k];
    }
  }
  if (threadIdx.x == 0) {
    const int done = atomicAdd(a.ticket + 1, 1);
    if (done == (int)gridDim.x - 1) {
      xchg_agent(a.ticket, 0);
      xchg_agent(a.ticket + 1, 0);
    }
  }
}

// Slot targets of the off-diagonal entries for a window partition: an entry
// whose row lies in the column's window gets -(row's window offset + 1); any
// other entry gets the next slot of its row (cursor zeroed, rptr = exclusive
// scan of need.y).
__global__ void k_lane_targets(const int *__restrict__ colPtr,
                               const int *__restrict__ rowIdx, int n,
                               int backward, const int *__restrict__ winPtr,
                               int nwin, int uniform_w,
                               const int *__restrict__ rptr, int *cursor,
                               int *etgt) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    const int p = backward ? n - 1 - c : c;
    const int w = uniform_w > 0 ? p / uniform_w
                                : upper_bound_minus1(winPtr, nwin + 1, p);
    const int wlo = winPtr[w], whi = winPtr[w + 1];
    const int f = colPtr[c], l = colPtr[c + 1];
    const int lo = backward ? f : f + 1, hi = backward ? l - 1 : l;
    for (int j = lo; j < hi; j++) {
      const int r = rowIdx[j];
      const int pr = backward ? n - 1 - r : r;
      etgt[j] = pr < whi ? -(pr - wlo) - 1 : rptr[r] + atomicAdd(&cursor[r], 1);
    }
  }
}

__global__ void k_need_remote(const int2 *__restrict__ need, int n, int *out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += gridDim.x * blockDim.x)
    out[r] = need[r].y;
}

// flag[p] = 1 when the column at solve position p has remote inputs.
__global__ void k_remote_flags(const int2 *__restrict__ need, int n,
                               int backward, int *flag) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += gridDim.x * blockDim.x)
    flag[p] = need[backward ? n - 1 - p : p].y > 0 ? 1 : 0;
}

// Receive lists: pos = exclusive scan of the flags (solve order), so each
// window's entries are contiguous and ascending; roff[w] = pos[winPtr[w]].
__global__ void k_receive_list(const int2 *__restrict__ need,
                               const int *__restrict__ rptr,
                               const int *__restrict__ pos,
                               const int *__restrict__ winPtr, int nwin, int n,
                               int backward, int uniform_w, int4 *rlist,
                               int *roff) {
  const int stride = gridDim.x * blockDim.x;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    const int c = backward ? n - 1 - p : p;
    const int m = need[c].y;
    if (m > 0) {
      const int w = uniform_w > 0 ? p / uniform_w
                                  : upper_bound_minus1(winPtr, nwin + 1, p);
      rlist[pos[p]] = make_int4(p - winPtr[w], rptr[c], m, 0);
    }
  }
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w <= nwin; w += stride)
    roff[w] = pos[winPtr[w]];
}

}  // namespace sptrsv
