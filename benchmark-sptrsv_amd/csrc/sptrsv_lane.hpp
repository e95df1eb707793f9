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
  const int2 *__restrict__ chain; // [n] by row: {in-batch predecessor lane
                                  // or -1, its entry index}
  const int *__restrict__ rptr;   // [n+1] slot range of each row
  const int *__restrict__ winPtr; // [nwin + 1] window bounds, solve order
  const int *__restrict__ bstart; // batch starts, window-local, by window
  const int *__restrict__ boff;   // [nwin + 1] each window's bstart range
  const int *__restrict__ bord;   // by window: its batches in processing
                                  // order (level in the batch DAG, index)
  const int4 *__restrict__ rlist; // columns with remote inputs, window order:
                                  // {local index, first slot, slot count, -}
  const int *__restrict__ roff;   // [nwin + 1] each window's rlist range
  const int *__restrict__ worder; // claim order of the windows, or null
                                  // (in order); a window's predecessors
                                  // always come earlier
  T *slots;                       // [ntiles][nslots][R]
  const T *__restrict__ b;        // [n][rhs]
  T *__restrict__ x;              // [n][rhs]
  int n, rhs, nwin, ntiles, maxw, nslots;
  int maxnb;                      // most batches in one window
  int nrecv;                      // receiver wavefronts per workgroup
  int *ticket;
  int *err;
  long long wait_cap;  // longest single wait, wall-clock ticks
  IrFuse ir;           // refinement fusions (FP32 solves)
};

constexpr int kPF = 4;  // off-diagonal entries per column held in registers
// entries past the registers loaded per group when a column publishes
constexpr int kLong = 4;

// Cross-lane move with DPP (no LDS): lanes whose source is outside the
// row or whose row is masked off get `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_t(double old, double v) {
  const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
  const unsigned long long s = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)dpp_i<CTRL, ROWMASK>((int)(unsigned)o, (int)(unsigned)s);
  const unsigned hi =
      (unsigned)dpp_i<CTRL, ROWMASK>((int)(unsigned)(o >> 32), (int)(unsigned)(s >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_t(float old, float v) {
  return __builtin_bit_cast(
      float, dpp_i<CTRL, ROWMASK>(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v)));
}

// One step of an inclusive wavefront scan of affine maps x -> a + c * x_prev
// for R right-hand sides sharing c: the element from the earlier lane
// (a1, c1) composes into this one as (a + c a1, c c1).  The identity (0, 1)
// fills lanes without a source.
template <typename T, int R, int CTRL, int ROWMASK>
__device__ __forceinline__ void scan_step(T (&a)[R], T &c) {
  T a1[R];
#pragma unroll
  for (int q = 0; q < R; q++) a1[q] = dpp_t<CTRL, ROWMASK>(T(0), a[q]);
  const T c1 = dpp_t<CTRL, ROWMASK>(T(1), c);
#pragma unroll
  for (int q = 0; q < R; q++) a[q] = fma_t(c, a1[q], a[q]);
  c = c * c1;
}

// A scan step whose coefficient c (this lane's composed coefficient before
// the step) was computed ahead.
template <typename T, int R, int CTRL, int ROWMASK>
__device__ __forceinline__ void scan_step_pc(T (&a)[R], T c) {
#pragma unroll
  for (int q = 0; q < R; q++) a[q] = fma_t(c, dpp_t<CTRL, ROWMASK>(T(0), a[q]), a[q]);
}

// Inclusive scan over the 64 lanes: row_shr 1, 2, 4, 8 inside each row of
// 16, then row_bcast:15 into rows 1 and 3, then row_bcast:31 into rows 2
// and 3 (the gfx9 DPP controls 0x111..0x118, 0x142, 0x143).
template <typename T, int R>
__device__ __forceinline__ void chain_scan(T (&a)[R], T &c) {
  scan_step<T, R, 0x111, 0xf>(a, c);
  scan_step<T, R, 0x112, 0xf>(a, c);
  scan_step<T, R, 0x114, 0xf>(a, c);
  scan_step<T, R, 0x118, 0xf>(a, c);
  scan_step<T, R, 0x142, 0xa>(a, c);
  scan_step<T, R, 0x143, 0xc>(a, c);
}

// Readiness along lane -> lane + 1 links: lane i can be solved when no lane
// of its segment (from the nearest lane at or below i that has no link, up
// to i) is unready.  linked: bit i set when lane i takes lane i - 1's value;
// ready: bit i set when lane i has its inputs (or is already solved).
__device__ __forceinline__ bool chain_ready(unsigned long long linked,
                                            unsigned long long ready,
                                            int lane) {
  const unsigned long long upto =
      lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
  const unsigned long long roots = ~linked & upto;  // bit 0 is always set
  const int s = 63 - __clzll((long long)roots);
  const unsigned long long seg = upto & ~((1ull << s) - 1);
  return (~ready & seg) == 0;
}

// An entry delivered inside its wavefront instead of through LDS or a slot
// (etgt value): the nearest earlier column of the same 64-column batch that
// row r depends on, when |L(r, c)| <= |L(r, r)|.
constexpr int kRegTarget = -2147483647 - 1;

// A batch that ends before a column starting a new chain holds at least this
// many columns (so a window has at most window / kMinChainBatch + 1 batches).
constexpr int kMinChainBatch = 16;

// LDS count reads per slot exchange while a direct row's exchange is in
// flight: ~16 reads span the exchange's round trip (8 leave part of it
// unused; 32-40 delay the next exchange).
constexpr int kDirectSpin = 16;

// LDS arrival counters of `rows` rows: one int each, or, compact (CC), one
// byte each, four to an int -- a row's count never exceeds its off-diagonal
// entries, so a byte holds it when no row has more than 255 entries.
// Compact counters let a window of one right-hand side hold 16,384 FP64
// rows in LDS (a whole plane of a 128^3 grid) instead of 13,600.
__host__ __device__ constexpr size_t cnt_ints(size_t rows, bool cc) {
  return cc ? (rows + 3) / 4 : rows;
}
template <bool CC>
__device__ __forceinline__ int cnt_load(const int *c, int k) {
  if constexpr (CC)
    return (int)(((unsigned)ld_wg(&c[k >> 2]) >> (8 * (k & 3))) & 0xffu);
  else
    return ld_wg(&c[k]);
}
template <bool CC>
__device__ __forceinline__ void cnt_add(int *c, int k, int v) {
  if constexpr (CC)
    add_wg(&c[k >> 2], (int)((unsigned)v << (8 * (k & 3))));
  else
    add_wg(&c[k], v);
}

// LDS of a lane-engine workgroup: window rows plus one dummy row per lane
// (R values and a counter each), then the window's batch starts and its
// batch processing order.
__host__ __device__ inline size_t lane_lds_bytes(size_t window, int R,
                                                 size_t tsize, bool cc = false) {
  return (window + kWave) * (size_t)R * tsize +
         cnt_ints(window + kWave, cc) * sizeof(int) +
         2 * sizeof(int) * (window / kMinChainBatch + 2);
}

// Most columns per lane in a wide batch: a run of pure links (see k_links)
// longer than a wavefront is one batch of up to kWave * kWideC columns.
constexpr int kWideC = 32;

// r / d from r * (1 / d) and one correction step.
template <typename T>
__device__ __forceinline__ T wide_quot(T r, T d, T inv) {
  const T qv = r * inv;
  return fma_t(fma_t(-qv, d, r), inv, qv);
}

// The last column of a wide batch publishes right-hand side q of its x:
// slots first, then LDS values (the counters follow a release).
template <typename T, int R>
__device__ __forceinline__ void wide_publish(const LaneArgs<T> &a, T *s_val,
                                             T *slots, const int (&tg)[kPF],
                                             const T (&vv)[kPF], int lo, int hi,
                                             T x, int q) {
#pragma unroll
  for (int i = 0; i < kPF; i++) {
    if (tg[i] >= 0)
      xchg_agent(&slots[(size_t)tg[i] * R + q], x * vv[i]);
    else if (tg[i] != kRegTarget)
      add_wg(&s_val[(-tg[i] - 1) * R + q], -(x * vv[i]));
  }
  for (int j = lo + kPF; j < hi; j++) {
    const int tj = a.etgt[j];
    const T v = a.val[j];
    if (tj >= 0)
      xchg_agent(&slots[(size_t)tj * R + q], x * v);
    else if (tj != kRegTarget)
      add_wg(&s_val[(-tj - 1) * R + q], -(x * v));
  }
}

// ... and then counts its LDS arrivals.
template <typename T, bool CC>
__device__ __forceinline__ void wide_count(const LaneArgs<T> &a, int *s_cnt,
                                           const int (&tg)[kPF], int lo, int hi) {
#pragma unroll
  for (int i = 0; i < kPF; i++)
    if (tg[i] < 0 && tg[i] != kRegTarget) cnt_add<CC>(s_cnt, -tg[i] - 1, 1);
  for (int j = lo + kPF; j < hi; j++) {
    const int tj = a.etgt[j];
    if (tj < 0 && tj != kRegTarget) cnt_add<CC>(s_cnt, -tj - 1, 1);
  }
}

// Wide batch [base, bend) (window-local, more than 64 columns): every column
// but the last has a single off-diagonal entry, into the next column, which
// has no other input, so
//     x_k = r_k / d_k + c_k * x_{k-1},   c_k = -L(k, k-1) / d_k,  |c_k| <= 1,
// with r_k = b_k for every column but the first (whose LDS arrivals are the
// batch's only wait).  Lane l holds C = ceil(len / 64) consecutive columns.
// Before the wait each lane replaces its columns' b by b / d in LDS and
// composes its maps, x_last = A + K * x_in (lane 0 leaves its first column
// out: A = A_rest + P * a_0, P the product of its other coefficients); after
// the wait lane 0 adds a_0, the wavefront scans the compositions with DPP,
// the batch's last column (the last of lane `owner`) publishes at once, and
// then each lane expands its columns from its predecessor lane's last x.
// Returns false when the stall watch fired.
template <typename T, bool BWD, int R, bool DIRECT, bool CC>
__device__ __forceinline__ bool lane_wide(const LaneArgs<T> &a, T *s_val,
                                          int *s_cnt, int base, int bend,
                                          int wlo, int tile, int rv) {
  const int lane = __lane_id();
  const int n = a.n;
  const int len = bend - base;
  const int C = (len + kWave - 1) / kWave;
  const int k0 = base + lane * C;
  const int mine = max(0, min(C, bend - k0));
  T d0 = T(1), iv0 = T(1);           // the batch's first column (lane 0)
  T A[R];
#pragma unroll
  for (int q = 0; q < R; q++) A[q] = T(0);
  T K = T(1), P = T(1);
  // coefficient of the previous column's x in column k (0 for the batch's
  // first column); loaded again for the expansion after the publish rather
  // than kept in kWideC registers per lane
  auto coef = [&](int k, T *d_out, T *inv_out) {
    const int p = wlo + k;
    const int c = BWD ? n - 1 - p : p;
    const T d = a.val[BWD ? a.colPtr[c + 1] - 1 : a.colPtr[c]];
    const T inv = T(1) / d;
    *d_out = d;
    *inv_out = inv;
    if (k == base) return T(0);
    const int cp = BWD ? n - p : p - 1;  // the pure link from the column at p - 1
    const T lv = a.val[BWD ? a.colPtr[cp + 1] - 2 : a.colPtr[cp] + 1];
    return -wide_quot(lv, d, inv);
  };
  for (int m = 0; m < mine; m++) {
    T d, inv;
    const T cm = coef(k0 + m, &d, &inv);
    if (k0 + m > base) {
#pragma unroll
      for (int q = 0; q < R; q++) {
        const T qv = wide_quot(s_val[(k0 + m) * R + q], d, inv);
        s_val[(k0 + m) * R + q] = qv;
        A[q] = fma_t(cm, A[q], qv);
      }
    } else {
      d0 = d;
      iv0 = inv;
    }
    K = cm * K;
    if (m > 0) P = cm * P;
  }
  // the last column's entries (they leave the batch), first kPF in registers
  const int kl = bend - 1;
  const int owner = (len - 1) / C;
  const int pl = wlo + kl;
  const int cl = BWD ? n - 1 - pl : pl;
  const int fl = a.colPtr[cl], ll = a.colPtr[cl + 1];
  const int lo = BWD ? fl : fl + 1, hi = BWD ? ll - 1 : ll;
  int tg[kPF];
  T vv[kPF];
#pragma unroll
  for (int i = 0; i < kPF; i++) {
    const bool has = lane == owner && lo + i < hi;
    tg[i] = has ? a.etgt[lo + i] : kRegTarget;
    vv[i] = has ? a.val[lo + i] : T(0);
  }
  // the first column's arrivals (its predecessor is in an earlier batch)
  const int pf = wlo + base;
  const int c0 = BWD ? n - 1 - pf : pf;
  const int2 nd = a.need[c0];
  const bool direct = DIRECT && nd.y == 1;
  const int need0 = nd.x + (direct ? 0 : nd.y * rv);
  T *slots = a.slots + (size_t)tile * a.nslots * R;
  T rem[R];
#pragma unroll
  for (int q = 0; q < R; q++) rem[q] = T(0);
  if (direct && lane == 0) {  // the first column's one remote slot
    const int sidx = a.rptr[c0];
    const T empty = slot_empty_value<T>();
    StallWatch w0;
    for (int q = 0; q < rv; q++) {
      T v = empty;
      while (slot_empty(v = xchg_take(&slots[(size_t)sidx * R + q], empty)))
        if (w0.step(false, a.err, a.wait_cap)) break;
      rem[q] = slot_empty(v) ? T(0) : v;
    }
  }
  StallWatch watch;
  while (cnt_load<CC>(s_cnt, base) != need0)
    if (watch.step(false, a.err, a.wait_cap)) return false;
  __builtin_amdgcn_s_setprio(3);
  fence_acq_wg();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < R; q++) {
      T r0 = s_val[base * R + q];
      if constexpr (DIRECT) r0 -= rem[q];
      const T a0 = wide_quot(r0, d0, iv0);
      s_val[base * R + q] = a0;
      A[q] = fma_t(P, a0, A[q]);
    }
  }
  chain_scan<T, R>(A, K);  // lane 0's K is 0: every A is now absolute
  // the batch's last column is the owner's last: publish it at once; the
  // other columns' x are only read at window end
  if (lane == owner)
    for (int q = 0; q < rv; q++)
      wide_publish<T, R>(a, s_val, slots, tg, vv, lo, hi, A[q], q);
  fence_rel_wg();
  if (lane == owner) wide_count<T, CC>(a, s_cnt, tg, lo, hi);
  __builtin_amdgcn_s_setprio(0);
  T x[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    x[q] = __shfl(A[q], lane > 0 ? lane - 1 : 0);
    if (lane == 0) x[q] = T(0);
  }
  for (int m = 0; m < mine; m++) {
    T d, inv;
    const T cm = coef(k0 + m, &d, &inv);
#pragma unroll
    for (int q = 0; q < R; q++) {
      x[q] = fma_t(cm, x[q], s_val[(k0 + m) * R + q]);
      s_val[(k0 + m) * R + q] = x[q];
    }
  }
  return true;
}

// Compute wavefront cw handles batches of up to 64 consecutive columns (lane
// = column) cw, cw + ncw, ..., each column for the R right-hand sides of its
// tile (rv of them valid).  Each row r may have one in-batch predecessor p(r)
// whose entry is delivered by register (analysis: `chain`); every other input
// arrives through LDS (same window) or a slot (earlier window, folded into
// LDS by the receivers).  Once a lane's LDS arrivals are complete its value
// is affine in its predecessor's:
//     x_r = r_r / d_r + c_r * x_p(r),   c_r = -L(r, p(r)) / d_r,  |c_r| <= 1,
// so the lanes whose inputs are complete along their whole predecessor chain
// are solved together: by a DPP scan when every link is lane -> lane + 1,
// otherwise by pointer jumping over the batch's forest (log2(chain depth)
// rounds of cross-lane permutes), instead of one lane per step.  Solved lanes
// publish at once: slots first (a slot's value is its own arrival flag), then
// LDS partial values, a workgroup release, then LDS arrival counters.  Lanes
// whose chain is not complete wait for the next round; the lowest unsolved
// lane always completes once its inputs from earlier batches arrive, so the
// batch always drains.
//
// DIRECT: a row with exactly one entry from an earlier window takes that slot
// in its own lane (an exchange poll issued beside each LDS count read) rather
// than through a receiver's LDS deposit and counter, one handoff less on a
// cross-window step.  It pays where windows hold whole grid planes, so that
// nearly every row is such a row (a 100^3 grid: 0.34 -> 0.23 ms, the 10^6
// chain's wide batches 0.46 -> 0.36 ms; a 128^3 grid, whose planes fit with
// byte-wide counters, 0.36 -> 0.29 ms); where windows split a plane and
// leave the receivers busy too, the extra pollers cost more than the
// handoff saves (128^3 with half-plane windows: +6%), and the configuration
// chooses.
template <typename T, bool BWD, int R, bool WIDE, bool DIRECT, bool CC>
__device__ __forceinline__ void lane_compute(const LaneArgs<T> &a, T *s_val,
                                            int *s_cnt, const int *s_bst,
                                            const int *s_bord, int nb, int wlo,
                                            int wlen, int tile, int rv, int cw,
                                            int ncw) {
  const int lane = __lane_id();
  const int n = a.n;
  T *slots = a.slots + (size_t)tile * a.nslots * R;
  bool stopped = false;  // the stall watch fired: leave the window
  int pf_bi = -1, pf_f = 0, pf_l = 0, pf_rp = 0;
  int2 pf_nd = make_int2(0, 0), pf_ch = make_int2(-1, 0);
  const T empty = slot_empty_value<T>();
  // the window's batches in level order, dealt round-robin: wavefront cw
  // takes positions cw, cw + ncw, ... of that order
  for (int pos = cw; pos < nb && !stopped; pos += ncw) {
    // a batch: up to 64 consecutive columns (window-local [base, bend)),
    // bounds staged in LDS
    const int bi = s_bord[pos];
    const int base = s_bst[bi];
    const int bend = bi + 1 < nb ? s_bst[bi + 1] : wlen;
    if constexpr (WIDE) {
      if (bend - base > kWave) {  // a run of pure links: one scan
        if (!lane_wide<T, BWD, R, DIRECT, CC>(a, s_val, s_cnt, base, bend, wlo, tile, rv))
          stopped = true;
        continue;
      }
    }
    const int k = base + lane;
    bool active = k < bend;
    int lo = 0, hi = 0, need_lds = 0, ne = 0, pl = lane;
    T diag = T(1), inv = T(1), coef = T(0);
    int tg[kPF];
    T vv[kPF];
    // a row with one remote entry takes that slot itself (sidx), its value
    // kept in rem and subtracted at the solve: no receiver deposit, no LDS
    // counter between the slot's arrival and the solve
    bool direct = false;
    int sidx = 0;
    if (active) {
      const int p = wlo + k;
      const int c = BWD ? n - 1 - p : p;
      // without wide batches b joins the column's LDS sum here (the window
      // starts from zeros): its load is issued first and lands with the
      // diagonal's
      T bq[R];
#pragma unroll
      for (int q = 0; q < R; q++)
        bq[q] = !WIDE && q < rv ? load_b(a.b, a.ir, (size_t)c * a.rhs + tile * R + q) : T(0);
      int f, l, rp;
      int2 nd, ch;
      if (pf_bi == bi) { f = pf_f; l = pf_l; nd = pf_nd; ch = pf_ch; rp = pf_rp; }
      else {
        f = a.colPtr[c]; l = a.colPtr[c + 1]; nd = a.need[c]; ch = a.chain[c];
        rp = DIRECT ? a.rptr[c] : 0;
      }
      direct = DIRECT && nd.y == 1;
      sidx = rp;
      diag = a.val[BWD ? l - 1 : f];
      inv = T(1) / diag;  // off the critical path
      if constexpr (!WIDE)
#pragma unroll
        for (int q = 0; q < R; q++) add_wg(&s_val[k * R + q], bq[q]);
      lo = BWD ? f : f + 1;
      hi = BWD ? l - 1 : l;
      ne = min(hi - lo, kPF);
      // receivers count one arrival per remote entry and right-hand side
      need_lds = nd.x + (direct ? 0 : nd.y * rv);
      if (ch.x >= 0 && ch.x < lane) {
        pl = ch.x;
        need_lds -= 1;
        const T lv = a.val[ch.y];
        T cq = lv * inv;
        cq = fma_t(fma_t(-cq, diag, lv), inv, cq);
        coef = -cq;
      }
#pragma unroll
      for (int i = 0; i < kPF; i++) {
        const int j = lo + i;
        tg[i] = j < hi ? a.etgt[j] : kRegTarget;
        vv[i] = j < hi ? a.val[j] : T(0);
      }
    }
    {  // the next batch's first-level metadata, in flight during this one
      pf_bi = -1;
      if (pos + ncw < nb) {
        const int nbi = s_bord[pos + ncw];
        const int nbase = s_bst[nbi];
        const int nbend = nbi + 1 < nb ? s_bst[nbi + 1] : wlen;
        const int kk = nbase + lane;
        pf_bi = nbi;
        if (kk < nbend && nbend - nbase <= kWave) {
          const int pp = wlo + kk;
          const int cc = BWD ? n - 1 - pp : pp;
          pf_f = a.colPtr[cc];
          pf_l = a.colPtr[cc + 1];
          pf_nd = a.need[cc];
          pf_ch = a.chain[cc];
          if (DIRECT) pf_rp = a.rptr[cc];
        }
      }
    }
    // LDS target of each register entry; entries that do not go to LDS add
    // their (unused) value and a count of 0 to this lane's own dummy row
    // past the window, so the publish is straight-line code with no
    // per-entry branches
    int ladr[kPF];
    int linc[kPF];
#pragma unroll
    for (int i = 0; i < kPF; i++) {
      const bool loc = i < ne && tg[i] < 0 && tg[i] != kRegTarget;
      ladr[i] = loc ? -tg[i] - 1 : a.maxw + lane;
      linc[i] = loc ? 1 : 0;
    }
    // every predecessor link is lane - 1 (stencil rows, chains): scan with
    // DPP instead of pointer jumping through LDS permutes
    const bool chain_batch =
        __ballot(active && pl != lane && pl != lane - 1) == 0;
    T xs[R];  // x of the lanes solved in earlier rounds
#pragma unroll
    for (int q = 0; q < R; q++) xs[q] = T(0);
    // the coefficient half of the chain scan depends only on the factor: its
    // six steps are composed while the batch waits, for the first round
    // (2D grids -3%, 3D -1%; not in the kernels with wide batches and
    // direct slots, whose chains it made 4% slower)
    constexpr bool kPre = !(WIDE && DIRECT);
    T pc[6];
    const unsigned long long act0 = __ballot(active);
    if (kPre && chain_batch) {
      T c = active ? coef : T(0);
      pc[0] = c;
      c = c * dpp_t<0x111, 0xf>(T(1), c);
      pc[1] = c;
      c = c * dpp_t<0x112, 0xf>(T(1), c);
      pc[2] = c;
      c = c * dpp_t<0x114, 0xf>(T(1), c);
      pc[3] = c;
      c = c * dpp_t<0x118, 0xf>(T(1), c);
      pc[4] = c;
      c = c * dpp_t<0x142, 0xa>(T(1), c);
      pc[5] = c;
    }
    // ready-lane mask of the last round that solved nothing: until it
    // changes no lane can become solvable, so the round is not repeated
    unsigned long long stale = 0;
    StallWatch watch;
    // DIRECT: the remote value taken (rem) and whether it is in (rgot)
    T rem = T(0);
    bool rgot = !direct;
    while (__ballot(active) != 0) {
      bool took = false;
      int cnt;
      if constexpr (DIRECT) {
        // The slot poll is issued first; while it is in flight the LDS
        // counts are read again, up to kDirectSpin times or until every
        // lane's in-window inputs are in (LDS reads wait on lgkmcnt, the
        // exchange on vmcnt).  An LDS arrival during the exchange's round
        // trip is then seen when the exchange returns, not one round trip
        // later (128^3 grid -3%, 100^3 -6%).
        T sv = empty;
        const bool pend = active && !rgot;
        if (pend) sv = xchg_take(&slots[sidx], empty);
        cnt = active ? cnt_load<CC>(s_cnt, k) : 0;
        if (__ballot(pend) != 0) {
          for (int sp = 0; sp < kDirectSpin; sp++) {
            if (__ballot(active && cnt != need_lds) == 0) break;
            cnt = active ? cnt_load<CC>(s_cnt, k) : 0;
          }
        }
        if (!slot_empty(sv)) {
          rem = sv;
          rgot = true;
          took = true;
        }
      } else {
        cnt = active ? cnt_load<CC>(s_cnt, k) : 0;
      }
      const bool ext = active && cnt == need_lds && (!DIRECT || rgot);
      bool progressed = false;  // some lane solved in this round
      const unsigned long long extm = __ballot(ext);
      if (extm != 0 && extm != stale) {
        // a solving wavefront takes issue priority over the SIMD's pollers
        __builtin_amdgcn_s_setprio(3);
        fence_acq_wg();
        // (av, cf): x = av + cf * x_pred; solved lanes are roots (cf = 0)
        T av[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
          T r = ext ? s_val[k * R + q] : T(0);
          if constexpr (DIRECT) r = ext ? r - rem : T(0);
          T qv = r * inv;
          qv = fma_t(fma_t(-qv, diag, r), inv, qv);
          av[q] = active ? qv : xs[q];
        }
        T cf = active ? coef : T(0);
        bool ok = active ? ext : true;
        if (chain_batch) {
          // lane 0 never has an in-batch predecessor, so after the full
          // scan every cf is 0 and av is x wherever the segment is ready
          const unsigned long long linked = __ballot(cf != T(0));
          const unsigned long long ready = __ballot(ok);
          if (kPre && __ballot(active) == act0) {
            scan_step_pc<T, R, 0x111, 0xf>(av, pc[0]);
            scan_step_pc<T, R, 0x112, 0xf>(av, pc[1]);
            scan_step_pc<T, R, 0x114, 0xf>(av, pc[2]);
            scan_step_pc<T, R, 0x118, 0xf>(av, pc[3]);
            scan_step_pc<T, R, 0x142, 0xa>(av, pc[4]);
            scan_step_pc<T, R, 0x143, 0xc>(av, pc[5]);
          } else {
            chain_scan<T, R>(av, cf);
          }
          ok = chain_ready(linked, ready, lane);
        } else {
          int pp = active ? pl : lane;
          // Every lane runs the rounds, so every permute reads a live lane.
          // A lane is done once it has absorbed a root (cf becomes 0): at
          // most 6 rounds for a 64-lane batch.  A product of coefficients
          // that underflows to 0 also ends a lane; the terms it leaves out
          // are below the smallest normal number times the ancestors' values.
          while (__ballot(ok && cf != T(0)) != 0) {
            const int qk = __shfl(pp | (ok ? 0 : 64), pp);
#pragma unroll
            for (int q = 0; q < R; q++) {
              const T ap = __shfl(av[q], pp);
              av[q] = fma_t(cf, ap, av[q]);
            }
            const T cp = __shfl(cf, pp);
            cf = cf * cp;
            ok = ok && (qk & 64) == 0;
            pp = qk & 63;
          }
        }
        const bool solved = active && ok;
        if (solved) {
#pragma unroll
          for (int q = 0; q < R; q++) xs[q] = av[q];
          active = false;
        }
        progressed = __ballot(solved) != 0;
        stale = progressed ? 0 : extm;
        if (progressed) {
          if (solved) {
            // slots first: the value is its own arrival flag
#pragma unroll
            for (int e = 0; e < kPF; e++)
              if (e < ne && tg[e] >= 0) {
#pragma unroll
                for (int q = 0; q < R; q++)
                  if (q < rv)
                    xchg_agent(&slots[(size_t)tg[e] * R + q], av[q] * vv[e]);
              }
#pragma unroll
            for (int q = 0; q < R; q++)  // the column is final: LDS holds x
              s_val[k * R + q] = av[q];
            // right-hand sides past rv hold x = 0 and add 0
#pragma unroll
            for (int e = 0; e < kPF; e++)
#pragma unroll
              for (int q = 0; q < R; q++)
                add_wg(&s_val[ladr[e] * R + q], -(av[q] * vv[e]));
          }
          // the register entries' rows (the next grid row of a stencil) are
          // released before the entries past the registers are loaded
          fence_rel_wg();
          if (solved) {
#pragma unroll
            for (int e = 0; e < kPF; e++)
              cnt_add<CC>(s_cnt, ladr[e], linc[e]);
          }
          // entries past the registers, kLong at a time: loads issued
          // together, one wait per group; each group's LDS values are
          // released (an LDS wait) before its counts.  A wavefront with no
          // long column skips the section on a scalar branch (without it
          // the 2D grids of 70-100 columns lose 2-3%)
          const bool longc = solved && hi - lo > kPF;
          if (__ballot(longc) != 0 && longc) {
            for (int j0 = lo + kPF; j0 < hi; j0 += kLong) {
              int tt[kLong];
              T tv[kLong];
#pragma unroll
              for (int i = 0; i < kLong; i++) {
                tt[i] = j0 + i < hi ? a.etgt[j0 + i] : kRegTarget;
                tv[i] = j0 + i < hi ? a.val[j0 + i] : T(0);
              }
#pragma unroll
              for (int i = 0; i < kLong; i++) {
                const int t = tt[i];
                if (t >= 0) {
#pragma unroll
                  for (int q = 0; q < R; q++)
                    if (q < rv)
                      xchg_agent(&slots[(size_t)t * R + q], av[q] * tv[i]);
                } else if (t != kRegTarget) {
#pragma unroll
                  for (int q = 0; q < R; q++)
                    if (q < rv)
                      add_wg(&s_val[(-t - 1) * R + q], -(av[q] * tv[i]));
                }
              }
              fence_rel_wg();
#pragma unroll
              for (int i = 0; i < kLong; i++)
                if (tt[i] < 0 && tt[i] != kRegTarget)
                  cnt_add<CC>(s_cnt, -tt[i] - 1, 1);
            }
          }
        }
      }
      __builtin_amdgcn_s_setprio(0);
      // Pollers spin on LDS without sleeping: the priority above keeps them
      // from delaying a solving wavefront, and a sleep would only delay the
      // poll that sees the next arrival.  A wavefront whose lanes have inputs
      // but wait on an unsolved predecessor counts as stalled.
      if constexpr (DIRECT) progressed = progressed || __ballot(took) != 0;
      if (watch.step(progressed, a.err, a.wait_cap)) {
        active = false;
        stopped = true;
      }
    }
  }
}

constexpr int kTake = 8;  // slots of one row taken per group

// Receivers: the first nrecv wavefronts walk the window's receive list (only
// the columns that have inputs from earlier windows, precomputed per window
// partition), one (column, right-hand side) pair per lane: lane l of
// receiver r takes pairs r*64 + l, stride nrecv*64.  A pair's slots are taken
// in groups of up to kTake with exchange(slot, EMPTY), which both polls and
// resets a slot.  Only the first slot of a group is polled until it arrives;
// the rest of the group is taken after it, and the group is summed in slot
// order once complete.  A row with more than kRecvChunk remote slots has one
// receive-list entry per kRecvChunk of them, so its chunks are taken by
// several lanes at once and added to the row's LDS sum (and count) one by
// one.  A receiver
// wavefront that took nothing in an iteration backs off.
template <typename T, int R, bool CC>
__device__ __forceinline__ void lane_receive(const LaneArgs<T> &a, T *s_val,
                                            int *s_cnt, int win, int tile,
                                            int rv, int rw) {
  T *slots = a.slots + (size_t)tile * a.nslots * R;
  const T empty = slot_empty_value<T>();
  const int e0 = a.roff[win];
  const int p_end = (a.roff[win + 1] - e0) * rv;
  const int stride = a.nrecv * kWave;
  int pi = rw * kWave + __lane_id();
  int k = 0, s = 0, r1 = 0, m = 0, q = 0;
  unsigned got = 0;
  T part[kTake];
  T sum = T(0);
  auto fetch = [&](int i) {
    const int e = R == 1 ? i : i / rv;
    return a.rlist[e0 + e];
  };
  int4 next = pi < p_end ? fetch(pi) : make_int4(0, 0, 0, 0);
  auto take_next = [&]() {
    k = next.x;
    s = next.y;
    m = next.z;
    q = R == 1 ? 0 : pi % rv;
    r1 = s + m;
    got = 0;
    sum = T(0);
    if (pi + stride < p_end) next = fetch(pi + stride);  // prefetch
  };
  bool active = pi < p_end;
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
      for (int i = 0; i < kTake; i++) {
        // slot 0 of the group first; the others once slot 0 is in
        if (i < g && !(got & (1u << i)) && (i == 0 || (got & 1u))) {
          const T v = xchg_take(&slots[(size_t)(s + i) * R + q], empty);
          if (!slot_empty(v)) {
            part[i] = v;
            got |= 1u << i;
          }
        }
      }
      progressed = got != before;
      if (got == (1u << g) - 1) {
#pragma unroll
        for (int i = 0; i < kTake; i++)
          if (i < g) sum += part[i];
        s += g;
        got = 0;
        if (s == r1) {
          add_wg(&s_val[k * R + q], -sum);
          fence_rel_wg();
          cnt_add<CC>(s_cnt, k, m);
          done = true;
        }
      }
    }
    if (done) {
      pi += stride;
      active = pi < p_end;
      if (active) take_next();
    }
    const bool any = __ballot(progressed) != 0;
    backoff(level, any);
    if (watch.step(any, a.err, a.wait_cap)) active = false;
  }
}

// Largest lane-engine workgroup for R right-hand sides per lane: 2 or more
// values per lane need the register budget of 512-thread workgroups (with
// 1024 threads, 128 VGPRs, the FP64 kernels spilled 2-4 registers at R = 2
// and 20 at R = 4).
constexpr int lane_max_block(int R) { return R >= 2 ? 512 : 1024; }

template <typename T, bool BWD, int R, bool WIDE, bool DIRECT, bool CC>
__global__ void __launch_bounds__(lane_max_block(R)) k_exec_lane(LaneArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *s_val = reinterpret_cast<T *>(smem_raw);
  // maxw window rows plus 64 dummy rows (one per lane) for branch-free
  // publishing, then the window's batch starts
  int *s_cnt = reinterpret_cast<int *>(s_val + (size_t)(a.maxw + kWave) * R);
  int *s_bst = s_cnt + cnt_ints(a.maxw + kWave, CC);
  int *s_bord = s_bst + a.maxnb;
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
    const int wi = t / a.ntiles;
    const int tile = t - wi * a.ntiles;
    const int win = a.worder ? a.worder[wi] : wi;
    const int rv = min(R, a.rhs - tile * R);  // valid right-hand sides
    const int wlo = a.winPtr[win];
    const int wlen = a.winPtr[win + 1] - wlo;
    // Partial sums start at zero and each column's b is added by its own
    // lane when its batch starts: staging a 128^3 grid's plane of b here
    // held the window's first batch back ~7 us (-2.6% on that grid).  Wide
    // batches (pure chains) read b from LDS before their wait, so windows
    // that have them stage b here (per batch it cost the 10^6 chain 3-5%).
    for (int i = threadIdx.x; i < wlen * R; i += blockDim.x) {
      T v = T(0);
      if constexpr (WIDE) {
        const int k = R == 1 ? i : i / R;
        const int q = R == 1 ? 0 : i - k * R;
        const int p = wlo + k;
        const int c = BWD ? n - 1 - p : p;
        if (q < rv) v = load_b(a.b, a.ir, (size_t)c * a.rhs + tile * R + q);
      }
      s_val[i] = v;
    }
    for (int k = threadIdx.x; k < (int)cnt_ints(wlen, CC); k += blockDim.x)
      s_cnt[k] = 0;
    const int b0 = a.boff[win], nb = a.boff[win + 1] - b0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      s_bst[i] = a.bstart[b0 + i];
      s_bord[i] = a.bord[b0 + i];
    }
    __syncthreads();
    if (wave < a.nrecv)
      lane_receive<T, R, CC>(a, s_val, s_cnt, win, tile, rv, wave);
    else
      lane_compute<T, BWD, R, WIDE, DIRECT, CC>(a, s_val, s_cnt, s_bst, s_bord, nb, wlo,
                                    wlen, tile, rv, wave - a.nrecv, ncw);
    __syncthreads();
    for (int i = threadIdx.x; i < wlen * R; i += blockDim.x) {
      const int k = R == 1 ? i : i / R;
      const int q = R == 1 ? 0 : i - k * R;
      const int p = wlo + k;
      const int c = BWD ? n - 1 - p : p;
      if (q < rv) {
        const size_t e = (size_t)c * a.rhs + tile * R + q;
        a.x[e] = s_val[i];
        ir_update(a.ir, e, s_val[i]);
      }
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

// Chain links by solve position (packed two bits per position, four per
// byte): link(p) = 1 when the column at position
// p - 1 has an entry in the row at position p (the nearest off-diagonal
// entry: first after the diagonal forward, last before it backward); 2 when
// that link is pure: the column at p - 1 has no other off-diagonal entry,
// the row at p no other input (inDeg counts the diagonal) and
// |L(p, p-1)| <= |L(p, p)|.  A run of pure links is a first-order
// recurrence that a wide batch solves with one wavefront scan.
template <typename T>
__global__ void k_links(const int *__restrict__ colPtr,
                        const int *__restrict__ rowIdx,
                        const T *__restrict__ val,
                        const int *__restrict__ inDeg, int n, int backward,
                        unsigned char *link) {
  // four positions per byte, two bits each (the host reads 1/4 the bytes)
  const int nq = (n + 3) / 4;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += gridDim.x * blockDim.x) {
    unsigned packed = 0;
    for (int k = 0; k < 4; k++) {
      const int p = 4 * i + k;
      unsigned v = 0;
      if (p > 0 && p < n) {
        const int c = backward ? n - p : p - 1;       // column at position p - 1
        const int want = backward ? n - 1 - p : p;    // row at position p
        const int f = colPtr[c], l = colPtr[c + 1];
        const int j = backward ? l - 2 : f + 1;       // the nearest entry
        if (j >= f && j < l && (backward ? j < l - 1 : j > f) && rowIdx[j] == want) {
          v = 1;
          const T d = val[backward ? colPtr[want + 1] - 1 : colPtr[want]];
          if (l - f == 2 && inDeg[want] == 2 && fabs(val[j]) <= fabs(d)) v = 2;
        }
      }
      packed |= v << (2 * k);
    }
    link[i] = (unsigned char)packed;
  }
}

// In-batch predecessor selection: for every row, the nearest earlier column
// of the same batch (solve order) among those with |L(r, c)| <= |L(r, r)|,
// as key = (position + 1) << 32 | entry index, kept by atomicMax (key
// zeroed).  The bound keeps the products of the affine coefficients of a
// jumped chain at most 1 in magnitude.
template <typename T>
__global__ void k_chain_key(const int *__restrict__ colPtr,
                            const int *__restrict__ rowIdx,
                            const T *__restrict__ val, int n, int backward,
                            const int *__restrict__ bid,
                            unsigned long long *key) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    const int p = backward ? n - 1 - c : c;
    const int lb = bid[p];
    const int f = colPtr[c], l = colPtr[c + 1];
    const int lo = backward ? f : f + 1, hi = backward ? l - 1 : l;
    for (int j = lo; j < hi; j++) {
      const int r = rowIdx[j];
      const int pr = backward ? n - 1 - r : r;
      if (bid[pr] != lb) continue;
      const T d = val[backward ? colPtr[r + 1] - 1 : colPtr[r]];
      const T v = val[j];
      if (!(fabs(v) <= fabs(d))) continue;
      atomicMax(&key[r], ((unsigned long long)(p + 1) << 32) | (unsigned)j);
    }
  }
}

// chain[r] = {lane of the selected predecessor, its entry} or {-1, 0}; the
// selected entry's target becomes kRegTarget.  bfirst: first solve position
// of each batch; bid: batch of each solve position.
__global__ void k_chain_mark(const unsigned long long *__restrict__ key, int n,
                             const int *__restrict__ bfirst, const int *__restrict__ bid,
                             int *etgt, int2 *chain) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += gridDim.x * blockDim.x) {
    const unsigned long long kk = key[r];
    if (kk == 0) {
      chain[r] = make_int2(-1, 0);
      continue;
    }
    const int pc = (int)(kk >> 32) - 1;
    const int j = (int)(unsigned)(kk & 0xffffffffull);
    chain[r] = make_int2(pc - bfirst[bid[pc]], j);  // lane in its batch
    etgt[j] = kRegTarget;
  }
}

// Reach of sampled columns: out[s] = the largest |row - column| among the
// entries of column s * stride (0 for a diagonal-only column).  The most
// common reach is the factor's period (a grid plane or row), to which the
// default windows are aligned.
__global__ void k_sample_reach(const int *__restrict__ colPtr,
                               const int *__restrict__ rowIdx, int n,
                               int stride, int nsamp, int *out) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nsamp;
       s += gridDim.x * blockDim.x) {
    const int c = s * stride;
    int reach = 0;
    if (c < n)
      for (int j = colPtr[c]; j < colPtr[c + 1]; j++)
        reach = max(reach, abs(rowIdx[j] - c));
    out[s] = reach;
  }
}

__global__ void k_need_remote(const int2 *__restrict__ need, int n, int *out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += gridDim.x * blockDim.x)
    out[r] = need[r].y;
}

// A receive-list entry covers at most kRecvChunk slots of a row, so the
// slots of a dense row are taken by several receiver lanes in parallel
// instead of one lane group by group.
constexpr int kRecvChunk = 2 * kTake;

// count[p] = receive-list entries of the column at solve position p (none
// for a row with one remote entry when such rows take their slot directly).
__global__ void k_remote_flags(const int2 *__restrict__ need, int n,
                               int backward, int direct, int *count) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += gridDim.x * blockDim.x)
  {
    const int m = need[backward ? n - 1 - p : p].y;
    count[p] = direct && m == 1 ? 0 : (m + kRecvChunk - 1) / kRecvChunk;
  }
}

// bid[p] = the batch holding solve position p (bfirst: nb + 1 batch starts).
__global__ void k_batch_ids(const int *__restrict__ bfirst, int nb, int *bid) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb;
       b += gridDim.x * blockDim.x)
    for (int p = bfirst[b]; p < bfirst[b + 1]; p++) bid[p] = b;
}

// bpos[b] = pos[first column of batch b] for b < nb; bpos[nb] = pos[n].
__global__ void k_batch_entry_pos(const int *__restrict__ bfirst,
                                  const int *__restrict__ pos, int nb, int n,
                                  int *bpos) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b <= nb;
       b += gridDim.x * blockDim.x)
    bpos[b] = pos[b < nb ? bfirst[b] : n];
}

// Receive lists: pos = exclusive scan of the entry counts (solve order), so
// each window's entries are contiguous; roff[w] = pos[winPtr[w]].  Within a
// window a batch's entries (its columns in solve order) move by shift[batch]
// to the batch's place in the processing order.  Entry i of a column covers
// its slots [rptr + i * kRecvChunk, ...), at most kRecvChunk of them; with
// `direct`, rows with a single remote entry have none.
__global__ void k_receive_list(const int2 *__restrict__ need,
                               const int *__restrict__ rptr,
                               const int *__restrict__ pos,
                               const int *__restrict__ bid,
                               const int *__restrict__ shift,
                               const int *__restrict__ winPtr, int nwin, int n,
                               int backward, int uniform_w, int direct,
                               int4 *rlist, int *roff) {
  const int stride = gridDim.x * blockDim.x;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    const int c = backward ? n - 1 - p : p;
    const int m = need[c].y;
    if (m > (direct ? 1 : 0)) {
      const int w = uniform_w > 0 ? p / uniform_w
                                  : upper_bound_minus1(winPtr, nwin + 1, p);
      const int e0 = pos[p] + shift[bid[p]];
      for (int i = 0, s = 0; s < m; i++, s += kRecvChunk)
        rlist[e0 + i] =
            make_int4(p - winPtr[w], rptr[c] + s, min(kRecvChunk, m - s), 0);
    }
  }
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w <= nwin; w += stride)
    roff[w] = pos[winPtr[w]];
}

}  // namespace sptrsv
