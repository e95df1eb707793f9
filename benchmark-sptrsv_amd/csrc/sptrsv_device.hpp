// sptrsv_device.hpp -- shared device helpers for the gfx950 SpTRSV kernels.
//
// Memory-model notes (CDNA3/CDNA4, 8 XCDs with one L2 each):
//  * Cross-workgroup data goes through agent-scope atomics.  A relaxed
//    agent-scope load compiles to a global load with sc1 (served coherently
//    across XCDs); an agent-scope release fence writes back the issuing XCD's
//    L2 (buffer_wbl2 sc1) and an acquire fence invalidates it (buffer_inv sc1).
//  * Within a workgroup, LDS atomics and workgroup-scope fences suffice.
//  * Lanes of one wavefront execute memory instructions together, so a fence
//    executed by a wavefront orders every lane's earlier accesses; the
//    executors rely on this when one lane publishes a counter after all lanes
//    of its group have added their partial sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sptrsv {

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Publish a value to another workgroup: an agent-scope read-modify-write is
// performed where all XCDs see it.  A plain store instead allocates the line
// in the writer's XCD L2; sc1 polls issued from that XCD can then keep hitting
// that copy (stale for every other word of the line) until it is evicted or
// an acquire fence invalidates it.  Every in-kernel write to a line that other
// workgroups poll is therefore an RMW.  The old value is unused, so the
// exchange is issued without a return and nothing waits on it.
template <typename T>
__device__ __forceinline__ void xchg_agent(T *p, T v) {
  (void)__hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Read the current value of a location other workgroups update, through a
// read-modify-write that leaves it unchanged.  Unlike a relaxed load (a sc1
// global load, which an L2 line already resident in this XCD can serve), an
// RMW is performed at the coherence point and returns the latest value in the
// location's modification order, so a spin-wait on it cannot stall on a stale
// cached copy.
__device__ __forceinline__ int rmw_read(int *p) {
  return __hip_atomic_fetch_or(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Take the value of a slot and leave `empty` behind, in one RMW.
template <typename T>
__device__ __forceinline__ T xchg_take(T *p, T empty) {
  return __hip_atomic_exchange(p, empty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Striped tickets: consecutive chunks of work are claimed from kStripes
// counters, chunk c from counter c % S, so claim atomics go to S addresses
// (on one address they serialise at ~11.5 ns each).  Claimant i (a
// wavefront or a workgroup, S <= the number of claimants) always uses
// counter i % S and gets its chunks in increasing order.  The lowest
// unfinished chunk is then always claimed or about to be: every chunk of its
// counter below it is finished, so that counter's claimants are free, and
// one of them takes it next, provided every counter has a resident claimant
// (workgroups are dispatched in order, so the first kStripes cover them).
// Counter s is int kStripeBase + 32 * s of the
// plan's sync words (its own 128-byte line); the last claimant to finish
// resets them.
constexpr int kStripes = 8;
constexpr int kStripeBase = 128;
__device__ __forceinline__ long long claim_chunk(int *sync, int stripe,
                                                 int nstripes) {
  const int k = atomicAdd(sync + kStripeBase + 32 * stripe, 1);
  return (long long)k * nstripes + stripe;
}
__device__ __forceinline__ void reset_stripes(int *sync) {
  for (int s = 0; s < kStripes; s++)
    __hip_atomic_exchange(sync + kStripeBase + 32 * s, 0, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void add_agent(T *p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Fused multiply-add in the operands' own precision: __builtin_fma is the
// double routine, which would run a float kernel's arithmetic in FP64 with
// conversions around every call.
__device__ __forceinline__ double fma_t(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fma_t(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
// Fusions for the FP32 solves of iterative refinement: the first solve reads
// b in FP64 and rounds it as it loads it (no conversion pass over b), the
// last stores x = xold + d in FP64 beside each d it solves (no update pass
// over x).  Null pointers in every other solve; FP64 kernels ignore them.
struct IrFuse {
  const double *b64;   // first solve: the FP64 right-hand sides, or null
  const double *xold;  // last solve: the iterate d corrects, or null (= 0)
  double *xout;        // last solve: where xold + d goes, or null
};
template <typename T>
__device__ __forceinline__ T load_b(const T *b, const IrFuse &f, size_t i) {
  if constexpr (std::is_same<T, float>::value)
    if (f.b64) return (float)f.b64[i];
  return b[i];
}
template <typename T>
__device__ __forceinline__ void ir_update(const IrFuse &f, size_t i, T d) {
  if constexpr (std::is_same<T, float>::value)
    if (f.xout) f.xout[i] = (f.xold ? f.xold[i] : 0.0) + (double)d;
}

template <typename T>
__device__ __forceinline__ T ld_wg(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void add_wg(T *p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void fence_acq_agent() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ __forceinline__ void fence_rel_agent() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}
__device__ __forceinline__ void fence_acq_wg() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ void fence_rel_wg() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}

// Exponential back-off for a wavefront whose spin-wait made no progress:
// sleeps 64 * level cycles, level doubling up to kBackoffMax, and resets when
// the wavefront makes progress.  Keeps thousands of waiting wavefronts from
// flooding the coherence point with polls.
constexpr int kBackoffMax = 16;
__device__ __forceinline__ void backoff(int &level, bool progressed) {
  if (progressed) {
    level = 0;
    return;
  }
  switch (level) {  // s_sleep takes an immediate
    case 0: break;
    case 1: __builtin_amdgcn_s_sleep(1); break;
    case 2: __builtin_amdgcn_s_sleep(2); break;
    case 4: __builtin_amdgcn_s_sleep(4); break;
    case 8: __builtin_amdgcn_s_sleep(8); break;
    default: __builtin_amdgcn_s_sleep(16); break;
  }
  level = level == 0 ? 1 : min(level * 2, kBackoffMax);
}

// Per-wavefront stall watch for spin-waits.  Only lane 0 touches the clock
// and the shared error word, and only every 1024 stalled iterations, so the
// watch adds no traffic to the coherence point (one shared word polled by
// every lane of every wavefront would serialise a memory channel).
// step() returns true when the wavefront must stop: another wavefront raised
// the error, or this one has made no progress for longer than cap ticks.
struct StallWatch {
  unsigned long long t_progress;
  unsigned stalled;
  __device__ explicit StallWatch() : t_progress(wall_clock64()), stalled(0) {}
  __device__ __forceinline__ bool step(bool progressed, int *err,
                                       long long cap) {
    if (progressed) {
      stalled = 0;
      t_progress = wall_clock64();
      return false;
    }
    if ((++stalled & 1023) != 0) return false;
    int stop = 0;
    if (__lane_id() == 0) {
      const unsigned long long now = wall_clock64();
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        stop = 1;
      else if ((long long)(now - t_progress) > cap) {
        __hip_atomic_exchange(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        stop = 1;
      }
    }
    return __shfl(stop, 0) != 0;
  }
};

// Largest index i in [0, len) with a[i] <= key (a ascending, a[0] <= key).
__device__ __forceinline__ int upper_bound_minus1(const int *a, int len,
                                                  int key) {
  int lo = 0, hi = len;  // invariant: a[lo] <= key, answer in [lo, hi)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= key)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

}  // namespace sptrsv
