// sptrsv_transpose.hpp -- device scan, histogram, stable CSR<->CSC transpose,
// in-degree analysis and factor validation.
//
// The transpose follows the result of matrix_transposition
// (SpTRSV_cuda/tranpose.h:6-43): bucket counts (:17-21), exclusive scan (:24),
// and an order-preserving scatter (:30-40).  On the GPU the scatter claims slots
// with atomics, which scrambles the order inside a bucket; every slot therefore
// carries its source position j (the CSR order the host loop walks) and each
// bucket is then sorted by j.  Source positions are unique, so the result is
// exactly the stable host transpose, duplicates and unsorted rows included.
#pragma once

#include "sptrsv_device.hpp"

namespace sptrsv {

// ---------------------------------------------------------------- scan -----
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048 ints per block

// Exclusive scan of one tile per block; writes the tile total to sums[block].
__global__ void __launch_bounds__(kScanThreads)
    k_scan_tile(int *data, int len, int *sums) {
  __shared__ int s_wave[kScanThreads / kWave];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = (base + i < len) ? data[base + i] : 0;
    local += v[i];
  }
  // inclusive scan of the per-thread totals across the wavefront
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  int incl = local;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int up = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += up;
  }
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  int wave_off = 0;
  for (int w = 0; w < wave; w++) wave_off += s_wave[w];
  int run = wave_off + incl - local;  // exclusive prefix of this thread
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < len) data[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == kScanThreads - 1) {
    int tot = 0;
    for (int w = 0; w < kScanThreads / kWave; w++) tot += s_wave[w];
    sums[blockIdx.x] = tot;
  }
}

__global__ void __launch_bounds__(kScanThreads)
    k_scan_add(int *data, int len, const int *offsets) {
  const int off = offsets[blockIdx.x];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
  for (int i = 0; i < kScanItems; i++)
    if (base + i < len) data[base + i] += off;
}

// ------------------------------------------------------ histogram/analyse --
// cnt[idx[j]] += 1 for every j (sptrsv_syncfree_cuda.h:8-19 for the row
// histogram; tranpose.h:17-21 for the column histogram).
__global__ void k_histogram(const int *__restrict__ idx, int nnz, int *cnt) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += gridDim.x * blockDim.x)
    atomicAdd(&cnt[idx[j]], 1);
}

// ---------------------------------------------------------- transpose ------
// keys[slot] = (j << 32) | row for every CSR entry j; slots claimed per column.
__global__ void k_scatter_keys(const int *__restrict__ rowPtr,
                               const int *__restrict__ colIdx, int m, int nnz,
                               int *cursor, unsigned long long *keys) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += gridDim.x * blockDim.x) {
    const int row = upper_bound_minus1(rowPtr, m + 1, j);
    const int slot = atomicAdd(&cursor[colIdx[j]], 1);
    keys[slot] = ((unsigned long long)(unsigned)j << 32) | (unsigned)row;
  }
}

constexpr int kSmallSeg = 16;    // sorted in registers by one thread
constexpr int kLdsSeg = 4096;    // bitonic sort in LDS by one workgroup
constexpr int kLdsThreads = 256;

// Sort buckets of <= kSmallSeg keys in place; list longer ones as segments for
// the LDS sorter (<= kLdsSeg) or as long buckets for chunked merging.
__global__ void k_sort_small(const int *__restrict__ ptr, int nb,
                             unsigned long long *keys, int *seg_start,
                             int *seg_len, int *n_seg, int *long_start,
                             int *long_len, int *n_long) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nb;
       c += gridDim.x * blockDim.x) {
    const int s = ptr[c], len = ptr[c + 1] - s;
    if (len <= 1) continue;
    if (len <= kSmallSeg) {
      unsigned long long v[kSmallSeg];
#pragma unroll
      for (int i = 0; i < kSmallSeg; i++)
        v[i] = i < len ? keys[s + i] : ~0ull;
      // odd-even transposition network on registers (fully unrolled)
#pragma unroll
      for (int pass = 0; pass < kSmallSeg; pass++) {
#pragma unroll
        for (int i = pass & 1; i + 1 < kSmallSeg; i += 2) {
          const unsigned long long a = v[i], b = v[i + 1];
          v[i] = a < b ? a : b;
          v[i + 1] = a < b ? b : a;
        }
      }
#pragma unroll
      for (int i = 0; i < kSmallSeg; i++)
        if (i < len) keys[s + i] = v[i];
    } else if (len <= kLdsSeg) {
      const int k = atomicAdd(n_seg, 1);
      seg_start[k] = s;
      seg_len[k] = len;
    } else {
      const int k = atomicAdd(n_long, 1);
      long_start[k] = s;
      long_len[k] = len;
    }
  }
}

// One workgroup sorts one segment of <= kLdsSeg keys with a bitonic network.
__global__ void __launch_bounds__(kLdsThreads)
    k_sort_lds(unsigned long long *keys, const int *__restrict__ seg_start,
               const int *__restrict__ seg_len) {
  __shared__ unsigned long long s[kLdsSeg];
  const int start = seg_start[blockIdx.x], len = seg_len[blockIdx.x];
  int p2 = 2;
  while (p2 < len) p2 <<= 1;
  for (int i = threadIdx.x; i < p2; i += blockDim.x)
    s[i] = i < len ? keys[start + i] : ~0ull;
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int i = threadIdx.x; i < (p2 >> 1); i += blockDim.x) {
        const int lo = 2 * i - (i & (stride - 1));
        const int hi = lo + stride;
        const bool asc = (lo & size) == 0;
        const unsigned long long a = s[lo], b = s[hi];
        if ((a > b) == asc) {
          s[lo] = b;
          s[hi] = a;
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < len; i += blockDim.x) keys[start + i] = s[i];
}

// One merge pass over a long bucket whose runs of `width` keys are sorted:
// each key lands at (its rank in its run) + (number of smaller keys in the
// partner run).  Keys are unique, so lower_bound gives a stable placement.
__global__ void k_merge_pass(const unsigned long long *__restrict__ in,
                             unsigned long long *__restrict__ out, int len,
                             int width) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += gridDim.x * blockDim.x) {
    const int run = i / width;
    const int partner = run ^ 1;
    const int p_lo = partner * width;
    const int pair_lo = (run < partner ? run : partner) * width;
    const unsigned long long key = in[i];
    if (p_lo >= len) {
      out[i] = key;
      continue;
    }
    const int p_hi = min(p_lo + width, len);
    int lo = p_lo, hi = p_hi;  // first index in partner with value >= key
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (in[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
    out[pair_lo + (i - run * width) + (lo - p_lo)] = key;
  }
}

// rowIdx/val of the transposed matrix from the sorted keys.
template <typename T>
__global__ void k_gather(const unsigned long long *__restrict__ keys, int nnz,
                         const T *__restrict__ val, int *__restrict__ outIdx,
                         T *__restrict__ outVal) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += gridDim.x * blockDim.x) {
    const unsigned long long k = keys[i];
    outIdx[i] = (int)(unsigned)(k & 0xffffffffull);
    if (outVal) outVal[i] = val[(unsigned)(k >> 32)];
  }
}

// ---------------------------------------------------------- validation -----
// CSR input of the transpose: row pointers non-decreasing, column indices in
// [0, n).  Sets *flags on any violation.
__global__ void k_check_csr(const int *__restrict__ rowPtr,
                            const int *__restrict__ colIdx, int m, int n,
                            int nnz, int *flags) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const int a = rowPtr[i], b = rowPtr[i + 1];
    if (a > b || a < 0 || b > nnz) atomicOr(flags, 1);
  }
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += stride) {
    const int c = colIdx[j];
    if (c < 0 || c >= n) atomicOr(flags, 1);
  }
}

enum : int {
  kBadIndex = 1,
  kBadUnsorted = 2,
  kBadTriangle = 4,
  kBadMissingDiag = 8,
  kBadZeroDiag = 16,
};

// Per column: pointers in range, rows in [0, n) and strictly ascending, all
// rows on the solved side of the diagonal, the diagonal stored first
// (forward) or last (backward) and nonzero.
template <typename T>
__global__ void k_validate(const int *__restrict__ colPtr,
                           const int *__restrict__ rowIdx,
                           const T *__restrict__ val, int n, int nnz,
                           int backward, int *flags) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    const int f = colPtr[c], l = colPtr[c + 1];
    int bad = 0;
    if (f < 0 || l > nnz || f > l) {
      atomicOr(flags, kBadIndex);
      continue;
    }
    if (f == l) {
      atomicOr(flags, kBadMissingDiag);
      continue;
    }
    int prev = -1;
    for (int j = f; j < l; j++) {
      const int r = rowIdx[j];
      if (r < 0 || r >= n) {
        bad |= kBadIndex;
        break;
      }
      if (r <= prev) bad |= kBadUnsorted;
      if (backward ? (r > c) : (r < c)) bad |= kBadTriangle;
      prev = r;
    }
    if (!(bad & kBadIndex)) {
      const int d = backward ? l - 1 : f;
      if (rowIdx[d] != c)
        bad |= kBadMissingDiag;
      else if (val != nullptr && val[d] == T(0))
        bad |= kBadZeroDiag;
    }
    if (bad) atomicOr(flags, bad);
  }
}

// ------------------------------------------------------ window analysis ----
// For the window partition winPtr (positions in solve order), split each row's
// off-diagonal count into contributions from columns of the same window
// (need.x, accumulated in LDS) and from earlier windows (need.y, accumulated
// in global memory).  need must be zeroed.
__global__ void k_need(const int *__restrict__ colPtr,
                       const int *__restrict__ rowIdx, int n, int backward,
                       const int *__restrict__ winPtr, int nwin, int uniform_w,
                       int2 *need) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    const int p = backward ? n - 1 - c : c;
    const int w = uniform_w > 0 ? p / uniform_w
                                : upper_bound_minus1(winPtr, nwin + 1, p);
    const int whi = winPtr[w + 1];
    const int f = colPtr[c], l = colPtr[c + 1];
    const int lo = backward ? f : f + 1, hi = backward ? l - 1 : l;
    for (int j = lo; j < hi; j++) {
      const int r = rowIdx[j];
      const int pr = backward ? n - 1 - r : r;
      if (pr < whi)
        atomicAdd(&need[r].x, 1);
      else
        atomicAdd(&need[r].y, 1);
    }
  }
}

__global__ void k_uniform_windows(int *winPtr, int nwin, int w, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= nwin;
       i += gridDim.x * blockDim.x)
    winPtr[i] = min(i * w, n);
}

template <typename T>
__global__ void k_fill(T *p, size_t len, T v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace sptrsv
