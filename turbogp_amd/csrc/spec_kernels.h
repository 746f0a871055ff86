// Bodies of the shape-specialised single-query kernels (spec_body: one query
// per launch, per-wave histograms folded into registers or a workgroup table;
// wave_body: replicated per-wave LDS histograms).  spec.hip wraps them for the
// shapes compiled ahead of time; jit.hip compiles the same bodies at run time
// for any other shape that extract_shape derives.
#pragma once

#include "wave_agg.h"

namespace tgp {

template <Shape S>
__device__ __forceinline__ void spec_body(const SpecArgs& A) {
  extern __shared__ __align__(16) unsigned long long lmem[];
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  // table 0: per-wave histograms of kHistWords words; table 1: one workgroup table
  // (counts[ngroups], then (lo, hi) sums) followed by nothing else.
  unsigned long long* hist = S.table == 0 ? lmem + wave * kHistWords : lmem;
  if constexpr (S.table == 0) {
    hist[lane] = 0;
  } else {
    const uint64_t words = A.ngroups * (1 + 2 * (uint64_t)S.naggs);
    for (uint64_t i = threadIdx.x; i < words; i += kBlock) lmem[i] = 0;
    __syncthreads();
  }
  // lane-owned exact accumulator of histogram word `lane` (table 0); a MIN / MAX
  // word keeps the largest order code instead
  uint64_t acc_lo = 0;
  int64_t acc_hi = 0;
  constexpr uint32_t kMM = shape_minmax_kinds(S);
  const uint32_t lk = lane % (1 + S.naggs);
  const bool lane_mm = kMM != 0 && lk > 0 && ((kMM >> (3 * (lk - 1))) & 7) != 0;
  uint32_t oor = 0;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t gw = (uint64_t)blockIdx.x * (kBlock / kWave) + wave;
  uint64_t min_len = A.col_len[0];
#pragma unroll
  for (int s = 1; s < S.nslots; ++s) min_len = A.col_len[s] < min_len ? A.col_len[s] : min_len;
  for (uint64_t k = gw; k < A.nslices; k += nwaves) {
    const uint64_t slice0 = A.base0 + k * kSliceRows;
    const uint64_t row = slice0 + lane * kSR;
    const bool fast = slice0 + kSliceRows <= min_len;
    SliceVals<S> v;
    load_all<S, 0>(A, row, fast, v);
    int32_t gid[kSR];
    const uint32_t sel = eval_slice<S>(A, row, v, gid, oor);
    if (!__ballot(sel != 0)) continue;
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if ((sel >> r) & 1) {
        if constexpr (S.table == 0) atomicAdd(&hist[gid[r] * (1 + S.naggs)], 1ull);
        else atomicAdd(&hist[gid[r]], 1ull);
      }
    bool big = false;
    accumulate_aggs<S, 0>(A, v, sel, gid, hist, big);
    if constexpr (S.table == 0) {
      // fold the slice's histogram into the lane-owned int128 accumulators
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      const unsigned long long w = hist[lane];
      hist[lane] = 0;
      if (lane_mm) {
        acc_lo = w > acc_lo ? w : acc_lo;
      } else {
        const uint64_t s = acc_lo + w;
        // count words are non-negative; sum words are signed 64-bit partials
        const bool is_count = lk == 0;
        acc_hi += (int64_t)(s < acc_lo) - ((!is_count && (int64_t)w < 0) ? 1 : 0);
        acc_lo = s;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  // ---- flush to the result ----
  if constexpr (S.table == 0) {
    // workgroup reduction of the 4 waves' lane accumulators, then device atomics
    __syncthreads();
    unsigned long long* red = lmem + 4 * kHistWords;  // [wave][lane] (lo, hi)
    red[(wave * 64 + lane) * 2] = acc_lo;
    red[(wave * 64 + lane) * 2 + 1] = (unsigned long long)acc_hi;
    __syncthreads();
    if (threadIdx.x < A.words) {
      const uint32_t g = threadIdx.x / (1 + S.naggs), a = threadIdx.x % (1 + S.naggs);
      const bool mm = kMM != 0 && a > 0 && ((kMM >> (3 * (a - 1))) & 7) != 0;
      uint64_t lo = 0;
      int64_t hi = 0;
      for (int w = 0; w < kBlock / kWave; ++w) {
        const uint64_t l2 = red[(w * 64 + threadIdx.x) * 2];
        if (mm) {
          lo = l2 > lo ? l2 : lo;
          continue;
        }
        const uint64_t s = lo + l2;
        hi += (int64_t)red[(w * 64 + threadIdx.x) * 2 + 1] + (int64_t)(s < lo);
        lo = s;
      }
      if (a == 0) {
        if (lo) atomicAdd(&A.out_count[g], (unsigned long long)lo);
      } else if (mm) {
        if (lo) atomicMax(&A.out_sums[((uint64_t)g * S.naggs + (a - 1)) * 2], lo);
      } else {
        atomic_add128(&A.out_sums[((uint64_t)g * S.naggs + (a - 1)) * 2], lo, hi);
      }
    }
  } else {
    __syncthreads();
    for (uint64_t g = threadIdx.x; g < A.ngroups; g += kBlock) {
      const unsigned long long c = lmem[g];
      if (!c) continue;
      atomicAdd(&A.out_count[g], c);
      for (int a = 0; a < S.naggs; ++a) {
        const uint64_t i = A.ngroups + (g * S.naggs + a) * 2;
        if (kMM != 0 && ((kMM >> (3 * a)) & 7) != 0) {
          if (lmem[i]) atomicMax(&A.out_sums[(g * S.naggs + a) * 2], lmem[i]);
        } else {
          atomic_add128(&A.out_sums[(g * S.naggs + a) * 2], lmem[i], (int64_t)lmem[i + 1]);
        }
      }
    }
  }
  if (oor && A.out_oor) atomicAdd(A.out_oor, (unsigned long long)oor);
}


// Single-query kernel on the wave-level blocks of wave_agg.h.  With chunk == 0
// wave w of the grid takes slices w, w + nwaves, ... (static interleave: all
// waves sweep the table together, no atomics); with chunk > 0 the waves take
// chunks of `chunk` slices from a device counter (a slow wave takes fewer).  A
// wave accumulates in its replicated LDS histogram and flushes into the result
// every 2^20 rows and at the end.  The LDS region of a wave is words << rshift
// words (dynamic shared memory of 4 such regions per workgroup).
template <Shape S>
__device__ __forceinline__ void wave_rows(const SpecArgs& A, uint64_t s, uint64_t min_len,
                                          unsigned long long* reg, uint32_t rshift,
                                          uint32_t& oor, uint64_t& since) {
  if (since + kSliceRows > kFlushRows) {
    wave_flush(A.out_count, A.out_sums, A.words, S.naggs, reg, rshift, shape_minmax_kinds(S));
    since = 0;
  }
  SliceVals<S> v;
  wave_load<S>(A, s, min_len, v);
  wave_slice<S>(A, s, v, reg, rshift, oor);
  since += kSliceRows;
}

template <Shape S>
__device__ __forceinline__ void wave_body(const SpecArgs& A, uint32_t rshift, uint32_t chunk,
                                          unsigned long long* counter) {
  extern __shared__ __align__(16) unsigned long long lreg[];
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  const uint32_t region = A.words << rshift;
  unsigned long long* reg = lreg + wave * region;
  for (uint32_t w = lane; w < region; w += 64) reg[w] = 0;
  __builtin_amdgcn_wave_barrier();
  uint64_t min_len = A.col_len[0];
#pragma unroll
  for (int s = 1; s < S.nslots; ++s) min_len = A.col_len[s] < min_len ? A.col_len[s] : min_len;
  uint32_t oor = 0;
  uint64_t since = 0;  // rows added to the region since its last flush
  if (chunk == 0) {
    const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
    for (uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + wave; s < A.nslices; s += nwaves)
      wave_rows<S>(A, s, min_len, reg, rshift, oor, since);
  } else {
    uint64_t k = grab_chunk(counter);
    while (k * chunk < A.nslices) {
      const uint64_t c0 = k * chunk;
      const uint64_t c1 = c0 + chunk < A.nslices ? c0 + chunk : A.nslices;
      const uint64_t next = grab_chunk(counter);  // in flight while this chunk runs
      for (uint64_t s = c0; s < c1; ++s) wave_rows<S>(A, s, min_len, reg, rshift, oor, since);
      k = next;
    }
  }
  // workgroup flush: every wave folds its copies, then one thread per word adds
  // the 4 waves' totals (as int128) into the result
  constexpr uint32_t kMM = shape_minmax_kinds(S);
  constexpr uint32_t W1 = 1 + S.naggs;
  wave_fold_copies(reg, A.words, rshift, W1, kMM);
  __syncthreads();
  for (uint32_t w = threadIdx.x; w < A.words; w += kBlock) {
    const uint32_t k = w % W1;
    const bool is_count = k == 0;
    const bool mm = kMM != 0 && k > 0 && ((kMM >> (3 * (k - 1))) & 7) != 0;
    uint64_t lo = 0;
    int64_t hi = 0;
    for (int v = 0; v < kBlock / kWave; ++v) {
      const uint64_t x = lreg[v * region + w];
      if (mm) {
        lo = x > lo ? x : lo;
        continue;
      }
      const uint64_t t = lo + x;
      hi += (int64_t)(t < lo) - ((!is_count && (int64_t)x < 0) ? 1 : 0);
      lo = t;
    }
    if (lo || hi) add_result_word(A.out_count, A.out_sums, w, S.naggs, lo, hi, kMM);
  }
  if (oor && A.out_oor) atomicAdd(A.out_oor, (unsigned long long)oor);
}

}  // namespace tgp
