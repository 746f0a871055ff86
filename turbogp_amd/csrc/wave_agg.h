// Wave-level scan-aggregate building blocks shared by the batched executor
// (batch.hip) and the single-query kernels (spec.hip).
//
// A wave works on 256-row slices (spec.h) of one query and accumulates COUNT(*)
// and the 64-bit partial sums per group in a wave-private LDS region, updated
// with 64-bit LDS atomics.  The region holds R = 2^rshift copies of the
// query's histogram (as many as fit in kWaveWords words, at most 64) and lane l
// updates copy l mod R, so that few lanes of one atomic instruction hit the
// same word even when the query has two or three groups.  The region is flushed
// into the query's exact int128 result with device atomics at least every
// kFlushRows rows: a copy then holds partial sums of at most 2^20 values of
// magnitude < 2^43, so the 64-bit sum of all copies is exact.  A slice with a
// larger value adds its values straight into the int128 result instead.
#pragma once

#include "spec.h"

namespace tgp {

constexpr int kWaveWords = 1024;             // per-wave region: R copies of groups * (1 + naggs)
constexpr uint64_t kFlushRows = 1ull << 20;  // rows between flushes
constexpr int kWaveBigBits = 43;             // |value| < 2^43 on the LDS path

// log2 of the number of histogram copies for a query of `words` words
// (groups * (1 + naggs)), or -1 when one copy does not fit.
inline int wave_rshift(uint64_t words) {
  if (words == 0 || words > (uint64_t)kWaveWords) return -1;
  int r = 6;
  while (r > 0 && (words << r) > (uint64_t)kWaveWords) --r;
  return r;
}

// Folds value v into m: (m >> 12) != 0 afterwards when some folded value lies
// outside [-2^43, 2^43).  v is inside exactly when its high word plus 2^11 is
// below 2^12 as an unsigned 32-bit value (2^43 has a zero low word, so the low
// word never carries into it): two plain VALU ops per value, where a 64-bit
// compare per value costs a compare, a hazard wait and a select.
static_assert(kWaveBigBits == 43, "big_fold tests the high word against 2^(43-32)");
__device__ __forceinline__ void big_fold(uint32_t& m, uint64_t v) {
  m |= (uint32_t)(v >> 32) + (1u << (kWaveBigBits - 32));
}
__device__ __forceinline__ bool big_any(uint32_t m) { return (m >> (kWaveBigBits - 31)) != 0; }

// One slice of a query of shape S into the wave-private region `reg`: the
// values `v` of the slice starting at query-relative slice `slice` are loaded.
template <Shape S, bool NT = true>
__device__ __forceinline__ void wave_load(const SpecArgs& A, uint64_t slice, uint64_t min_len,
                                          SliceVals<S>& v) {
  const uint64_t slice0 = A.base0 + slice * kSliceRows;
  const uint64_t row = slice0 + (threadIdx.x & 63) * kSR;
  load_all<S, 0, NT>(A, row, slice0 + kSliceRows <= min_len, v);
}

template <Shape S>
__device__ __forceinline__ void wave_slice(const SpecArgs& A, uint64_t slice,
                                            const SliceVals<S>& v, unsigned long long* reg,
                                            uint32_t rshift, uint32_t& oor) {
  const uint32_t lane = threadIdx.x & 63;
  // this lane's copy of the histogram: word w at reg[(w << rshift) + copy]
  unsigned long long* mine = reg + (lane & ((1u << rshift) - 1));
  const uint64_t row = A.base0 + slice * kSliceRows + lane * kSR;
  int32_t gid[kSR];
  const uint32_t sel = eval_slice<S>(A, row, v, gid, oor);
  if (!__ballot(sel != 0)) return;
  constexpr int W1 = 1 + S.naggs;
  // word w of the lane's copy is mine[w << rshift]: a row's words start at
  // rb = (gid * W1) << rshift and aggregate a's word is rb + ((1 + a) << rshift),
  // the second term the same for the whole wave
  uint32_t rb[kSR];
#pragma unroll
  for (int r = 0; r < kSR; ++r) {
    rb[r] = (uint32_t)(gid[r] * W1) << rshift;
    if ((sel >> r) & 1) atomicAdd(&mine[rb[r]], 1ull);
  }
  static_for<S.naggs>([&](auto AI) {
    constexpr int a = decltype(AI)::value;
    const uint32_t off = (uint32_t)(1 + a) << rshift;
    uint64_t val[kSR];
    agg_values<S, a>(A, v, val);
    if constexpr (is_minmax(S.kind[a])) {  // the largest order code; no exact path needed
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1)
          atomicMax(&mine[rb[r] + off], (unsigned long long)mm_encode(S.kind[a], val[r]));
      return;
    }
    // every row of the slice is tested, selected or not: a large unselected
    // value only sends the slice to the exact path
    uint32_t m = 0;
#pragma unroll
    for (int r = 0; r < kSR; ++r) big_fold(m, val[r]);
    if (__ballot(big_any(m))) {
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1)
          atomic_add128_i64(&A.out_sums[((uint64_t)gid[r] * S.naggs + a) * 2], (int64_t)val[r]);
    } else {
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1) atomicAdd(&mine[rb[r] + off], (unsigned long long)val[r]);
    }
  });
}

// True when word w of a region of groups x (count, aggregates) is a MIN / MAX word
// (kinds: 3 bits per aggregate, a plan's minmax_kinds; W1 = 1 + aggregates).
__device__ __forceinline__ bool minmax_word(uint32_t w, uint32_t W1, uint32_t kinds) {
  if (!kinds) return false;
  const uint32_t k = w % W1;
  return k > 0 && ((kinds >> (3 * (k - 1))) & 7) != 0;
}

// Sums the R = 2^rshift copies of every word of the wave region in place (the
// largest for a MIN / MAX word): afterwards reg[w] (w < words) holds word w's
// 64-bit total and the rest of the region is zero.  Lanes read consecutive words
// (no bank conflicts); the copies of a word then sit in R consecutive lanes and
// are combined with a butterfly.  Every copy holds partials of at most 2^20
// values of magnitude < 2^43, so the 64-bit total is exact.
__device__ __forceinline__ void wave_fold_copies(unsigned long long* reg, uint32_t words,
                                                 uint32_t rshift, uint32_t W1 = 1,
                                                 uint32_t kinds = 0) {
  const uint32_t lane = threadIdx.x & 63;
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const uint32_t T = words << rshift, R = 1u << rshift;
  // sums of iteration `base` go to words below (base + 64) / R <= base + 64, i.e.
  // never to a word a later iteration still has to read
  for (uint32_t base = 0; base < T; base += 64) {
    const uint32_t idx = base + lane;
    unsigned long long x = idx < T ? reg[idx] : 0ull;
    if (kinds && minmax_word(idx >> rshift, W1, kinds)) {
      for (uint32_t m = 1; m < R; m <<= 1) {
        const unsigned long long y = __shfl_xor(x, (int)m);
        x = y > x ? y : x;
      }
    } else {
      for (uint32_t m = 1; m < R; m <<= 1) x += __shfl_xor(x, (int)m);
    }
    if (idx < T) reg[idx] = 0;
    if ((lane & (R - 1)) == 0 && idx < T) reg[idx >> rshift] = x;
  }
  __builtin_amdgcn_wave_barrier();
}

// Adds word w's total (a count when w % (1 + naggs) == 0, else a signed 64-bit
// partial sum of aggregate w % (1 + naggs) - 1, or a MIN / MAX code in lo) given as
// int128 (lo, hi) into the query's result.
__device__ __forceinline__ void add_result_word(unsigned long long* out_count,
                                                unsigned long long* out_sums, uint32_t w,
                                                int naggs, uint64_t lo, int64_t hi,
                                                uint32_t kinds = 0) {
  const uint32_t W1 = 1 + naggs;
  const uint32_t g = w / W1, k = w % W1;
  if (k == 0) atomicAdd(&out_count[g], (unsigned long long)lo);
  else if (minmax_word(w, W1, kinds)) atomicMax(&out_sums[((uint64_t)g * naggs + (k - 1)) * 2], lo);
  else atomic_add128(&out_sums[((uint64_t)g * naggs + (k - 1)) * 2], lo, hi);
}

// Adds the wave region (`words` words, groups x (count, naggs sums), in 2^rshift
// copies) into the query's exact result and clears it.
__device__ __forceinline__ void wave_flush(unsigned long long* out_count,
                                           unsigned long long* out_sums, uint32_t words,
                                           int naggs, unsigned long long* reg, uint32_t rshift,
                                           uint32_t kinds = 0) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t W1 = 1 + naggs;
  wave_fold_copies(reg, words, rshift, W1, kinds);
  for (uint32_t w = lane; w < words; w += 64) {
    const unsigned long long x = reg[w];
    reg[w] = 0;
    if (!x) continue;
    const bool is_count = w % W1 == 0;
    const bool signed_word = !is_count && !minmax_word(w, W1, kinds);
    add_result_word(out_count, out_sums, w, naggs, x,
                    (signed_word && (int64_t)x < 0) ? -1 : 0, kinds);
  }
  __builtin_amdgcn_wave_barrier();
}

// Takes `n` consecutive work indices from a device counter (one atomic per
// wave); returns the first.
__device__ __forceinline__ uint64_t grab_chunks(unsigned long long* counter, uint32_t n) {
  unsigned long long k = 0;
  if ((threadIdx.x & 63) == 0) k = atomicAdd(counter, (unsigned long long)n);
  return __builtin_amdgcn_readfirstlane((uint32_t)k) |
         ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(k >> 32)) << 32);
}

// Takes the next chunk index from a device counter (one atomic per wave).
__device__ __forceinline__ uint64_t grab_chunk(unsigned long long* counter) {
  unsigned long long k = 0;
  if ((threadIdx.x & 63) == 0) k = atomicAdd(counter, 1ull);
  return __builtin_amdgcn_readfirstlane((uint32_t)k) |
         ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(k >> 32)) << 32);
}

}  // namespace tgp
