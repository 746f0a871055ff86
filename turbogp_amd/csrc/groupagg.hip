// Fused filter + GROUP BY + exact aggregation.
//
// Reference path replaced (query_interpreter):
//   aggregate.rs:507-697   execute_grouped: mask -> iter_set_bits -> per-chunk
//                          FxHashMap<keyhash, Vec<usize>> -> per group sum_vec
//   aggregate.rs:1485-1792 sum_vec: exact i64 for SUM(col), f64 for expressions
//   accum_grouped.rs:55-409 try_accumulator_grouped: wrapping u64 accumulators
//   aggregate.rs:1055-1289 scalar (no GROUP BY) aggregates
// Here one pass reads each referenced column once at its native width, the
// group key is a dense composite id (the reference keys groups by a combined
// 64-bit hash, accum_grouped.rs:213-218; for dense small keys both identify
// the same groups), and every sum is an exact 128-bit integer.
//
// Accumulation strategies (chosen by the planner from the group count):
//   Reg1 / Reg8  per-thread registers, a wave64 shuffle reduction and one set
//                of 64-bit device atomics per workgroup (<= 8 groups);
//   Lds          a per-workgroup table in LDS updated with 64-bit LDS atomics,
//                flushed to the result with device atomics;
//   Global       64-bit device atomics straight into the result table.
// int128 sums are kept as (low word, high word); see atomic_add128.
#include "plan.h"

namespace tgp {

// ---- register strategy ----------------------------------------------------------------
template <int GMAX, int AMAX>
__global__ void __launch_bounds__(kBlock, 2) k_agg_reg(DevPlan P) {
  extern __shared__ __align__(16) uint8_t lds[];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = wave_index();
  uint32_t cnt[GMAX];
  uint64_t lo[GMAX][AMAX];
  int32_t hi[GMAX][AMAX];
#pragma unroll
  for (int g = 0; g < GMAX; ++g) {
    cnt[g] = 0;
#pragma unroll
    for (int a = 0; a < AMAX; ++a) { lo[g][a] = 0; hi[g][a] = 0; }
  }
  uint32_t oor = 0;
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    uint32_t sel = eval_select(P, buf, row0);
    if (__ballot(sel != 0) == 0) return;  // nothing selected in this wave's rows
    int64_t gid[kRows];
    if (P.nparts) {
      eval_groups(P, buf, gid);
    } else {
#pragma unroll
      for (int r = 0; r < kRows; ++r) gid[r] = 0;
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      if (!((sel >> r) & 1)) continue;
      if (gid[r] < 0 || gid[r] >= GMAX) {
        ++oor;
        sel &= ~(1u << r);
        continue;
      }
#pragma unroll
      for (int g = 0; g < GMAX; ++g) cnt[g] += gid[r] == g;
    }
#pragma unroll
    for (int a = 0; a < AMAX; ++a) {
      if (a >= P.naggs) break;
      uint64_t v[kRows];
      eval_agg(P.aggs[a], buf, v);
      const int32_t mk = (P.minmax_kinds >> (3 * a)) & 7;
      if (mk) {
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          if (!((sel >> r) & 1)) continue;
          const uint64_t c = mm_encode(mk, v[r]);
#pragma unroll
          for (int g = 0; g < GMAX; ++g)
            if (gid[r] == g && c > lo[g][a]) lo[g][a] = c;
        }
        continue;
      }
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        if (!((sel >> r) & 1)) continue;
#pragma unroll
        for (int g = 0; g < GMAX; ++g)
          if (gid[r] == g) add128(lo[g][a], hi[g][a], (int64_t)v[r]);
      }
    }
  });
  // ---- flush: wave reduction, then one workgroup reduction through LDS ----
  constexpr int W = 1 + 2 * AMAX;  // words per group: count, (lo, hi) per agg
  unsigned long long* red = reinterpret_cast<unsigned long long*>(lds);
#pragma unroll
  for (int g = 0; g < GMAX; ++g) {
    unsigned long long c = cnt[g];
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) red[(wave * GMAX + g) * W] = c;
#pragma unroll
    for (int a = 0; a < AMAX; ++a) {
      unsigned long long l = lo[g][a];
      long long h = hi[g][a];
      if ((P.minmax_kinds >> (3 * a)) & 7) {
        for (int off = 32; off > 0; off >>= 1) {
          const unsigned long long l2 = __shfl_xor(l, off);
          l = l2 > l ? l2 : l;
        }
      } else {
        for (int off = 32; off > 0; off >>= 1) {
          const unsigned long long l2 = __shfl_xor(l, off);
          const long long h2 = __shfl_xor(h, off);
          const unsigned long long s = l + l2;
          h += h2 + (long long)(s < l);
          l = s;
        }
      }
      if (lane == 0) {
        red[(wave * GMAX + g) * W + 1 + 2 * a] = l;
        red[(wave * GMAX + g) * W + 2 + 2 * a] = (unsigned long long)h;
      }
    }
  }
  __syncthreads();
  constexpr int kWaves = kBlock / kWave;
  for (uint32_t i = threadIdx.x; i < (uint32_t)(GMAX * (1 + AMAX)); i += kBlock) {
    const int g = i / (1 + AMAX), a = (int)(i % (1 + AMAX)) - 1;
    if ((uint64_t)g >= P.ngroups || a >= P.naggs) continue;
    if (a < 0) {
      unsigned long long c = 0;
      for (int w = 0; w < kWaves; ++w) c += red[(w * GMAX + g) * W];
      if (c) atomicAdd(&P.out_count[g], c);
    } else if ((P.minmax_kinds >> (3 * a)) & 7) {
      unsigned long long l = 0;
      for (int w = 0; w < kWaves; ++w) {
        const unsigned long long l2 = red[(w * GMAX + g) * W + 1 + 2 * a];
        l = l2 > l ? l2 : l;
      }
      if (l) atomicMax(&P.out_sums[((uint64_t)g * P.naggs + a) * 2], l);
    } else {
      unsigned long long l = 0;
      long long h = 0;
      for (int w = 0; w < kWaves; ++w) {
        const unsigned long long l2 = red[(w * GMAX + g) * W + 1 + 2 * a];
        const unsigned long long s = l + l2;
        h += (long long)red[(w * GMAX + g) * W + 2 + 2 * a] + (long long)(s < l);
        l = s;
      }
      atomic_add128(&P.out_sums[((uint64_t)g * P.naggs + a) * 2], l, h);
    }
  }
  if (oor && P.out_oor) atomicAdd(P.out_oor, (unsigned long long)oor);
}

// ---- table strategies (LDS table or global table) ----------------------------------------
// Where a table's words live.  An LDS table holds group g's words (the count, then
// per aggregate the low and high split sums or a MIN/MAX code) at g * W + k, every
// word in 2^rs copies (lane l adds to copy l mod 2^rs, so that the lanes of one
// atomic rarely meet on a word when groups are few); the result table keeps the
// result's layout (count[g], sums[(g * naggs + a) * 2]).
struct TableRef {
  unsigned long long* cnt;  // LDS: the table; result: the counts
  unsigned long long* sum;  // result: the sums
  uint32_t W, rs, copy;
  bool lds;
  __device__ __forceinline__ unsigned long long* count_at(uint64_t g) const {
    return lds ? cnt + ((g * W) << rs) + copy : cnt + g;
  }
  __device__ __forceinline__ unsigned long long* sum_at(uint64_t g, int a) const {
    return lds ? cnt + ((g * W + 1 + 2 * a) << rs) + copy : sum + (g * ((W - 1) / 2) + a) * 2;
  }
  __device__ __forceinline__ uint32_t hi_step() const { return lds ? 1u << rs : 1u; }
};

__device__ __forceinline__ TableRef lds_table(const DevPlan& P, unsigned long long* base) {
  return TableRef{base, nullptr, 1 + 2 * (uint32_t)P.naggs, P.lds_rshift,
                  (threadIdx.x & 63) & ((1u << P.lds_rshift) - 1), true};
}

__device__ __forceinline__ TableRef result_table(const DevPlan& P) {
  return TableRef{P.out_count, P.out_sums, 1 + 2 * (uint32_t)P.naggs, 0, 0, false};
}

__device__ __forceinline__ uint64_t lds_table_words(const DevPlan& P) {
  return (P.ngroups * (1 + 2 * (uint64_t)P.naggs)) << P.lds_rshift;
}

// One tile's rows into a table of counts and split sums / MIN-MAX codes (sum of
// the low 32 bits, sum of the high 32 bits, with non-returning atomics).  A table
// holds fewer than 2^32 rows' additions per word: an LDS table is flushed per
// workgroup (or per batch item), the result holds the split sums of fewer than
// 2^32 rows until the finalisation turns them into int128 words; without `split`
// (the result of a range of 2^32 rows or more) the sums are int128 words.
__device__ __forceinline__ void table_tile(const DevPlan& P, const uint8_t* buf, uint64_t row0,
                                           const TableRef& T, uint32_t& oor, bool split) {
  uint32_t sel = eval_select(P, buf, row0);
  if (__ballot(sel != 0) == 0) return;
  int64_t gid[kRows];
  eval_groups(P, buf, gid);
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (!((sel >> r) & 1)) continue;
    if (gid[r] < 0) {
      ++oor;
      sel &= ~(1u << r);
      continue;
    }
    atomicAdd(T.count_at((uint64_t)gid[r]), 1ull);
  }
  const uint32_t step = T.hi_step();
  for (int a = 0; a < P.naggs; ++a) {
    uint64_t v[kRows];
    eval_agg(P.aggs[a], buf, v);
    const int32_t mk = (P.minmax_kinds >> (3 * a)) & 7;
    if (mk) {
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if ((sel >> r) & 1)
          atomicMax(T.sum_at((uint64_t)gid[r], a), (unsigned long long)mm_encode(mk, v[r]));
      continue;
    }
    if (split) {
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if ((sel >> r) & 1) {
          unsigned long long* p = T.sum_at((uint64_t)gid[r], a);
          atomicAdd(p, (unsigned long long)(uint32_t)v[r]);
          const int64_t h = (int64_t)v[r] >> 32;
          if (h) atomicAdd(p + step, (unsigned long long)h);
        }
    } else {
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if ((sel >> r) & 1) atomic_add128_i64(T.sum_at((uint64_t)gid[r], a), (int64_t)v[r]);
    }
  }
}

// Adds an LDS table (copies summed, split sums rebuilt) into the result's words.
__device__ __forceinline__ void table_flush(const DevPlan& P, const unsigned long long* tab) {
  const uint32_t W = 1 + 2 * (uint32_t)P.naggs, rs = P.lds_rshift, R = 1u << rs;
  for (uint64_t g = threadIdx.x; g < P.ngroups; g += kBlock) {
    const unsigned long long* w0 = tab + ((g * W) << rs);
    unsigned long long c = 0;
    for (uint32_t k = 0; k < R; ++k) c += w0[k];
    if (!c) continue;
    atomicAdd(&P.out_count[g], c);
    for (int a = 0; a < P.naggs; ++a) {
      const unsigned long long* wl = w0 + ((uint64_t)(1 + 2 * a) << rs);
      const uint64_t i = (g * P.naggs + a) * 2;
      if ((P.minmax_kinds >> (3 * a)) & 7) {
        unsigned long long m = 0;
        for (uint32_t k = 0; k < R; ++k) m = wl[k] > m ? wl[k] : m;
        atomicMax(&P.out_sums[i], m);
      } else {
        unsigned long long slo = 0, shi = 0;
        for (uint32_t k = 0; k < R; ++k) {
          slo += wl[k];
          shi += wl[k + R];
        }
        uint64_t lo;
        int64_t hi;
        split_to_i128(slo, (int64_t)shi, &lo, &hi);
        atomic_add128(&P.out_sums[i], lo, hi);
      }
    }
  }
}

// USE_LDS: an LDS table of split sums; else the result table, of split sums when
// SPLIT (fewer than 2^32 rows), of int128 words otherwise.
template <bool USE_LDS, bool SPLIT>
__global__ void __launch_bounds__(kBlock, 2) k_agg_table(DevPlan P) {
  extern __shared__ __align__(16) uint8_t lds[];
  unsigned long long* tab = reinterpret_cast<unsigned long long*>(lds + 2 * P.stage_bytes);
  if (USE_LDS) {
    const uint64_t words = lds_table_words(P);
    for (uint64_t i = threadIdx.x; i < words; i += kBlock) tab[i] = 0;
    // made visible by the barrier at the top of the first tile
  }
  const TableRef T = USE_LDS ? lds_table(P, tab) : result_table(P);
  uint32_t oor = 0;
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    table_tile(P, buf, row0, T, oor, USE_LDS || SPLIT);
  });
  if (USE_LDS) table_flush(P, tab);
  if (oor && P.out_oor) atomicAdd(P.out_oor, (unsigned long long)oor);
}

// ---- radix-partitioned aggregation (kAggPart) ---------------------------------------------
// A table too large for LDS costs one device atomic per row for the count and one or
// two per aggregate, all at random words.  kAggPart instead sorts the rows into
// partitions of 2^part_shift consecutive groups and aggregates each partition in an
// LDS table:
//   (1) k_part_hist_tiles: through a plan that stages only the predicate and group
//       columns, workgroup h of k G (k per scatter workgroup, for occupancy: its LDS
//       is small) counts its selected rows per partition p = gid >> part_shift into
//       mat[p * k G + (h mod G) k + h / G]: the k counts of the tiles scatter
//       workgroup w will take are adjacent;
//   (2) an exclusive scan of mat (partition-major) gives every (partition, scatter
//       workgroup) pair its range of records;
//   (3) k_part_scatter_tiles: the same tiles on the same G workgroups (so the
//       counts match) through the full plan; each selected row takes a rank among
//       its tile's rows of the same partition (LDS counter) and is written at the
//       pair's cursor plus that rank, as the local group (16 bits) and one 64-bit
//       value per aggregate (structure of arrays);
//   (4) k_part_agg_recs: workgroup c takes records [c * kPartChunk, ...), a run of
//       whole or partial partitions since the records are sorted by partition,
//       aggregates each partition's run in an LDS table of 2^part_shift groups
//       (count, split sums or MIN/MAX codes) and adds its nonzero groups to the
//       result with carry-exact device atomics.
// The order of records inside a range varies from run to run; the totals do not.
constexpr int kPartAggBlock = 1024;

__global__ void __launch_bounds__(kBlock) k_part_hist_tiles(DevPlan P, uint32_t shift,
                                                            uint32_t nparts, uint32_t k,
                                                            unsigned long long* __restrict__ mat) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t* h = reinterpret_cast<uint32_t*>(lds + 2 * P.stage_bytes);
  for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) h[p] = 0;
  // made visible by the barrier at the top of the first tile
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    const uint32_t sel = eval_select(P, buf, row0);
    if (__ballot(sel != 0) == 0) return;
    int64_t gid[kRows];
    eval_groups(P, buf, gid);
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      if (((sel >> r) & 1) && gid[r] >= 0) atomicAdd(&h[(uint64_t)gid[r] >> shift], 1u);
  });
  __syncthreads();
  const uint32_t G = gridDim.x / k;
  const uint64_t col = (uint64_t)(blockIdx.x % G) * k + blockIdx.x / G;
  for (uint32_t p = threadIdx.x; p < nparts; p += kBlock)
    mat[(uint64_t)p * gridDim.x + col] = h[p];
}

// mat: exclusive offsets after the scan (k columns per workgroup).  cap: the length
// of each record array.
__global__ void __launch_bounds__(kBlock) k_part_scatter_tiles(
    DevPlan P, uint32_t shift, uint32_t nparts, uint32_t k,
    const unsigned long long* __restrict__ mat, uint64_t cap, uint16_t* __restrict__ rgrp,
    uint64_t* __restrict__ rval) {
  extern __shared__ __align__(16) uint8_t lds[];
  unsigned long long* cursor = reinterpret_cast<unsigned long long*>(lds + 2 * P.stage_bytes);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cursor + nparts);
  for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) {
    cursor[p] = mat[((uint64_t)p * gridDim.x + blockIdx.x) * k];
    cnt[p] = 0;
  }
  // made visible by the barrier at the top of the first tile
  const uint32_t lmask = (1u << shift) - 1;
  uint32_t oor = 0;
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    const uint32_t sel = eval_select(P, buf, row0);
    if (!__syncthreads_or(sel != 0)) return;  // the same for the whole workgroup
    int64_t gid[kRows];
    eval_groups(P, buf, gid);
    uint32_t pr[kRows], rk[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      pr[r] = 0xffffffffu;
      rk[r] = 0;
      if (!((sel >> r) & 1)) continue;
      if (gid[r] < 0) {
        ++oor;
        continue;
      }
      pr[r] = (uint32_t)((uint64_t)gid[r] >> shift);
      rk[r] = atomicAdd(&cnt[pr[r]], 1u);
    }
    __syncthreads();  // every rank taken
    uint64_t dst[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      dst[r] = 0;
      if (pr[r] == 0xffffffffu) continue;
      dst[r] = cursor[pr[r]] + rk[r];
      rgrp[dst[r]] = (uint16_t)((uint64_t)gid[r] & lmask);
    }
    for (int a = 0; a < P.naggs; ++a) {
      uint64_t v[kRows];
      eval_agg(P.aggs[a], buf, v);
      uint64_t* out = rval + (uint64_t)a * cap;
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if (pr[r] != 0xffffffffu) out[dst[r]] = v[r];
    }
    __syncthreads();  // every cursor read
    // the row ranked first in its partition advances that partition's cursor; the
    // barrier at the top of the next tile orders this before the next ranks
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      if (pr[r] != 0xffffffffu && rk[r] == 0) {
        cursor[pr[r]] += cnt[pr[r]];
        cnt[pr[r]] = 0;
      }
  });
  if (oor && P.out_oor) atomicAdd(P.out_oor, (unsigned long long)oor);
}

// Aggregates one partition's run of n records (local groups grp[0..n), value k of
// record i = val(i, k)) in the LDS table tab of gp groups and adds its nonzero groups
// to the result at groups g0.. with carry-exact device atomics.  Called by the whole
// workgroup (kPartAggBlock threads); ends with the table free again.
template <typename Val>
__device__ __forceinline__ void agg_run(int32_t naggs, uint32_t minmax_kinds,
                                        unsigned long long* out_count,
                                        unsigned long long* out_sums, uint64_t g0, uint32_t gp,
                                        unsigned long long* tab, const uint16_t* grp, uint64_t n,
                                        Val&& val) {
  const uint32_t W = 1 + 2 * (uint32_t)naggs;
  for (uint32_t i = threadIdx.x; i < gp * W; i += kPartAggBlock) tab[i] = 0;
  __syncthreads();
  // U records per thread per step: their loads are issued before their atomics
  constexpr int U = 4;
  for (uint64_t i0 = threadIdx.x; i0 < n; i0 += (uint64_t)U * kPartAggBlock) {
    uint32_t g[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + (uint64_t)u * kPartAggBlock;
      ok[u] = i < n;
      g[u] = ok[u] ? grp[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (ok[u]) atomicAdd(tab + g[u] * W, 1ull);
    for (int k = 0; k < naggs; ++k) {
      uint64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = ok[u] ? val(i0 + (uint64_t)u * kPartAggBlock, k) : 0;
      const int32_t mk = (minmax_kinds >> (3 * k)) & 7;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        unsigned long long* w = tab + g[u] * W + 1 + 2 * k;
        if (mk) {
          atomicMax(w, (unsigned long long)mm_encode(mk, v[u]));
        } else {
          atomicAdd(w, (unsigned long long)(uint32_t)v[u]);
          const int64_t h = (int64_t)v[u] >> 32;
          if (h) atomicAdd(w + 1, (unsigned long long)h);
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t gl = threadIdx.x; gl < gp; gl += kPartAggBlock) {
    const unsigned long long* w = tab + gl * W;
    if (!w[0]) continue;
    const uint64_t g = g0 + gl;
    atomicAdd(&out_count[g], w[0]);
    for (int k = 0; k < naggs; ++k) {
      unsigned long long* o = &out_sums[(g * naggs + k) * 2];
      if ((minmax_kinds >> (3 * k)) & 7) {
        atomicMax(o, w[1 + 2 * k]);
      } else {
        uint64_t l;
        int64_t h;
        split_to_i128(w[1 + 2 * k], (int64_t)w[2 + 2 * k], &l, &h);
        atomic_add128(o, l, h);
      }
    }
  }
  __syncthreads();  // every flush read before the table is cleared again
}

// G: mat's columns (the histogram's workgroups).  LDS: 2^shift x (1 + 2 naggs) words.
// Results (zeroed) receive int128 sums and MIN/MAX codes.
__global__ void __launch_bounds__(kPartAggBlock) k_part_agg_recs(
    DevPlan P, uint32_t shift, uint32_t nparts, uint64_t G,
    const unsigned long long* __restrict__ mat, uint64_t cap,
    const uint16_t* __restrict__ rgrp, const uint64_t* __restrict__ rval) {
  extern __shared__ __align__(16) unsigned long long tab[];
  const uint64_t total = mat[(uint64_t)nparts * G];
  uint64_t lo = (uint64_t)blockIdx.x * kPartChunk;
  if (lo >= total) return;
  const uint64_t hi = lo + kPartChunk < total ? lo + kPartChunk : total;
  const uint32_t gp = 1u << shift;
  // the partition holding record lo: the last p with mat[p * G] <= lo
  uint32_t pa = 0, pb = nparts - 1;
  while (pa < pb) {
    const uint32_t m = (pa + pb + 1) >> 1;
    if (mat[(uint64_t)m * G] <= lo) pa = m;
    else pb = m - 1;
  }
  for (uint32_t p = pa; lo < hi && p < nparts; ++p) {
    const uint64_t pe = mat[(uint64_t)(p + 1) * G];
    const uint64_t end = pe < hi ? pe : hi;
    if (end <= lo) continue;
    agg_run(P.naggs, P.minmax_kinds, P.out_count, P.out_sums, (uint64_t)p << shift, gp, tab,
            rgrp + lo, end - lo, [&](uint64_t i, int k) { return rval[(uint64_t)k * cap + lo + i]; });
    lo = end;
  }
}

constexpr uint32_t kLdsMax = 160 * 1024;

template <typename K>
static int32_t allow_lds(K kern, uint32_t bytes) {
  if (bytes > 65536)
    TGP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
  return TGP_OK;
}

// ---- kAggPart over many queries at once (the batch's partitioned queries) -----------------
// The passes of kAggPart, each one persistent launch over the work of every such
// query: histogram items (query, tiles [t0, t1), its item j) through the query's
// histogram plan, one exclusive scan of all queries' counts, scatter items (the same
// items through the full plans), then aggregation chunks (query, 32K records).  A
// query's counts sit partition-major (mat[mat_base + p * nitems + j]), so after the
// scan its records are contiguous, sorted by partition, starting at mat[mat_base].

// Copies plan q into the workgroup's LDS plan (all threads; a barrier must follow).
__device__ __forceinline__ void load_plan(const DevPlan* __restrict__ plans, uint32_t q,
                                          DevPlan* sp) {
  const uint64_t* src = reinterpret_cast<const uint64_t*>(&plans[q]);
  uint64_t* dst = reinterpret_cast<uint64_t*>(sp);
  for (uint32_t i = threadIdx.x; i < sizeof(DevPlan) / 8; i += blockDim.x) dst[i] = src[i];
}

// Takes the next work item (whole workgroup); UINT32_MAX-free: returns the index.
__device__ __forceinline__ uint32_t next_item(unsigned long long* counter, uint32_t* s_item) {
  __syncthreads();  // the previous item is done; s_item and the LDS plan are free
  if (threadIdx.x == 0) *s_item = (uint32_t)atomicAdd(counter, 1ull);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*s_item);
}

__global__ void __launch_bounds__(kBlock) k_bpart_hist(const DevPlan* __restrict__ hplans,
                                                       const PartQuery* __restrict__ pqs,
                                                       const PartItem* __restrict__ items,
                                                       uint64_t nitems,
                                                       unsigned long long* counter,
                                                       uint32_t stage_max,
                                                       unsigned long long* __restrict__ mat) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ DevPlan sp;
  __shared__ uint32_t s_item;
  uint32_t* h = reinterpret_cast<uint32_t*>(lds + 2 * stage_max);
  for (;;) {
    const uint32_t it = next_item(counter, &s_item);
    if (it >= nitems) break;
    const PartItem I = items[it];
    const PartQuery Q = pqs[I.pq];
    load_plan(hplans, Q.hplan, &sp);
    for (uint32_t p = threadIdx.x; p < Q.nparts; p += kBlock) h[p] = 0;
    __syncthreads();
    const DevPlan& P = sp;
    stage_issue(P, I.t0, lds);
    for (uint32_t t = I.t0, i = 0; t < I.t1; ++t, ++i) {
      stage_wait();
      __syncthreads();  // tile t is in LDS; every wave is done with the other buffer
      uint8_t* cur = lds + (i & 1) * stage_max;
      uint8_t* nxt = lds + ((i + 1) & 1) * stage_max;
      if (t + 1 < I.t1) stage_issue(P, t + 1, nxt);
      const uint32_t sel = eval_select(P, cur, P.base0 + (uint64_t)t * kTileRows);
      if (__ballot(sel != 0) == 0) continue;
      int64_t gid[kRows];
      eval_groups(P, cur, gid);
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if (((sel >> r) & 1) && gid[r] >= 0) atomicAdd(&h[(uint64_t)gid[r] >> Q.shift], 1u);
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < Q.nparts; p += kBlock)
      mat[Q.mat_base + (uint64_t)p * Q.nitems + I.j] = h[p];
  }
}

__global__ void __launch_bounds__(kBlock) k_bpart_scatter(const DevPlan* __restrict__ plans,
                                                          const PartQuery* __restrict__ pqs,
                                                          const PartItem* __restrict__ items,
                                                          uint64_t nitems,
                                                          unsigned long long* counter,
                                                          uint32_t stage_max, uint32_t nparts_max,
                                                          const unsigned long long* __restrict__ mat,
                                                          uint16_t* __restrict__ grp,
                                                          uint64_t* __restrict__ val) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ DevPlan sp;
  __shared__ uint32_t s_item;
  unsigned long long* cursor = reinterpret_cast<unsigned long long*>(lds + 2 * stage_max);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cursor + nparts_max);
  for (;;) {
    const uint32_t it = next_item(counter, &s_item);
    if (it >= nitems) break;
    const PartItem I = items[it];
    const PartQuery Q = pqs[I.pq];
    load_plan(plans, Q.plan, &sp);
    const uint64_t rec0 = mat[Q.mat_base];  // the query's first record
    for (uint32_t p = threadIdx.x; p < Q.nparts; p += kBlock) {
      cursor[p] = mat[Q.mat_base + (uint64_t)p * Q.nitems + I.j] - rec0;
      cnt[p] = 0;
    }
    __syncthreads();
    const DevPlan& P = sp;
    const uint32_t lmask = (1u << Q.shift) - 1;
    uint16_t* qgrp = grp + Q.grp_base;
    uint64_t* qval = val + Q.val_base;
    uint32_t oor = 0;
    stage_issue(P, I.t0, lds);
    for (uint32_t t = I.t0, i = 0; t < I.t1; ++t, ++i) {
      stage_wait();
      __syncthreads();  // tile t is in LDS; every wave is done with the other buffer
      uint8_t* cur = lds + (i & 1) * stage_max;
      uint8_t* nxt = lds + ((i + 1) & 1) * stage_max;
      if (t + 1 < I.t1) stage_issue(P, t + 1, nxt);
      const uint32_t sel = eval_select(P, cur, P.base0 + (uint64_t)t * kTileRows);
      if (!__syncthreads_or(sel != 0)) continue;  // the same for the whole workgroup
      int64_t gid[kRows];
      eval_groups(P, cur, gid);
      uint32_t pr[kRows], rk[kRows];
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        pr[r] = 0xffffffffu;
        rk[r] = 0;
        if (!((sel >> r) & 1)) continue;
        if (gid[r] < 0) {
          ++oor;
          continue;
        }
        pr[r] = (uint32_t)((uint64_t)gid[r] >> Q.shift);
        rk[r] = atomicAdd(&cnt[pr[r]], 1u);
      }
      __syncthreads();  // every rank taken
      uint64_t dst[kRows];
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        dst[r] = 0;
        if (pr[r] == 0xffffffffu) continue;
        dst[r] = cursor[pr[r]] + rk[r];
        qgrp[dst[r]] = (uint16_t)((uint64_t)gid[r] & lmask);
      }
      for (int a = 0; a < P.naggs; ++a) {
        uint64_t v[kRows];
        eval_agg(P.aggs[a], cur, v);
#pragma unroll
        for (int r = 0; r < kRows; ++r)
          if (pr[r] != 0xffffffffu) qval[dst[r] * Q.naggs + a] = v[r];
      }
      __syncthreads();  // every cursor read
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if (pr[r] != 0xffffffffu && rk[r] == 0) {
          cursor[pr[r]] += cnt[pr[r]];
          cnt[pr[r]] = 0;
        }
      // the next tile's first barrier orders these before its ranks
    }
    if (oor && P.out_oor) atomicAdd(P.out_oor, (unsigned long long)oor);
  }
}

__global__ void __launch_bounds__(kPartAggBlock) k_bpart_agg(
    const DevPlan* __restrict__ plans, const PartQuery* __restrict__ pqs,
    const PartChunk* __restrict__ chunks, uint64_t nchunks, unsigned long long* counter,
    const unsigned long long* __restrict__ mat, const uint16_t* __restrict__ grp,
    const uint64_t* __restrict__ val) {
  extern __shared__ __align__(16) unsigned long long tab[];
  __shared__ uint32_t s_item;
  for (;;) {
    const uint32_t it = next_item(counter, &s_item);
    if (it >= nchunks) break;
    const PartChunk C = chunks[it];
    const PartQuery Q = pqs[C.pq];
    const DevPlan& P = plans[Q.plan];
    const uint64_t rec0 = mat[Q.mat_base];
    const uint64_t rend = mat[Q.mat_base + (uint64_t)Q.nparts * Q.nitems];
    uint64_t lo = rec0 + (uint64_t)C.c * kPartChunk;
    if (lo >= rend) continue;
    const uint64_t hi = lo + kPartChunk < rend ? lo + kPartChunk : rend;
    // the partition holding record lo: the last p with its first record <= lo
    uint32_t pa = 0, pb = Q.nparts - 1;
    while (pa < pb) {
      const uint32_t m = (pa + pb + 1) >> 1;
      if (mat[Q.mat_base + (uint64_t)m * Q.nitems] <= lo) pa = m;
      else pb = m - 1;
    }
    const uint16_t* qgrp = grp + Q.grp_base;
    const uint64_t* qval = val + Q.val_base;
    for (uint32_t p = pa; lo < hi && p < Q.nparts; ++p) {
      const uint64_t pe = p + 1 < Q.nparts ? mat[Q.mat_base + (uint64_t)(p + 1) * Q.nitems] : rend;
      const uint64_t end = pe < hi ? pe : hi;
      if (end <= lo) continue;
      const uint64_t l0 = lo - rec0;  // local index of the run's first record
      agg_run(P.naggs, P.minmax_kinds, P.out_count, P.out_sums, (uint64_t)p << Q.shift,
              1u << Q.shift, tab, qgrp + l0, end - lo,
              [&](uint64_t i, int k) { return qval[(l0 + i) * Q.naggs + k]; });
      lo = end;
    }
  }
}

int32_t launch_batch_part(tgp_ctx* ctx, const DevPlan* d_plans, const DevPlan* d_hplans,
                          const PartQuery* d_pq, const PartItem* d_items, uint64_t nitems,
                          const PartChunk* d_chunks, uint64_t nchunks, uint32_t hstage_max,
                          uint32_t stage_max, uint32_t nparts_max, uint32_t table_max,
                          unsigned long long* counters, unsigned long long* mat, uint64_t mat_n,
                          uint64_t* bsum, uint16_t* grp, uint64_t* val, hipStream_t s) {
  if (!nitems) return TGP_OK;
  const uint32_t lh = 2 * hstage_max + 4 * nparts_max;
  const uint32_t ls = 2 * stage_max + 12 * nparts_max;
  TGP_TRY(allow_lds(k_bpart_hist, lh));
  TGP_TRY(allow_lds(k_bpart_scatter, ls));
  TGP_TRY(allow_lds(k_bpart_agg, table_max));
  auto grid_of = [&](auto kern, uint32_t block, uint32_t lds, uint64_t units) -> unsigned {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    uint64_t g = (uint64_t)persistent_grid(ctx, per_cu);
    return (unsigned)(g < units ? g : units);
  };
  k_bpart_hist<<<grid_of(k_bpart_hist, kBlock, lh, nitems), kBlock, lh, s>>>(
      d_hplans, d_pq, d_items, nitems, counters, hstage_max, mat);
  TGP_HIP(hipGetLastError());
  TGP_TRY(scan_u64(s, reinterpret_cast<uint64_t*>(mat), mat_n, bsum));
  k_bpart_scatter<<<grid_of(k_bpart_scatter, kBlock, ls, nitems), kBlock, ls, s>>>(
      d_plans, d_pq, d_items, nitems, counters + 1, stage_max, nparts_max, mat, grp, val);
  TGP_HIP(hipGetLastError());
  if (nchunks) {
    k_bpart_agg<<<grid_of(k_bpart_agg, kPartAggBlock, table_max, nchunks), kPartAggBlock,
                  table_max, s>>>(d_plans, d_pq, d_chunks, nchunks, counters + 2, mat, grp, val);
    TGP_HIP(hipGetLastError());
  }
  return TGP_OK;
}

// ---- many queries of any form in one launch (the batch's generic items) -----------------
// Persistent workgroups take items (query, tiles [t0, t1)) from a counter, copy
// the item's plan into LDS, stage its tiles double-buffered and aggregate them
// into an LDS table (flushed after the item) or straight into the result.
__global__ void __launch_bounds__(kBlock) k_batch_generic(const DevPlan* __restrict__ plans,
                                                          const GenericItem* __restrict__ items,
                                                          uint64_t nitems,
                                                          unsigned long long* counter,
                                                          uint32_t stage_max) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ DevPlan sp;
  __shared__ uint32_t s_item;
  static_assert(sizeof(DevPlan) % 8 == 0, "plan copied as 64-bit words");
  unsigned long long* table = reinterpret_cast<unsigned long long*>(lds + 2 * stage_max);
  for (;;) {
    __syncthreads();  // the previous item is flushed; sp and s_item are free
    if (threadIdx.x == 0) s_item = (uint32_t)atomicAdd(counter, 1ull);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= nitems) break;
    const GenericItem I = items[it];
    {
      const uint64_t* src = reinterpret_cast<const uint64_t*>(&plans[I.q]);
      uint64_t* dst = reinterpret_cast<uint64_t*>(&sp);
      for (uint32_t i = threadIdx.x; i < sizeof(DevPlan) / 8; i += kBlock) dst[i] = src[i];
    }
    __syncthreads();
    const DevPlan& P = sp;
    if (I.mode == 1) {
      const uint64_t words = lds_table_words(P);
      for (uint64_t i = threadIdx.x; i < words; i += kBlock) table[i] = 0;
    }
    const TableRef T = I.mode == 1 ? lds_table(P, table) : result_table(P);
    uint32_t oor = 0;
    stage_issue(P, I.t0, lds);
    for (uint32_t t = I.t0, i = 0; t < I.t1; ++t, ++i) {
      stage_wait();
      __syncthreads();  // tile t is in LDS (and the table clear is visible)
      uint8_t* cur = lds + (i & 1) * stage_max;
      uint8_t* nxt = lds + ((i + 1) & 1) * stage_max;
      if (t + 1 < I.t1) stage_issue(P, t + 1, nxt);
      table_tile(P, cur, P.base0 + (uint64_t)t * kTileRows, T, oor, I.mode != 2);
    }
    __syncthreads();
    if (I.mode == 1) table_flush(P, table);
    if (oor && P.out_oor) atomicAdd(P.out_oor, (unsigned long long)oor);
  }
}

// Result word pair i = (group, aggregate): a MIN/MAX code becomes the value (low
// word; high word its sign extension for the signed kinds; 0 for groups without
// rows), and, when `split`, a split sum becomes int128 words.
__device__ __forceinline__ void finalize_pair(const unsigned long long* cnt,
                                              unsigned long long* sums, uint64_t i, uint64_t g,
                                              int32_t a, uint32_t kinds, bool split) {
  const int32_t mk = (kinds >> (3 * a)) & 7;
  if (mk) {
    const uint64_t v = cnt[g] ? mm_decode(mk, sums[2 * i]) : 0;
    const bool sgn = mk == TGP_AGG_MIN || mk == TGP_AGG_MAX;
    sums[2 * i] = v;
    sums[2 * i + 1] = sgn && (int64_t)v < 0 ? ~0ull : 0ull;
  } else if (split) {
    uint64_t lo;
    int64_t hi;
    split_to_i128(sums[2 * i], (int64_t)sums[2 * i + 1], &lo, &hi);
    sums[2 * i] = lo;
    sums[2 * i + 1] = (uint64_t)hi;
  }
}

__global__ void __launch_bounds__(256) k_agg_finalize(const unsigned long long* __restrict__ cnt,
                                                      unsigned long long* __restrict__ sums,
                                                      uint64_t ngroups, int32_t naggs,
                                                      uint32_t kinds, int32_t split) {
  const uint64_t n = ngroups * (uint64_t)naggs;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = i / (uint32_t)naggs;
    finalize_pair(cnt, sums, i, g, (int32_t)(i - g * (uint32_t)naggs), kinds, split != 0);
  }
}

// The batch's generic queries finalised: blockIdx.y picks list entry y = plan
// index | split << 31 (the plan's sums went straight into the result, split).
__global__ void __launch_bounds__(256) k_batch_finalize(const DevPlan* __restrict__ plans,
                                                        const uint32_t* __restrict__ list) {
  const uint32_t e = list[blockIdx.y];
  const DevPlan& P = plans[e & 0x7fffffffu];
  const uint64_t n = P.ngroups * (uint64_t)P.naggs;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = i / (uint32_t)P.naggs;
    finalize_pair(P.out_count, P.out_sums, i, g, (int32_t)(i - g * (uint32_t)P.naggs),
                  P.minmax_kinds, (e >> 31) != 0);
  }
}

int32_t finalize_minmax(hipStream_t s, unsigned long long* cnt, unsigned long long* sums,
                        uint64_t ngroups, int32_t naggs, uint32_t kinds) {
  if (!kinds || !ngroups || naggs <= 0) return TGP_OK;
  const uint64_t items = ngroups * (uint64_t)naggs;
  const uint64_t grid = std::min<uint64_t>((items + 255) / 256, 65536);
  hipLaunchKernelGGL(k_agg_finalize, dim3((uint32_t)grid), dim3(256), 0, s, cnt, sums, ngroups,
                     naggs, kinds, 0);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t launch_batch_generic(tgp_ctx* ctx, const DevPlan* d_plans, const GenericItem* d_items,
                             uint64_t nitems, unsigned long long* counter, uint32_t stage_max,
                             uint32_t table_max, const uint32_t* d_minmax, uint32_t nminmax,
                             uint64_t minmax_words, hipStream_t s) {
  if (nitems) {
    const uint32_t lds = 2 * stage_max + table_max;
    if (lds > 65536)
      TGP_HIP(hipFuncSetAttribute((const void*)k_batch_generic,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_generic, kBlock, lds));
    if (per_cu < 1) per_cu = 1;
    uint64_t grid = (uint64_t)persistent_grid(ctx, per_cu);
    if (grid > nitems) grid = nitems;
    k_batch_generic<<<dim3((unsigned)grid), dim3(kBlock), lds, s>>>(d_plans, d_items, nitems,
                                                                    counter, stage_max);
    TGP_HIP(hipGetLastError());
  }
  if (nminmax) {
    const uint64_t gx = std::min<uint64_t>((minmax_words + 255) / 256, 1024);
    k_batch_finalize<<<dim3((unsigned)gx, nminmax), dim3(256), 0, s>>>(d_plans, d_minmax);
    TGP_HIP(hipGetLastError());
  }
  return TGP_OK;
}

int32_t try_spec(tgp_ctx* ctx, const tgp_agg_query* q, const tgp_group_result* out,
                 hipStream_t st, bool* launched);
int32_t spec_shape_index(const tgp_agg_query* q);

// The histogram's plan: q with only its predicate and group-key columns (the same
// row range, selection bitmap, tiles and group ids) and no aggregates.
int32_t plan_part_hist(const tgp_agg_query* q, DevPlan* H) {
  tgp_agg_query rq = *q;
  int map[TGP_MAX_COLS];
  for (int c = 0; c < TGP_MAX_COLS; ++c) map[c] = -1;
  rq.ncols = 0;
  auto use = [&](int c) {
    if (map[c] < 0) {
      map[c] = rq.ncols;
      rq.cols[rq.ncols++] = q->cols[c];
    }
    return map[c];
  };
  for (int p = 0; p < q->npreds; ++p) rq.preds[p].col = use(q->preds[p].col);
  for (int k = 0; k < q->nparts; ++k) rq.parts[k].col = use(q->parts[k].col);
  rq.naggs = 0;
  PlanInfo info;
  return plan_agg(&rq, H, &info);
}

// LDS of the partitioned passes' tile kernels.
static uint32_t part_hist_lds(const DevPlan& H, uint32_t nparts) {
  return 2 * H.stage_bytes + nparts * 4;
}
static uint32_t part_scatter_lds(const DevPlan& P, uint32_t nparts) {
  return 2 * P.stage_bytes + nparts * 12;
}

// kAggPart: histogram, scan, scatter and per-partition aggregation on stream s
// (the result already zeroed).
static int32_t run_part(tgp_ctx* ctx, const DevPlan& P, const DevPlan& H, const PlanInfo& info,
                        hipStream_t s) {
  if (P.ntiles == 0) return TGP_OK;
  const uint32_t np = info.nparts, sh = info.part_shift;
  const uint64_t rows = P.row_end - P.row_begin;
  uint64_t G = std::min<uint64_t>(P.ntiles, (uint64_t)persistent_grid(ctx, 2));
  const int64_t gmax = ctx->tune_grid_max.load();
  if (gmax > 0 && G > (uint64_t)gmax) G = (uint64_t)gmax;
  // histogram workgroups per scatter workgroup: up to 8 per CU while mat stays small
  uint32_t k = 1;
  while (k < 4 && G * 2 * k <= P.ntiles && (uint64_t)np * G * 2 * k <= (1u << 22)) k *= 2;
  const uint64_t mat_n = (uint64_t)np * G * k;
  const uint64_t nb = (mat_n + 4095) / 4096;
  // scratch: mat[mat_n + 1] | scan block sums[nb + 1] | values[naggs][rows] | groups[rows]
  const uint64_t words = mat_n + 1 + nb + 1 + (uint64_t)P.naggs * rows;
  void* scr = nullptr;
  TGP_TRY(get_scratch(ctx, s, words * 8 + rows * 2, &scr));
  auto* mat = static_cast<unsigned long long*>(scr);
  auto* bsum = reinterpret_cast<uint64_t*>(mat + mat_n + 1);
  auto* rval = bsum + nb + 1;
  auto* rgrp = reinterpret_cast<uint16_t*>(rval + (uint64_t)P.naggs * rows);
  const uint32_t lh = part_hist_lds(H, np), ls = part_scatter_lds(P, np);
  TGP_TRY(allow_lds(k_part_hist_tiles, lh));
  TGP_TRY(allow_lds(k_part_scatter_tiles, ls));
  k_part_hist_tiles<<<dim3((unsigned)(G * k)), dim3(kBlock), lh, s>>>(H, sh, np, k, mat);
  TGP_HIP(hipGetLastError());
  TGP_TRY(scan_u64(s, reinterpret_cast<uint64_t*>(mat), mat_n, bsum));
  k_part_scatter_tiles<<<dim3((unsigned)G), dim3(kBlock), ls, s>>>(P, sh, np, k, mat, rows,
                                                                   rgrp, rval);
  TGP_HIP(hipGetLastError());
  const uint32_t la = (uint32_t)((1ull << sh) * (1 + 2 * (uint64_t)P.naggs) * 8);
  TGP_TRY(allow_lds(k_part_agg_recs, la));
  const uint64_t ga = (rows + kPartChunk - 1) / kPartChunk;
  k_part_agg_recs<<<dim3((unsigned)ga), dim3(kPartAggBlock), la, s>>>(P, sh, np, G * k, mat,
                                                                      rows, rgrp, rval);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t run_agg(tgp_ctx* ctx, const DevPlan& P, const PlanInfo& info, hipStream_t s,
                const tgp_agg_query* q = nullptr, const tgp_group_result* out = nullptr) {
  if (q) {
    // the specialised kernels zero the result themselves
    bool launched = false;
    TGP_TRY(try_spec(ctx, q, out, s, &launched));
    if (launched) return TGP_OK;
  }
  // Results are accumulated with atomics: clear them first (one fill kernel, so
  // the sequence can be captured in a graph).
  uint64_t* p[3] = {reinterpret_cast<uint64_t*>(P.out_count),
                    reinterpret_cast<uint64_t*>(P.out_sums),
                    reinterpret_cast<uint64_t*>(P.out_oor)};
  const uint64_t n[3] = {P.ngroups, P.ngroups * P.naggs * 2, 1};
  TGP_TRY(zero_ranges(s, 3, p, n));
  int strategy = info.strategy;
  DevPlan H;
  if (info.nparts && (strategy == kAggGlobal || strategy == kAggPart)) {
    const int64_t mode = ctx->tune_agg_part.load();
    bool part = mode == 2 || (mode == 0 && strategy == kAggPart);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    TGP_HIP(hipStreamIsCapturing(s, &cs));
    if (!q || cs != hipStreamCaptureStatusNone) part = false;  // scratch may be allocated
    if (part) {
      TGP_TRY(plan_part_hist(q, &H));
      part = part_hist_lds(H, info.nparts) <= kLdsMax &&
             part_scatter_lds(P, info.nparts) <= kLdsMax;
    }
    strategy = part ? kAggPart : kAggGlobal;
  }
  switch (strategy) {
    case kAggReg1: TGP_TRY(launch_tiles(ctx, k_agg_reg<1, TGP_MAX_AGGS>, P, info.lds_bytes, s)); break;
    case kAggReg8: TGP_TRY(launch_tiles(ctx, k_agg_reg<8, 4>, P, info.lds_bytes, s)); break;
    case kAggLds: TGP_TRY(launch_tiles(ctx, k_agg_table<true, true>, P, info.lds_bytes, s)); break;
    case kAggGlobal:
      TGP_TRY(launch_tiles(ctx, k_agg_table<false, true>, P, info.lds_bytes, s));
      break;
    case kAggPart: TGP_TRY(run_part(ctx, P, H, info, s)); break;
    default: TGP_TRY(launch_tiles(ctx, k_agg_table<false, false>, P, info.lds_bytes, s)); break;
  }
  // the global table holds split sums; MIN/MAX hold codes
  const bool split = strategy == kAggGlobal;
  if ((P.minmax_kinds || split) && P.ngroups && P.naggs && P.ntiles) {
    const uint64_t items = P.ngroups * (uint64_t)P.naggs;
    const uint64_t grid = std::min<uint64_t>((items + 255) / 256, 65536);
    hipLaunchKernelGGL(k_agg_finalize, dim3((uint32_t)grid), dim3(256), 0, s, P.out_count,
                       P.out_sums, P.ngroups, P.naggs, P.minmax_kinds, (int32_t)split);
    TGP_HIP(hipGetLastError());
  }
  return TGP_OK;
}

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_group_agg_explain(const tgp_agg_query* q, int32_t* strategy, int32_t* spec_shape) {
  TGP_REQUIRE(q && strategy && spec_shape, "tgp_group_agg_explain: null argument");
  DevPlan P;
  PlanInfo info;
  TGP_TRY(plan_agg(q, &P, &info));
  *strategy = info.strategy;
  *spec_shape = spec_shape_index(q);
  return TGP_OK;
}

int32_t tgp_group_agg_ngroups(const tgp_agg_query* q, uint64_t* out_ngroups) {
  TGP_REQUIRE(q && out_ngroups, "tgp_group_agg_ngroups: null argument");
  uint64_t g = 1;
  for (int k = 0; k < q->nparts && k < TGP_MAX_GROUP_PARTS; ++k) g *= (uint64_t)q->parts[k].card;
  *out_ngroups = g;
  return TGP_OK;
}

int32_t tgp_group_agg(tgp_ctx* ctx, const tgp_agg_query* q, const tgp_group_result* out,
                      void* stream) {
  TGP_REQUIRE(ctx && q && out, "tgp_group_agg: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_REQUIRE(out->d_count && (q->naggs == 0 || out->d_sums), "tgp_group_agg: null result buffer");
  DevPlan P;
  PlanInfo info;
  TGP_TRY(plan_agg(q, &P, &info));
  P.out_count = reinterpret_cast<unsigned long long*>(out->d_count);
  P.out_sums = reinterpret_cast<unsigned long long*>(out->d_sums);
  P.out_oor = reinterpret_cast<unsigned long long*>(out->d_n_out_of_range);
  return run_agg(ctx, P, info, pick_stream(ctx, stream), q, out);
}

}  // extern "C"
