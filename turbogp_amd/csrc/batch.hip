// Batched executor: thousands of filter + GROUP BY queries in one persistent
// kernel launch, captured in a hipGraph.
//
// Every query of a batch is planned into its specialised shape (spec.h) and cut
// into 256-row slices.  The slices of all queries form one flattened list, cut
// into chunks of a few dozen slices that the waves of the persistent grid take
// from a device counter (a chunk may cross query boundaries), so that waves
// that run slower take fewer chunks.  A wave accumulates into a wave-private LDS
// region (counts and 64-bit partial sums per group) and flushes it into the
// query's exact int128 result with device atomics at the end of each segment
// (the part of one query inside one chunk) and after every 2^20 rows (the
// partials stay exact for |value| < 2^43; larger values take an exact slow path
// per slice).  The region holds R = 2^k copies of the
// query's histogram (as many as fit, at most 64) and lane l updates copy
// l mod R, so that few lanes of one LDS atomic instruction hit the same word
// even when the query has only two or three groups.
//
// The one-launch-per-query form (tgp_group_agg per query) is kept as the
// baseline mode.  The reference has no batched executor; its per-query entry is
// QueryEngine::execute (engine/mod.rs:1487).
#include <algorithm>
#include <map>
#include <utility>
#include <vector>

#include "jit.h"
#include "wave_agg.h"

namespace tgp {

constexpr uint64_t kGranSlices = 16;  // granularity of the slice -> query table

struct BatchQuery {
  SpecArgs a;           // arguments; a.out_* point into the batch's result arena
  uint64_t slice0;      // first flattened slice of the query
  int32_t shape;        // index into kShapes
  int32_t words;        // groups * (1 + naggs)
  int32_t naggs;
  int32_t rshift;       // log2 of the number of histogram copies R
};

// Query plans are read through the constant address space: the plan is not
// written while the kernel runs, so uniform reads become scalar loads into SGPRs
// (as kernel arguments are) instead of per-lane vector loads.
typedef const __attribute__((address_space(4))) BatchQuery ConstQuery;

// Slices [s_begin, s_end) (query-relative) of one query of shape S, with a
// flush every kFlushRows rows and at the end.  NT: streaming loads (each row is
// read once) or cached loads (the table-ordered schedule, where the queries
// that cover a region read it at about the same time).
template <Shape S, bool NT>
__device__ __forceinline__ void batch_segment(const SpecArgs& A, uint64_t s_begin,
                                              uint64_t s_end, unsigned long long* reg,
                                              uint32_t rshift, uint32_t& oor) {
  uint64_t min_len = A.col_len[0];
#pragma unroll
  for (int s = 1; s < S.nslots; ++s) min_len = A.col_len[s] < min_len ? A.col_len[s] : min_len;
  constexpr uint64_t kFlushSlices = kFlushRows / kSliceRows;
  for (uint64_t c = s_begin; c < s_end; c += kFlushSlices) {
    const uint64_t e = c + kFlushSlices < s_end ? c + kFlushSlices : s_end;
    for (uint64_t s = c; s < e; ++s) {
      SliceVals<S> v;
      wave_load<S, NT>(A, s, min_len, v);
      wave_slice<S>(A, s, v, reg, rshift, oor);
    }
    wave_flush(A.out_count, A.out_sums, A.words, S.naggs, reg, rshift);
  }
}

template <int I, bool NT>
__device__ __forceinline__ void dispatch_segment(int shape, const SpecArgs& A, uint64_t s_begin,
                                                 uint64_t s_end, unsigned long long* reg,
                                                 uint32_t rshift, uint32_t& oor) {
  if constexpr (I < kNumShapes) {
    if constexpr (kShapes[I].nslots > 0 && kShapes[I].has_sel == 0) {
      if (shape == I) {
        batch_segment<kShapes[I], NT>(A, s_begin, s_end, reg, rshift, oor);
        return;
      }
    }
    dispatch_segment<I + 1, NT>(shape, A, s_begin, s_end, reg, rshift, oor);
  }
}

// One query-relative slice range of query q.
template <bool NT>
__device__ __forceinline__ void batch_run_segment(ConstQuery* qs, uint32_t q, uint64_t b,
                                                  uint64_t e, unsigned long long* reg) {
  const SpecArgs& A = *(const SpecArgs*)&qs[q].a;
  uint32_t oor = 0;
  dispatch_segment<0, NT>(qs[q].shape, A, b, e, reg, (uint32_t)qs[q].rshift, oor);
  if (oor) atomicAdd(A.out_oor, (unsigned long long)oor);
}

// Persistent kernel, query order: every wave takes chunks of `chunk`
// consecutive slices of the flattened slice list (all slices of query 0, then
// of query 1, ...) from a device counter until none are left, so a slow wave
// takes fewer chunks and all waves end together.  granule_q[g] is the query
// that contains flattened slice g * kGranSlices.
template <bool NT>
__global__ void __launch_bounds__(kBlock) k_batch(const BatchQuery* __restrict__ plan, uint32_t nq,
                                                  uint64_t total_slices,
                                                  const uint32_t* __restrict__ granule_q,
                                                  uint32_t chunk, unsigned long long* counter) {
  __shared__ unsigned long long lreg[(kBlock / kWave) * kWaveWords];
  ConstQuery* qs = (ConstQuery*)plan;
  const __attribute__((address_space(4))) uint32_t* gq =
      (const __attribute__((address_space(4))) uint32_t*)granule_q;
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  unsigned long long* reg = lreg + wave * kWaveWords;
  for (int w = lane; w < kWaveWords; w += 64) reg[w] = 0;
  __builtin_amdgcn_wave_barrier();
  uint64_t k = grab_chunk(counter);
  while (k * chunk < total_slices) {
    const uint64_t c0 = k * chunk;
    const uint64_t c1 = c0 + chunk < total_slices ? c0 + chunk : total_slices;
    const uint64_t next = grab_chunk(counter);  // in flight while this chunk runs
    uint32_t q = gq[c0 / kGranSlices];
    while (q + 1 < nq && qs[q + 1].slice0 <= c0) ++q;
    // the chunk as segments of consecutive queries
    for (; q < nq; ++q) {
      const uint64_t q0 = qs[q].slice0;
      if (q0 >= c1) break;
      const uint64_t q1 = q + 1 < nq ? qs[q + 1].slice0 : total_slices;
      const uint64_t b = c0 > q0 ? c0 : q0, e = c1 < q1 ? c1 : q1;
      if (b < e) batch_run_segment<NT>(qs, q, b - q0, e - q0, reg);
    }
    k = next;
  }
}

// Persistent kernel, table order: work item i = (query items[3i], slices
// [items[3i+1], items[3i+2]) of it).  The items form `nq` queues, each sorted by
// the table row its items start at, so the waves sweep the table and the
// queries that cover a region read it within a short time of each other (from
// the caches after the first).  With one queue every wave takes from it; with
// kQueues queues, table chunk k's items are in queue k % kQueues and
// workgroup w takes from queue w % kQueues first: workgroups are dispatched to
// the 8 XCDs of an MI355X round robin, so the queries of one region then share
// one XCD's L2.  A wave whose queue is empty moves on to the next queue, so
// all waves end together.  The mapping only affects locality, not results.
constexpr uint32_t kQueues = 8;

struct ItemQueues {
  uint64_t off[kQueues + 1];  // items of queue i: [off[i], off[i + 1])
  uint32_t nq;
  uint32_t _pad;
};

template <bool NT>
__global__ void __launch_bounds__(kBlock) k_batch_items(const BatchQuery* __restrict__ plan,
                                                        const uint32_t* __restrict__ items,
                                                        ItemQueues Q,
                                                        unsigned long long* counters,
                                                        uint32_t per_grab) {
  __shared__ unsigned long long lreg[(kBlock / kWave) * kWaveWords];
  ConstQuery* qs = (ConstQuery*)plan;
  const __attribute__((address_space(4))) uint32_t* it =
      (const __attribute__((address_space(4))) uint32_t*)items;
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  unsigned long long* reg = lreg + wave * kWaveWords;
  for (int w = lane; w < kWaveWords; w += 64) reg[w] = 0;
  __builtin_amdgcn_wave_barrier();
  const uint32_t home = blockIdx.x % Q.nq;
  for (uint32_t t = 0; t < Q.nq; ++t) {
    const uint32_t qi = home + t < Q.nq ? home + t : home + t - Q.nq;
    const uint64_t base = Q.off[qi], n = Q.off[qi + 1] - base;
    // `per_grab` consecutive items per counter atomic: the same table chunk for
    // several queries, so a wave re-reads lines it just loaded
    uint64_t k = grab_chunks(&counters[qi], per_grab);
    while (k < n) {
      const uint64_t next = grab_chunks(&counters[qi], per_grab);  // in flight meanwhile
      const uint64_t kend = k + per_grab < n ? k + per_grab : n;
      for (uint64_t j = k; j < kend; ++j) {
        const uint64_t i = base + j;
        batch_run_segment<NT>(qs, it[3 * i], it[3 * i + 1], it[3 * i + 2], reg);
      }
      k = next;
    }
  }
}

// ---- shared scan (order 3) ----------------------------------------------------------
// The queries of a batch that read the same columns with the same aggregate forms
// (the Q1 family: date range, one of the five key forms, the four sums) differ only
// in their row range, date bounds and key parts.  A workgroup takes one table chunk
// and the family queries overlapping it; each wave loads a 256-row slice of the
// chunk ONCE, computes the row values (the four summed values, the key sources
// flagA, flagB and dt / D, the big-value flags) once, and then applies every query
// of the chunk to them from registers: its range and date test, its group id, and
// 64-bit LDS atomics into the query's part of the workgroup's LDS region (the
// planner gives each query an offset and 2^rshift copies so the region fits).  The
// region is added to the exact int128 results once per chunk.  Against one pass per
// query over the chunk, every row is read and multiplied once instead of once per
// covering query.  A chunk has at most 2^20 rows, so with |value| < 2^43 the 64-bit
// partials stay exact; a slice holding a larger value adds that aggregate straight
// into the int128 result.
constexpr int kSharedWords = 4992;           // LDS region per workgroup (39 KiB): 4 per CU
static_assert(kShapes[0].dt[0] == TGP_I32 && kShapes[0].dt[5] == TGP_U8 &&
              kShapes[0].dt[6] == TGP_U8, "the shared scan packs the Q1 key columns");
constexpr uint32_t kSharedMaxSlices = 4096;  // rows per chunk <= 2^20

struct SharedQuery {
  uint64_t row_begin, row_end;
  int64_t plo, phi;         // dt in [plo, phi]
  int64_t gbase[2], gcard[2];
  int32_t nparts;           // 0..2 key parts
  int32_t src[2];           // key source: 0 dt / divisor, 1 flagA, 2 flagB
  int32_t stride1;          // group id = part0 + part1 * stride1
  int32_t words;            // groups * 5
  int32_t kind;             // key form: the Q1 shape index (0 flagA x flagB, 1 flagA,
                            // 2 flagB, 3 dt / divisor, 4 no key)
  uint32_t sig;             // equal for members with equal date bounds and key parts
  uint32_t pstride;         // 0: a member; else a signature group's partials query:
                            // out_count is its partial rows, pstride words each
  uint32_t pk0;             // the group's first chunk (partials query)
  uint32_t live;            // 0: no dt passes the date test
  int32_t dlo;              // the date test on int32 dt: (uint32)(dt - dlo) <= dspan
  uint32_t dspan;
  uint32_t ksig;            // equal for members whose key parts below are equal
  uint32_t klo[2], kspan[2];  // key part k in range: (uint32)(x_k - klo[k]) < kspan[k]
  uint32_t kc, st5;         // group id x 5 = 5 (x_0 - klo[0]) + st5 (x_1 - klo[1]) + kc
  unsigned long long* out_count;
  unsigned long long* out_sums;
  unsigned long long* out_oor;
};
typedef const __attribute__((address_space(4))) SharedQuery ConstShared;
typedef const __attribute__((address_space(4))) uint32_t ConstU32;

struct SharedFamily {
  SpecArgs a;         // the family's columns and aggregate arguments (Q1 layout)
  uint32_t magic;     // dt / divisor (Granlund-Montgomery), when a member groups by it
  int32_t shift;
  int32_t divisor;    // 0: no member divides
  uint32_t chunk;     // slices per chunk (items start inside chunk first slice / chunk)
};

// Row values of one slice shared by the queries of a chunk (this lane's 4 rows).
struct SharedRows {
  int32_t dt[kSR];
  uint32_t fab[kSR];             // flagA | flagB << 8 (u8 columns)
  int32_t dq[kSR];               // dt / divisor (|dt / divisor| < 2^31 for int32 dt)
  uint64_t val[4][kSR];          // the four summed values
  bool big[4];                   // wave-uniform: the slice has |value| >= 2^43
  bool any_big;
  int32_t mode;                  // wave-uniform: 0 no big value and every first value
                                 // in [0, 2^16) (val[0] then holds kCountOne + value),
                                 // 1 no big value, 2 some big value
};

// Word 0 of a group holds its row count in bits 40..63 and, for slices whose first
// value lies in [0, 2^16), the sum of those values in bits 0..39 (one atomic for
// both): a region sees at most 2^20 rows, so the count stays below 2^24 and the
// packed sum below 2^36.  Other slices add their first value to word 1.
constexpr unsigned long long kCountOne = 1ull << 40;
static_assert((uint64_t)kSharedMaxSlices * kSliceRows <= (1ull << 20),
              "a region's count and packed sum need chunks of at most 2^20 rows");

// Adds (lo, hi) to an int128 whose words are `st` words apart (64-bit atomics;
// the carries out of the low word are counted exactly, as in atomic_add128).
__device__ __forceinline__ void atomic_add128_split(unsigned long long* p, uint32_t st,
                                                    uint64_t lo, int64_t hi) {
  if (lo == 0 && hi == 0) return;
  const uint64_t old = atomicAdd(p, (unsigned long long)lo);
  const int64_t h = hi + (int64_t)((old + lo) < old);
  if (h != 0) atomicAdd(p + st, (unsigned long long)h);
}

// Where one entry of a chunk adds its results.  A member adds into its own result
// (count[g], sums[(g * 4 + a) * 2], oor, consecutive words); a signature group's
// partials query adds into the chunk's column of the group's partial rows: row g
// the counts, rows G + 2 (g * 4 + a) and the next the low and high sum words, row
// 9 G the out-of-range count, position 1 + (chunk - first chunk) of each row.
struct EntryOut {
  unsigned long long* count;
  unsigned long long* sums;
  unsigned long long* oor;
  uint32_t st;  // word distance of consecutive result words
};

template <typename SQ>
__device__ __forceinline__ EntryOut entry_out(SQ& D, uint64_t kchunk) {
  if (!D.pstride) return EntryOut{D.out_count, D.out_sums, D.out_oor, 1u};
  const uint32_t st = D.pstride, G = (uint32_t)D.words / 5;
  unsigned long long* base = D.out_count + 1 + (kchunk - D.pk0);
  return EntryOut{base, base + (uint64_t)G * st, base + (uint64_t)9 * G * st, st};
}

// An item's entries are held one per lane (lane j: entry j of the item, at most
// kWave per item) and read back with v_readlane, so the loop over a slice's entries
// issues no scalar or LDS loads: a scalar load there would wait for the previous
// entry's LDS atomics, which share its completion counter.
struct EntryLane {
  uint32_t pack;           // region offset | rshift << 16 | key form << 20
  uint32_t rb, re;         // the item's rows [rb, re) the query covers (item-relative)
  uint32_t dlo, dspan;     // the date test: (uint32)(dt - dlo) <= dspan
  uint32_t ksig;           // key signature: equal ids give equal group ids and tests
  uint32_t klo0, kspan0, klo1, kspan1, kc, st5;  // the key parts (SharedQuery)
  uint32_t sums_lo, sums_hi, oor_lo, oor_hi, st;  // EntryOut of the item's chunk
};

__device__ __forceinline__ EntryLane load_entry_lane(const SharedQuery* __restrict__ fq,
                                                     const uint32_t* __restrict__ entries,
                                                     uint32_t e, uint64_t irow0, uint32_t nrows,
                                                     uint64_t kchunk) {
  EntryLane E;
  const SharedQuery& D = fq[entries[2 * e]];
  E.pack = entries[2 * e + 1] | ((uint32_t)D.kind << 20);
  const uint64_t b = D.row_begin > irow0 ? D.row_begin - irow0 : 0;
  const uint64_t x = D.row_end > irow0 ? D.row_end - irow0 : 0;
  E.rb = (uint32_t)(b < nrows ? b : nrows);
  E.re = (uint32_t)(x < nrows ? x : nrows);
  if (!D.live) E.rb = E.re = 0;  // no row passes the date test
  E.dlo = (uint32_t)D.dlo;
  E.dspan = D.dspan;
  E.ksig = D.ksig;
  E.klo0 = D.klo[0];
  E.kspan0 = D.kspan[0];
  E.klo1 = D.klo[1];
  E.kspan1 = D.kspan[1];
  E.kc = D.kc;
  E.st5 = D.st5;
  const EntryOut O = entry_out(D, kchunk);
  E.sums_lo = (uint32_t)(uintptr_t)O.sums;
  E.sums_hi = (uint32_t)((uintptr_t)O.sums >> 32);
  E.oor_lo = (uint32_t)(uintptr_t)O.oor;
  E.oor_hi = (uint32_t)((uintptr_t)O.oor >> 32);
  E.st = O.st;
  return E;
}

__device__ __forceinline__ uint32_t lane_of(uint32_t v, uint32_t j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
}

__device__ __forceinline__ unsigned long long* lane_ptr(uint32_t lo, uint32_t hi, uint32_t j) {
  return (unsigned long long*)(uintptr_t)(((uint64_t)lane_of(hi, j) << 32) | lane_of(lo, j));
}

// Group ids (x 5, the words of a group) and key range tests of this lane's rows for
// the key form of lane j's entry: part k's source x_k is in range when
// (uint32)(x_k - klo_k) < kspan_k, and then 5 group = 5 t0 + st5 t1 + kc with
// t_k = x_k - klo_k (exact modulo 2^32: the planner keeps 5 groups below 2^31).
__device__ __forceinline__ void key_ids(const EntryLane& E, uint32_t j, const SharedRows& v,
                                        uint32_t (&gid5)[kSR], bool (&ok)[kSR]) {
  const uint32_t kind = lane_of(E.pack, j) >> 20;
  const uint32_t lo0 = lane_of(E.klo0, j), sp0 = lane_of(E.kspan0, j);
  const uint32_t lo1 = lane_of(E.klo1, j), sp1 = lane_of(E.kspan1, j);
  const uint32_t kc = lane_of(E.kc, j), st5 = lane_of(E.st5, j);
#pragma unroll
  for (int r = 0; r < kSR; ++r) {
    const uint32_t x0 = kind == 3 ? (uint32_t)v.dq[r] : kind == 2 ? v.fab[r] >> 8 : v.fab[r] & 0xffu;
    const uint32_t t0 = x0 - lo0;
    bool in = t0 < sp0;
    uint32_t g = (t0 << 2) + t0 + kc;
    if (kind == 0) {
      const uint32_t t1 = (v.fab[r] >> 8) - lo1;
      in = in && t1 < sp1;
      g += __umul24(t1, st5);
    }
    if (kind == 4) {
      in = true;
      g = 0;
    }
    ok[r] = in;
    gid5[r] = g;
  }
}

// The selected rows' LDS atomics (lane l on copy l mod 2^RS, compile-time word
// offsets); a slice holding a big value adds that aggregate straight into the
// int128 result.
template <int RS>
__device__ __forceinline__ void shared_add(const EntryLane& E, uint32_t j, const SharedRows& v,
                                           const bool (&sel)[kSR], const uint32_t (&gid5)[kSR],
                                           unsigned long long* reg) {
  unsigned long long* mine = reg + ((threadIdx.x & 63) & ((1u << RS) - 1));
  if (v.mode == 0) {  // four atomics: word 0 holds the count and value 0 (kCountOne added)
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if (sel[r]) {
        unsigned long long* p = mine + (gid5[r] << RS);
        atomicAdd(p, (unsigned long long)v.val[0][r]);
        atomicAdd(p + (2 << RS), (unsigned long long)v.val[1][r]);
        atomicAdd(p + (3 << RS), (unsigned long long)v.val[2][r]);
        atomicAdd(p + (4 << RS), (unsigned long long)v.val[3][r]);
      }
    return;
  }
  if (v.mode == 1) {
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if (sel[r]) {
        unsigned long long* p = mine + (gid5[r] << RS);
        atomicAdd(p, kCountOne);
        atomicAdd(p + (1 << RS), (unsigned long long)v.val[0][r]);
        atomicAdd(p + (2 << RS), (unsigned long long)v.val[1][r]);
        atomicAdd(p + (3 << RS), (unsigned long long)v.val[2][r]);
        atomicAdd(p + (4 << RS), (unsigned long long)v.val[3][r]);
      }
    return;
  }
#pragma unroll
  for (int r = 0; r < kSR; ++r)
    if (sel[r]) atomicAdd(mine + (gid5[r] << RS), kCountOne);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (v.big[a]) {
      unsigned long long* sums = lane_ptr(E.sums_lo, E.sums_hi, j);
      const uint32_t st = lane_of(E.st, j);
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if (sel[r]) {
          const uint64_t g = gid5[r] / 5u;
          const int64_t x = (int64_t)v.val[a][r];
          atomic_add128_split(sums + (g * 4 + a) * 2 * st, st, (uint64_t)x, x < 0 ? -1 : 0);
        }
    } else {
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if (sel[r]) atomicAdd(mine + (gid5[r] << RS) + ((1 + a) << RS), (unsigned long long)v.val[a][r]);
    }
  }
}

// The common case of one entry on one slice: every key of the slice in range and
// no big value (mode 0 or 1).  Each row's date test (and, on a slice the query
// covers in part, its row-range test) guards that row's atomics directly.
template <int MODE, int RS>
__device__ __forceinline__ void shared_rows(const SharedRows& v, const uint32_t (&gid5)[kSR],
                                            uint32_t dlo, uint32_t dspan, bool part,
                                            uint32_t rb, uint32_t re, uint32_t rel0,
                                            unsigned long long* reg) {
  unsigned long long* mine = reg + ((threadIdx.x & 63) & ((1u << RS) - 1));
#pragma unroll
  for (int r = 0; r < kSR; ++r) {
    bool c = (uint32_t)v.dt[r] - dlo <= dspan;
    if (part) c = c && rel0 + r - rb < re - rb;
    if (c) {
      unsigned long long* p = mine + (gid5[r] << RS);
      if constexpr (MODE == 0) {  // word 0: kCountOne + value 0
        atomicAdd(p, (unsigned long long)v.val[0][r]);
      } else {
        atomicAdd(p, kCountOne);
        atomicAdd(p + (1 << RS), (unsigned long long)v.val[0][r]);
      }
      atomicAdd(p + (2 << RS), (unsigned long long)v.val[1][r]);
      atomicAdd(p + (3 << RS), (unsigned long long)v.val[2][r]);
      atomicAdd(p + (4 << RS), (unsigned long long)v.val[3][r]);
    }
  }
}

template <int MODE>
__device__ __forceinline__ void shared_rows_rs(uint32_t rs, const SharedRows& v,
                                               const uint32_t (&gid5)[kSR], uint32_t dlo,
                                               uint32_t dspan, bool part, uint32_t rb, uint32_t re,
                                               uint32_t rel0, unsigned long long* reg) {
#define TGP_SR_CASE(K) \
  case K: shared_rows<MODE, K>(v, gid5, dlo, dspan, part, rb, re, rel0, reg); break;
  switch (rs) {
    TGP_SR_CASE(0) TGP_SR_CASE(1) TGP_SR_CASE(2) TGP_SR_CASE(3) TGP_SR_CASE(4) TGP_SR_CASE(5)
    TGP_SR_CASE(6)
    default: break;
  }
#undef TGP_SR_CASE
}

// One entry on one slice where some key is out of range or a value is big: the
// selected rows (date and row-range tests, in-range keys), the out-of-range ones
// counted, and the atomics of shared_add.
__device__ __forceinline__ void shared_entry_slow(const EntryLane& E, uint32_t j,
                                                  const SharedRows& v, const uint32_t (&gid5)[kSR],
                                                  const bool (&ok)[kSR], bool all_ok, uint32_t s_rel,
                                                  uint32_t rb, uint32_t re, uint32_t dlo,
                                                  uint32_t dspan, uint32_t pack,
                                                  unsigned long long* region) {
  const uint32_t lane = threadIdx.x & 63;
  bool sel[kSR];
#pragma unroll
  for (int r = 0; r < kSR; ++r) sel[r] = (uint32_t)v.dt[r] - dlo <= dspan;
  if (rb > s_rel || re < s_rel + kSliceRows) {
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      const uint32_t rel = s_rel + lane * kSR + r;
      sel[r] = sel[r] && rel - rb < re - rb;
    }
  }
  if (!all_ok) {
    uint32_t nbad = 0;
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      nbad += (uint32_t)__popcll(__ballot(sel[r] && !ok[r]));
      sel[r] = sel[r] && ok[r];
    }
    if (nbad && lane == 0) atomicAdd(lane_ptr(E.oor_lo, E.oor_hi, j), (unsigned long long)nbad);
  }
  if (!__ballot(sel[0] || sel[1] || sel[2] || sel[3])) return;
  unsigned long long* reg = region + (pack & 0xffffu);
  switch ((pack >> 16) & 15u) {
    case 0: shared_add<0>(E, j, v, sel, gid5, reg); break;
    case 1: shared_add<1>(E, j, v, sel, gid5, reg); break;
    case 2: shared_add<2>(E, j, v, sel, gid5, reg); break;
    case 3: shared_add<3>(E, j, v, sel, gid5, reg); break;
    case 4: shared_add<4>(E, j, v, sel, gid5, reg); break;
    case 5: shared_add<5>(E, j, v, sel, gid5, reg); break;
    case 6: shared_add<6>(E, j, v, sel, gid5, reg); break;
    default: break;
  }
}

// Every entry of the item on one slice (item-relative first row s_rel): the date
// and row-range tests, out-of-range keys counted, and the selected rows added into
// the entry's part of the region.  Entries arrive sorted by key signature, so the
// group ids are computed once per signature.
__device__ __forceinline__ void shared_entries(const EntryLane& E, uint32_t ne, const SharedRows& v,
                                               uint32_t s_rel, unsigned long long* region) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t cur = ~0u;
  uint32_t gid5[kSR];
  bool ok[kSR];
  bool all_ok = true;
#pragma unroll
  for (int r = 0; r < kSR; ++r) {
    gid5[r] = 0;
    ok[r] = true;
  }
  for (uint32_t j = 0; j < ne; ++j) {
    const uint32_t rb = lane_of(E.rb, j), re = lane_of(E.re, j);
    if (re <= s_rel || rb >= s_rel + kSliceRows) continue;
    const uint32_t ks = lane_of(E.ksig, j);
    if (ks != cur) {
      cur = ks;
      key_ids(E, j, v, gid5, ok);
      all_ok = __ballot(!(ok[0] && ok[1] && ok[2] && ok[3])) == 0;
    }
    const uint32_t dlo = lane_of(E.dlo, j), dspan = lane_of(E.dspan, j);
    const uint32_t pack = lane_of(E.pack, j);
    if (all_ok && v.mode != 2) {
      const bool part = rb > s_rel || re < s_rel + kSliceRows;
      const uint32_t rs = (pack >> 16) & 15u;
      unsigned long long* reg = region + (pack & 0xffffu);
      if (v.mode == 0)
        shared_rows_rs<0>(rs, v, gid5, dlo, dspan, part, rb, re, s_rel + lane * kSR, reg);
      else
        shared_rows_rs<1>(rs, v, gid5, dlo, dspan, part, rb, re, s_rel + lane * kSR, reg);
      continue;
    }
    shared_entry_slow(E, j, v, gid5, ok, all_ok, s_rel, rb, re, dlo, dspan, pack, region);
  }
}

// The row values every query of a chunk shares for this lane's rows of one slice:
// the key sources flagA, flagB and dt / divisor, the summed values, their big-value
// flags and the slice's accumulation mode.
template <bool NT>
__device__ __forceinline__ void shared_rows_of(const SharedFamily& F, uint64_t row0,
                                               uint64_t min_len, SharedRows& sr) {
  constexpr Shape S = kShapes[0];
  const uint64_t row = row0 + (threadIdx.x & 63) * kSR;
  SliceVals<S> v;
  load_all<S, 0, NT>(F.a, row, row0 + kSliceRows <= min_len, v);
#pragma unroll
  for (int r = 0; r < kSR; ++r) {
    sr.dt[r] = (int32_t)v.x[0][r];
    sr.fab[r] = (uint32_t)v.x[5][r] | ((uint32_t)v.x[6][r] << 8);
    int64_t x = v.x[0][r];
    if (F.divisor > 1) {
      if (x >= 0 && x <= 0xffffffffll) {
        const uint32_t nn = (uint32_t)x;
        const uint32_t q = __umulhi(F.magic, nn);
        x = (int64_t)((q + ((nn - q) >> 1)) >> F.shift);
      } else {
        x = x / F.divisor;
      }
    }
    sr.dq[r] = (int32_t)x;
  }
  sr.any_big = false;
  static_for<S.naggs>([&](auto AI) {
    constexpr int a = decltype(AI)::value;
    agg_values<S, a>(F.a, v, sr.val[a]);
    uint32_t m = 0;
#pragma unroll
    for (int r = 0; r < kSR; ++r) big_fold(m, sr.val[a][r]);
    sr.big[a] = __ballot(big_any(m)) != 0;
    sr.any_big = sr.any_big || sr.big[a];
  });
  bool small = true;
#pragma unroll
  for (int r = 0; r < kSR; ++r) small = small && sr.val[0][r] < (1ull << 16);
  sr.mode = sr.any_big ? 2 : __ballot(!small) == 0 ? 0 : 1;
  if (sr.mode == 0) {
#pragma unroll
    for (int r = 0; r < kSR; ++r) sr.val[0][r] += kCountOne;
  }
}

// items: 4 words each (first absolute slice, slices, first entry, entries);
// entries: 2 words each (query index into the family, region offset | rshift << 16)
template <bool NT>
__global__ void __launch_bounds__(kBlock, 4) k_batch_shared(SharedFamily F,
                                                         const SharedQuery* __restrict__ fq,
                                                         const uint32_t* __restrict__ items,
                                                         const uint32_t* __restrict__ entries,
                                                         ItemQueues Q,
                                                         unsigned long long* counters) {
  constexpr Shape S = kShapes[0];  // column dtypes and aggregate forms of the family
  constexpr int W1 = 1 + S.naggs;
  __shared__ unsigned long long region[kSharedWords];
  __shared__ uint32_t s_item;
  ConstShared* qs = (ConstShared*)fq;
  ConstU32* it = (ConstU32*)items;
  ConstU32* en = (ConstU32*)entries;
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  for (uint32_t w = threadIdx.x; w < (uint32_t)kSharedWords; w += kBlock) region[w] = 0;
  uint64_t min_len = F.a.col_len[0];
#pragma unroll
  for (int c = 1; c < S.nslots; ++c) min_len = F.a.col_len[c] < min_len ? F.a.col_len[c] : min_len;
  const uint32_t home = blockIdx.x % Q.nq;
  for (uint32_t t = 0; t < Q.nq; ++t) {
    const uint32_t qi = home + t < Q.nq ? home + t : home + t - Q.nq;
    const uint64_t base = Q.off[qi], n = Q.off[qi + 1] - base;
    for (;;) {
      __syncthreads();  // the region is clear and s_item free for the next grab
      if (threadIdx.x == 0) s_item = (uint32_t)atomicAdd(&counters[qi], 1ull);
      __syncthreads();
      const uint64_t k = s_item;
      if (k >= n) break;
      const uint64_t i = base + k;
      const uint64_t sl0 = it[4 * i];
      const uint32_t nsl = it[4 * i + 1], e0 = it[4 * i + 2], ne = it[4 * i + 3];
      const uint64_t kchunk = sl0 / F.chunk;
      // the item's entry descriptors, one per lane (the planner keeps ne <= kWave)
      EntryLane E{};
      if (lane < ne)
        E = load_entry_lane(fq, entries, e0 + lane, sl0 * kSliceRows, nsl * kSliceRows, kchunk);
      for (uint32_t sj = wave; sj < nsl; sj += kBlock / kWave) {
        SharedRows sr;
        shared_rows_of<NT>(F, (sl0 + sj) * kSliceRows, min_len, sr);
        shared_entries(E, ne, sr, sj * kSliceRows, region);
      }
      __syncthreads();
      // every query's copies summed into its exact result; the region cleared
      for (uint32_t e = e0; e < e0 + ne; ++e) {
        const uint32_t q = en[2 * e], pack = en[2 * e + 1];
        ConstShared& D = qs[q];
        const uint32_t rs = pack >> 16, R = 1u << rs;
        unsigned long long* reg = region + (pack & 0xffffu);
        const EntryOut O = entry_out(D, kchunk);
        for (uint32_t w = threadIdx.x; w < (uint32_t)D.words; w += kBlock) {
          unsigned long long x = 0;
          for (uint32_t c = 0; c < R; ++c) {
            x += reg[(w << rs) + c];
            reg[(w << rs) + c] = 0;
          }
          if (!x) continue;
          const uint32_t g = w / W1, k = w % W1;
          if (k == 0) {  // the count, and the packed part of value 0's sum
            if (x >> 40) atomicAdd(O.count + (uint64_t)g * O.st, x >> 40);
            const uint64_t packed = x & (kCountOne - 1);
            if (packed)
              atomic_add128_split(O.sums + (uint64_t)g * S.naggs * 2 * O.st, O.st, packed, 0);
          } else
            atomic_add128_split(O.sums + ((uint64_t)g * S.naggs + (k - 1)) * 2 * O.st, O.st, x,
                                (int64_t)x < 0 ? -1 : 0);
        }
      }
    }
  }
}

// Signature groups: members of the shared scan with equal date bounds and key parts
// share one partials query over the chunks any of them covers whole.  Its partial
// rows hold, per result word, position 0 = 0 and position 1 + c = chunk c's part.
struct PartialGroup {
  unsigned long long* rows;
  uint32_t nk1;     // positions per row: chunks + 1
  uint32_t groups;  // G: rows G + 2 G naggs + 1
};

// Turns every partial row into its running sums: position p then holds the sum of
// chunks [0, p).  One workgroup per (group, result word) (units[blockIdx.x] = group |
// word << 16); sum words are int128 (row pairs), count and out-of-range words 64-bit.
__global__ void __launch_bounds__(256) k_partials_scan(const PartialGroup* __restrict__ pg,
                                                       const uint32_t* __restrict__ units,
                                                       uint32_t naggs) {
  const PartialGroup P = pg[units[blockIdx.x] & 0xffffu];
  const uint32_t G = P.groups, u = units[blockIdx.x] >> 16;
  const bool wide = u >= G && u < G + G * naggs;
  const uint32_t row = u < G ? u : wide ? G + 2 * (u - G) : G + 2 * G * naggs;
  unsigned long long* lo_row = P.rows + (uint64_t)row * P.nk1;
  unsigned long long* hi_row = lo_row + P.nk1;
  const uint32_t n = P.nk1, per = (n + 255) / 256;
  const uint32_t b = threadIdx.x * per, e = b + per < n ? b + per : n;
  uint64_t lo = 0;
  int64_t hi = 0;
  for (uint32_t i = b; i < e; ++i) {
    const uint64_t x = lo_row[i];
    const uint64_t s = lo + x;
    if (wide) hi += (int64_t)hi_row[i] + (int64_t)(s < lo);
    lo = s;
  }
  __shared__ uint64_t s_lo[256];
  __shared__ int64_t s_hi[256];
  s_lo[threadIdx.x] = lo;
  s_hi[threadIdx.x] = hi;
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive prefix of the 256 block totals
    uint64_t rl = 0;
    int64_t rh = 0;
    for (int t = 0; t < 256; ++t) {
      const uint64_t xl = s_lo[t];
      const int64_t xh = s_hi[t];
      s_lo[t] = rl;
      s_hi[t] = rh;
      const uint64_t sl = rl + xl;
      rh += xh + (int64_t)(sl < rl);
      rl = sl;
    }
  }
  __syncthreads();
  lo = s_lo[threadIdx.x];
  hi = s_hi[threadIdx.x];
  for (uint32_t i = b; i < e; ++i) {
    const uint64_t x = lo_row[i];
    const uint64_t s = lo + x;
    if (wide) hi += (int64_t)hi_row[i] + (int64_t)(s < lo);
    lo = s;
    lo_row[i] = lo;
    if (wide) hi_row[i] = (uint64_t)hi;
  }
}

// One member of a signature group: positions p0 < p1 of the chunks it covers whole.
struct PartialLink {
  const unsigned long long* rows;
  unsigned long long* out_count;
  unsigned long long* out_sums;
  unsigned long long* out_oor;
  uint32_t nk1, groups, p0, p1;
};

// Adds the group's sums over chunks [p0, p1) (running sum at p1 minus at p0) into
// every member's result.  One workgroup per member.
__global__ void __launch_bounds__(256) k_partials_links(const PartialLink* __restrict__ pl,
                                                        uint32_t naggs) {
  const PartialLink L = pl[blockIdx.x];
  const uint32_t G = L.groups, units = G + G * naggs + 1;
  for (uint32_t u = threadIdx.x; u < units; u += blockDim.x) {
    const bool wide = u >= G && u < G + G * naggs;
    const uint32_t row = u < G ? u : wide ? G + 2 * (u - G) : G + 2 * G * naggs;
    const unsigned long long* r = L.rows + (uint64_t)row * L.nk1;
    const uint64_t a = r[L.p1], c = r[L.p0];
    const uint64_t lo = a - c;
    if (u < G) {
      if (lo) atomicAdd(&L.out_count[u], lo);
    } else if (wide) {
      const int64_t hi = (int64_t)(r[L.nk1 + L.p1] - r[L.nk1 + L.p0]) - (int64_t)(a < c);
      atomic_add128(&L.out_sums[(uint64_t)(u - G) * 2], lo, hi);
    } else if (lo) {
      atomicAdd(L.out_oor, lo);
    }
  }
}

}  // namespace tgp

using namespace tgp;

struct tgp_batch {
  tgp_ctx* ctx = nullptr;
  uint32_t max_queries = 0, n = 0;
  // host copies
  std::vector<tgp_agg_query> queries;
  std::vector<BatchQuery> plan;
  std::vector<uint64_t> count_off, sums_off, oor_off, ngroups;
  std::vector<int32_t> fallback;  // set_queries: the queries of no batch shape (generic items)
  uint64_t result_words = 0, total_slices = 0;
  // device
  BatchQuery* d_plan = nullptr;
  uint64_t* d_results = nullptr;
  uint64_t plan_cap = 0, results_cap = 0;
  // tuning / graph
  uint32_t wg_per_cu = 0;  // 0 = occupancy
  uint32_t chunk = 32;     // slices per dynamically scheduled chunk / work item
  uint32_t per_grab = 1;   // work items per counter atomic (table orders)
  int32_t order = 0;       // 0: query order; 1: table order; 2: table order, XCD-affine
  bool cached = false;     // cached instead of streaming column loads
  uint32_t* d_items = nullptr;
  uint64_t items_cap = 0, nitems = 0;
  uint64_t queue_off[kQueues + 1] = {};
  uint32_t nqueues = 1;
  uint32_t items_chunk = 0;
  int32_t items_order = 0;
  bool items_valid = false;
  uint32_t* d_granule = nullptr;
  uint64_t granule_cap = 0;
  bool granule_valid = false;
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  hipStream_t graph_stream = nullptr;
  std::vector<void*> retired;  // outgrown device buffers, freed with the batch
  // shared scan (order 3): the family members and their chunk items
  std::vector<uint8_t> in_family;  // per plan entry
  std::vector<SharedQuery> fam;    // members, in plan order
  std::vector<uint32_t> fam_plan;  // plan index of each member
  SharedFamily family{};
  SharedQuery* d_fam = nullptr;
  uint64_t fam_cap = 0;
  uint32_t* d_sitems = nullptr;
  uint32_t* d_sentries = nullptr;
  uint64_t sitems_cap = 0, sentries_cap = 0, nsitems = 0;
  uint64_t squeue_off[kQueues + 1] = {};
  uint32_t sitems_chunk = 0;
  bool sitems_valid = false;
  // signature groups: their partials queries follow the members in d_fam
  std::vector<SharedQuery> sfam;
  std::vector<PartialGroup> pgroups;
  std::vector<PartialLink> plinks;
  PartialGroup* d_pgroups = nullptr;
  PartialLink* d_plinks = nullptr;
  uint64_t* d_partials = nullptr;
  std::vector<uint32_t> punits;  // per scan workgroup: group | result word << 16
  // generic items: the queries of no batch shape, through the tile engine
  std::vector<DevPlan> gplans;
  std::vector<GenericItem> gitems;
  std::vector<uint32_t> gminmax;  // plans to finalise: index | split sums << 31
  DevPlan* d_gplans = nullptr;
  GenericItem* d_gitems = nullptr;
  uint32_t* d_gminmax = nullptr;
  uint64_t gplans_cap = 0, gitems_cap = 0, gminmax_cap = 0, gminmax_words = 0;
  uint32_t gstage_max = 0, gtable_max = 0;
  // generic queries whose groups exceed an LDS table: kAggPart's passes over all of
  // them (histogram plans, queries, items, chunks; counters | counts | scan sums;
  // records)
  std::vector<DevPlan> hplans;
  std::vector<PartQuery> pqs;
  std::vector<PartItem> pitems;
  std::vector<PartChunk> pchunks;
  DevPlan* d_hplans = nullptr;
  PartQuery* d_pqs = nullptr;
  PartItem* d_pitems = nullptr;
  PartChunk* d_pchunks = nullptr;
  unsigned long long* d_pmat = nullptr;
  uint16_t* d_pgrp = nullptr;
  uint64_t* d_pval = nullptr;
  uint64_t hplans_cap = 0, pqs_cap = 0, pitems_cap = 0, pchunks_cap = 0, pmat_cap = 0,
           pgrp_cap = 0, pval_cap = 0, pmat_n = 0;
  uint32_t phstage_max = 0, pstage_max = 0, pnparts_max = 0, ptable_max = 0;
  // generic queries with a specialised shape whose kernels are ready (compiled ahead
  // of time or at run time): one launch of that kernel each, in the batch's graph
  struct SpecRun {
    Shape shape;
    ShapeKernels k;
    SpecArgs a;
  };
  std::vector<SpecRun> spec_runs;
  unsigned long long* d_spec_counter = nullptr;  // one chunk counter per run
  uint64_t spec_counter_cap = 0;
  // the runs go round-robin to side streams (default priority) forked from the
  // batch stream after the result fill and joined into it at the launch's end, so
  // small runs overlap each other and the batch's other kernels
  static constexpr int kSpecStreams = 3;
  hipStream_t spec_streams[kSpecStreams] = {};
  hipEvent_t spec_fork = nullptr;
  hipEvent_t spec_join[kSpecStreams] = {};
  uint32_t* d_punits = nullptr;
  uint64_t pgroups_cap = 0, plinks_cap = 0, partials_cap = 0, partial_words = 0, punits_cap = 0;
};

// Grows a device buffer to at least `need` elements.  The old buffer is not
// freed now: hipFree synchronises the whole device, which would stall every
// other stream's work (a server plans one batch while others run).  Buffers
// grow at least twofold, so the retired ones add up to less than the last.
template <typename T>
static int32_t ensure_cap(tgp_batch* b, T** p, uint64_t* cap, uint64_t need) {
  if (*cap >= need) return TGP_OK;
  const uint64_t n = need > 2 * *cap ? need : 2 * *cap;
  void* q = nullptr;
  TGP_TRY(hbm_alloc(b->ctx, n * sizeof(T), &q));
  if (*p) b->retired.push_back(*p);
  *p = static_cast<T*>(q);
  *cap = n;
  return TGP_OK;
}

static void drop_graph(tgp_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
  b->graph = nullptr;
  b->graph_def = nullptr;
}

template <typename K>
static int32_t persistent_launch_grid(tgp_batch* b, K kern, uint64_t units, uint64_t* grid_out) {
  int per_cu = 0;
  TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
  if (per_cu < 1) per_cu = 1;
  if (b->wg_per_cu && (int)b->wg_per_cu < per_cu) per_cu = (int)b->wg_per_cu;
  uint64_t grid = (uint64_t)persistent_grid(b->ctx, per_cu);
  const uint64_t need = (units + (kBlock / kWave) - 1) / (kBlock / kWave);
  if (grid > need) grid = need;
  const int64_t gmax = b->ctx->tune_grid_max.load();
  if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
  *grid_out = grid == 0 ? 1 : grid;
  return TGP_OK;
}

template <bool NT>
static int32_t launch_variant(tgp_batch* b, hipStream_t s) {
  auto* counter = reinterpret_cast<unsigned long long*>(b->d_results + b->result_words);
  // counter[0]: chunk counter of the query-ordered kernel; counter[0..nq): queue counters;
  // counter[kQueues..2 kQueues): the shared scan's queue counters
  uint64_t grid = 0;
  if (b->order == 0) {
    auto kern = k_batch<NT>;
    const uint64_t chunks = (b->total_slices + b->chunk - 1) / b->chunk;
    TGP_TRY(persistent_launch_grid(b, kern, chunks, &grid));
    kern<<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(b->d_plan, (uint32_t)b->plan.size(),
                                                       b->total_slices, b->d_granule, b->chunk,
                                                       counter);
  } else {
    if (b->order == 3 && b->nsitems) {
      auto skern = k_batch_shared<NT>;
      int per_cu = 0;
      TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, skern, kBlock, 0));
      if (per_cu < 1) per_cu = 1;
      if (b->wg_per_cu && (int)b->wg_per_cu < per_cu) per_cu = (int)b->wg_per_cu;
      uint64_t sgrid = (uint64_t)persistent_grid(b->ctx, per_cu);
      if (sgrid > b->nsitems) sgrid = b->nsitems;
      ItemQueues SQ{};
      SQ.nq = kQueues;
      for (uint32_t i = 0; i <= kQueues; ++i) SQ.off[i] = b->squeue_off[i];
      skern<<<dim3((unsigned)sgrid), dim3(kBlock), 0, s>>>(b->family, b->d_fam, b->d_sitems,
                                                          b->d_sentries, SQ, counter + kQueues);
      TGP_HIP(hipGetLastError());
      if (!b->pgroups.empty()) {
        k_partials_scan<<<dim3((unsigned)b->punits.size()), dim3(256), 0, s>>>(
            b->d_pgroups, b->d_punits, (uint32_t)kShapes[0].naggs);
        TGP_HIP(hipGetLastError());
        k_partials_links<<<dim3((unsigned)b->plinks.size()), dim3(256), 0, s>>>(
            b->d_plinks, (uint32_t)kShapes[0].naggs);
        TGP_HIP(hipGetLastError());
      }
    }
    if (b->nitems) {
      auto kern = k_batch_items<NT>;
      TGP_TRY(persistent_launch_grid(b, kern, (b->nitems + b->per_grab - 1) / b->per_grab,
                                     &grid));
      ItemQueues Q{};
      Q.nq = b->nqueues;
      for (uint32_t i = 0; i <= kQueues; ++i) Q.off[i] = b->queue_off[i];
      kern<<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(b->d_plan, b->d_items, Q, counter,
                                                         b->per_grab);
    }
  }
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

// A member's date and key tests in the 32-bit forms the shared scan evaluates (the
// int64 bounds clamped to what int32 dt, the u8 flags and int32 dt / divisor can
// hold), and its group id as 5 t0 + st5 t1 + kc.  False: the query stays out of
// the family (a key part of more than 2^30 values, a stride past 24-bit products).
static bool shared_tests(SharedQuery& m) {
  const int64_t lo = std::max<int64_t>(m.plo, INT32_MIN), hi = std::min<int64_t>(m.phi, INT32_MAX);
  m.live = lo <= hi ? 1u : 0u;
  m.dlo = m.live ? (int32_t)lo : 0;
  m.dspan = m.live ? (uint32_t)(hi - lo) : 0u;
  static constexpr int kParts[5] = {2, 1, 1, 1, 0};  // key parts of key forms 0..4
  if (m.kind < 0 || m.kind > 4 || m.nparts != kParts[m.kind]) return false;
  uint32_t c[2] = {0, 0};
  m.klo[0] = m.klo[1] = m.kspan[0] = m.kspan[1] = 0;
  for (int k = 0; k < m.nparts; ++k) {
    if (m.gcard[k] < 1 || m.gcard[k] > (1 << 30)) return false;
    const bool wide = m.kind == 3;  // dt / divisor (int32); else a u8 flag
    const __int128 smin = wide ? (__int128)INT32_MIN : 0;
    const __int128 send = wide ? (__int128)INT32_MAX + 1 : 256;
    const __int128 l = std::max<__int128>(m.gbase[k], smin);
    const __int128 h = std::min<__int128>((__int128)m.gbase[k] + m.gcard[k], send);
    if (h > l) {  // else no value is in range: kspan 0
      m.klo[k] = (uint32_t)(int64_t)l;
      m.kspan[k] = (uint32_t)(h - l);
      c[k] = (uint32_t)(l - m.gbase[k]);  // below gcard
    }
  }
  m.st5 = 0;
  if (m.kind == 0) {
    if (m.stride1 < 0 || m.stride1 >= (1 << 22)) return false;
    m.st5 = (uint32_t)m.stride1 * 5u;
  }
  m.kc = 5u * c[0] + m.st5 * c[1];
  return true;
}

// The shared scan's items for chunks of `chunk` slices: chunk k lists the family
// members overlapping it, with each member's part of the workgroup LDS region
// (offset, 2^rshift copies: up to 64 for one group, fewer as groups grow, fewer
// again while the chunk's members do not fit), sorted by key signature; a chunk
// whose members still do not fit with one copy each, or number more than kWave,
// becomes several items over the same slices.  Chunk k's items go to queue
// k % kQueues.
static int32_t build_shared_items(tgp_batch* b, hipStream_t s) {
  const uint64_t C = b->chunk < kSharedMaxSlices ? b->chunk : kSharedMaxSlices;
  const size_t nf = b->fam.size();
  constexpr int NA = kShapes[0].naggs;
  b->family.chunk = (uint32_t)C;
  // chunks each member covers whole, [f0, f1): chunk k holds rows
  // [k C 256, min((k + 1) C 256, T)) of the T-row table
  uint64_t T = b->family.a.col_len[0];
  for (int c = 1; c < kShapes[0].nslots; ++c) T = std::min(T, b->family.a.col_len[c]);
  const uint64_t CR = C * kSliceRows, K = (T + CR - 1) / CR;
  std::vector<uint64_t> f0(nf, 0), f1(nf, 0);
  std::map<uint32_t, std::vector<uint32_t>> by_sig;
  for (size_t m = 0; m < nf; ++m) {
    const SharedQuery& M = b->fam[m];
    if (M.row_end <= M.row_begin) continue;
    f0[m] = (M.row_begin + CR - 1) / CR;
    f1[m] = M.row_end >= T ? K : M.row_end / CR;
    if (f1[m] > f0[m]) by_sig[M.sig].push_back((uint32_t)m);
  }
  // a signature with two or more such members gets a partials query over the
  // chunks they cover whole; each of them then runs only on its other chunks
  b->sfam = b->fam;
  b->pgroups.clear();
  b->plinks.clear();
  b->punits.clear();
  std::vector<int32_t> grp(nf, -1);
  std::vector<std::vector<uint8_t>> gcover;
  std::vector<uint64_t> goff;
  std::vector<uint32_t> link_group;
  uint64_t pw = 0;
  for (auto& kv : by_sig) {
    const std::vector<uint32_t>& ms = kv.second;
    if (ms.size() < 2) continue;
    if (b->pgroups.size() >= 0xffffu) break;  // group indices take 16 bits
    uint64_t F0 = ~0ull, F1 = 0, covered = 0;
    for (uint32_t m : ms) {
      F0 = std::min(F0, f0[m]);
      F1 = std::max(F1, f1[m]);
      covered += f1[m] - f0[m];
    }
    // worth it when the members cover their chunks twice over on average: the
    // partials query then scans each chunk once instead of twice or more
    std::vector<uint8_t> cover(F1 - F0, 0);
    for (uint32_t m : ms)
      for (uint64_t k = f0[m]; k < f1[m]; ++k) cover[k - F0] = 1;
    uint64_t uni = 0;
    for (uint8_t c : cover) uni += c;
    if (covered < 2 * uni) continue;
    const uint32_t G = (uint32_t)b->fam[ms[0]].words / (1 + NA);
    const uint64_t nk1 = F1 - F0 + 1;
    SharedQuery v = b->fam[ms[0]];
    v.row_begin = F0 * CR;
    v.row_end = std::min(F1 * CR, T);
    v.pstride = (uint32_t)nk1;
    v.pk0 = (uint32_t)F0;
    v.out_sums = v.out_oor = nullptr;
    const int32_t vi = (int32_t)b->sfam.size();
    b->sfam.push_back(v);
    goff.push_back(pw);
    b->pgroups.push_back(PartialGroup{nullptr, (uint32_t)nk1, G});
    for (uint32_t u = 0; u < G + G * NA + 1; ++u)
      b->punits.push_back(((uint32_t)b->pgroups.size() - 1) | (u << 16));
    for (uint32_t m : ms) {
      grp[m] = vi;
      const SharedQuery& M = b->fam[m];
      b->plinks.push_back(PartialLink{nullptr, M.out_count, M.out_sums, M.out_oor,
                                      (uint32_t)nk1, G, (uint32_t)(f0[m] - F0),
                                      (uint32_t)(f1[m] - F0)});
      link_group.push_back((uint32_t)b->pgroups.size() - 1);
    }
    gcover.push_back(std::move(cover));
    pw += (uint64_t)(G + 2 * G * NA + 1) * nk1;
  }
  b->partial_words = pw;
  if (pw) {
    TGP_TRY(ensure_cap(b, &b->d_partials, &b->partials_cap, pw));
    for (size_t g = 0; g < b->pgroups.size(); ++g) {
      auto* rows = reinterpret_cast<unsigned long long*>(b->d_partials + goff[g]);
      b->pgroups[g].rows = rows;
      b->sfam[nf + g].out_count = rows;
    }
    for (size_t li = 0; li < b->plinks.size(); ++li)
      b->plinks[li].rows = b->pgroups[link_group[li]].rows;
  }
  const size_t ns = b->sfam.size();
  // member m's chunk k is one of its entries unless the group's partials query
  // takes it; the partials query g's chunks are those its members cover whole
  std::vector<uint64_t> sb(ns), se(ns);
  uint64_t kend = 0;
  for (size_t m = 0; m < ns; ++m) {
    sb[m] = b->sfam[m].row_begin / kSliceRows;
    se[m] = b->sfam[m].row_end > b->sfam[m].row_begin
                ? (b->sfam[m].row_end + kSliceRows - 1) / kSliceRows : sb[m];
    if (se[m] > sb[m]) kend = std::max(kend, (se[m] + C - 1) / C);
  }
  // (chunk, query) pairs placed by counting in (queue, chunk, query) order: chunk k
  // is in queue k % kQueues
  const uint64_t per_q = (kend + kQueues - 1) / kQueues;
  auto pos_of = [&](uint64_t k) { return (k % kQueues) * per_q + k / kQueues; };
  std::vector<uint32_t> at(kQueues * per_q + 1, 0);
  auto each_pair = [&](auto&& fn) {
    for (size_t m = 0; m < nf; ++m) {
      if (se[m] <= sb[m]) continue;
      const uint64_t k0 = sb[m] / C, k1 = (se[m] + C - 1) / C;
      if (grp[m] < 0) {
        for (uint64_t k = k0; k < k1; ++k) fn(k, (uint32_t)m);
      } else {  // the chunks before and after those the partials query takes
        for (uint64_t k = k0; k < std::min(k1, f0[m]); ++k) fn(k, (uint32_t)m);
        for (uint64_t k = std::max(k0, f1[m]); k < k1; ++k) fn(k, (uint32_t)m);
      }
    }
    for (size_t g = 0; g < b->pgroups.size(); ++g) {
      const uint64_t F0 = b->sfam[nf + g].pk0;
      for (uint64_t k = 0; k < gcover[g].size(); ++k)
        if (gcover[g][k]) fn(F0 + k, (uint32_t)(nf + g));
    }
  };
  each_pair([&](uint64_t k, uint32_t) { at[pos_of(k) + 1] += 1; });
  for (size_t i = 1; i < at.size(); ++i) at[i] += at[i - 1];
  std::vector<std::pair<uint64_t, uint32_t>> pairs(at.back());
  // queries in order: each chunk's list stays ascending
  each_pair([&](uint64_t k, uint32_t m) { pairs[at[pos_of(k)]++] = {k, m}; });
  std::vector<uint32_t> items, entries;
  uint64_t total = 0;
  // copies per word: lanes of one atomic spread over groups x copies words.  Queries
  // of at most 8 groups keep at least 8 copies (fewer made their atomics conflict);
  // a chunk whose queries then do not fit one region becomes several items over the
  // same slices (its later items read the slices from the L2)
  auto rcap = [](int32_t words) {
    const int32_t groups = words / 5;
    return groups <= 3 ? 6 : groups <= 8 ? 5 : groups <= 32 ? 2 : 1;
  };
  auto rmin = [](int32_t words) { return words / 5 <= 8 ? 3 : 0; };
  size_t p = 0;
  uint32_t cur_queue = 0;
  b->squeue_off[0] = 0;
  // the previous chunk's member list and its items' entry ranges: a chunk with the
  // same members (the chunks inside the same long queries) reuses them
  std::vector<uint32_t> prevL;
  std::vector<std::pair<uint32_t, uint32_t>> prev_ranges;
  std::vector<uint32_t> L;
  std::vector<int32_t> r;
  while (p < pairs.size()) {
    const uint64_t k = pairs[p].first;
    size_t pe = p;
    while (pe < pairs.size() && pairs[pe].first == k) ++pe;
    const uint32_t q = (uint32_t)(k % kQueues);
    while (cur_queue < q) b->squeue_off[++cur_queue] = total;
    L.clear();
    for (size_t j = p; j < pe; ++j) L.push_back(pairs[j].second);
    // by key signature: the kernel computes group ids once per signature and slice
    std::sort(L.begin(), L.end(), [&](uint32_t x, uint32_t y) {
      return b->sfam[x].ksig != b->sfam[y].ksig ? b->sfam[x].ksig < b->sfam[y].ksig : x < y;
    });
    p = pe;
    uint64_t s0 = (k + 1) * C, s1 = k * C;
    for (uint32_t m : L) {
      s0 = std::min(s0, std::max(k * C, sb[m]));
      s1 = std::max(s1, std::min((k + 1) * C, se[m]));
    }
    if (L != prevL) {
      r.assign(L.size(), 0);
      uint64_t need = 0;
      for (size_t j = 0; j < L.size(); ++j) {
        const int32_t w = b->sfam[L[j]].words;
        r[j] = rcap(w);
        while (r[j] > 0 && ((uint64_t)w << r[j]) > (uint64_t)kSharedWords) --r[j];
        need += (uint64_t)w << r[j];
      }
      // fewer copies for the largest parts: every part of the largest size drops one
      // level per pass, down to its minimum
      while (need > (uint64_t)kSharedWords) {
        uint64_t big = 0;
        for (size_t j = 0; j < L.size(); ++j)
          if (r[j] > rmin(b->sfam[L[j]].words))
            big = std::max(big, (uint64_t)b->sfam[L[j]].words << r[j]);
        if (big == 0) break;
        for (size_t j = 0; j < L.size() && need > (uint64_t)kSharedWords; ++j)
          if (r[j] > rmin(b->sfam[L[j]].words) && ((uint64_t)b->sfam[L[j]].words << r[j]) == big) {
            need -= (uint64_t)b->sfam[L[j]].words << (r[j] - 1);
            --r[j];
          }
      }
      prev_ranges.clear();
      size_t j = 0;
      while (j < L.size()) {
        const uint32_t e0 = (uint32_t)(entries.size() / 2);
        uint64_t off = 0;
        while (j < L.size()) {
          const uint64_t sz = (uint64_t)b->sfam[L[j]].words << r[j];
          if (off + sz > (uint64_t)kSharedWords) break;
          if (entries.size() / 2 - e0 >= (uint64_t)kWave) break;  // one entry per lane
          entries.push_back(L[j]);
          entries.push_back((uint32_t)off | ((uint32_t)r[j] << 16));
          off += sz;
          ++j;
        }
        prev_ranges.push_back({e0, (uint32_t)(entries.size() / 2) - e0});
      }
      prevL = L;
    }
    for (const auto& er : prev_ranges) {
      items.push_back((uint32_t)s0);
      items.push_back((uint32_t)(s1 - s0));
      items.push_back(er.first);
      items.push_back(er.second);
      ++total;
    }
  }
  while (cur_queue < kQueues - 1) b->squeue_off[++cur_queue] = total;
  b->squeue_off[kQueues] = total;
  if (items.empty()) items.assign(4, 0);
  if (entries.empty()) entries.assign(2, 0);
  if (!b->sfam.empty()) {
    TGP_TRY(ensure_cap(b, &b->d_fam, &b->fam_cap, (uint64_t)b->sfam.size()));
    TGP_HIP(hipMemcpyAsync(b->d_fam, b->sfam.data(), b->sfam.size() * sizeof(SharedQuery),
                           hipMemcpyHostToDevice, s));
  }
  if (!b->pgroups.empty()) {
    TGP_TRY(ensure_cap(b, &b->d_pgroups, &b->pgroups_cap, (uint64_t)b->pgroups.size()));
    TGP_TRY(ensure_cap(b, &b->d_plinks, &b->plinks_cap, (uint64_t)b->plinks.size()));
    TGP_HIP(hipMemcpyAsync(b->d_pgroups, b->pgroups.data(),
                           b->pgroups.size() * sizeof(PartialGroup), hipMemcpyHostToDevice, s));
    TGP_HIP(hipMemcpyAsync(b->d_plinks, b->plinks.data(), b->plinks.size() * sizeof(PartialLink),
                           hipMemcpyHostToDevice, s));
    TGP_TRY(ensure_cap(b, &b->d_punits, &b->punits_cap, (uint64_t)b->punits.size()));
    TGP_HIP(hipMemcpyAsync(b->d_punits, b->punits.data(), b->punits.size() * 4,
                           hipMemcpyHostToDevice, s));
  }
  TGP_TRY(ensure_cap(b, &b->d_sitems, &b->sitems_cap, (uint64_t)items.size()));
  TGP_TRY(ensure_cap(b, &b->d_sentries, &b->sentries_cap, (uint64_t)entries.size()));
  TGP_HIP(hipMemcpyAsync(b->d_sitems, items.data(), items.size() * 4, hipMemcpyHostToDevice, s));
  TGP_HIP(hipMemcpyAsync(b->d_sentries, entries.data(), entries.size() * 4,
                         hipMemcpyHostToDevice, s));
  TGP_HIP(hipStreamSynchronize(s));
  b->nsitems = total;
  b->sitems_chunk = b->chunk;
  b->sitems_valid = true;
  return TGP_OK;
}

// Work items of the table-ordered schedules for the current chunk size
// (synchronous upload; called outside graph capture).
//   order 1: every query's slices in chunks of `chunk` slices, one queue sorted by
//            the table row each item starts at;
//   order 2: the table cut into absolute chunks of chunk * 256 rows; every query
//            has one item per chunk it overlaps, so the items of all queries on
//            chunk k cover the same rows; chunk k's items go to queue k % kQueues.
// The query-ordered kernel's slice -> query table: granule g holds the plan index
// of the query that contains slice g * kGranSlices (synchronous upload).
static int32_t build_granules(tgp_batch* b, hipStream_t s) {
  const uint64_t ngran = (b->total_slices + kGranSlices - 1) / kGranSlices;
  std::vector<uint32_t> gran(ngran ? ngran : 1, 0);
  for (uint64_t g = 0, q = 0; g < ngran; ++g) {
    while (q + 1 < b->plan.size() && b->plan[q + 1].slice0 <= g * kGranSlices) ++q;
    gran[g] = (uint32_t)q;
  }
  TGP_TRY(ensure_cap(b, &b->d_granule, &b->granule_cap, gran.size()));
  TGP_HIP(hipMemcpyAsync(b->d_granule, gran.data(), gran.size() * 4, hipMemcpyHostToDevice, s));
  TGP_HIP(hipStreamSynchronize(s));
  b->granule_valid = true;
  return TGP_OK;
}

// The queries of no batch shape run as generic items of kGenericTiles tiles in one
// more persistent launch: their plans (results in the arena), items and the LDS
// they need (two stage buffers; a table for every plan whose groups fit beside
// them, the others add straight into the result).
constexpr uint32_t kGenericTiles = 16;
constexpr uint32_t kGenericLds = 160 * 1024 - (uint32_t)sizeof(DevPlan) - 1024;
// Generic queries of at least this many rows whose groups exceed an LDS table take
// the partitioned passes (smaller ones add into the result as items).
constexpr uint64_t kBatchPartRows = 1ull << 16;

// Generic queries of at least kSpecRunRows rows whose specialised shape has kernels
// ready run those kernels, one launch each (smaller ones cost more in launches than
// as items); under TGP_TUNE_JIT 0 a shape not compiled yet starts compiling in the
// background (later batches use it) and its queries stay generic items; 2 compiles
// it now; 1 never.
constexpr uint64_t kSpecRunRows = 1ull << 20;
static int32_t plan_spec_runs(tgp_batch* b, const tgp_agg_query* qs) {
  b->spec_runs.clear();
  const int64_t mode = b->ctx->tune_jit.load();
  if (mode == 1 || b->ctx->tune_no_spec.load()) return TGP_OK;
  std::vector<int32_t> rest;
  for (int32_t i : b->fallback) {
    tgp_batch::SpecRun r;
    bool ok = false;
    if (qs[i].row_end - qs[i].row_begin >= kSpecRunRows &&
        extract_shape(&qs[i], &r.shape, &r.a))
      TGP_TRY(find_shape_kernels(r.shape, mode == 2, &r.k, &ok));
    if (!ok) {
      rest.push_back(i);
      continue;
    }
    r.a.out_count = reinterpret_cast<unsigned long long*>(b->d_results + b->count_off[i]);
    r.a.out_sums = reinterpret_cast<unsigned long long*>(b->d_results + b->sums_off[i]);
    r.a.out_oor = reinterpret_cast<unsigned long long*>(b->d_results + b->oor_off[i]);
    b->spec_runs.push_back(r);
  }
  b->fallback.swap(rest);
  if (!b->spec_runs.empty()) {
    TGP_TRY(ensure_cap(b, &b->d_spec_counter, &b->spec_counter_cap, b->spec_runs.size()));
    // created here, outside any graph capture of the launch
    if (!b->spec_fork) {
      TGP_HIP(hipEventCreateWithFlags(&b->spec_fork, hipEventDisableTiming));
      for (int i = 0; i < tgp_batch::kSpecStreams; ++i) {
        TGP_HIP(hipStreamCreateWithFlags(&b->spec_streams[i], hipStreamNonBlocking));
        TGP_HIP(hipEventCreateWithFlags(&b->spec_join[i], hipEventDisableTiming));
      }
    }
  }
  return TGP_OK;
}

static int32_t plan_generic(tgp_batch* b, const tgp_agg_query* qs, hipStream_t s) {
  TGP_TRY(plan_spec_runs(b, qs));
  b->gplans.clear();
  b->gitems.clear();
  b->gminmax.clear();
  b->gstage_max = b->gtable_max = 0;
  b->gminmax_words = 0;
  b->hplans.clear();
  b->pqs.clear();
  b->pitems.clear();
  b->pchunks.clear();
  b->phstage_max = b->pstage_max = b->pnparts_max = b->ptable_max = 0;
  std::vector<PlanInfo> infos;
  std::vector<int32_t> orig;
  for (int32_t i : b->fallback) {
    DevPlan P;
    PlanInfo info;
    TGP_TRY(plan_agg(&qs[i], &P, &info));
    P.out_count = reinterpret_cast<unsigned long long*>(b->d_results + b->count_off[i]);
    P.out_sums = reinterpret_cast<unsigned long long*>(b->d_results + b->sums_off[i]);
    P.out_oor = reinterpret_cast<unsigned long long*>(b->d_results + b->oor_off[i]);
    b->gstage_max = std::max(b->gstage_max, P.stage_bytes);
    b->gplans.push_back(P);
    infos.push_back(info);
    orig.push_back(i);
  }
  uint64_t mat_n = 0, grp_n = 0, val_n = 0;
  for (uint32_t q = 0; q < b->gplans.size(); ++q) {
    DevPlan& P = b->gplans[q];
    const uint64_t tb = P.ngroups * (1 + 2 * (uint64_t)P.naggs) * 8;
    const bool lds = 2ull * b->gstage_max + tb <= kGenericLds;
    while (P.lds_rshift > 0 && 2ull * b->gstage_max + (tb << P.lds_rshift) > kGenericLds)
      --P.lds_rshift;
    if (lds) b->gtable_max = std::max<uint32_t>(b->gtable_max, (uint32_t)(tb << P.lds_rshift));
    const bool split = !lds && P.row_end - P.row_begin < (1ull << 32);
    const uint64_t rows = P.row_end - P.row_begin;
    if (split && infos[q].nparts && rows >= kBatchPartRows && P.ntiles) {
      DevPlan H;
      TGP_TRY(plan_part_hist(&qs[orig[q]], &H));
      const uint32_t np = infos[q].nparts;
      const uint32_t hst = std::max(b->phstage_max, H.stage_bytes);
      const uint32_t st = std::max(b->pstage_max, P.stage_bytes);
      const uint32_t npm = std::max(b->pnparts_max, np);
      if (2ull * hst + 4ull * npm <= kGenericLds && 2ull * st + 12ull * npm <= kGenericLds) {
        b->phstage_max = hst;
        b->pstage_max = st;
        b->pnparts_max = npm;
        const uint32_t nit = (uint32_t)((P.ntiles + kGenericTiles - 1) / kGenericTiles);
        const uint32_t pq = (uint32_t)b->pqs.size();
        b->pqs.push_back(PartQuery{q, (uint32_t)b->hplans.size(), infos[q].part_shift, np, nit,
                                   (uint32_t)P.naggs, mat_n, grp_n, val_n});
        b->hplans.push_back(H);
        mat_n += (uint64_t)np * nit;
        grp_n += rows;
        val_n += rows * (uint64_t)P.naggs;
        for (uint32_t j = 0; j < nit; ++j)
          b->pitems.push_back(PartItem{pq, j, j * kGenericTiles,
                                       (uint32_t)std::min<uint64_t>((uint64_t)(j + 1) * kGenericTiles,
                                                                    P.ntiles)});
        for (uint64_t c = 0; c * kPartChunk < rows; ++c)
          b->pchunks.push_back(PartChunk{pq, (uint32_t)c});
        b->ptable_max = std::max<uint32_t>(
            b->ptable_max, (uint32_t)((1ull << infos[q].part_shift) * (1 + 2 * P.naggs) * 8));
        if (P.minmax_kinds && P.naggs) {  // the partitioned sums are int128 words already
          b->gminmax.push_back(q);
          b->gminmax_words = std::max(b->gminmax_words, P.ngroups * (uint64_t)P.naggs);
        }
        continue;
      }
    }
    if ((P.minmax_kinds || split) && P.ngroups && P.naggs && P.ntiles) {
      b->gminmax.push_back(q | (split ? 1u << 31 : 0u));
      b->gminmax_words = std::max(b->gminmax_words, P.ngroups * (uint64_t)P.naggs);
    }
    for (uint64_t t = 0; t < P.ntiles; t += kGenericTiles)
      b->gitems.push_back(GenericItem{q, (uint32_t)t,
                                      (uint32_t)std::min<uint64_t>(t + kGenericTiles, P.ntiles),
                                      lds ? 1u : split ? 0u : 2u});
  }
  b->fallback.clear();  // every query runs in the batch launch
  b->pmat_n = mat_n;
  if (!b->pqs.empty()) {
    // d_pmat: 3 counters, a pad word, mat_n counts + the total, the scan's block sums
    const uint64_t words = 4 + mat_n + 1 + (mat_n + 4095) / 4096 + 1;
    TGP_TRY(ensure_cap(b, &b->d_pmat, &b->pmat_cap, words));
    TGP_TRY(ensure_cap(b, &b->d_pgrp, &b->pgrp_cap, std::max<uint64_t>(grp_n, 1)));
    TGP_TRY(ensure_cap(b, &b->d_pval, &b->pval_cap, std::max<uint64_t>(val_n, 1)));
    TGP_TRY(ensure_cap(b, &b->d_hplans, &b->hplans_cap, (uint64_t)b->hplans.size()));
    TGP_TRY(ensure_cap(b, &b->d_pqs, &b->pqs_cap, (uint64_t)b->pqs.size()));
    TGP_TRY(ensure_cap(b, &b->d_pitems, &b->pitems_cap, (uint64_t)b->pitems.size()));
    TGP_TRY(ensure_cap(b, &b->d_pchunks, &b->pchunks_cap, (uint64_t)b->pchunks.size()));
    TGP_HIP(hipMemcpyAsync(b->d_hplans, b->hplans.data(), b->hplans.size() * sizeof(DevPlan),
                           hipMemcpyHostToDevice, s));
    TGP_HIP(hipMemcpyAsync(b->d_pqs, b->pqs.data(), b->pqs.size() * sizeof(PartQuery),
                           hipMemcpyHostToDevice, s));
    TGP_HIP(hipMemcpyAsync(b->d_pitems, b->pitems.data(), b->pitems.size() * sizeof(PartItem),
                           hipMemcpyHostToDevice, s));
    TGP_HIP(hipMemcpyAsync(b->d_pchunks, b->pchunks.data(),
                           b->pchunks.size() * sizeof(PartChunk), hipMemcpyHostToDevice, s));
  }
  if (b->gplans.empty()) return TGP_OK;
  TGP_TRY(ensure_cap(b, &b->d_gplans, &b->gplans_cap, (uint64_t)b->gplans.size()));
  TGP_HIP(hipMemcpyAsync(b->d_gplans, b->gplans.data(), b->gplans.size() * sizeof(DevPlan),
                         hipMemcpyHostToDevice, s));
  if (!b->gitems.empty()) {
    TGP_TRY(ensure_cap(b, &b->d_gitems, &b->gitems_cap, (uint64_t)b->gitems.size()));
    TGP_HIP(hipMemcpyAsync(b->d_gitems, b->gitems.data(), b->gitems.size() * sizeof(GenericItem),
                           hipMemcpyHostToDevice, s));
  }
  if (!b->gminmax.empty()) {
    TGP_TRY(ensure_cap(b, &b->d_gminmax, &b->gminmax_cap, (uint64_t)b->gminmax.size()));
    TGP_HIP(hipMemcpyAsync(b->d_gminmax, b->gminmax.data(), b->gminmax.size() * 4,
                           hipMemcpyHostToDevice, s));
  }
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

static int32_t build_items(tgp_batch* b, hipStream_t s) {
  if (b->order == 0 && !b->granule_valid) TGP_TRY(build_granules(b, s));
  if (b->order == 3 && !(b->sitems_valid && b->sitems_chunk == b->chunk))
    TGP_TRY(build_shared_items(b, s));
  if (b->order == 0 || (b->items_chunk == b->chunk && b->items_order == b->order &&
                        b->items_valid))
    return TGP_OK;
  const uint32_t nq = b->order >= 2 ? kQueues : 1;
  // order 3: the family members run in the shared scan, the rest as in order 2
  auto skip = [&](uint32_t q) { return b->order == 3 && b->in_family[q]; };
  std::vector<uint32_t> items;
  uint64_t total = 0;
  if (b->order == 1) {
    // one queue sorted by the table row each item starts at
    struct Item {
      uint64_t key;
      uint32_t q, sb, se;
    };
    std::vector<Item> qv;
    for (uint32_t q = 0; q < b->plan.size(); ++q) {
      const SpecArgs& A = b->plan[q].a;
      for (uint64_t c = 0; c < A.nslices; c += b->chunk)
        qv.push_back(Item{A.base0 + c * kSliceRows, q, (uint32_t)c,
                          (uint32_t)(c + b->chunk < A.nslices ? c + b->chunk : A.nslices)});
    }
    std::sort(qv.begin(), qv.end(), [](const Item& x, const Item& y) {
      return x.key != y.key ? x.key < y.key : x.q < y.q;
    });
    total = qv.size();
    items.reserve(3 * (total ? total : 1));
    b->queue_off[0] = 0;
    for (const Item& it : qv) {
      items.push_back(it.q);
      items.push_back(it.sb);
      items.push_back(it.se);
    }
  } else {
    // XCD-affine: absolute table chunk k goes to queue k % 8, items in (k, query)
    // order.  Placed by counting (no sort): a chunk's items are the queries that
    // overlap it, visited in query order.
    const uint64_t C = (uint64_t)b->chunk * kSliceRows;  // rows per absolute chunk
    uint64_t kmax = 0;
    for (uint32_t q = 0; q < b->plan.size(); ++q) {
      const BatchQuery& bq = b->plan[q];
      if (bq.a.nslices && !skip(q)) {
        const uint64_t end = bq.a.base0 + bq.a.nslices * kSliceRows;
        kmax = (end + C - 1) / C > kmax ? (end + C - 1) / C : kmax;
      }
    }
    // position of chunk k's first item: queues in order, chunks k = i, i + 8, ... in queue i
    std::vector<uint64_t> at(kmax + 1, 0);
    for (uint32_t q = 0; q < b->plan.size(); ++q) {
      const BatchQuery& bq = b->plan[q];
      if (bq.a.nslices && !skip(q)) {
        const uint64_t end = bq.a.base0 + bq.a.nslices * kSliceRows;
        for (uint64_t k = bq.a.base0 / C; k * C < end; ++k) at[k] += 1;
      }
    }
    for (uint32_t i = 0; i < nq; ++i) {
      b->queue_off[i] = total;
      for (uint64_t k = i; k < kmax; k += nq) {
        const uint64_t c = at[k];
        at[k] = total;
        total += c;
      }
    }
    items.resize(3 * (total ? total : 1));
    for (uint32_t q = 0; q < b->plan.size(); ++q) {
      const SpecArgs& A = b->plan[q].a;
      if (A.nslices == 0 || skip(q)) continue;
      const uint64_t end = A.base0 + A.nslices * kSliceRows;
      for (uint64_t k = A.base0 / C; k * C < end; ++k) {
        const uint64_t r0 = k * C > A.base0 ? k * C : A.base0;
        const uint64_t r1 = (k + 1) * C < end ? (k + 1) * C : end;
        const uint64_t pos = at[k]++;
        items[3 * pos] = q;
        items[3 * pos + 1] = (uint32_t)((r0 - A.base0) / kSliceRows);
        items[3 * pos + 2] = (uint32_t)((r1 - A.base0) / kSliceRows);
      }
    }
  }
  for (uint32_t i = nq; i <= kQueues; ++i) b->queue_off[i] = total;
  if (items.empty()) items.assign(3, 0);
  TGP_TRY(ensure_cap(b, &b->d_items, &b->items_cap, (uint64_t)items.size()));
  TGP_HIP(hipMemcpyAsync(b->d_items, items.data(), items.size() * 4, hipMemcpyHostToDevice, s));
  TGP_HIP(hipStreamSynchronize(s));
  b->nitems = total;
  b->nqueues = nq;
  b->items_chunk = b->chunk;
  b->items_order = b->order;
  b->items_valid = true;
  return TGP_OK;
}

static int32_t batch_launch_rest(tgp_batch* b, hipStream_t s);

// Side streams used by the specialised runs: min(runs, kSpecStreams).
static int spec_stream_count(const tgp_batch* b) {
  return (int)std::min<size_t>(tgp_batch::kSpecStreams, b->spec_runs.size());
}

static int32_t batch_launch(tgp_batch* b, hipStream_t s) {
  // results and the counters behind them start at zero
  TGP_TRY(fill_u64(s, b->d_results, b->result_words + 2 * kQueues + 1, 0));
  const int ns = spec_stream_count(b);
  if (ns) {
    TGP_HIP(hipEventRecord(b->spec_fork, s));
    for (int i = 0; i < ns; ++i) TGP_HIP(hipStreamWaitEvent(b->spec_streams[i], b->spec_fork, 0));
    for (size_t k = 0; k < b->spec_runs.size(); ++k) {
      const auto& r = b->spec_runs[k];
      TGP_TRY(launch_shape(b->ctx, r.shape, r.k, r.a, b->spec_streams[k % ns],
                           b->d_spec_counter + k, true));
    }
  }
  const int32_t rc = batch_launch_rest(b, s);
  for (int i = 0; i < ns; ++i) {  // joined on every path: a capture must end joined
    TGP_HIP(hipEventRecord(b->spec_join[i], b->spec_streams[i]));
    TGP_HIP(hipStreamWaitEvent(s, b->spec_join[i], 0));
  }
  return rc;
}

// The batch's kernels on its own stream after the result fill.
static int32_t batch_launch_rest(tgp_batch* b, hipStream_t s) {
  if (!b->pqs.empty()) {  // before the generic launch, whose finalisation covers them
    TGP_TRY(fill_u64(s, reinterpret_cast<uint64_t*>(b->d_pmat), 3, 0));
    unsigned long long* mat = b->d_pmat + 4;
    TGP_TRY(launch_batch_part(b->ctx, b->d_gplans, b->d_hplans, b->d_pqs, b->d_pitems,
                              b->pitems.size(), b->d_pchunks, b->pchunks.size(), b->phstage_max,
                              b->pstage_max, b->pnparts_max, b->ptable_max, b->d_pmat, mat,
                              b->pmat_n, reinterpret_cast<uint64_t*>(mat + b->pmat_n + 1),
                              b->d_pgrp, b->d_pval, s));
  }
  if (!b->gplans.empty()) {
    auto* counter = reinterpret_cast<unsigned long long*>(b->d_results + b->result_words +
                                                          2 * kQueues);
    TGP_TRY(launch_batch_generic(b->ctx, b->d_gplans, b->d_gitems, b->gitems.size(), counter,
                                 b->gstage_max, b->gtable_max, b->d_gminmax,
                                 (uint32_t)b->gminmax.size(), b->gminmax_words, s));
  }
  if (!b->total_slices) return TGP_OK;
  if (b->order == 3 && b->partial_words)
    TGP_TRY(fill_u64(s, b->d_partials, b->partial_words, 0));
  return b->cached ? launch_variant<false>(b, s) : launch_variant<true>(b, s);
}

extern "C" {

int32_t tgp_batch_create(tgp_ctx* ctx, uint32_t max_queries, tgp_batch** out) {
  TGP_REQUIRE(ctx && out && max_queries > 0, "tgp_batch_create: bad argument");
  auto* b = new tgp_batch();
  b->ctx = ctx;
  b->max_queries = max_queries;
  *out = b;
  return TGP_OK;
}

int32_t tgp_batch_destroy(tgp_batch* b) {
  if (!b) return TGP_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  drop_graph(b);
  for (int i = 0; i < tgp_batch::kSpecStreams; ++i) {
    if (b->spec_streams[i]) {
      (void)hipStreamSynchronize(b->spec_streams[i]);
      (void)hipStreamDestroy(b->spec_streams[i]);
    }
    if (b->spec_join[i]) (void)hipEventDestroy(b->spec_join[i]);
  }
  if (b->spec_fork) (void)hipEventDestroy(b->spec_fork);
  if (b->d_plan) hbm_free(b->ctx, b->d_plan);
  if (b->d_results) hbm_free(b->ctx, b->d_results);
  if (b->d_granule) hbm_free(b->ctx, b->d_granule);
  if (b->d_items) hbm_free(b->ctx, b->d_items);
  if (b->d_fam) hbm_free(b->ctx, b->d_fam);
  if (b->d_sitems) hbm_free(b->ctx, b->d_sitems);
  if (b->d_sentries) hbm_free(b->ctx, b->d_sentries);
  if (b->d_pgroups) hbm_free(b->ctx, b->d_pgroups);
  if (b->d_plinks) hbm_free(b->ctx, b->d_plinks);
  if (b->d_punits) hbm_free(b->ctx, b->d_punits);
  if (b->d_gplans) hbm_free(b->ctx, b->d_gplans);
  if (b->d_gitems) hbm_free(b->ctx, b->d_gitems);
  if (b->d_gminmax) hbm_free(b->ctx, b->d_gminmax);
  if (b->d_partials) hbm_free(b->ctx, b->d_partials);
  for (void* p : {(void*)b->d_hplans, (void*)b->d_pqs, (void*)b->d_pitems, (void*)b->d_pchunks,
                  (void*)b->d_pmat, (void*)b->d_pgrp, (void*)b->d_pval,
                  (void*)b->d_spec_counter})
    if (p) hbm_free(b->ctx, p);
  for (void* p : b->retired) hbm_free(b->ctx, p);
  delete b;
  return TGP_OK;
}

int32_t tgp_batch_set_queries(tgp_batch* b, const tgp_agg_query* qs, uint32_t n, void* stream) {
  TGP_REQUIRE(b && (n == 0 || qs), "tgp_batch_set_queries: null argument");
  TGP_REQUIRE(n <= b->max_queries, "batch of %u queries above its capacity %u", n,
              b->max_queries);
  TGP_TRY(check_injected(b->ctx));
  hipStream_t s = pick_stream(b->ctx, stream);
  drop_graph(b);
  b->n = n;
  b->queries.assign(qs, qs + n);
  b->plan.clear();
  b->fallback.clear();
  b->count_off.assign(n, 0);
  b->sums_off.assign(n, 0);
  b->oor_off.assign(n, 0);
  b->ngroups.assign(n, 0);
  // result arena layout per query: count[ngroups], sums[ngroups*naggs*2], oor
  uint64_t words = 0;
  for (uint32_t i = 0; i < n; ++i) {
    DevPlan P;
    PlanInfo info;
    TGP_TRY(plan_agg(&qs[i], &P, &info));  // validation (same rules as tgp_group_agg)
    b->ngroups[i] = P.ngroups;
    b->count_off[i] = words;
    words += P.ngroups;
    b->sums_off[i] = words;
    words += P.ngroups * qs[i].naggs * 2;
    b->oor_off[i] = words;
    words += 1;
  }
  b->result_words = words;
  // the arena ends with the persistent kernels' work counters
  TGP_TRY(ensure_cap(b, &b->d_results, &b->results_cap, words + 2 * kQueues + 1));
  uint64_t slices = 0;
  for (uint32_t i = 0; i < n; ++i) {
    Shape sh;
    BatchQuery bq{};
    int idx = -1;
    if (extract_shape(&qs[i], &sh, &bq.a)) {
      for (int k = 0; k < kNumShapes; ++k)
        if (shape_eq(sh, kShapes[k])) idx = k;
    }
    if (idx < 0 || sh.has_sel || wave_rshift(bq.a.words) < 0) {
      b->fallback.push_back((int32_t)i);
      continue;
    }
    bq.shape = idx;
    bq.words = (int32_t)bq.a.words;
    bq.naggs = qs[i].naggs;
    bq.rshift = wave_rshift(bq.a.words);
    bq.slice0 = slices;
    bq.a.out_count = reinterpret_cast<unsigned long long*>(b->d_results + b->count_off[i]);
    bq.a.out_sums = reinterpret_cast<unsigned long long*>(b->d_results + b->sums_off[i]);
    bq.a.out_oor = reinterpret_cast<unsigned long long*>(b->d_results + b->oor_off[i]);
    slices += bq.a.nslices;
    b->plan.push_back(bq);
  }
  b->total_slices = slices;
  const uint64_t bytes = b->plan.size() * sizeof(BatchQuery);
  TGP_TRY(ensure_cap(b, &b->d_plan, &b->plan_cap, b->plan.size() ? b->plan.size() : 1));
  b->granule_valid = false;  // built when the query-ordered kernel runs
  TGP_TRY(plan_generic(b, qs, s));
  if (bytes) {
    TGP_HIP(hipMemcpyAsync(b->d_plan, b->plan.data(), bytes, hipMemcpyHostToDevice, s));
    TGP_HIP(hipStreamSynchronize(s));
  }
  // the shared-scan family: Q1-layout queries on the first such query's columns
  // with its aggregate arguments (and, grouping by dt / divisor, its divisor)
  b->in_family.assign(b->plan.size(), 0);
  b->fam.clear();
  b->fam_plan.clear();
  b->family = SharedFamily{};
  int32_t ref = -1;
  // signature ids: members with equal date bounds, key form and key parts
  std::map<std::vector<int64_t>, uint32_t> sigs;
  auto sig_of = [&](const SharedQuery& m) {
    const std::vector<int64_t> k{m.plo, m.phi, m.kind, m.nparts, m.gbase[0], m.gbase[1],
                                 m.gcard[0], m.gcard[1], m.stride1, m.src[0], m.src[1]};
    auto it = sigs.find(k);
    if (it != sigs.end()) return it->second;
    const uint32_t id = (uint32_t)sigs.size();
    sigs.emplace(k, id);
    return id;
  };
  // key signature ids: members whose group ids and key tests are equal
  std::map<std::vector<uint32_t>, uint32_t> ksigs;
  auto ksig_of = [&](const SharedQuery& m) {
    const std::vector<uint32_t> k{(uint32_t)m.kind, m.klo[0], m.kspan[0], m.klo[1], m.kspan[1],
                                  m.kc, m.st5};
    auto it = ksigs.find(k);
    if (it != ksigs.end()) return it->second;
    const uint32_t id = (uint32_t)ksigs.size();
    ksigs.emplace(k, id);
    return id;
  };
  for (uint32_t q = 0; q < b->plan.size(); ++q) {
    const BatchQuery& bq = b->plan[q];
    if (bq.shape > 4 || bq.naggs != kShapes[0].naggs) continue;
    const SpecArgs& A = bq.a;
    if (ref < 0) {
      ref = (int32_t)q;
      b->family.a = A;
    }
    const SpecArgs& R = b->family.a;
    bool same = true;
    for (int c = 0; c < kShapes[0].nslots; ++c)
      same = same && A.col[c] == R.col[c] && A.col_len[c] == R.col_len[c];
    for (int a = 0; a < kShapes[0].naggs; ++a) {
      same = same && A.coef[a] == R.coef[a];
      for (int k = 0; k < TGP_MAX_FACTORS; ++k)
        same = same && A.fa[a][k] == R.fa[a][k] && A.fb[a][k] == R.fb[a][k];
    }
    const Shape& S = kShapes[bq.shape];
    for (int k = 0; k < S.nparts; ++k)
      if (S.gdiv[k]) {
        if (b->family.divisor == 0) {
          b->family.divisor = A.gdivisor[k];
          b->family.magic = A.gmagic[k];
          b->family.shift = A.gshift[k];
        }
        same = same && A.gdivisor[k] == b->family.divisor;
      }
    if (!same) continue;
    SharedQuery m{};
    m.row_begin = A.row_begin;
    m.row_end = A.row_end;
    m.plo = A.plo[0];
    m.phi = A.phi[0];
    m.nparts = S.nparts;
    for (int k = 0; k < S.nparts; ++k) {
      m.src[k] = S.gslot[k] == 5 ? 1 : S.gslot[k] == 6 ? 2 : 0;
      m.gbase[k] = A.gbase[k];
      m.gcard[k] = A.gcard[k];
    }
    m.stride1 = S.nparts == 2 ? (int32_t)A.gstride[1] : 0;
    m.words = (int32_t)A.words;
    m.kind = bq.shape;
    if (!shared_tests(m)) continue;
    m.sig = sig_of(m);
    m.ksig = ksig_of(m);
    m.out_count = A.out_count;
    m.out_sums = A.out_sums;
    m.out_oor = A.out_oor;
    b->fam.push_back(m);
    b->fam_plan.push_back(q);
    b->in_family[q] = 1;
  }
  b->items_valid = false;
  b->sitems_valid = false;
  return build_items(b, s);
}

int32_t tgp_batch_set_tuning(tgp_batch* b, const tgp_batch_tuning* t) {
  TGP_REQUIRE(b && t, "tgp_batch_set_tuning: null argument");
  TGP_REQUIRE(t->chunk_slices <= (1u << 16), "chunk of %u slices above 65536", t->chunk_slices);
  TGP_REQUIRE(t->order >= 0 && t->order <= 3, "order %d not 0 (queries), 1 (table), 2 "
              "(table, XCD-affine) or 3 (shared scan)", t->order);
  TGP_REQUIRE(t->items_per_grab <= 64, "items_per_grab %u above 64", t->items_per_grab);
  drop_graph(b);
  b->wg_per_cu = t->wg_per_cu;
  b->chunk = t->chunk_slices == 0 ? 32 : t->chunk_slices;
  b->order = t->order;
  b->cached = t->cached_loads != 0;
  b->use_graph = t->use_graph != 0;
  b->per_grab = t->items_per_grab == 0 ? 1 : t->items_per_grab;
  return build_items(b, b->ctx->stream);
}

int32_t tgp_batch_get_tuning(tgp_batch* b, tgp_batch_tuning* t) {
  TGP_REQUIRE(b && t, "tgp_batch_get_tuning: null argument");
  t->wg_per_cu = b->wg_per_cu;
  t->chunk_slices = b->chunk;
  t->order = b->order;
  t->cached_loads = b->cached ? 1 : 0;
  t->use_graph = b->use_graph ? 1 : 0;
  t->items_per_grab = b->per_grab;
  return TGP_OK;
}

int32_t tgp_batch_run(tgp_batch* b, int32_t mode, void* stream) {
  TGP_REQUIRE(b, "tgp_batch_run: null batch");
  TGP_REQUIRE(mode == TGP_BATCH_PERSISTENT || mode == TGP_BATCH_PER_QUERY, "bad mode %d", mode);
  TGP_TRY(check_injected(b->ctx));
  hipStream_t s = pick_stream(b->ctx, stream);
  if (mode == TGP_BATCH_PER_QUERY) {
    // the one-launch-per-query form: tgp_group_agg for every query
    for (uint32_t i = 0; i < b->n; ++i) {
      tgp_group_result r{b->d_results + b->count_off[i], b->d_results + b->sums_off[i],
                         b->d_results + b->oor_off[i]};
      TGP_TRY(tgp_group_agg(b->ctx, &b->queries[i], &r, s));
    }
    return TGP_OK;
  }
  if (b->use_graph) {
    if (!b->graph || b->graph_stream != s) {
      drop_graph(b);
      TGP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int32_t rc = batch_launch(b, s);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      TGP_TRY(rc);
      if (e != hipSuccess) {
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return TGP_ERR_HIP;
      }
      b->graph_def = g;
      TGP_HIP(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
      b->graph_stream = s;
    }
    TGP_HIP(hipGraphLaunch(b->graph, s));
  } else {
    TGP_TRY(batch_launch(b, s));
  }
  return TGP_OK;
}

int32_t tgp_batch_run_timed(tgp_batch* b, int32_t mode, void* stream, float* h_query_ms,
                            float* h_total_ms) {
  TGP_REQUIRE(b && h_total_ms, "tgp_batch_run_timed: null argument");
  hipStream_t s = pick_stream(b->ctx, stream);
  std::vector<hipEvent_t> ev;
  const uint32_t nev = mode == TGP_BATCH_PER_QUERY ? b->n + 1 : 2;
  ev.resize(nev);
  for (auto& e : ev) TGP_HIP(hipEventCreate(&e));
  int32_t rc = TGP_OK;
  TGP_HIP(hipEventRecord(ev[0], s));
  if (mode == TGP_BATCH_PER_QUERY) {
    for (uint32_t i = 0; i < b->n && rc == TGP_OK; ++i) {
      tgp_group_result r{b->d_results + b->count_off[i], b->d_results + b->sums_off[i],
                         b->d_results + b->oor_off[i]};
      rc = tgp_group_agg(b->ctx, &b->queries[i], &r, s);
      if (rc == TGP_OK) TGP_HIP(hipEventRecord(ev[i + 1], s));
    }
  } else {
    rc = tgp_batch_run(b, mode, s);
    if (rc == TGP_OK) TGP_HIP(hipEventRecord(ev[1], s));
  }
  if (rc == TGP_OK) {
    TGP_HIP(hipEventSynchronize(ev[nev - 1]));
    float total = 0;
    TGP_HIP(hipEventElapsedTime(&total, ev[0], ev[nev - 1]));
    *h_total_ms = total;
    if (h_query_ms)
      for (uint32_t i = 0; i < b->n; ++i) {
        if (mode == TGP_BATCH_PER_QUERY)
          TGP_HIP(hipEventElapsedTime(&h_query_ms[i], ev[0], ev[i + 1]));
        else
          h_query_ms[i] = total;  // every query of a persistent batch completes at its end
      }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int32_t tgp_batch_info(tgp_batch* b, uint64_t* result_words, uint64_t* total_slices,
                       uint32_t* n_batched, uint32_t* n_fallback) {
  TGP_REQUIRE(b, "tgp_batch_info: null batch");
  if (result_words) *result_words = b->result_words;
  if (total_slices) *total_slices = b->total_slices;
  if (n_batched) *n_batched = (uint32_t)(b->plan.size() + b->gplans.size() + b->spec_runs.size());
  if (n_fallback) *n_fallback = (uint32_t)b->fallback.size();
  return TGP_OK;
}

int32_t tgp_batch_result_layout(tgp_batch* b, uint32_t i, uint64_t* count_off, uint64_t* sums_off,
                                uint64_t* oor_off, uint64_t* ngroups) {
  TGP_REQUIRE(b && i < b->n, "tgp_batch_result_layout: bad query index");
  if (count_off) *count_off = b->count_off[i];
  if (sums_off) *sums_off = b->sums_off[i];
  if (oor_off) *oor_off = b->oor_off[i];
  if (ngroups) *ngroups = b->ngroups[i];
  return TGP_OK;
}

int32_t tgp_batch_results_sync(tgp_batch* b, uint64_t* h_words, uint64_t nwords, void* stream) {
  TGP_REQUIRE(b && (nwords == 0 || h_words), "tgp_batch_results_sync: null argument");
  TGP_REQUIRE(nwords <= b->result_words, "asked for %llu result words, arena holds %llu",
              (unsigned long long)nwords, (unsigned long long)b->result_words);
  hipStream_t s = pick_stream(b->ctx, stream);
  if (nwords) TGP_HIP(hipMemcpyAsync(h_words, b->d_results, nwords * 8, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_batch_autotune(tgp_batch* b, void* stream, tgp_batch_tuning* best_out,
                           float* best_ms) {
  TGP_REQUIRE(b, "tgp_batch_autotune: null batch");
  hipStream_t s = pick_stream(b->ctx, stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TGP_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("tgp_batch_autotune: hipEventCreate failed");
    return TGP_ERR_HIP;
  }
  int32_t rc = TGP_OK;
  // best of 3 timed launches (after one warm-up) of the current settings
  auto time_it = [&](float* out) -> int32_t {
    TGP_TRY(batch_launch(b, s));
    float ms_min = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      float ms = 0;
      TGP_HIP(hipEventRecord(e0, s));
      TGP_TRY(batch_launch(b, s));
      TGP_HIP(hipEventRecord(e1, s));
      TGP_HIP(hipEventSynchronize(e1));
      TGP_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms_min = ms < ms_min ? ms : ms_min;
    }
    *out = ms_min;
    return TGP_OK;
  };
  tgp_batch_tuning best_t{};
  float best = 1e30f;
  auto trial = [&](uint32_t wg, uint32_t chunk, int32_t order, bool cached, uint32_t grab) {
    if (rc != TGP_OK) return;
    b->wg_per_cu = wg;
    b->chunk = chunk;
    b->order = order;
    b->cached = cached;
    b->per_grab = grab;
    rc = build_items(b, s);
    float ms = 0;
    if (rc == TGP_OK) rc = time_it(&ms);
    if (rc == TGP_OK && ms < best) {
      best = ms;
      best_t = tgp_batch_tuning{wg, chunk, order, cached ? 1 : 0, b->use_graph ? 1 : 0, grab};
    }
  };
  // coordinate search: the schedules' chunk sizes at full occupancy, items per
  // grab for the best table-ordered setting, then workgroups per CU
  for (uint32_t chunk : {16u, 32u, 64u, 128u}) trial(0, chunk, 0, false, 1);
  for (int cached = 0; cached < 2; ++cached)
    for (uint32_t chunk : {8u, 16u, 32u, 64u}) trial(0, chunk, 1, cached != 0, 1);
  for (uint32_t chunk : {8u, 16u, 32u, 64u}) trial(0, chunk, 2, true, 1);
  if (!b->fam.empty())
    for (int cached = 0; cached < 2; ++cached)
      for (uint32_t chunk : {16u, 32u, 64u, 128u, 256u}) trial(0, chunk, 3, cached != 0, 1);
  if (best_t.order == 1 || best_t.order == 2) {
    const tgp_batch_tuning t1 = best_t;
    for (uint32_t chunk : {4u, 8u, 16u, 32u})
      for (uint32_t grab : {2u, 4u, 8u}) trial(0, chunk, t1.order, t1.cached_loads != 0, grab);
  }
  const tgp_batch_tuning at_occ = best_t;
  for (uint32_t wg = 2; wg <= 4; ++wg)
    trial(wg, at_occ.chunk_slices, at_occ.order, at_occ.cached_loads != 0,
          at_occ.items_per_grab);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->wg_per_cu = best_t.wg_per_cu;
  b->chunk = best_t.chunk_slices;
  b->order = best_t.order;
  b->cached = best_t.cached_loads != 0;
  b->per_grab = best_t.items_per_grab == 0 ? 1 : best_t.items_per_grab;
  drop_graph(b);
  if (rc == TGP_OK) rc = build_items(b, s);
  if (rc != TGP_OK) return rc;
  if (best_out) *best_out = best_t;
  if (best_ms) *best_ms = best;
  return TGP_OK;
}

}  // extern "C"
