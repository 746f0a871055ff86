// GROUP BY arbitrary integer key tuples through a device hash table.
//
// Reference path replaced:
//   query_interpreter/accum_grouped.rs:198-335  per-chunk FxHashMap keyed by the
//       combined key hash (:213-218), merged per key (:291-335)
//   query_interpreter/aggregate.rs:582-627      HashMap grouping of row indices
//   engine/dispatch.rs:1125-1140, :1377-1385     dispatch GROUP BY, one or several keys
// The dense-id kernels (groupagg.hip, spec.hip) need the key space up front and
// small; here the groups are whatever distinct key tuples the selected rows hold.
//
// Table.  Open addressing over `cap` = 2^k slots of one 64-bit word, linear
// probing from a SplitMix64 hash, claimed with a 64-bit CAS.  The word holds:
//   raw form (one key column)      the key cell itself; empty = INT64_MIN bits.
//                                  The key INT64_MIN has slot cap, NULL slot cap + 1.
//   packed form (several columns whose value ranges multiply to < 2^63)
//                                  sum_k (v_k - min_k) * mul_k, NULL as one more
//                                  value per nullable column; empty = ~0.
//   verified form (wider tuples)   (32-bit hash tag << 32) | (representative row + 1);
//                                  empty = 0.  A tag match is confirmed by comparing
//                                  the whole tuple with the representative row's
//                                  cells, as the joins confirm a packed key.
// So no two distinct tuples ever share a group (the reference's combined hash can
// merge them).  Slot s owns the accumulators at acc[s * W]: COUNT(*), then per
// aggregate an int128 (lo, hi) or a MIN/MAX order code, all updated with 64-bit
// atomics that keep the int128 carries exact (atomic_add128*).
//
// Hot keys.  Rows of one key would all hit one slot's atomics; a per-workgroup LDS
// cache of kHgCache entries, keyed by slot, takes the rows of the first keys the
// workgroup meets (int128 partials with returning LDS atomics, exact for any
// count) and is added to the table once at the end.  Keys that find no cache entry
// within kHgCacheProbe steps go straight to the table.
//
// Filter and aggregate values come from the LDS-staged tile engine
// (tile_engine.h), the same per-slot program the dense generic kernels run.
#include "plan.h"

namespace tgp {

namespace {

constexpr int kHgRaw = 0, kHgPacked = 1, kHgVerified = 2;
constexpr uint64_t kEmptyRaw = 0x8000000000000000ull;
constexpr uint64_t kEmptyPacked = ~0ull;
constexpr int kHgCache = 256;  // LDS cache entries per workgroup (power of two)
constexpr int kHgCacheProbe = 4;

__device__ __forceinline__ uint64_t hg_mix(uint64_t z) {
  // SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct HgDev {
  DevCol c[TGP_MAX_GROUP_KEYS];               // staged key columns (tile engine slots)
  const uint8_t* data[TGP_MAX_GROUP_KEYS];    // the same columns in HBM (verified form)
  const uint64_t* nulls[TGP_MAX_GROUP_KEYS];  // NULL bitmaps, bit i <-> row row_begin + i
  int64_t kmin[TGP_MAX_GROUP_KEYS];           // packed form: smallest cell
  uint64_t kcard[TGP_MAX_GROUP_KEYS];         // packed form: values (+1 when nullable)
  uint64_t kmul[TGP_MAX_GROUP_KEYS];          // packed form: place value
  int32_t nkeys, mode;
  int32_t W;                                  // accumulator words per slot
  uint32_t tag_mask;                          // verified form: bits of the tag kept
  uint64_t* slot;                             // [cap]
  unsigned long long* acc;                    // [(cap + 2) * W]
  uint64_t mask;                              // cap - 1
  unsigned int* overflow;                     // set when a row found no free slot
};

// A cell of `data` (native width) widened as the tile engine widens it.
__device__ __forceinline__ int64_t hg_cell(const uint8_t* data, const DevCol& c, uint64_t row) {
  switch (c.width) {
    case 1: return c.is_signed ? (int64_t)(int8_t)data[row] : (int64_t)data[row];
    case 2: {
      const uint16_t v = reinterpret_cast<const uint16_t*>(data)[row];
      return c.is_signed ? (int64_t)(int16_t)v : (int64_t)v;
    }
    case 4: {
      const uint32_t v = reinterpret_cast<const uint32_t*>(data)[row];
      return c.is_signed ? (int64_t)(int32_t)v : (int64_t)v;
    }
    default: return reinterpret_cast<const int64_t*>(data)[row];
  }
}

__device__ __forceinline__ bool hg_null(const uint64_t* nb, uint64_t i) {
  return nb && ((nb[i >> 6] >> (i & 63)) & 1);
}

// Slot of key word K (raw / packed forms): the slot holding K, or an empty slot
// claimed for it; -1 when the table is full.
__device__ __forceinline__ int64_t hg_probe_word(const HgDev& H, uint64_t K, uint64_t empty) {
  uint64_t pos = hg_mix(K) & H.mask;
  for (uint64_t step = 0; step <= H.mask; ++step) {
    // a stale read can only show `empty` (a claimed word never changes): the CAS
    // then returns the word that is there
    uint64_t cur = H.slot[pos];
    if (cur == K) return (int64_t)pos;
    if (cur == empty) {
      cur = atomicCAS(reinterpret_cast<unsigned long long*>(&H.slot[pos]),
                      (unsigned long long)empty, (unsigned long long)K);
      if (cur == empty || cur == K) return (int64_t)pos;
    }
    pos = (pos + 1) & H.mask;
  }
  return -1;
}

// Does the representative row `rep` hold the tuple (x, nf)?
__device__ __forceinline__ bool hg_same_tuple(const HgDev& H, const int64_t (&x)[TGP_MAX_GROUP_KEYS],
                                              uint32_t nf, uint64_t rep, uint64_t row_begin) {
#pragma unroll
  for (int k = 0; k < TGP_MAX_GROUP_KEYS; ++k) {
    if (k >= H.nkeys) break;
    const bool rn = hg_null(H.nulls[k], rep - row_begin);
    if (rn != (((nf >> k) & 1) != 0)) return false;
    if (!rn && hg_cell(H.data[k], H.c[k], rep) != x[k]) return false;
  }
  return true;
}

// Slot of the tuple (x, nf) of row `row`, inserting it when new; -1 when full.
__device__ __forceinline__ int64_t hg_find(const HgDev& H, const int64_t (&x)[TGP_MAX_GROUP_KEYS],
                                           uint32_t nf, uint64_t row, uint64_t row_begin) {
  const uint64_t cap = H.mask + 1;
  if (H.mode == kHgRaw) {
    if (nf & 1) return (int64_t)(cap + 1);
    if ((uint64_t)x[0] == kEmptyRaw) return (int64_t)cap;
    return hg_probe_word(H, (uint64_t)x[0], kEmptyRaw);
  }
  if (H.mode == kHgPacked) {
    uint64_t K = 0;
#pragma unroll
    for (int k = 0; k < TGP_MAX_GROUP_KEYS; ++k) {
      if (k >= H.nkeys) break;
      const uint64_t part =
          ((nf >> k) & 1) ? H.kcard[k] - 1 : (uint64_t)x[k] - (uint64_t)H.kmin[k];
      K += part * H.kmul[k];
    }
    return hg_probe_word(H, K, kEmptyPacked);
  }
  uint64_t h = 0x9E3779B97F4A7C15ull ^ nf;
#pragma unroll
  for (int k = 0; k < TGP_MAX_GROUP_KEYS; ++k) {
    if (k >= H.nkeys) break;
    h = hg_mix(h ^ (((nf >> k) & 1) ? 0 : (uint64_t)x[k]));
  }
  const uint64_t tag = (h >> 32) & H.tag_mask;
  const uint64_t mine = (tag << 32) | (row + 1);
  uint64_t pos = h & H.mask;
  for (uint64_t step = 0; step <= H.mask; ++step) {
    uint64_t cur = H.slot[pos];
    if (cur == 0) {
      cur = atomicCAS(reinterpret_cast<unsigned long long*>(&H.slot[pos]), 0ull,
                      (unsigned long long)mine);
      if (cur == 0) return (int64_t)pos;
    }
    if ((cur >> 32) == tag && hg_same_tuple(H, x, nf, (cur & 0xffffffffull) - 1, row_begin))
      return (int64_t)pos;
    pos = (pos + 1) & H.mask;
  }
  return -1;
}

// LDS cache entry of table slot s (claimed when free), or -1.
__device__ __forceinline__ int32_t hg_cache(uint32_t* ctag, uint64_t s) {
  const uint32_t tag = (uint32_t)s + 1;
  uint32_t pos = (uint32_t)(s * 0x9E3779B1u) >> (32 - 8);  // kHgCache = 2^8
  for (int p = 0; p < kHgCacheProbe; ++p) {
    uint32_t cur = __atomic_load_n(&ctag[pos], __ATOMIC_RELAXED);
    if (cur == tag) return (int32_t)pos;
    if (cur == 0) {
      cur = atomicCAS(&ctag[pos], 0u, tag);
      if (cur == 0 || cur == tag) return (int32_t)pos;
    }
    pos = (pos + 1) & (kHgCache - 1);
  }
  return -1;
}

static_assert(kHgCache == 256, "hg_cache hashes to 8 bits");

__global__ void __launch_bounds__(kBlock) k_hash_group(DevPlan P, HgDev H) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t* ctag = reinterpret_cast<uint32_t*>(lds + 2 * P.stage_bytes);
  unsigned long long* cacc = reinterpret_cast<unsigned long long*>(ctag + kHgCache);
  const int W = H.W;
  for (uint32_t i = threadIdx.x; i < (uint32_t)kHgCache; i += kBlock) ctag[i] = 0;
  for (uint32_t i = threadIdx.x; i < (uint32_t)(kHgCache * W); i += kBlock) cacc[i] = 0;
  __syncthreads();
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    uint32_t sel = eval_select(P, buf, row0);
    if (__ballot(sel != 0) == 0) return;
    int64_t kx[TGP_MAX_GROUP_KEYS][kRows];
#pragma unroll
    for (int k = 0; k < TGP_MAX_GROUP_KEYS; ++k)
      if (k < H.nkeys) load_col(H.c[k], buf, kx[k]);
    int64_t slot[kRows];
    int32_t ce[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      slot[r] = -1;
      ce[r] = -1;
      if (!((sel >> r) & 1)) continue;
      const uint64_t row = row0 + threadIdx.x + r * kBlock;
      uint32_t nf = 0;
      int64_t x[TGP_MAX_GROUP_KEYS];
#pragma unroll
      for (int k = 0; k < TGP_MAX_GROUP_KEYS; ++k) {
        x[k] = k < H.nkeys ? kx[k][r] : 0;
        if (k < H.nkeys && hg_null(H.nulls[k], row - P.row_begin)) nf |= 1u << k;
      }
      slot[r] = hg_find(H, x, nf, row, P.row_begin);
      if (slot[r] < 0) {
        atomicOr(H.overflow, 1u);
        sel &= ~(1u << r);
        continue;
      }
      ce[r] = hg_cache(ctag, (uint64_t)slot[r]);
    }
    auto word = [&](int r, int w) -> unsigned long long* {
      return ce[r] >= 0 ? &cacc[(uint32_t)ce[r] * W + w]
                        : &H.acc[(uint64_t)slot[r] * W + w];
    };
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      if ((sel >> r) & 1) atomicAdd(word(r, 0), 1ull);
    for (int a = 0; a < P.naggs; ++a) {
      uint64_t v[kRows];
      eval_agg(P.aggs[a], buf, v);
      const int32_t mk = (P.minmax_kinds >> (3 * a)) & 7;
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        if (!((sel >> r) & 1)) continue;
        if (mk) atomicMax(word(r, 1 + 2 * a), (unsigned long long)mm_encode(mk, v[r]));
        else atomic_add128_i64(word(r, 1 + 2 * a), (int64_t)v[r]);
      }
    }
  });
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < (uint32_t)kHgCache; e += kBlock) {
    const uint32_t t = ctag[e];
    if (!t) continue;
    const uint64_t s = t - 1;
    const unsigned long long* c = &cacc[e * W];
    atomicAdd(&H.acc[s * W], c[0]);
    for (int a = 0; a < P.naggs; ++a) {
      unsigned long long* g = &H.acc[s * W + 1 + 2 * a];
      if ((P.minmax_kinds >> (3 * a)) & 7) atomicMax(g, c[1 + 2 * a]);
      else atomic_add128(g, c[1 + 2 * a], (int64_t)c[2 + 2 * a]);
    }
  }
}

// Occupied slots (COUNT > 0: a slot is claimed only by a row that then counts
// itself) as a bitmap of nslots bits.
__global__ void __launch_bounds__(256) k_hg_occupied(const unsigned long long* __restrict__ acc,
                                                     int32_t W, uint64_t nslots,
                                                     uint64_t* __restrict__ bitmap) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const bool occ = i < nslots && acc[i * W] != 0;
  const uint64_t b = __ballot(occ);
  if ((threadIdx.x & 63) == 0 && i < nslots) bitmap[i >> 6] = b;
}

// Group i of the output from slot idx[i]: key cells, NULL flags, count, sums.
__global__ void __launch_bounds__(256) k_hg_emit(HgDev H, const uint32_t* __restrict__ idx,
                                                 uint64_t ngroups, int32_t naggs, uint32_t kinds,
                                                 uint64_t row_begin, uint64_t* __restrict__ okeys,
                                                 uint32_t* __restrict__ onulls,
                                                 uint64_t* __restrict__ ocount,
                                                 uint64_t* __restrict__ osums) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= ngroups) return;
  const uint64_t s = idx[i], cap = H.mask + 1;
  const unsigned long long* a = &H.acc[s * H.W];
  ocount[i] = a[0];
  for (int j = 0; j < naggs; ++j) {
    const int32_t mk = (kinds >> (3 * j)) & 7;
    uint64_t lo = a[1 + 2 * j], hi = a[2 + 2 * j];
    if (mk) {
      lo = mm_decode(mk, lo);
      hi = (mk == TGP_AGG_MIN || mk == TGP_AGG_MAX) && (int64_t)lo < 0 ? ~0ull : 0ull;
    }
    osums[(i * naggs + j) * 2] = lo;
    osums[(i * naggs + j) * 2 + 1] = hi;
  }
  uint32_t nf = 0;
  for (int k = 0; k < H.nkeys; ++k) {
    uint64_t v = 0;
    if (H.mode == kHgRaw) {
      if (s == cap + 1) nf = 1;
      else v = s == cap ? kEmptyRaw : H.slot[s];
    } else if (H.mode == kHgPacked) {
      const uint64_t part = (H.slot[s] / H.kmul[k]) % H.kcard[k];
      if (H.nulls[k] && part == H.kcard[k] - 1) nf |= 1u << k;
      else v = (uint64_t)H.kmin[k] + part;
    } else {
      const uint64_t rep = (H.slot[s] & 0xffffffffull) - 1;
      if (hg_null(H.nulls[k], rep - row_begin)) nf |= 1u << k;
      else v = (uint64_t)hg_cell(H.data[k], H.c[k], rep);
    }
    okeys[i * H.nkeys + k] = v;
  }
  if (onulls) onulls[i] = nf;
}

// Smallest and largest cell (in the column's order) of rows [b, e).
__global__ void __launch_bounds__(256) k_hg_minmax(const uint8_t* __restrict__ data, DevCol c,
                                                   uint64_t b, uint64_t e, long long* out) {
  long long mn = LLONG_MAX, mx = LLONG_MIN;
  unsigned long long umn = ~0ull, umx = 0;
  for (uint64_t i = b + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < e;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t v = hg_cell(data, c, i);
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
    umn = (uint64_t)v < umn ? (uint64_t)v : umn;
    umx = (uint64_t)v > umx ? (uint64_t)v : umx;
  }
  if (c.is_signed) {
    atomicMin(&out[0], mn);
    atomicMax(&out[1], mx);
  } else {
    atomicMin(reinterpret_cast<unsigned long long*>(&out[0]), umn);
    atomicMax(reinterpret_cast<unsigned long long*>(&out[1]), umx);
  }
}

// ---- partitioned form (raw and packed key words, many groups) ------------------------------
// Each row of the global table's form costs a random probe and random atomics in a
// table of millions of slots.  Instead the rows are partitioned by the top bits of
// their key word's hash (the raw form's NULL and INT64_MIN keys into one more
// partition), as kAggPart partitions dense ids: a histogram through a plan of the
// predicate and key columns, a scan, a scatter of records (key word, one int64 per
// aggregate) through the full plan, then workgroups of kPartChunk records that
// aggregate each partition's run in an LDS hash table and merge its groups into the
// global table (one probe and the atomics per group, not per row).  A key that
// finds no LDS slot goes to the global table directly, so the result never
// depends on the LDS table fitting; the global table's slots, accumulators and
// capacity rules are those of the one-pass form, and so is everything after it.
constexpr int kHgpProbe = 32;  // LDS probe steps before a key goes to the global table
constexpr uint64_t kHgPartMinGroups = 1ull << 15;  // capacities below: the one-pass form

// Key words and partitions of the thread's rows of a staged tile (part = nparts for
// the raw form's special keys, ~0u for rows not selected); each key column is folded
// into the word as it is loaded.
__device__ __forceinline__ void hg_tile_words(const DevPlan& P, const HgDev& H, const uint8_t* buf,
                                              uint64_t row0, uint32_t sel, uint32_t pbits,
                                              uint32_t nparts, uint64_t (&K)[kRows],
                                              uint32_t (&part)[kRows]) {
  uint32_t special = 0;  // bit r: row r has a raw special key (K[r] 0: NULL, 1: INT64_MIN)
#pragma unroll
  for (int r = 0; r < kRows; ++r) K[r] = 0;
  for (int k = 0; k < H.nkeys; ++k) {
    int64_t x[kRows];
    load_col(H.c[k], buf, x);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint64_t row = row0 + threadIdx.x + r * kBlock;
      const bool isnull = ((sel >> r) & 1) && hg_null(H.nulls[k], row - P.row_begin);
      if (H.mode == kHgRaw) {
        if (isnull || (uint64_t)x[r] == kEmptyRaw) {
          special |= 1u << r;
          K[r] = isnull ? 0 : 1;
        } else {
          K[r] = (uint64_t)x[r];
        }
      } else {
        K[r] += (isnull ? H.kcard[k] - 1 : (uint64_t)x[r] - (uint64_t)H.kmin[k]) * H.kmul[k];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r)
    part[r] = !((sel >> r) & 1) ? ~0u
              : ((special >> r) & 1) ? nparts : (uint32_t)(hg_mix(K[r]) >> (64 - pbits));
}

// mat[p * gridDim.x + (h mod G) k + h / G] = workgroup h's rows of partition p
// (p <= nparts), as k_part_hist_tiles lays them out.
__global__ void __launch_bounds__(kBlock) k_hgp_hist(DevPlan P, HgDev H, uint32_t pbits,
                                                     uint32_t nparts, uint32_t k,
                                                     unsigned long long* __restrict__ mat) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t* h = reinterpret_cast<uint32_t*>(lds + 2 * P.stage_bytes);
  for (uint32_t p = threadIdx.x; p <= nparts; p += kBlock) h[p] = 0;
  // made visible by the barrier at the top of the first tile
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    const uint32_t sel = eval_select(P, buf, row0);
    if (__ballot(sel != 0) == 0) return;
    uint64_t K[kRows];
    uint32_t part[kRows];
    hg_tile_words(P, H, buf, row0, sel, pbits, nparts, K, part);
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      if (part[r] != ~0u) atomicAdd(&h[part[r]], 1u);
  });
  __syncthreads();
  const uint32_t G = gridDim.x / k;
  const uint64_t col = (uint64_t)(blockIdx.x % G) * k + blockIdx.x / G;
  for (uint32_t p = threadIdx.x; p <= nparts; p += kBlock)
    mat[(uint64_t)p * gridDim.x + col] = h[p];
}

// Records: rkey[r] (the key word, or 0 / 1 for the special keys) and
// rval[a * cap + r]; mat holds the scan's exclusive offsets.
__global__ void __launch_bounds__(kBlock) k_hgp_scatter(DevPlan P, HgDev H, uint32_t pbits,
                                                        uint32_t nparts, uint32_t k,
                                                        const unsigned long long* __restrict__ mat,
                                                        uint64_t cap, uint64_t* __restrict__ rkey,
                                                        uint64_t* __restrict__ rval) {
  extern __shared__ __align__(16) uint8_t lds[];
  unsigned long long* cursor = reinterpret_cast<unsigned long long*>(lds + 2 * P.stage_bytes);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cursor + nparts + 1);
  for (uint32_t p = threadIdx.x; p <= nparts; p += kBlock) {
    cursor[p] = mat[((uint64_t)p * gridDim.x + blockIdx.x) * k];
    cnt[p] = 0;
  }
  // made visible by the barrier at the top of the first tile
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    const uint32_t sel = eval_select(P, buf, row0);
    if (!__syncthreads_or(sel != 0)) return;  // the same for the whole workgroup
    uint64_t K[kRows];
    uint32_t part[kRows], rk[kRows];
    hg_tile_words(P, H, buf, row0, sel, pbits, nparts, K, part);
#pragma unroll
    for (int r = 0; r < kRows; ++r) rk[r] = part[r] != ~0u ? atomicAdd(&cnt[part[r]], 1u) : 0u;
    __syncthreads();  // every rank taken
    uint64_t dst[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      dst[r] = 0;
      if (part[r] == ~0u) continue;
      dst[r] = cursor[part[r]] + rk[r];
      rkey[dst[r]] = K[r];
    }
    for (int a = 0; a < P.naggs; ++a) {
      uint64_t v[kRows];
      eval_agg(P.aggs[a], buf, v);
      uint64_t* out = rval + (uint64_t)a * cap;
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if (part[r] != ~0u) out[dst[r]] = v[r];
    }
    __syncthreads();  // every cursor read
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      if (part[r] != ~0u && rk[r] == 0) {
        cursor[part[r]] += cnt[part[r]];
        cnt[part[r]] = 0;
      }
    // the next tile's first barrier orders these before its ranks
  });
}

// Adds the accumulators `a` (count, then per aggregate split sums or a MIN/MAX code)
// into global slot s (int128 sums, codes).
__device__ __forceinline__ void hg_merge(const HgDev& H, int32_t naggs, uint32_t kinds,
                                         uint64_t s, const unsigned long long* a) {
  unsigned long long* g = &H.acc[s * H.W];
  atomicAdd(g, a[0]);
  for (int j = 0; j < naggs; ++j) {
    if ((kinds >> (3 * j)) & 7) {
      atomicMax(&g[1 + 2 * j], a[1 + 2 * j]);
    } else {
      uint64_t lo;
      int64_t hi;
      split_to_i128(a[1 + 2 * j], (int64_t)a[2 + 2 * j], &lo, &hi);
      atomic_add128(&g[1 + 2 * j], lo, hi);
    }
  }
}

// Global slot of key word K (special keys: their own slots), or -1 (table full).
__device__ __forceinline__ int64_t hg_global_slot(const HgDev& H, uint64_t K, bool special,
                                                  uint64_t empty) {
  const uint64_t cap = H.mask + 1;
  if (special) return (int64_t)(K == 0 ? cap + 1 : cap);
  return hg_probe_word(H, K, empty);
}

// kPartAggBlock threads; LDS: S key words, then S x W accumulators.
constexpr int kHgpBlock = 1024;
__global__ void __launch_bounds__(kHgpBlock) k_hgp_agg(HgDev H, int32_t naggs, uint32_t kinds,
                                                       uint32_t nparts, uint64_t G,
                                                       const unsigned long long* __restrict__ mat,
                                                       uint64_t cap, const uint64_t* __restrict__ rkey,
                                                       const uint64_t* __restrict__ rval,
                                                       uint32_t S, uint64_t empty) {
  extern __shared__ __align__(16) unsigned long long lmem[];
  unsigned long long* key = lmem;
  unsigned long long* acc = lmem + S;
  const uint32_t W = (uint32_t)H.W;
  const uint64_t total = mat[(uint64_t)(nparts + 1) * G];
  uint64_t lo = (uint64_t)blockIdx.x * kPartChunk;
  if (lo >= total) return;
  const uint64_t hi = lo + kPartChunk < total ? lo + kPartChunk : total;
  uint32_t pa = 0, pb = nparts;  // partitions 0..nparts (nparts: the special keys)
  while (pa < pb) {
    const uint32_t m = (pa + pb + 1) >> 1;
    if (mat[(uint64_t)m * G] <= lo) pa = m;
    else pb = m - 1;
  }
  for (uint32_t p = pa; lo < hi && p <= nparts; ++p) {
    const uint64_t pe = mat[(uint64_t)(p + 1) * G];
    const uint64_t end = pe < hi ? pe : hi;
    if (end <= lo) continue;
    const bool special = p == nparts;
    for (uint32_t i = threadIdx.x; i < S; i += kHgpBlock) key[i] = empty;
    for (uint32_t i = threadIdx.x; i < S * W; i += kHgpBlock) acc[i] = 0;
    __syncthreads();
    for (uint64_t i = lo + threadIdx.x; i < end; i += kHgpBlock) {
      const uint64_t K = rkey[i];
      // the LDS slot of K (claimed when free), or none after kHgpProbe steps
      uint32_t pos = (uint32_t)hg_mix(K) & (S - 1);
      int32_t e = -1;
      for (int step = 0; step < kHgpProbe; ++step) {
        unsigned long long cur = key[pos];
        if (cur == empty)
          cur = atomicCAS(&key[pos], (unsigned long long)empty, (unsigned long long)K);
        if (cur == empty || cur == K) {
          e = (int32_t)pos;
          break;
        }
        pos = (pos + 1) & (S - 1);
      }
      if (e >= 0) {
        unsigned long long* a = &acc[(uint32_t)e * W];
        atomicAdd(a, 1ull);
        for (int j = 0; j < naggs; ++j) {
          const uint64_t v = rval[(uint64_t)j * cap + i];
          const int32_t mk = (kinds >> (3 * j)) & 7;
          if (mk) {
            atomicMax(&a[1 + 2 * j], (unsigned long long)mm_encode(mk, v));
          } else {
            atomicAdd(&a[1 + 2 * j], (unsigned long long)(uint32_t)v);
            const int64_t h = (int64_t)v >> 32;
            if (h) atomicAdd(&a[2 + 2 * j], (unsigned long long)h);
          }
        }
      } else {  // straight into the global table
        const int64_t s = hg_global_slot(H, K, special, empty);
        if (s < 0) {
          atomicOr(H.overflow, 1u);
          continue;
        }
        unsigned long long* g = &H.acc[(uint64_t)s * W];
        atomicAdd(g, 1ull);
        for (int j = 0; j < naggs; ++j) {
          const uint64_t v = rval[(uint64_t)j * cap + i];
          const int32_t mk = (kinds >> (3 * j)) & 7;
          if (mk) atomicMax(&g[1 + 2 * j], (unsigned long long)mm_encode(mk, v));
          else atomic_add128_i64(&g[1 + 2 * j], (int64_t)v);
        }
      }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < S; e += kHgpBlock) {
      const uint64_t K = key[e];
      if (K == empty) continue;
      const int64_t s = hg_global_slot(H, K, special, empty);
      if (s < 0) {
        atomicOr(H.overflow, 1u);
        continue;
      }
      hg_merge(H, naggs, kinds, (uint64_t)s, &acc[e * W]);
    }
    __syncthreads();  // every merge read before the next partition clears the table
    lo = end;
  }
}

// The histogram's plan of the partitioned form: q's predicate and key columns only,
// and H's key columns placed in that plan's stage.
int32_t plan_hg_hist(const tgp_agg_query* q, const tgp_hash_group_query* hq, const HgDev& H,
                     DevPlan* Ph, HgDev* Hh) {
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
  int kc[TGP_MAX_GROUP_KEYS] = {};
  for (int k = 0; k < hq->nkeys; ++k) kc[k] = use(hq->key_col[k]);
  rq.nparts = 0;
  rq.naggs = 0;
  PlanInfo info;
  TGP_TRY(plan_agg(&rq, Ph, &info));
  *Hh = H;
  for (int k = 0; k < hq->nkeys; ++k) Hh->c[k].lds_off = Ph->slots[kc[k]].lds_off;
  return TGP_OK;
}

// Runs the partitioned form's passes into the (initialised) global table H.
// *done false: the query's stage and partition counters do not fit LDS; nothing ran.
int32_t run_hg_part(tgp_ctx* ctx, const tgp_agg_query* q, const tgp_hash_group_query* hq,
                    const DevPlan& P, const HgDev& H, uint64_t capacity, uint64_t empty,
                    hipStream_t s, bool* done) {
  *done = false;
  if (P.ntiles == 0) {
    *done = true;
    return TGP_OK;
  }
  DevPlan Ph;
  HgDev Hh;
  TGP_TRY(plan_hg_hist(q, hq, H, &Ph, &Hh));
  // an LDS hash table of S slots (key word + W accumulator words) in
  // kPartTableBytes; partitions of about S / 2 groups each, at most 4096
  const uint32_t W = (uint32_t)H.W;
  uint32_t S = 64;
  while ((uint64_t)S * 2 * (8 + 8 * W) <= kPartTableBytes) S *= 2;
  const uint64_t want = (capacity + S / 2 - 1) / (S / 2);
  uint32_t pbits = 1;
  while ((1ull << pbits) < want && pbits < 12) ++pbits;
  const uint32_t np = 1u << pbits;
  const uint64_t rows = P.row_end - P.row_begin;
  const uint32_t lh = 2 * Ph.stage_bytes + 4 * (np + 1);
  const uint32_t ls = 2 * P.stage_bytes + 12 * (np + 1);
  const uint32_t la = S * 8 * (1 + W);
  if (lh > 160 * 1024 || ls > 160 * 1024) return TGP_OK;
  uint64_t G = std::min<uint64_t>(P.ntiles, (uint64_t)persistent_grid(ctx, 2));
  const int64_t gmax = ctx->tune_grid_max.load();
  if (gmax > 0 && G > (uint64_t)gmax) G = (uint64_t)gmax;
  uint32_t k = 1;  // histogram workgroups per scatter workgroup (its LDS is small)
  while (k < 4 && G * 2 * k <= P.ntiles && (uint64_t)(np + 1) * G * 2 * k <= (1u << 22)) k *= 2;
  const uint64_t mat_n = (uint64_t)(np + 1) * G * k;
  const uint64_t nb = (mat_n + 4095) / 4096;
  // scratch: mat[mat_n + 1] | scan block sums[nb + 1] | keys[rows] | values[naggs][rows]
  const uint64_t words = mat_n + 1 + nb + 1 + rows * (1 + (uint64_t)q->naggs);
  void* scr = nullptr;
  TGP_TRY(get_scratch(ctx, s, words * 8, &scr));
  auto* mat = static_cast<unsigned long long*>(scr);
  auto* bsum = reinterpret_cast<uint64_t*>(mat + mat_n + 1);
  uint64_t* rkey = bsum + nb + 1;
  uint64_t* rval = rkey + rows;
  for (auto kern : {(const void*)k_hgp_hist, (const void*)k_hgp_scatter, (const void*)k_hgp_agg}) {
    const uint32_t b = kern == (const void*)k_hgp_hist ? lh
                     : kern == (const void*)k_hgp_scatter ? ls : la;
    if (b > 65536)
      TGP_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b));
  }
  k_hgp_hist<<<dim3((unsigned)(G * k)), dim3(kBlock), lh, s>>>(Ph, Hh, pbits, np, k, mat);
  TGP_HIP(hipGetLastError());
  TGP_TRY(scan_u64(s, reinterpret_cast<uint64_t*>(mat), mat_n, bsum));
  k_hgp_scatter<<<dim3((unsigned)G), dim3(kBlock), ls, s>>>(P, H, pbits, np, k, mat, rows, rkey,
                                                            rval);
  TGP_HIP(hipGetLastError());
  const uint64_t ga = (rows + kPartChunk - 1) / kPartChunk;
  k_hgp_agg<<<dim3((unsigned)ga), dim3(kHgpBlock), la, s>>>(H, q->naggs, P.minmax_kinds, np,
                                                            G * k, mat, rows, rkey, rval, S,
                                                            empty);
  TGP_HIP(hipGetLastError());
  *done = true;
  return TGP_OK;
}

// The one-pass form: every row into the global table through the LDS cache of hot keys.
int32_t run_hg_one_pass(tgp_ctx* ctx, const DevPlan& P, const HgDev& H, hipStream_t s) {
  // one workgroup may meet any number of rows: the LDS partials are int128
  const uint32_t lds = 2 * P.stage_bytes + kHgCache * 4 + kHgCache * H.W * 8;
  if (lds > 65536)
    TGP_HIP(hipFuncSetAttribute((const void*)k_hash_group,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hash_group, kBlock, lds));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 4) per_cu = 4;
  uint64_t grid = (uint64_t)persistent_grid(ctx, per_cu);
  if (grid > P.ntiles) grid = P.ntiles;
  if (grid == 0) return TGP_OK;
  k_hash_group<<<(unsigned)grid, kBlock, lds, s>>>(P, H);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

unsigned grid_for(uint64_t items) {
  uint64_t g = (items + 255) / 256;
  if (g > 8192) g = 8192;
  return (unsigned)(g ? g : 1);
}

uint64_t next_pow2(uint64_t x) {
  uint64_t c = 1;
  while (c < x) c <<= 1;
  return c;
}

}  // namespace

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_hash_group_agg_sync(tgp_ctx* ctx, const tgp_hash_group_query* hq,
                                const tgp_hash_group_result* out, uint64_t* h_ngroups,
                                void* stream) {
  TGP_REQUIRE(ctx && hq && out && h_ngroups, "tgp_hash_group_agg_sync: null argument");
  TGP_TRY(check_injected(ctx));
  *h_ngroups = 0;
  const tgp_agg_query* q = &hq->agg;
  TGP_REQUIRE(q->nparts == 0, "tgp_hash_group_agg_sync: the keys replace the dense parts "
                              "(nparts must be 0)");
  TGP_REQUIRE(hq->nkeys >= 1 && hq->nkeys <= TGP_MAX_GROUP_KEYS, "nkeys %d outside [1,%d]",
              hq->nkeys, TGP_MAX_GROUP_KEYS);
  TGP_REQUIRE(out->capacity == 0 || (out->d_keys && out->d_count &&
                                     (q->naggs == 0 || out->d_sums)),
              "tgp_hash_group_agg_sync: null result buffer");
  TGP_REQUIRE(out->capacity < (1ull << 30), "capacity %llu above 2^30 groups",
              (unsigned long long)out->capacity);
  DevPlan P;
  PlanInfo info;
  TGP_TRY(plan_agg(q, &P, &info));
  HgDev H;
  std::memset(&H, 0, sizeof(H));
  H.nkeys = hq->nkeys;
  H.W = 1 + 2 * q->naggs;
  for (int k = 0; k < hq->nkeys; ++k) {
    const int kc = hq->key_col[k];
    TGP_REQUIRE(kc >= 0 && kc < q->ncols, "key %d: column %d out of range", k, kc);
    H.c[k].lds_off = P.slots[kc].lds_off;
    H.c[k].width = P.slots[kc].width;
    H.c[k].is_signed = dtype_signed(q->cols[kc].dtype) ? 1 : 0;
    H.data[k] = static_cast<const uint8_t*>(q->cols[kc].data);
    H.nulls[k] = hq->d_key_nulls[k];
  }
  const uint64_t n = q->row_end - q->row_begin;
  if (n == 0) return TGP_OK;
  hipStream_t s = pick_stream(ctx, stream);
  // the form of the key word
  const int64_t tag_bits = ctx->tune_group_tag_bits.load();
  H.mode = hq->nkeys == 1 ? kHgRaw : kHgPacked;
  H.tag_mask = 0xffffffffu;
  void* mm = nullptr;
  if (H.mode == kHgPacked && tag_bits == 0) {
    TGP_TRY(get_scratch(ctx, s, 16 * TGP_MAX_GROUP_KEYS, &mm));
    long long init[2 * TGP_MAX_GROUP_KEYS];
    for (int k = 0; k < hq->nkeys; ++k) {
      init[2 * k] = H.c[k].is_signed ? LLONG_MAX : (long long)~0ull;
      init[2 * k + 1] = H.c[k].is_signed ? LLONG_MIN : 0;
    }
    TGP_HIP(hipMemcpyAsync(mm, init, 16 * hq->nkeys, hipMemcpyHostToDevice, s));
    for (int k = 0; k < hq->nkeys; ++k)
      k_hg_minmax<<<grid_for(n), 256, 0, s>>>(H.data[k], H.c[k], q->row_begin, q->row_end,
                                              static_cast<long long*>(mm) + 2 * k);
    TGP_HIP(hipGetLastError());
    long long r[2 * TGP_MAX_GROUP_KEYS];
    TGP_HIP(hipMemcpyAsync(r, mm, 16 * hq->nkeys, hipMemcpyDeviceToHost, s));
    TGP_HIP(hipStreamSynchronize(s));
    unsigned __int128 space = 1;
    for (int k = 0; k < hq->nkeys && H.mode == kHgPacked; ++k) {
      // values span max - min + 1 in the column's order; unsigned arithmetic
      // gives the span for either order
      const unsigned __int128 span =
          (unsigned __int128)((uint64_t)r[2 * k + 1] - (uint64_t)r[2 * k]) + 1 +
          (H.nulls[k] ? 1 : 0);
      H.kmin[k] = r[2 * k];
      H.kmul[k] = (uint64_t)space;
      space *= span;
      if (space >= ((unsigned __int128)1 << 63)) H.mode = kHgVerified;
      else H.kcard[k] = (uint64_t)span;
    }
  } else if (H.mode == kHgPacked) {
    H.mode = kHgVerified;
    H.tag_mask = tag_bits >= 32 ? 0xffffffffu : (uint32_t)((1ull << tag_bits) - 1);
  }
  if (H.mode == kHgVerified)
    TGP_REQUIRE(q->row_end < 0xffffffffull,
                "tgp_hash_group_agg_sync: a hashed multi-column key needs row_end < 2^32 - 1");
  // table: at most half full when the groups fit the capacity
  uint64_t want = out->capacity > 32 ? 2 * out->capacity : 64;
  if (want > 2 * n) want = 2 * n > 64 ? 2 * n : 64;  // never more groups than rows
  const uint64_t cap = next_pow2(want);
  const uint64_t nslots = cap + 2;
  const uint64_t words = (nslots + 63) / 64;
  const uint64_t off_acc = cap * 8;
  const uint64_t off_bits = off_acc + nslots * H.W * 8;
  const uint64_t off_flag = off_bits + words * 8;
  const uint64_t off_idx = (off_flag + 8 + 255) & ~255ull;
  const uint64_t bytes = off_idx + nslots * 4;
  void* scr = nullptr;
  TGP_TRY(get_scratch(ctx, s, bytes, &scr, 1));
  uint8_t* base = static_cast<uint8_t*>(scr);
  H.slot = reinterpret_cast<uint64_t*>(base);
  H.acc = reinterpret_cast<unsigned long long*>(base + off_acc);
  H.mask = cap - 1;
  H.overflow = reinterpret_cast<unsigned int*>(base + off_flag);
  uint64_t* bits = reinterpret_cast<uint64_t*>(base + off_bits);
  uint32_t* idx = reinterpret_cast<uint32_t*>(base + off_idx);
  const uint64_t empty = H.mode == kHgRaw ? kEmptyRaw : H.mode == kHgPacked ? kEmptyPacked : 0;
  TGP_TRY(fill_u64(s, H.slot, cap, empty));
  {
    uint64_t* zp[2] = {reinterpret_cast<uint64_t*>(H.acc), reinterpret_cast<uint64_t*>(H.overflow)};
    const uint64_t zn[2] = {nslots * H.W, 1};
    TGP_TRY(zero_ranges(s, 2, zp, zn));
  }
  // many groups (raw and packed forms): the partitioned form, unless its LDS does not fit
  const int64_t part_mode = ctx->tune_agg_part.load();
  bool parted = false;
  if ((H.mode == kHgRaw || H.mode == kHgPacked) &&
      (part_mode == 2 ||
       (part_mode == 0 && n >= kPartMinRows && out->capacity >= kHgPartMinGroups)))
    TGP_TRY(run_hg_part(ctx, q, hq, P, H, out->capacity ? out->capacity : 1, empty, s, &parted));
  if (!parted) TGP_TRY(run_hg_one_pass(ctx, P, H, s));
  k_hg_occupied<<<(unsigned)((words * 64 + 255) / 256), 256, 0, s>>>(H.acc, H.W, nslots, bits);

  TGP_HIP(hipGetLastError());
  uint64_t G = 0;
  TGP_TRY(compact_begin(ctx, s, bits, nslots, &G));  // synchronous
  unsigned int ovf = 0;
  TGP_HIP(hipMemcpyAsync(&ovf, H.overflow, 4, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  if (ovf) {
    *h_ngroups = (G > out->capacity ? G : out->capacity) + 1;
    set_error("tgp_hash_group_agg_sync: more groups than the table for capacity %llu holds",
              (unsigned long long)out->capacity);
    return TGP_ERR_CAPACITY;
  }
  *h_ngroups = G;
  if (G > out->capacity) {
    set_error("tgp_hash_group_agg_sync: %llu groups > capacity %llu", (unsigned long long)G,
              (unsigned long long)out->capacity);
    return TGP_ERR_CAPACITY;
  }
  if (G == 0) return TGP_OK;
  TGP_TRY(compact_finish(ctx, s, bits, nslots, 0, 4, idx));
  k_hg_emit<<<(unsigned)((G + 255) / 256), 256, 0, s>>>(
      H, idx, G, q->naggs, P.minmax_kinds, q->row_begin, out->d_keys, out->d_key_nulls,
      out->d_count, out->d_sums);
  TGP_HIP(hipGetLastError());
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

}  // extern "C"
