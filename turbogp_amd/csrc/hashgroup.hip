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
  k_hash_group<<<(unsigned)grid, kBlock, lds, s>>>(P, H);
  TGP_HIP(hipGetLastError());
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
