// Hash join: open-addressing build with 64-bit CAS, probe -> count / marks /
// pairs, and the fused probe -> GROUP BY -> exact SUM, COUNT.
//
// Reference path replaced:
//   exec/join_hash_table.rs:38-241  JoinHashTable (CRC32 hash, 16-bit bloom tag
//                                   directory, chained entries, serial insert)
//   kernel/hash.rs:84-225           HashBuild / HashProbe kernels (match count)
//   query_interpreter/join.rs:544-939 hash_join_with_keys: serial build, rayon
//                                   probe_all in 65 536-row chunks, then full
//                                   materialisation of the joined rows
//   exec/join.rs:56-199             hash_join pairs, semi_join, anti_join
//   accum_grouped.rs:198-335        grouping of the joined rows
//
// Table layout (HBM): capacity = next_pow2(max(16, 2n)) slots of 16 bytes
// {int64 key, uint32 build row, uint32 payload}; a slot is free while its key
// holds the sentinel INT64_MIN.  Build threads claim a slot with one 64-bit CAS
// on the key; duplicate build keys each take their own slot (multiset
// semantics, like the reference's chains), and the table remembers whether any
// duplicate exists so that probes of a unique-key table stop at the first
// match.  Build rows whose key equals the sentinel go to a side list.  A probe
// is one 16-byte load per visited slot (linear probing, load factor <= 0.5).
#include "tgp_internal.h"
#include "tile_engine.h"

namespace tgp {

constexpr unsigned long long kEmpty = 0x8000000000000000ull;

struct Slot {
  unsigned long long key;
  uint32_t row;
  uint32_t payload;
};
static_assert(sizeof(Slot) == 16, "slot must be 16 bytes");

struct JoinTableDev {
  Slot* slots;
  uint64_t mask;
  const uint2* side;  // (row, payload) of build rows whose key is the sentinel
  uint64_t n_side;
  int32_t has_dup;
  int32_t _pad;
  const unsigned long long* bloom;  // optional blocked bloom filter (nullptr: none)
  uint64_t bloom_mask;              // its words - 1
};

__device__ __forceinline__ uint64_t hmix(uint64_t z) {
  // SplitMix64 finaliser: a bijection with full avalanche, so the low bits of
  // the mixed key index the table.
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Blocked bloom filter over the table's keys: a key sets 4 bits of one 64-bit word
// (word and bits from a second mix of the key), so a probe tests it with one load
// that the L2 holds (2M keys: 2 MiB at 8 bits per key).  A key whose bits are not
// all set is not in the table; the others are probed as usual.
__device__ __forceinline__ void bloom_pos(unsigned long long k, uint64_t mask, uint64_t* word,
                                          unsigned long long* bits) {
  const uint64_t h = hmix(k ^ 0x6A09E667F3BCC909ull);
  *word = (h >> 32) & mask;
  *bits = (1ull << (h & 63)) | (1ull << ((h >> 6) & 63)) | (1ull << ((h >> 12) & 63)) |
          (1ull << ((h >> 18) & 63));
}

__device__ __forceinline__ int64_t load_i64(const void* base, int32_t dtype, uint64_t i) {
  switch (dtype) {
    case TGP_I8: return ((const int8_t*)base)[i];
    case TGP_U8: return ((const uint8_t*)base)[i];
    case TGP_I16: return ((const int16_t*)base)[i];
    case TGP_U16: return ((const uint16_t*)base)[i];
    case TGP_I32: return ((const int32_t*)base)[i];
    case TGP_U32: return ((const uint32_t*)base)[i];
    default: return ((const int64_t*)base)[i];
  }
}

// Loads rows idx[u] of a column as int64 (0 where !ok[u]), for U rows at once.
// The dtype switch sits around all U loads so that they are in flight
// together: with a switch per load the compiler waits for each load inside its
// branch.  idx[u] must be a valid row even where !ok[u] (callers pass a row of
// the range), so that every load is unconditional.
// The loads are non-temporal: probe columns stream through once, and cached
// loads would evict the hash table's lines from the L2.
template <typename T, int U>
__device__ __forceinline__ void load_rows_as(const void* base, const uint64_t (&idx)[U],
                                             const bool (&ok)[U], int64_t (&out)[U]) {
  const T* p = static_cast<const T*>(base);
  T raw[U];
  static_for<U>([&](auto I) {
    raw[decltype(I)::value] = __builtin_nontemporal_load(p + idx[decltype(I)::value]);
  });
  static_for<U>([&](auto I) {
    constexpr int u = decltype(I)::value;
    out[u] = ok[u] ? (int64_t)raw[u] : 0;
  });
}

template <int U>
__device__ __forceinline__ void load_rows(const void* base, int32_t dtype,
                                          const uint64_t (&idx)[U], const bool (&ok)[U],
                                          int64_t (&out)[U]) {
  switch (dtype) {
    case TGP_I8: load_rows_as<int8_t>(base, idx, ok, out); break;
    case TGP_U8: load_rows_as<uint8_t>(base, idx, ok, out); break;
    case TGP_I16: load_rows_as<int16_t>(base, idx, ok, out); break;
    case TGP_U16: load_rows_as<uint16_t>(base, idx, ok, out); break;
    case TGP_I32: load_rows_as<int32_t>(base, idx, ok, out); break;
    case TGP_U32: load_rows_as<uint32_t>(base, idx, ok, out); break;
    default: load_rows_as<int64_t>(base, idx, ok, out); break;
  }
}

__global__ void __launch_bounds__(256) k_table_init(Slot* slots, uint64_t cap) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap;
       i += (uint64_t)gridDim.x * blockDim.x) {
    Slot s;
    s.key = kEmpty;
    s.row = 0;
    s.payload = 0;
    slots[i] = s;
  }
}

// flags[0]: duplicate seen (u32 as u64 word), flags[1]: sentinel-key rows
__global__ void __launch_bounds__(256) k_build(const void* keys, int32_t kdt, const void* pay,
                                               int32_t pdt, uint64_t n, Slot* slots,
                                               uint64_t mask, unsigned long long* flags) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = (unsigned long long)load_i64(keys, kdt, i);
    if (k == kEmpty) {
      atomicAdd(&flags[1], 1ull);
      continue;
    }
    const uint32_t p = pay ? (uint32_t)load_i64(pay, pdt, i) : (uint32_t)i;
    uint64_t s = hmix(k) & mask;
    bool dup = false;
    for (;;) {
      const unsigned long long old = atomicCAS(&slots[s].key, kEmpty, k);
      if (old == kEmpty) {
        slots[s].row = (uint32_t)i;
        slots[s].payload = p;
        break;
      }
      dup |= (old == k);
      s = (s + 1) & mask;
    }
    if (dup) atomicOr(&flags[0], 1ull);
  }
}

__global__ void __launch_bounds__(256) k_build_side(const void* keys, int32_t kdt, const void* pay,
                                                    int32_t pdt, uint64_t n, uint2* side,
                                                    unsigned long long* counter) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if ((unsigned long long)load_i64(keys, kdt, i) != kEmpty) continue;
    const unsigned long long at = atomicAdd(counter, 1ull);
    side[at] = make_uint2((uint32_t)i, pay ? (uint32_t)load_i64(pay, pdt, i) : (uint32_t)i);
  }
}

// Calls f(row, payload) for every build row matching key k (visit order: the
// probe sequence, then the side list).
template <typename F>
__device__ __forceinline__ uint32_t probe_key(const JoinTableDev& T, unsigned long long k, F&& f) {
  uint32_t m = 0;
  if (k == kEmpty) {
    for (uint64_t j = 0; j < T.n_side; ++j) {
      f(T.side[j].x, T.side[j].y);
      ++m;
    }
    return m;
  }
  if (T.bloom) {
    uint64_t wi;
    unsigned long long b;
    bloom_pos(k, T.bloom_mask, &wi, &b);
    if ((T.bloom[wi] & b) != b) return 0;
  }
  uint64_t s = hmix(k) & T.mask;
  for (;;) {
    const Slot e = T.slots[s];
    if (e.key == kEmpty) break;
    if (e.key == k) {
      f(e.row, e.payload);
      ++m;
      if (!T.has_dup) break;
    }
    s = (s + 1) & T.mask;
  }
  return m;
}

// Probes U keys with their first slot loads issued together (U random loads in
// flight per lane instead of one), then finishes each key's linear probe.
// ok[u] false skips key u.  f(std::integral_constant<int, u>, build row,
// payload) is called per match; the key index is a compile-time constant, so
// per-key arrays of the caller stay in registers.
template <typename F>
__device__ __forceinline__ void probe_from(const JoinTableDev& T, unsigned long long k, uint64_t s,
                                           Slot e, F&& f) {
  if (k == kEmpty) {
    for (uint64_t j = 0; j < T.n_side; ++j) f(T.side[j].x, T.side[j].y);
    return;
  }
  for (;;) {
    if (e.key == kEmpty) break;
    if (e.key == k) {
      f(e.row, e.payload);
      if (!T.has_dup) break;
    }
    s = (s + 1) & T.mask;
    e = T.slots[s];
  }
}

// ---- hot-key cache -------------------------------------------------------------------------
// Foreign keys are often skewed (a few build keys take most of the probes), and
// every probe of the 16-byte-slot table is a random read of a table far larger
// than an XCD's L2.  Before a large probe of a unique-key table, a sample of the
// probe keys is counted, the keys seen at least kHotThreshold times are looked
// up once in the table, and their exact entries (build row and payload, or
// "absent") go into a small bucketed image that every probe workgroup copies
// into LDS.  A probe whose key is in the image takes its match from LDS and
// skips the table; any other key probes the table as before, so results do not
// depend on what the image holds.  A key goes into way 0..3 of bucket
// (hmix(key) >> 40) mod kHotBuckets, or is left out when the bucket is full.
constexpr uint32_t kHotBuckets = 256, kHotWays = 4, kHotEntries = kHotBuckets * kHotWays;
constexpr uint32_t kHotAbsent = 0xffffffffu;  // image entry of a key with no build row
constexpr uint32_t kHotSampleChunks = 256, kHotSampleRows = 1024;  // 256K sampled probe keys
constexpr uint32_t kHotThreshold = 16;     // sampled 16 times: about 1 probe in 16K or more
constexpr uint32_t kHotMapSlots = 1u << 19;  // sample count map: at most 256K keys, load <= 0.5
constexpr uint32_t kHotMapProbes = 64;      // longest map probe (a key past it is not counted)
constexpr uint64_t kHotLdsBytes = kHotEntries * 16ull;  // keys + entries in LDS
static_assert((uint64_t)kHotSampleChunks * kHotSampleRows < (1ull << 32), "sample counts fit 32 bits");
// the image is used when its keys took at least 1/kHotMinShare of the sample
constexpr uint64_t kHotMinShare = 8;

// Header words of the image: [0] keys that reached the threshold, [1] sampled
// occurrences of the keys placed in the image, [2] sampled rows (sentinel keys
// excluded).
struct HotDev {
  const unsigned long long* key;  // kHotEntries keys (kEmpty: free way)
  const uint2* val;               // kHotEntries (build row | kHotAbsent, payload)
  const unsigned long long* hdr;
};

// The image in LDS (filled by hot_to_lds).
struct HotLds {
  const unsigned long long* key;
  const uint2* val;
};

// True when the image covers enough of the sample to be worth its LDS lookups
// (the same answer in every thread: two scalar loads).
__device__ __forceinline__ bool hot_in_use(const HotDev& H) {
  if (!H.hdr) return false;
  // both at most kHotSampleChunks * kHotSampleRows (2^18): 32 bits hold them
  const uint64_t cov = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)H.hdr[1]);
  const uint64_t tot = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)H.hdr[2]);
  return cov > 0 && cov * kHotMinShare >= tot;
}

// Copies the image into `lds` (kHotLdsBytes) and returns its LDS view; every
// thread of the workgroup calls it, and a barrier follows before the lookups.
__device__ __forceinline__ HotLds hot_to_lds(const HotDev& H, unsigned long long* lds) {
  uint2* lv = reinterpret_cast<uint2*>(lds + kHotEntries);
  for (uint32_t i = threadIdx.x; i < kHotEntries; i += blockDim.x) {
    lds[i] = H.key[i];
    lv[i] = H.val[i];
  }
  return HotLds{lds, lv};
}

__device__ __forceinline__ uint32_t hot_bucket(uint64_t h) {
  return (uint32_t)(h >> 40) & (kHotBuckets - 1);
}

template <bool HOT, int U, typename F>
__device__ __forceinline__ void probe_keys(const JoinTableDev& T, const unsigned long long (&k)[U],
                                           const bool (&ok)[U], F&& f, const HotLds& H) {
  uint64_t s[U];
  Slot e[U];
  bool go[U];
  uint64_t h[U];
  static_for<U>([&](auto I) {
    constexpr int u = decltype(I)::value;
    go[u] = ok[u];
    h[u] = hmix(k[u]);
  });
  if constexpr (HOT) {  // keys in the LDS image take their entry from it
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      if (!go[u] || k[u] == kEmpty) return;
      const uint32_t b = hot_bucket(h[u]) * kHotWays;
      const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(H.key + b);
      const ulonglong2 c = *reinterpret_cast<const ulonglong2*>(H.key + b + 2);
      const int w = a.x == k[u] ? 0 : a.y == k[u] ? 1 : c.x == k[u] ? 2 : c.y == k[u] ? 3 : -1;
      if (w < 0) return;
      go[u] = false;
      const uint2 x = H.val[b + w];
      if (x.x != kHotAbsent) f(I, x.x, x.y);
    });
  }
  if (T.bloom) {  // keys whose bloom bits are not all set skip the table
    unsigned long long w[U], b[U];
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      uint64_t wi;
      bloom_pos(k[u], T.bloom_mask, &wi, &b[u]);
      w[u] = (ok[u] && k[u] != kEmpty) ? T.bloom[wi] : ~0ull;
    });
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      if (k[u] != kEmpty && (w[u] & b[u]) != b[u]) go[u] = false;
    });
  }
  static_for<U>([&](auto I) {
    constexpr int u = decltype(I)::value;
    s[u] = h[u] & T.mask;
    e[u] = Slot{kEmpty, 0, 0};
    if (go[u] && k[u] != kEmpty) e[u] = T.slots[s[u]];
  });
  static_for<U>([&](auto I) {
    constexpr int u = decltype(I)::value;
    if (go[u]) probe_from(T, k[u], s[u], e[u], [&](uint32_t row, uint32_t pay) { f(I, row, pay); });
  });
}

template <int U, typename F>
__device__ __forceinline__ void probe_keys(const JoinTableDev& T, const unsigned long long (&k)[U],
                                           const bool (&ok)[U], F&& f) {
  probe_keys<false>(T, k, ok, f, HotLds{nullptr, nullptr});
}

constexpr int kProbeUnroll = 4;

// Scratch of one image: count map keys, count map counts, hot list, image keys,
// image entries, header.
struct HotScratch {
  unsigned long long* map_key;  // kHotMapSlots (kEmpty: free)
  uint32_t* map_cnt;            // kHotMapSlots
  unsigned long long* list;     // kHotEntries keys that reached the threshold
  unsigned long long* key;      // kHotEntries image keys
  uint2* val;                   // kHotEntries image entries
  unsigned long long* hdr;      // 4 words
};
constexpr uint64_t kHotScratchBytes =
    kHotMapSlots * 8ull + kHotMapSlots * 4ull + kHotEntries * 8ull * 2 + kHotEntries * 8ull + 64;

__global__ void __launch_bounds__(256) k_hot_clear(HotScratch S) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (uint64_t i = t; i < kHotMapSlots; i += stride) {
    S.map_key[i] = kEmpty;
    S.map_cnt[i] = 0;
  }
  for (uint64_t i = t; i < kHotEntries; i += stride) S.key[i] = kEmpty;
  if (t < 4) S.hdr[t] = 0;
}

// Counts kHotSampleRows consecutive probe keys from each of kHotSampleChunks
// evenly spaced places of [b, e) (one workgroup per place).  The workgroup
// counts its rows in an LDS map first and adds each distinct key's count to
// the device map once (the hottest keys fill a tenth of a sample, and device
// atomics on one word serialise); a key whose device count crosses
// kHotThreshold joins the hot list (once: only the addition that crosses it
// appends).
constexpr uint32_t kHotLdsMap = 2 * kHotSampleRows;
__global__ void __launch_bounds__(256) k_hot_sample(HotScratch S, const void* keys, int32_t kdt,
                                                    uint64_t b, uint64_t e) {
  __shared__ unsigned long long lkey[kHotLdsMap];
  __shared__ uint32_t lcnt[kHotLdsMap];
  for (uint32_t i = threadIdx.x; i < kHotLdsMap; i += blockDim.x) {
    lkey[i] = kEmpty;
    lcnt[i] = 0;
  }
  __syncthreads();
  const uint64_t c0 = b + (e - b) * blockIdx.x / gridDim.x;
  uint32_t sampled = 0;
  for (uint32_t r = threadIdx.x; r < kHotSampleRows; r += blockDim.x) {
    const uint64_t i = c0 + r;
    if (i >= e) break;
    const unsigned long long k = (unsigned long long)load_i64(keys, kdt, i);
    if (k == kEmpty) continue;  // the side list's key: never cached
    ++sampled;
    // at most kHotSampleRows keys in 2 x as many slots: a free slot or the key
    // is always found
    uint32_t s = (uint32_t)(hmix(k) >> 24) & (kHotLdsMap - 1);
    for (;;) {
      const unsigned long long cur = atomicCAS(&lkey[s], kEmpty, k);
      if (cur == kEmpty || cur == k) {
        atomicAdd(&lcnt[s], 1u);
        break;
      }
      s = (s + 1) & (kHotLdsMap - 1);
    }
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < kHotLdsMap; j += blockDim.x) {
    const uint32_t c = lcnt[j];
    if (!c) continue;
    const unsigned long long k = lkey[j];
    uint32_t s = (uint32_t)(hmix(k) >> 24) & (kHotMapSlots - 1);
    for (uint32_t p = 0; p < kHotMapProbes; ++p, s = (s + 1) & (kHotMapSlots - 1)) {
      unsigned long long cur = __atomic_load_n(&S.map_key[s], __ATOMIC_RELAXED);
      if (cur == kEmpty) {
        cur = atomicCAS(&S.map_key[s], kEmpty, k);
        if (cur == kEmpty) cur = k;  // claimed
      }
      if (cur != k) continue;
      const uint32_t old = atomicAdd(&S.map_cnt[s], c);
      if (old < kHotThreshold && old + c >= kHotThreshold) {
        const unsigned long long at = atomicAdd(&S.hdr[0], 1ull);
        if (at < kHotEntries) S.list[at] = k;
      }
      break;
    }
  }
  for (int off = 32; off > 0; off >>= 1) sampled += __shfl_xor(sampled, off);
  if ((threadIdx.x & 63) == 0 && sampled) atomicAdd(&S.hdr[2], (unsigned long long)sampled);
}

// Looks every hot key up in the (unique-key) table and places (key, build row
// and payload or kHotAbsent) in its bucket's first free way; adds the key's
// sampled count to hdr[1] when it is placed.
__global__ void __launch_bounds__(256) k_hot_image(HotScratch S, JoinTableDev T) {
  const unsigned long long nh = S.hdr[0] < kHotEntries ? S.hdr[0] : kHotEntries;
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= nh) return;
  const unsigned long long k = S.list[i];
  uint2 x = make_uint2(kHotAbsent, 0u);
  T.bloom = nullptr;
  probe_key(T, k, [&](uint32_t row, uint32_t pay) { x = make_uint2(row, pay); });
  const uint32_t b = hot_bucket(hmix(k)) * kHotWays;
  for (uint32_t w = 0; w < kHotWays; ++w) {
    if (atomicCAS(&S.key[b + w], kEmpty, k) != kEmpty) continue;
    S.val[b + w] = x;
    // the key's count in the sample (it is in the map: it reached the threshold)
    uint32_t s = (uint32_t)(hmix(k) >> 24) & (kHotMapSlots - 1);
    for (uint32_t p = 0; p < kHotMapProbes; ++p, s = (s + 1) & (kHotMapSlots - 1))
      if (S.map_key[s] == k) {
        atomicAdd(&S.hdr[1], (unsigned long long)S.map_cnt[s]);
        break;
      }
    return;
  }
}

// ---- multi-column keys -------------------------------------------------------------------
// A table over several key columns is keyed by a 64-bit packing of the row's
// cells (not collision-free), and every candidate match is verified against
// all key columns of its build row, as exec/join.rs:51-66 and :102-111 do.
constexpr int kMaxJoinKeys = 4;
constexpr uint32_t kNoBuildRow = 0xffffffffu;  // LEFT join: a probe row without a match

struct KeyCols {
  const void* p[kMaxJoinKeys];
  int32_t dt[kMaxJoinKeys];
  int32_t n;
  uint32_t pack_mask_bits;  // 0: all 64 bits of the packing; b: its low b bits (tests)
};

__device__ __forceinline__ unsigned long long pack_row(const KeyCols& K, uint64_t i) {
  uint64_t h = 0x9E3779B97F4A7C15ull * (uint64_t)K.n;
  for (int j = 0; j < K.n; ++j) h = hmix(h ^ (uint64_t)load_i64(K.p[j], K.dt[j], i));
  if (K.pack_mask_bits) h &= (1ull << K.pack_mask_bits) - 1;
  return h;
}

// Every key column of probe row i equals that of build row r (cells compared
// as 64-bit values, narrow columns widened by their signedness).
__device__ __forceinline__ bool keys_equal(const KeyCols& P, uint64_t i, const KeyCols& B,
                                           uint64_t r) {
  for (int j = 0; j < P.n; ++j)
    if (load_i64(P.p[j], P.dt[j], i) != load_i64(B.p[j], B.dt[j], r)) return false;
  return true;
}

__global__ void __launch_bounds__(256) k_pack_keys(KeyCols K, uint64_t n,
                                                   unsigned long long* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = pack_row(K, i);
}

// ---- match count (HashProbe) ----------------------------------------------------------
template <bool HOT>
__device__ __forceinline__ unsigned long long count_rows(const JoinTableDev& T, const void* keys,
                                                         int32_t kdt, uint64_t b, uint64_t e,
                                                         const HotLds& H) {
  unsigned long long c = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = b + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i0 < e;
       i0 += stride * kProbeUnroll) {
    uint64_t idx[kProbeUnroll];
    bool ok[kProbeUnroll];
    int64_t kv[kProbeUnroll];
    unsigned long long k[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
      ok[u] = i0 + u * stride < e;
      idx[u] = ok[u] ? i0 + u * stride : i0;
    }
    load_rows(keys, kdt, idx, ok, kv);
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) k[u] = (unsigned long long)kv[u];
    probe_keys<HOT>(T, k, ok, [&](auto, uint32_t, uint32_t) { ++c; }, H);
  }
  return c;
}

// HD.hdr non-null: the launch gives kHotLdsBytes of dynamic LDS for the image.
__global__ void __launch_bounds__(256) k_probe_count(JoinTableDev T, HotDev HD, const void* keys,
                                                     int32_t kdt, uint64_t b, uint64_t e,
                                                     unsigned long long* out) {
  extern __shared__ __align__(16) unsigned long long dyn[];
  unsigned long long c;
  if (hot_in_use(HD)) {
    const HotLds H = hot_to_lds(HD, dyn);
    __syncthreads();
    c = count_rows<true>(T, keys, kdt, b, e, H);
  } else {
    c = count_rows<false>(T, keys, kdt, b, e, HotLds{nullptr, nullptr});
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// ---- semi-join marks: one bitmap word per wave-row-group of 64 -------------------------
// MULTI: P holds the probe's key columns, B the build's, and a candidate counts
// only when every column matches.
template <bool MULTI>
__global__ void __launch_bounds__(256) k_probe_mark(JoinTableDev T, KeyCols P, KeyCols B,
                                                    uint64_t b, uint64_t e, uint64_t* bitmap) {
  const uint64_t n = e - b;
  for (uint64_t base = blockIdx.x * 256ull; base < n; base += gridDim.x * 256ull) {
    const uint64_t r = base + threadIdx.x;
    bool hit = false;
    if (r < n) {
      const uint64_t i = b + r;
      if constexpr (MULTI) {
        probe_key(T, pack_row(P, i), [&](uint32_t row, uint32_t) {
          hit = hit || keys_equal(P, i, B, row);
        });
      } else {
        const unsigned long long k = (unsigned long long)load_i64(P.p[0], P.dt[0], i);
        T.has_dup = 0;  // existence only: stop at the first match
        hit = probe_key(T, k, [](uint32_t, uint32_t) {}) > 0;
      }
    }
    const unsigned long long w = __ballot(hit);
    const uint64_t wbase = base + (threadIdx.x & ~63u);
    if ((threadIdx.x & 63) == 0 && wbase < n) bitmap[wbase >> 6] = w;
  }
}

// ---- match bits of a unique-key table --------------------------------------------------
// Bit (i - b) of `bits` is set when probe row i matches a build row below
// row_limit (a unique-key table: at most one).  Rows are taken as in
// k_probe_count (lane-consecutive, kProbeUnroll rows a grid stride apart); the
// grid stride is a multiple of 64, so each of a wave's kProbeUnroll groups of
// 64 rows is one bitmap word, written with one ballot.  LIMIT false: no build
// prefix (the slot's build row is not read).
template <bool HOT, bool LIMIT>
__device__ __forceinline__ void probe_bits_rows(const JoinTableDev& T, const void* keys,
                                                int32_t kdt, uint64_t b, uint64_t e,
                                                uint32_t row_limit, uint64_t* bits,
                                                const HotLds& H) {
  const uint64_t n = e - b;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t r0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r0 - lane < n;
       r0 += stride * kProbeUnroll) {
    uint64_t idx[kProbeUnroll];
    bool ok[kProbeUnroll], hit[kProbeUnroll];
    int64_t kv[kProbeUnroll];
    unsigned long long k[kProbeUnroll];
    static_for<kProbeUnroll>([&](auto I) {
      constexpr int u = decltype(I)::value;
      const uint64_t r = r0 + u * stride;
      ok[u] = r < n;
      idx[u] = b + (ok[u] ? r : r0 - lane);  // r0 - lane < n: a row of the range
      hit[u] = false;
    });
    load_rows(keys, kdt, idx, ok, kv);
    static_for<kProbeUnroll>([&](auto I) {
      k[decltype(I)::value] = (unsigned long long)kv[decltype(I)::value];
    });
    probe_keys<HOT>(T, k, ok, [&](auto I, uint32_t row, uint32_t) {
      if (!LIMIT || row < row_limit) hit[decltype(I)::value] = true;
    }, H);
    static_for<kProbeUnroll>([&](auto I) {
      constexpr int u = decltype(I)::value;
      const uint64_t w0 = r0 - lane + u * stride;  // the wave's first row of group u
      const unsigned long long word = __ballot(hit[u]);
      if (lane == 0 && w0 < n) bits[w0 >> 6] = word;
    });
  }
}

// HD.hdr non-null: the launch gives kHotLdsBytes of dynamic LDS for the image.
template <bool LIMIT>
__global__ void __launch_bounds__(256) k_probe_bits(JoinTableDev T, HotDev HD, const void* keys,
                                                    int32_t kdt, uint64_t b, uint64_t e,
                                                    uint32_t row_limit, uint64_t* bits) {
  extern __shared__ __align__(16) unsigned long long dyn[];
  T.has_dup = 0;  // unique keys: stop at the first match
  if (hot_in_use(HD)) {
    const HotLds H = hot_to_lds(HD, dyn);
    __syncthreads();
    probe_bits_rows<true, LIMIT>(T, keys, kdt, b, e, row_limit, bits, H);
  } else {
    probe_bits_rows<false, LIMIT>(T, keys, kdt, b, e, row_limit, bits, HotLds{nullptr, nullptr});
  }
}

// ---- fused probe -> group -> exact sum ---------------------------------------------------
struct ProbeAggDev {
  const void* keys;
  const void* val;
  const void* grp;
  int32_t kdt, vdt, gdt, group_by;
  int64_t key_base;
  uint64_t ngroups;
  uint64_t b, e;
  uint32_t row_limit;  // build rows at or above it do not match (~0u: none excluded)
  uint32_t _pad;
  unsigned long long* out_count;
  unsigned long long* out_sums;
  unsigned long long* out_oor;
};

// Exact sums without returning atomics: a value v is added as its unsigned low
// 32 bits into one 64-bit word and its signed high 32 bits into another.  Each
// word stays exact for fewer than 2^32 additions, and
//   sum(v) = sum(lo32) + 2^32 * sum(hi32)
// is rebuilt as int128 when the partial table is flushed.
__device__ __forceinline__ void acc_add(unsigned long long* cnt, unsigned long long* sum,
                                        uint64_t g, int64_t v) {
  atomicAdd(&cnt[g], 1ull);
  atomicAdd(&sum[2 * g], (unsigned long long)(uint32_t)v);
  const int64_t h = v >> 32;
  if (h) atomicAdd(&sum[2 * g + 1], (unsigned long long)h);
}


// LDS form: every wave owns a table of ngroups x (count, sum lo32, sum hi32)
// words in R = 2^rshift copies, word w of copy c at (w << rshift) + c, and lane
// l updates copy l mod R, so that the lanes of one atomic rarely collide when
// there are few groups.  The copies and the 4 waves are added up at the end.
struct LdsTab {
  unsigned long long* base;
  uint32_t rshift, copy;
  __device__ __forceinline__ unsigned long long* at(uint64_t w) const {
    return base + ((uint32_t)w << rshift) + copy;
  }
};

__device__ __forceinline__ void acc_add_lds(const LdsTab& t, uint64_t ng, uint64_t g, int64_t v) {
  atomicAdd(t.at(g), 1ull);
  atomicAdd(t.at(ng + 2 * g), (unsigned long long)(uint32_t)v);
  const int64_t h = v >> 32;
  if (h) atomicAdd(t.at(ng + 2 * g + 1), (unsigned long long)h);
}

// Exact for any number of additions: the (lo, hi) words hold an int128 and the
// carry out of the low word is taken from the value the atomic returns.  A table
// with duplicate build keys gives a probe row several matches, so the number of
// additions per word is not bounded by the probe rows; these tables use this form.
__device__ __forceinline__ void i128_atomic_add(unsigned long long* lo, unsigned long long* hi,
                                                int64_t v) {
  const unsigned long long old = atomicAdd(lo, (unsigned long long)v);
  const int64_t h = (v >> 63) + (int64_t)((old + (unsigned long long)v) < old);
  if (h) atomicAdd(hi, (unsigned long long)h);
}

// Probes rows i_start, i_start + stride, ... below e (kProbeUnroll rows per lane at
// a time) and calls add(group, value) per match inside the query's build prefix;
// matches whose group falls outside [0, ngroups) count in oor.
template <bool HOT, typename Add>
__device__ __forceinline__ void probe_rows(const JoinTableDev& T, const ProbeAggDev& Q,
                                           uint64_t i_start, uint64_t e, uint64_t stride,
                                           Add&& add, uint32_t& oor, const HotLds& H) {
  for (uint64_t i0 = i_start; i0 < e; i0 += stride * kProbeUnroll) {
    uint64_t idx[kProbeUnroll];
    bool ok[kProbeUnroll];
    unsigned long long k[kProbeUnroll];
    int64_t kv[kProbeUnroll], v[kProbeUnroll], g[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
      ok[u] = i0 + u * stride < e;
      idx[u] = ok[u] ? i0 + u * stride : i0;
    }
    load_rows(Q.keys, Q.kdt, idx, ok, kv);
    load_rows(Q.val, Q.vdt, idx, ok, v);
    if (Q.group_by == TGP_JOIN_GROUP_PROBE_COL) load_rows(Q.grp, Q.gdt, idx, ok, g);
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
      k[u] = (unsigned long long)kv[u];
      if (Q.group_by != TGP_JOIN_GROUP_PROBE_COL) g[u] = 0;
    }
    if (Q.group_by == TGP_JOIN_GROUP_BUILD_PAYLOAD) {
      probe_keys<HOT>(T, k, ok, [&](auto I, uint32_t row, uint32_t pay) {
        if (row >= Q.row_limit) return;
        const int64_t gg = (int64_t)(int32_t)pay - Q.key_base;
        if (gg < 0 || (uint64_t)gg >= Q.ngroups) { ++oor; return; }
        add((uint64_t)gg, v[decltype(I)::value]);
      }, H);
    } else {
      probe_keys<HOT>(T, k, ok, [&](auto I, uint32_t row, uint32_t) {
        if (row >= Q.row_limit) return;
        constexpr int u = decltype(I)::value;
        const int64_t gg = Q.group_by == TGP_JOIN_GROUP_PROBE_COL ? g[u] - Q.key_base : 0;
        if (gg < 0 || (uint64_t)gg >= Q.ngroups) { ++oor; return; }
        add((uint64_t)gg, v[u]);
      }, H);
    }
  }
}

// Adds the workgroup's LDS tables (nwt tables of `region` words, each group's
// words in 2^rshift copies) into the query's result and clears them.  Called by
// every thread after a barrier; each thread reads and clears only its groups.
template <bool EXACT>
__device__ __forceinline__ void flush_lds_tables(unsigned long long* ltab, uint64_t region,
                                                 uint32_t nwt, uint32_t rshift,
                                                 const ProbeAggDev& Q) {
  // word w summed over the tables and copies: lo32 sums and counts stay below
  // 2^64 (fewer than 2^32 additions per word between flushes), hi32 sums are
  // signed 64-bit
  auto total = [&](uint64_t w) {
    unsigned long long t = 0;
    for (uint32_t wv = 0; wv < nwt; ++wv)
      for (uint32_t c = 0; c < (1u << rshift); ++c) {
        unsigned long long& x = ltab[wv * region + (w << rshift) + c];
        t += x;
        x = 0;
      }
    return t;
  };
  for (uint64_t gi = threadIdx.x; gi < Q.ngroups; gi += blockDim.x) {
    const unsigned long long c = total(gi);
    uint64_t lo;
    int64_t hi;
    if constexpr (EXACT) {
      // int128 copies: add the low words with their carries, then the high words
      lo = 0;
      hi = 0;
      for (uint32_t wv = 0; wv < nwt; ++wv)
        for (uint32_t c2 = 0; c2 < (1u << rshift); ++c2) {
          const uint64_t base = wv * region + c2;
          unsigned long long& l = ltab[base + ((Q.ngroups + 2 * gi) << rshift)];
          unsigned long long& h = ltab[base + ((Q.ngroups + 2 * gi + 1) << rshift)];
          const uint64_t nl = lo + l;
          hi += (int64_t)h + (int64_t)(nl < lo);
          lo = nl;
          l = 0;
          h = 0;
        }
    } else {
      const unsigned long long slo = total(Q.ngroups + 2 * gi);
      split_to_i128(slo, (int64_t)total(Q.ngroups + 2 * gi + 1), &lo, &hi);
    }
    if (!c) continue;
    atomicAdd(&Q.out_count[gi], c);
    if (lo || hi) {
      const unsigned long long old = atomicAdd(&Q.out_sums[2 * gi], (unsigned long long)lo);
      const int64_t h = hi + (int64_t)((old + lo) < old);
      if (h) atomicAdd(&Q.out_sums[2 * gi + 1], (unsigned long long)h);
    }
  }
}

template <bool USE_LDS, bool EXACT>
__device__ __forceinline__ void lds_or_global_add(const LdsTab& lt, const ProbeAggDev& Q,
                                                  uint64_t g, int64_t v) {
  if constexpr (EXACT) {
    if constexpr (USE_LDS) {
      atomicAdd(lt.at(g), 1ull);
      i128_atomic_add(lt.at(Q.ngroups + 2 * g), lt.at(Q.ngroups + 2 * g + 1), v);
    } else {  // the result buffers hold int128 words
      atomicAdd(&Q.out_count[g], 1ull);
      i128_atomic_add(&Q.out_sums[2 * g], &Q.out_sums[2 * g + 1], v);
    }
  } else if constexpr (USE_LDS) {
    acc_add_lds(lt, Q.ngroups, g, v);
  } else {  // the result buffers hold the split sums until k_split_finalize
    acc_add(Q.out_count, Q.out_sums, g, v);
  }
}

template <bool USE_LDS, bool EXACT>
__global__ void __launch_bounds__(256) k_probe_agg(JoinTableDev T, ProbeAggDev Q,
                                                   uint32_t rshift, uint32_t shared_table) {
  extern __shared__ __align__(16) unsigned long long ltab[];
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  const uint64_t words = 3 * Q.ngroups;                  // per copy
  const uint64_t region = words << rshift;               // per wave (or per workgroup)
  const uint32_t nwt = shared_table ? 1 : 4;             // tables in the workgroup
  LdsTab lt{ltab + (shared_table ? 0 : wave) * region, rshift, lane & ((1u << rshift) - 1)};
  if (USE_LDS) {
    for (uint64_t i = threadIdx.x; i < nwt * region; i += blockDim.x) ltab[i] = 0;
    __syncthreads();
  }
  uint32_t oor = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  probe_rows<false>(T, Q, Q.b + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, Q.e, stride,
                    [&](uint64_t g, int64_t v) { lds_or_global_add<USE_LDS, EXACT>(lt, Q, g, v); },
                    oor, HotLds{nullptr, nullptr});
  if (USE_LDS) {
    __syncthreads();
    flush_lds_tables<EXACT>(ltab, region, nwt, rshift, Q);
  }
  if (oor && Q.out_oor) atomicAdd(Q.out_oor, (unsigned long long)oor);
}

// ---- many probe aggregations in one launch -----------------------------------------------
// Each query of a batch probes its own rows of its own table into its own result;
// work items of up to kBatchItemRows probe rows are taken from a counter by
// persistent workgroups, whose four waves keep per-wave LDS tables of the item's
// query (2^rshift copies, as in k_probe_agg) and add them to the query's result
// after every item.
constexpr uint64_t kBatchItemRows = 16384;
constexpr uint32_t kBatchWaveWords = 640;  // 5 KiB per wave, 20 KiB per workgroup: 8 per CU

struct ProbeBatchQuery {
  JoinTableDev T;
  ProbeAggDev Q;
  uint32_t rshift;
  uint32_t _pad;
};

struct ProbeItem {
  uint64_t b;      // first probe row
  uint32_t q, n;   // query, rows
};

template <bool EXACT>
__global__ void __launch_bounds__(256) k_probe_agg_batch(const ProbeBatchQuery* __restrict__ qs,
                                                         const ProbeItem* __restrict__ items,
                                                         uint64_t nitems,
                                                         unsigned long long* counter) {
  __shared__ __align__(16) unsigned long long ltab[4 * kBatchWaveWords];
  __shared__ uint32_t s_item;
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  for (uint32_t i = threadIdx.x; i < 4 * kBatchWaveWords; i += blockDim.x) ltab[i] = 0;
  for (;;) {
    __syncthreads();  // tables clear, s_item free
    if (threadIdx.x == 0) s_item = (uint32_t)atomicAdd(counter, 1ull);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= nitems) break;
    const uint64_t b = items[it].b;
    const uint32_t q = items[it].q, n = items[it].n;
    const JoinTableDev T = qs[q].T;
    const ProbeAggDev Q = qs[q].Q;
    const uint32_t rshift = qs[q].rshift;
    const uint64_t region = (3 * Q.ngroups) << rshift;
    LdsTab lt{ltab + wave * region, rshift, lane & ((1u << rshift) - 1)};
    uint32_t oor = 0;
    probe_rows<false>(T, Q, b + threadIdx.x, b + n, blockDim.x,
                      [&](uint64_t g, int64_t v) { lds_or_global_add<true, EXACT>(lt, Q, g, v); },
                      oor, HotLds{nullptr, nullptr});
    if (oor && Q.out_oor) atomicAdd(Q.out_oor, (unsigned long long)oor);
    __syncthreads();
    flush_lds_tables<EXACT>(ltab, region, 4, rshift, Q);
  }
}

// Zeroes every query's result (count, sums, out-of-range word): one workgroup each.
__global__ void __launch_bounds__(256) k_probe_batch_zero(const ProbeBatchQuery* __restrict__ qs) {
  const ProbeAggDev& Q = qs[blockIdx.x].Q;
  for (uint64_t i = threadIdx.x; i < Q.ngroups; i += blockDim.x) Q.out_count[i] = 0;
  for (uint64_t i = threadIdx.x; i < 2 * Q.ngroups; i += blockDim.x) Q.out_sums[i] = 0;
  if (threadIdx.x == 0 && Q.out_oor) *Q.out_oor = 0;
}

// In place: (sum of lo32, sum of hi32) pairs -> int128 (lo, hi) words.
__global__ void __launch_bounds__(256) k_split_finalize(unsigned long long* sums, uint64_t n) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < n;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo;
    int64_t hi;
    split_to_i128(sums[2 * g], (int64_t)sums[2 * g + 1], &lo, &hi);
    sums[2 * g] = lo;
    sums[2 * g + 1] = (unsigned long long)hi;
  }
}

__global__ void __launch_bounds__(1024) k_scan_blocks(uint64_t* v, uint64_t n) {
  __shared__ unsigned long long part[1024];
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t lo = threadIdx.x * per, hi = min(lo + per, n);
  unsigned long long s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += v[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned long long x = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x ? part[threadIdx.x - 1] : 0;
  for (uint64_t i = lo; i < hi; ++i) {
    const unsigned long long c = v[i];
    v[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) v[n] = part[1023];
}

// Exclusive prefix sum of x over the 256 threads of a workgroup (wave
// shuffles, then the 4 wave totals); *total receives the workgroup's sum.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t* total) {
  __shared__ unsigned long long wave_tot[4];
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  uint64_t incl = x;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  uint64_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += wave_tot[w];
  *total = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
  __syncthreads();
  return before + incl - x;
}

// Exclusive scan of n u64 words in place, the total stored at v[n]:
// per-block sums, a scan of the block sums, then each block scans its chunk.
constexpr int kScanPer = 16;                      // words per thread
constexpr int kScanChunk = 256 * kScanPer;        // words per block

__global__ void __launch_bounds__(256) k_scan_sums(const uint64_t* v, uint64_t n, uint64_t* bsum) {
  const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  uint64_t x = 0;
  for (int j = 0; j < kScanPer; ++j)
    if (c0 + j < n) x += v[c0 + j];
  uint64_t tot = 0;
  (void)block_exclusive_scan(x, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_scan_chunks(uint64_t* v, uint64_t n, const uint64_t* bsum,
                                                     uint64_t nblocks) {
  const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  uint64_t w[kScanPer];
  uint64_t x = 0;
  for (int j = 0; j < kScanPer; ++j) {
    w[j] = c0 + j < n ? v[c0 + j] : 0;
    x += w[j];
  }
  uint64_t tot = 0;
  uint64_t run = bsum[blockIdx.x] + block_exclusive_scan(x, &tot);
  for (int j = 0; j < kScanPer; ++j) {
    if (c0 + j < n) v[c0 + j] = run;
    run += w[j];
  }
  if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) v[n] = bsum[nblocks];
}

// bsum holds nblocks + 1 words.
int32_t scan_u64(hipStream_t s, uint64_t* v, uint64_t n, uint64_t* bsum) {
  const uint64_t nb = (n + kScanChunk - 1) / kScanChunk;
  if (nb == 0) return fill_u64(s, v, 1, 0);
  k_scan_sums<<<(unsigned)nb, 256, 0, s>>>(v, n, bsum);
  k_scan_blocks<<<1, 1024, 0, s>>>(bsum, nb);
  k_scan_chunks<<<(unsigned)nb, 256, 0, s>>>(v, n, bsum, nb);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

// ---- GROUP BY a probe column with many groups: radix partitioning -------------------
// Random device atomics on a table of millions of groups are the cost of the
// one-pass form.  Instead the probe rows are partitioned by
// p = (group - key_base) >> kPartShift (groups outside [0, ngroups) go to an
// extra partition P) and each partition is aggregated in LDS:
//   (1) k_part_hist: workgroup w of G counts its contiguous share of the rows
//       per partition into mat[p * G + w];
//   (2) an exclusive scan of mat gives every (partition, workgroup) pair its
//       output range, so the scatter needs no device atomics (the order of
//       records inside a range varies, the integer totals do not);
//   (3) k_part_scatter: each workgroup probes a tile of its rows (first slot
//       loads of 4 rows in flight per lane), ranks the rows per partition with
//       LDS counters and writes one 12-byte record per row, (value, local
//       group | matched), at its partition's cursor plus its rank.  A row with
//       several matches (duplicate build keys) adds matches x value to the
//       zeroed result with device atomics and writes an unmatched record;
//   (4) k_part_agg: one workgroup per partition aggregates its matched records
//       in an LDS table of 2^kPartShift groups and adds its groups' totals to
//       the result with plain stores (each group belongs to exactly one
//       partition, and the scatter's atomics finished in the kernel before).
// Probing in the scatter keeps the key out of the records and lets the
// aggregation read them without touching the hash table.
constexpr int kPartShift = 12;          // 4096 groups per partition: 80 KB of LDS
constexpr int kScatterRows = 16;        // rows per thread per scatter tile
constexpr int kPartTile = 256 * kScatterRows;  // rows per scatter tile
constexpr uint32_t kMaxParts = 1024;     // scatter LDS <= 12 KB; 4M groups
constexpr uint32_t kRecMatched = 0x8000; // record flag: the row has exactly one match

// One scatter record: a probe row's value and its flagged local group, written
// and read as one 12-byte access.
struct Rec12 {
  uint32_t lo, hi, grp;
};
static_assert(sizeof(Rec12) == 12, "12-byte record");

// Partition of a group-column value (nparts: outside [key_base, key_base + ngroups)).
__device__ __forceinline__ uint32_t part_of(const ProbeAggDev& Q, int64_t cell, uint32_t nparts) {
  const int64_t g = cell - Q.key_base;
  return (g < 0 || (uint64_t)g >= Q.ngroups) ? nparts : (uint32_t)((uint64_t)g >> kPartShift);
}

// Rows [Q.b, Q.e) split into G contiguous shares; share w = [row_lo(w), row_lo(w+1)).
__device__ __forceinline__ uint64_t row_lo(const ProbeAggDev& Q, uint64_t w, uint64_t G) {
  return Q.b + (Q.e - Q.b) * w / G;
}

// bits non-null: only rows whose bit (i - Q.b) is set are counted (k_part_scatter_bits).
__global__ void __launch_bounds__(256) k_part_hist(ProbeAggDev Q, uint32_t nparts,
                                                   const uint64_t* bits,
                                                   unsigned long long* mat) {
  extern __shared__ __align__(16) uint32_t h[];  // nparts + 1 counters
  const uint64_t G = gridDim.x;
  for (uint32_t p = threadIdx.x; p <= nparts; p += blockDim.x) h[p] = 0;
  __syncthreads();
  const uint64_t lo = row_lo(Q, blockIdx.x, G), hi = row_lo(Q, blockIdx.x + 1, G);
  constexpr int U = 8;
  for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += U * (uint64_t)blockDim.x) {
    uint64_t idx[U];
    bool ok[U];
    int64_t g[U];
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      ok[u] = i0 + (uint64_t)u * blockDim.x < hi;
      idx[u] = ok[u] ? i0 + (uint64_t)u * blockDim.x : i0;
    });
    load_rows(Q.grp, Q.gdt, idx, ok, g);
    if (bits)
      static_for<U>([&](auto I) {
        constexpr int u = decltype(I)::value;
        const uint64_t r = idx[u] - Q.b;
        ok[u] = ok[u] && ((bits[r >> 6] >> (r & 63)) & 1);
      });
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      if (ok[u]) atomicAdd(&h[part_of(Q, g[u], nparts)], 1u);
    });
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p <= nparts; p += blockDim.x)
    mat[(uint64_t)p * G + blockIdx.x] = h[p];
}

// mat: exclusive offsets after the scan ((nparts + 1) * G entries, then the total).
// Per tile of kPartTile rows (kScatterRows per thread, kept in registers): the
// rows are loaded with coalesced loads and probed, ranked per partition with
// LDS counters, and each record is written at its partition's cursor plus its
// rank.  With hundreds of partitions a tile holds only a few rows of each, so
// sorting the tile first would not make the writes contiguous; the small LDS
// footprint keeps 8 workgroups per CU to hide the probes' latency instead.
// BITS: the matches come from `bits` (bit i - Q.b set: probe row i has exactly one
// match, a unique-key table probed beforehand by k_probe_bits) and only matched
// rows get records; else every row is probed here and gets a record.
template <bool BITS>
__device__ __forceinline__ void scatter_rows(const JoinTableDev& T, const ProbeAggDev& Q,
                                             uint32_t nparts, unsigned long long* cursor,
                                             uint32_t* cnt, Rec12* recs, const uint64_t* bits) {
  constexpr int U = 4, R = kScatterRows;
  static_assert(R % U == 0, "rows per thread: whole probe groups");
  const uint64_t G = gridDim.x;
  const uint64_t lo = row_lo(Q, blockIdx.x, G), hi = row_lo(Q, blockIdx.x + 1, G);
  unsigned long long oor_multi = 0;  // matches of out-of-range rows with several matches
  for (uint64_t t0 = lo; t0 < hi; t0 += (uint64_t)R * blockDim.x) {
    for (uint32_t p = threadIdx.x; p <= nparts; p += blockDim.x) cnt[p] = 0;
    __syncthreads();
    int64_t v[R];
    uint32_t pr[R], rk[R];
    uint32_t rg[R];
    static_for<R / U>([&](auto H) {
      constexpr int h = decltype(H)::value;
      unsigned long long k[U];
      int64_t kv[U], vv[U], g[U];
      uint64_t idx[U];
      bool ok[U];
      uint32_t m[U];
      static_for<U>([&](auto I) {
        constexpr int u = decltype(I)::value;
        const uint64_t i = t0 + (uint64_t)(h * U + u) * blockDim.x + threadIdx.x;
        ok[u] = i < hi;
        idx[u] = ok[u] ? i : t0;  // t0 < hi: a row of the share
        m[u] = 0;
      });
      if constexpr (BITS) {
        static_for<U>([&](auto I) {
          constexpr int u = decltype(I)::value;
          const uint64_t r = idx[u] - Q.b;
          m[u] = ok[u] ? (uint32_t)((bits[r >> 6] >> (r & 63)) & 1) : 0u;
        });
      } else {
        load_rows(Q.keys, Q.kdt, idx, ok, kv);
      }
      load_rows(Q.val, Q.vdt, idx, ok, vv);
      load_rows(Q.grp, Q.gdt, idx, ok, g);
      static_for<U>([&](auto I) {
        constexpr int u = decltype(I)::value;
        if constexpr (!BITS) k[u] = (unsigned long long)kv[u];
        v[h * U + u] = vv[u];
        g[u] -= Q.key_base;
      });
      if constexpr (!BITS) {
        probe_keys(T, k, ok, [&](auto I, uint32_t row, uint32_t) {
          if (row < Q.row_limit) ++m[decltype(I)::value];
        });
      }
      static_for<U>([&](auto I) {
        constexpr int u = decltype(I)::value;
        constexpr int r = h * U + u;
        pr[r] = 0xffffffffu;
        if (!ok[u] || (BITS && m[u] == 0)) return;
        const bool in = g[u] >= 0 && (uint64_t)g[u] < Q.ngroups;
        const uint32_t p = in ? (uint32_t)((uint64_t)g[u] >> kPartShift) : nparts;
        if (m[u] > 1) {  // duplicate build keys: matches x value straight into the result
          if (in) {
            atomicAdd(&Q.out_count[g[u]], (unsigned long long)m[u]);
            const __int128 s = (__int128)v[r] * (__int128)m[u];
            atomic_add128(&Q.out_sums[2 * g[u]], (uint64_t)s, (int64_t)(s >> 64));
          } else {
            oor_multi += m[u];
          }
        }
        rg[r] = (uint32_t)((in ? ((uint64_t)g[u] & ((1u << kPartShift) - 1)) : 0) |
                           (m[u] == 1 ? kRecMatched : 0));
        pr[r] = p;
        rk[r] = atomicAdd(&cnt[p], 1u);
      });
    });
    __syncthreads();
    static_for<R>([&](auto I) {
      constexpr int r = decltype(I)::value;
      if (pr[r] == 0xffffffffu) return;
      const uint64_t dst = cursor[pr[r]] + rk[r];
      recs[dst] = Rec12{(uint32_t)(uint64_t)v[r], (uint32_t)((uint64_t)v[r] >> 32), rg[r]};
    });
    __syncthreads();
    for (uint32_t p = threadIdx.x; p <= nparts; p += blockDim.x) cursor[p] += cnt[p];
  }
  if (oor_multi && Q.out_oor) atomicAdd(Q.out_oor, oor_multi);
}

__global__ void __launch_bounds__(256) k_part_scatter(JoinTableDev T, ProbeAggDev Q,
                                                      uint32_t nparts,
                                                      const unsigned long long* mat,
                                                      Rec12* recs) {
  extern __shared__ __align__(16) unsigned long long cursor[];        // nparts + 1
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cursor + nparts + 1);  // nparts + 1
  for (uint32_t p = threadIdx.x; p <= nparts; p += blockDim.x)
    cursor[p] = mat[(uint64_t)p * gridDim.x + blockIdx.x];
  // the tile loop's first barrier orders these LDS writes before any read
  scatter_rows<false>(T, Q, nparts, cursor, cnt, recs, nullptr);
}

__global__ void __launch_bounds__(256) k_part_scatter_bits(ProbeAggDev Q, uint32_t nparts,
                                                           const uint64_t* bits,
                                                           const unsigned long long* mat,
                                                           Rec12* recs) {
  extern __shared__ __align__(16) unsigned long long cursor[];        // nparts + 1
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cursor + nparts + 1);  // nparts + 1
  for (uint32_t p = threadIdx.x; p <= nparts; p += blockDim.x)
    cursor[p] = mat[(uint64_t)p * gridDim.x + blockIdx.x];
  JoinTableDev none{};
  scatter_rows<true>(none, Q, nparts, cursor, cnt, recs, bits);
}

// One workgroup per partition (partition nparts: the out-of-range rows, whose
// matches are only counted).  Partition p's records are
// [mat[p * G], mat[(p + 1) * G]).
__global__ void __launch_bounds__(1024) k_part_agg(ProbeAggDev Q, uint32_t nparts, uint64_t G,
                                                  const unsigned long long* mat,
                                                  const Rec12* recs) {
  // counts fit 32 bits: a partitioned call takes fewer than 2^32 rows
  __shared__ uint32_t cnt[1 << kPartShift];
  __shared__ unsigned long long slo[1 << kPartShift], shi[1 << kPartShift];
  constexpr int U = 4;
  const uint32_t p = blockIdx.x;
  const uint64_t r0 = mat[(uint64_t)p * G], r1 = mat[(uint64_t)(p + 1) * G];
  const bool oor_part = p == nparts;
  if (!oor_part) {
    for (uint32_t g = threadIdx.x; g < (1u << kPartShift); g += blockDim.x) {
      cnt[g] = 0;
      slo[g] = 0;
      shi[g] = 0;
    }
    __syncthreads();
  }
  unsigned long long m = 0;
  for (uint64_t rb = r0 + threadIdx.x; rb < r1; rb += (uint64_t)blockDim.x * U) {
    int64_t v[U];
    uint32_t g[U];
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      const uint64_t r = rb + (uint64_t)u * blockDim.x;
      const bool ok = r < r1;
      const Rec12 x = recs[ok ? r : r0];  // r0 < r1 inside the loop
      v[u] = (int64_t)(((uint64_t)x.hi << 32) | x.lo);
      g[u] = ok ? x.grp : 0;
    });
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      if (!(g[u] & kRecMatched)) return;
      if (oor_part) {
        ++m;
        return;
      }
      const uint32_t gl = g[u] & (kRecMatched - 1);
      atomicAdd(&cnt[gl], 1u);
      atomicAdd(&slo[gl], (unsigned long long)(uint32_t)v[u]);
      const int64_t h = v[u] >> 32;
      if (h) atomicAdd(&shi[gl], (unsigned long long)h);
    });
  }
  if (oor_part) {
    if (m && Q.out_oor) atomicAdd(Q.out_oor, m);
    return;
  }
  __syncthreads();
  // the result was zeroed and holds the scatter's several-match rows: add to it
  const uint64_t g0 = (uint64_t)p << kPartShift;
  for (uint32_t gl = threadIdx.x; gl < (1u << kPartShift); gl += blockDim.x) {
    if (g0 + gl >= Q.ngroups) break;
    if (!cnt[gl]) continue;
    uint64_t lo;
    int64_t hi;
    split_to_i128(slo[gl], (int64_t)shi[gl], &lo, &hi);
    const uint64_t g = g0 + gl;
    Q.out_count[g] += (unsigned long long)cnt[gl];
    const uint64_t l0 = Q.out_sums[2 * g];
    const uint64_t s = l0 + lo;
    Q.out_sums[2 * g] = s;
    Q.out_sums[2 * g + 1] += (unsigned long long)(hi + (int64_t)(s < l0));
  }
}

// ---- inner-join pairs in probe order ---------------------------------------------------------
constexpr int kPairBlock = 1024;  // probe rows per block (256 threads x 4)

// Output rows per probe row: its (verified, for MULTI) matches, and 1 for a
// row without any under LEFT.
template <bool MULTI>
__global__ void __launch_bounds__(256) k_pairs_count(JoinTableDev T, KeyCols P, KeyCols B,
                                                     int32_t left, uint64_t b, uint64_t e,
                                                     uint32_t* per_row, uint64_t* block_tot) {
  __shared__ unsigned long long red[4];
  const uint64_t r0 = b + blockIdx.x * (uint64_t)kPairBlock;
  constexpr int U = kPairBlock / 256;
  uint64_t idx[U];
  bool ok[U];
  int64_t kv[U];
  unsigned long long k[U];
  uint32_t m[U];
  static_for<U>([&](auto I) {
    constexpr int u = decltype(I)::value;
    const uint64_t i = r0 + u * 256 + threadIdx.x;
    ok[u] = i < e;
    idx[u] = ok[u] ? i : r0;  // r0 < e: the grid covers the range
    m[u] = 0;
  });
  if constexpr (MULTI) {
    static_for<U>([&](auto I) {
      constexpr int u = decltype(I)::value;
      k[u] = pack_row(P, idx[u]);
    });
    probe_keys(T, k, ok, [&](auto I, uint32_t row, uint32_t) {
      constexpr int u = decltype(I)::value;
      if (keys_equal(P, idx[u], B, row)) ++m[u];
    });
  } else {
    load_rows(P.p[0], P.dt[0], idx, ok, kv);
    static_for<U>([&](auto I) { k[decltype(I)::value] = (unsigned long long)kv[decltype(I)::value]; });
    probe_keys(T, k, ok, [&](auto I, uint32_t, uint32_t) { ++m[decltype(I)::value]; });
  }
  unsigned long long c = 0;
  static_for<U>([&](auto I) {
    constexpr int u = decltype(I)::value;
    if (!ok[u]) return;
    const uint32_t out_rows = (left && m[u] == 0) ? 1u : m[u];
    per_row[idx[u] - b] = out_rows;
    c += out_rows;
  });
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) block_tot[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}


template <bool MULTI>
__global__ void __launch_bounds__(256) k_pairs_write(JoinTableDev T, KeyCols P, KeyCols B,
                                                     uint64_t b, uint64_t e,
                                                     const uint32_t* per_row,
                                                     const uint64_t* block_off,
                                                     unsigned long long* out_probe,
                                                     uint32_t* out_pay) {
  __shared__ unsigned long long pre[kPairBlock];
  const uint64_t r0 = b + blockIdx.x * (uint64_t)kPairBlock;
  for (int u = 0; u < kPairBlock / 256; ++u) {
    const uint64_t i = r0 + u * 256 + threadIdx.x;
    pre[u * 256 + threadIdx.x] = i < e ? per_row[i - b] : 0;
  }
  __syncthreads();
  // inclusive scan of the block's 1024 counts (Hillis-Steele, 4 per thread)
  for (int off = 1; off < kPairBlock; off <<= 1) {
    unsigned long long x[4];
    for (int u = 0; u < 4; ++u) {
      const int j = u * 256 + threadIdx.x;
      x[u] = j >= off ? pre[j - off] : 0;
    }
    __syncthreads();
    for (int u = 0; u < 4; ++u) pre[u * 256 + threadIdx.x] += x[u];
    __syncthreads();
  }
  const uint64_t boff = block_off[blockIdx.x];
  for (int u = 0; u < kPairBlock / 256; ++u) {
    const int j = u * 256 + threadIdx.x;
    const uint64_t i = r0 + j;
    if (i >= e) continue;
    const uint32_t m = per_row[i - b];
    if (!m) continue;
    const uint64_t start = boff + pre[j] - m;
    uint32_t w = 0;
    auto emit = [&](uint32_t pay) {
      if (w < m) {  // the count pass found m rows: the bound keeps writes in the range
        out_probe[start + w] = i;
        out_pay[start + w] = pay;
      }
      ++w;
    };
    if constexpr (MULTI) {
      probe_key(T, pack_row(P, i), [&](uint32_t row, uint32_t pay) {
        if (keys_equal(P, i, B, row)) emit(pay);
      });
    } else {
      probe_key(T, (unsigned long long)load_i64(P.p[0], P.dt[0], i),
                [&](uint32_t, uint32_t pay) { emit(pay); });
    }
    if (w == 0) {  // LEFT: the row without a match (m == 1 was counted for it)
      emit(kNoBuildRow);
      continue;
    }
    // duplicates: ascending build payload within the probe row (insertion sort)
    for (uint32_t x = 1; x < m; ++x) {
      const uint32_t key = out_pay[start + x];
      uint32_t y = x;
      while (y > 0 && out_pay[start + y - 1] > key) {
        out_pay[start + y] = out_pay[start + y - 1];
        --y;
      }
      out_pay[start + y] = key;
    }
  }
}

// Sets every stored key's bloom bits (the keys of the filled slots; the sentinel key
// lives on the side list and never consults the filter).
__global__ void __launch_bounds__(256) k_bloom_build(const Slot* __restrict__ slots, uint64_t cap,
                                                     unsigned long long* bloom, uint64_t mask) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = slots[i].key;
    if (k == kEmpty) continue;
    uint64_t w;
    unsigned long long b;
    bloom_pos(k, mask, &w, &b);
    if ((bloom[w] & b) != b) atomicOr(&bloom[w], b);
  }
}

static unsigned grid_for(uint64_t items, uint64_t cap = 8192) {
  uint64_t g = (items + 255) / 256;
  if (g > cap) g = cap;
  return (unsigned)(g ? g : 1);
}

// Hot-key image sizes: tables below kHotMinTableBytes stay in the L2s and the
// MALL anyway, and below kHotMinProbes the sample costs more than it saves.
constexpr uint64_t kHotMinTableBytes = 8ull << 20;
constexpr uint64_t kHotMinProbes = 4ull << 20;

// Queues the hot-key image of probe keys [b, e) of t into `region`
// (kHotScratchBytes, 256-byte aligned) and returns the kernels' view of it.
static int32_t build_hot(hipStream_t s, const JoinTableDev& T, const void* keys, int32_t kdt,
                         uint64_t b, uint64_t e, void* region, HotDev* out) {
  uint8_t* p = static_cast<uint8_t*>(region);
  HotScratch S;
  S.map_key = reinterpret_cast<unsigned long long*>(p);
  p += kHotMapSlots * 8ull;
  S.map_cnt = reinterpret_cast<uint32_t*>(p);
  p += kHotMapSlots * 4ull;
  S.list = reinterpret_cast<unsigned long long*>(p);
  p += kHotEntries * 8ull;
  S.key = reinterpret_cast<unsigned long long*>(p);
  p += kHotEntries * 8ull;
  S.val = reinterpret_cast<uint2*>(p);
  p += kHotEntries * 8ull;
  S.hdr = reinterpret_cast<unsigned long long*>(p);
  k_hot_clear<<<256, 256, 0, s>>>(S);
  k_hot_sample<<<kHotSampleChunks, 256, 0, s>>>(S, keys, kdt, b, e);
  k_hot_image<<<kHotEntries / 256, 256, 0, s>>>(S, T);
  TGP_HIP(hipGetLastError());
  *out = HotDev{S.key, S.val, S.hdr};
  return TGP_OK;
}

}  // namespace tgp

using namespace tgp;

struct tgp_hash_table {
  tgp_ctx* ctx = nullptr;
  void* mem = nullptr;       // slots, then 2 flag words
  bool owns_mem = true;      // false: caller memory (tgp_hash_build_in)
  void* side_mem = nullptr;  // side list (only when sentinel keys exist)
  uint64_t cap = 0, n_rows = 0, n_side = 0, bytes = 0;
  int32_t has_dup = 0;
  KeyCols bkeys{};  // build key columns (n > 1: a packed multi-column table)
  unsigned long long* bloom = nullptr;  // tgp_hash_table_add_bloom
  uint64_t bloom_words = 0;
  JoinTableDev dev() const {
    JoinTableDev d;
    d.slots = static_cast<Slot*>(mem);
    d.mask = cap - 1;
    d.side = static_cast<const uint2*>(side_mem);
    d.n_side = n_side;
    d.has_dup = has_dup;
    d._pad = 0;
    d.bloom = bloom;
    d.bloom_mask = bloom_words ? bloom_words - 1 : 0;
    return d;
  }
};

static uint64_t table_capacity(uint64_t n) {
  uint64_t c = 16;
  while (c < 2 * n) c <<= 1;  // join_hash_table.rs:76-85 sizes its directory the same way
  return c;
}

static int32_t check_key_column(const tgp_column* c, const char* what) {
  TGP_REQUIRE(c != nullptr, "%s: null column", what);
  TGP_REQUIRE(dtype_width(c->dtype) > 0, "%s: unknown dtype", what);
  TGP_REQUIRE(c->len == 0 || c->data != nullptr, "%s: null data", what);
  return TGP_OK;
}

// Whether a probe of n rows of t takes the hot-key image (unique-key single-column
// tables only: an image entry is then a key's only match).
static bool want_hot(tgp_ctx* ctx, const tgp_hash_table* t, uint64_t n) {
  const int64_t mode = ctx->tune_join_hot.load();
  if (mode == 1 || t->has_dup || t->bkeys.n > 1 || n == 0) return false;
  if (mode == 2) return true;
  return n >= kHotMinProbes && t->cap * sizeof(Slot) >= kHotMinTableBytes;
}

static int32_t require_single_key(const tgp_hash_table* t, const char* what) {
  TGP_REQUIRE(t->bkeys.n <= 1,
              "%s: the table has %d key columns (use tgp_hash_join_pairs_sync / tgp_hash_join_mark)",
              what, t->bkeys.n);
  return TGP_OK;
}

// The argument rules of a probe aggregation (tgp_hash_probe_agg and each query
// of tgp_hash_probe_agg_batch).
static int32_t check_probe_agg(const tgp_join_agg_query* q, const tgp_group_result* out) {
  TGP_REQUIRE(q->table, "tgp_hash_probe_agg: null table");
  TGP_TRY(require_single_key(q->table, "tgp_hash_probe_agg"));
  TGP_REQUIRE(out->d_count && out->d_sums, "tgp_hash_probe_agg: null result buffer");
  TGP_TRY(check_key_column(&q->probe_keys, "probe keys"));
  TGP_TRY(check_key_column(&q->value, "value column"));
  TGP_REQUIRE(q->group_by >= TGP_JOIN_GROUP_BUILD_PAYLOAD && q->group_by <= TGP_JOIN_GROUP_NONE,
              "bad group_by %d", q->group_by);
  if (q->group_by == TGP_JOIN_GROUP_PROBE_COL) {
    TGP_TRY(check_key_column(&q->group_col, "group column"));
    TGP_REQUIRE(q->group_col.len >= q->row_end, "group column shorter than row_end");
  }
  TGP_REQUIRE(q->ngroups >= 1 && q->ngroups <= (1ull << 32), "ngroups out of range");
  TGP_REQUIRE(q->group_by != TGP_JOIN_GROUP_NONE || q->ngroups == 1,
              "GROUP_NONE needs ngroups == 1");
  TGP_REQUIRE(q->row_begin <= q->row_end && q->row_end <= q->probe_keys.len &&
                  q->row_end <= q->value.len,
              "bad probe row range");
  return TGP_OK;
}

// The kernel arguments of one probe aggregation.
static ProbeAggDev probe_agg_dev(const tgp_join_agg_query* q, const tgp_group_result* out) {
  ProbeAggDev Q;
  Q.keys = q->probe_keys.data;
  Q.kdt = q->probe_keys.dtype;
  Q.val = q->value.data;
  Q.vdt = q->value.dtype;
  Q.grp = q->group_col.data;
  Q.gdt = q->group_col.dtype;
  Q.group_by = q->group_by;
  Q.key_base = q->group_by == TGP_JOIN_GROUP_NONE ? 0 : q->key_base;
  Q.ngroups = q->ngroups;
  Q.b = q->row_begin;
  Q.e = q->row_end;
  Q.row_limit = q->build_rows ? q->build_rows : ~0u;
  Q._pad = 0;
  Q.out_count = reinterpret_cast<unsigned long long*>(out->d_count);
  Q.out_sums = reinterpret_cast<unsigned long long*>(out->d_sums);
  Q.out_oor = reinterpret_cast<unsigned long long*>(out->d_n_out_of_range);
  return Q;
}

extern "C" {

uint64_t tgp_hash_table_bytes_for(uint64_t n_build_rows) {
  return table_capacity(n_build_rows) * sizeof(Slot) + 16;
}

static int32_t build_table(tgp_ctx* ctx, const tgp_column* keys, const tgp_column* payload,
                           void* d_mem, uint64_t mem_bytes, tgp_hash_table** out, void* stream);

int32_t tgp_hash_build(tgp_ctx* ctx, const tgp_column* keys, const tgp_column* payload,
                       tgp_hash_table** out, void* stream) {
  return build_table(ctx, keys, payload, nullptr, 0, out, stream);
}

int32_t tgp_hash_build_in(tgp_ctx* ctx, const tgp_column* keys, const tgp_column* payload,
                          void* d_mem, uint64_t mem_bytes, tgp_hash_table** out, void* stream) {
  TGP_REQUIRE(d_mem && (reinterpret_cast<uintptr_t>(d_mem) & 15) == 0,
              "tgp_hash_build_in: table memory must be non-null and 16-byte aligned");
  return build_table(ctx, keys, payload, d_mem, mem_bytes, out, stream);
}

int32_t tgp_hash_build_multi(tgp_ctx* ctx, const tgp_column* keys, int32_t nkeys,
                             const tgp_column* payload, tgp_hash_table** out, void* stream) {
  TGP_REQUIRE(ctx && keys && out, "tgp_hash_build_multi: null argument");
  TGP_REQUIRE(nkeys >= 1 && nkeys <= kMaxJoinKeys, "tgp_hash_build_multi: %d key columns (1..%d)",
              nkeys, kMaxJoinKeys);
  *out = nullptr;
  if (nkeys == 1) return build_table(ctx, keys, payload, nullptr, 0, out, stream);
  KeyCols K{};
  K.n = nkeys;
  K.pack_mask_bits = (uint32_t)ctx->tune_join_pack_bits.load();
  for (int j = 0; j < nkeys; ++j) {
    TGP_TRY(check_key_column(&keys[j], "tgp_hash_build_multi key"));
    TGP_REQUIRE(keys[j].len == keys[0].len, "tgp_hash_build_multi: key columns of %llu and %llu rows",
                (unsigned long long)keys[0].len, (unsigned long long)keys[j].len);
    K.p[j] = keys[j].data;
    K.dt[j] = keys[j].dtype;
  }
  const uint64_t n = keys[0].len;
  hipStream_t s = pick_stream(ctx, stream);
  // the packed keys, in the stream's outer scratch region while the table is built
  void* packed = nullptr;
  TGP_TRY(get_scratch(ctx, s, n * 8 + 8, &packed, 1));
  if (n) {
    k_pack_keys<<<grid_for(n), 256, 0, s>>>(K, n, static_cast<unsigned long long*>(packed));
    TGP_HIP(hipGetLastError());
  }
  tgp_column pk{};
  pk.data = packed;
  pk.len = n;
  pk.dtype = TGP_U64;
  TGP_TRY(build_table(ctx, &pk, payload, nullptr, 0, out, stream));
  (*out)->bkeys = K;
  return TGP_OK;
}

static int32_t build_table(tgp_ctx* ctx, const tgp_column* keys, const tgp_column* payload,
                           void* d_mem, uint64_t mem_bytes, tgp_hash_table** out, void* stream) {
  TGP_REQUIRE(ctx && out, "tgp_hash_build: null argument");
  *out = nullptr;
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_key_column(keys, "tgp_hash_build keys"));
  if (payload) {
    TGP_TRY(check_key_column(payload, "tgp_hash_build payload"));
    TGP_REQUIRE(payload->len >= keys->len, "tgp_hash_build: payload shorter than keys");
  }
  TGP_REQUIRE(keys->len < (1ull << 32), "tgp_hash_build: build rows must fit in 32 bits");
  hipStream_t s = pick_stream(ctx, stream);
  tgp_hash_table* t = new tgp_hash_table();
  t->ctx = ctx;
  t->n_rows = keys->len;
  t->cap = table_capacity(keys->len);
  t->bytes = t->cap * sizeof(Slot) + 16;
  int32_t rc = TGP_OK;
  if (d_mem) {
    if (mem_bytes < t->bytes) {
      set_error("tgp_hash_build_in: %llu bytes given, the table needs %llu",
                (unsigned long long)mem_bytes, (unsigned long long)t->bytes);
      delete t;
      return TGP_ERR_INVALID;
    }
    t->mem = d_mem;
    t->owns_mem = false;
  } else {
    rc = hbm_alloc(ctx, t->bytes, &t->mem);
    if (rc != TGP_OK) {
      delete t;
      return rc;
    }
  }
  Slot* slots = static_cast<Slot*>(t->mem);
  unsigned long long* flags = reinterpret_cast<unsigned long long*>(slots + t->cap);
  auto fail = [&](int32_t code) {
    if (t->owns_mem) hbm_free(ctx, t->mem);
    delete t;
    return code;
  };
  if (fill_u64(s, reinterpret_cast<uint64_t*>(flags), 2, 0) != TGP_OK) return fail(TGP_ERR_HIP);
  k_table_init<<<grid_for(t->cap), 256, 0, s>>>(slots, t->cap);
  const void* pd = payload ? payload->data : nullptr;
  const int32_t pdt = payload ? payload->dtype : TGP_U32;
  if (keys->len)
    k_build<<<grid_for(keys->len), 256, 0, s>>>(keys->data, keys->dtype, pd, pdt, keys->len, slots,
                                                t->cap - 1, flags);
  if (hipGetLastError() != hipSuccess) {
    set_error("tgp_hash_build: kernel launch failed");
    return fail(TGP_ERR_HIP);
  }
  unsigned long long hflags[2] = {0, 0};
  if (hipMemcpyAsync(hflags, flags, 16, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("tgp_hash_build: reading build flags failed");
    return fail(TGP_ERR_HIP);
  }
  t->has_dup = hflags[0] ? 1 : 0;
  t->n_side = hflags[1];
  t->bkeys.n = 1;
  t->bkeys.p[0] = keys->data;
  t->bkeys.dt[0] = keys->dtype;
  if (t->n_side) {
    rc = hbm_alloc(ctx, t->n_side * sizeof(uint2) + 8, &t->side_mem);
    if (rc != TGP_OK) return fail(rc);
    t->bytes += t->n_side * sizeof(uint2) + 8;
    unsigned long long* counter =
        reinterpret_cast<unsigned long long*>(static_cast<uint2*>(t->side_mem) + t->n_side);
    fill_u64(s, reinterpret_cast<uint64_t*>(counter), 1, 0);
    k_build_side<<<grid_for(keys->len), 256, 0, s>>>(keys->data, keys->dtype, pd, pdt, keys->len,
                                                     static_cast<uint2*>(t->side_mem), counter);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      hbm_free(ctx, t->side_mem);
      set_error("tgp_hash_build: side list failed");
      return fail(TGP_ERR_HIP);
    }
    if (t->n_side > 1) t->has_dup = 1;
  }
  *out = t;
  return TGP_OK;
}

int32_t tgp_hash_table_destroy(tgp_ctx* ctx, tgp_hash_table* t) {
  if (!t) return TGP_OK;
  TGP_REQUIRE(ctx && t->ctx == ctx, "tgp_hash_table_destroy: table belongs to another context");
  int32_t rc = TGP_OK;
  if (t->owns_mem || t->side_mem) {
    TGP_HIP(hipStreamSynchronize(ctx->stream));
    if (t->owns_mem) rc = hbm_free(ctx, t->mem);
    if (t->side_mem) hbm_free(ctx, t->side_mem);
  }
  if (t->bloom) {
    TGP_HIP(hipStreamSynchronize(ctx->stream));
    hbm_free(ctx, t->bloom);
  }
  delete t;
  return rc;
}

int32_t tgp_hash_table_footprint(const tgp_hash_table* t, uint64_t* out_bytes) {  // incl. caller memory
  TGP_REQUIRE(t && out_bytes, "tgp_hash_table_footprint: null argument");
  *out_bytes = t->bytes;
  return TGP_OK;
}

int32_t tgp_hash_table_info(const tgp_hash_table* t, uint64_t* n_rows, uint64_t* capacity,
                            int32_t* has_duplicates) {
  TGP_REQUIRE(t, "tgp_hash_table_info: null table");
  if (n_rows) *n_rows = t->n_rows;
  if (capacity) *capacity = t->cap;
  if (has_duplicates) *has_duplicates = t->has_dup;
  return TGP_OK;
}

// The probe's key columns as kernel arguments, checked against the table.
static int32_t probe_key_cols(const tgp_hash_table* t, const tgp_column* keys, int32_t nkeys,
                              uint64_t row_end, KeyCols* P) {
  const int32_t want = t->bkeys.n > 1 ? t->bkeys.n : 1;
  TGP_REQUIRE(keys && nkeys == want, "join probe: %d key columns for a table of %d", nkeys, want);
  *P = KeyCols{};
  P->n = nkeys;
  P->pack_mask_bits = t->bkeys.pack_mask_bits;
  for (int j = 0; j < nkeys; ++j) {
    TGP_TRY(check_key_column(&keys[j], "join probe key"));
    TGP_REQUIRE(row_end <= keys[j].len, "join probe: rows up to %llu, key column %d has %llu",
                (unsigned long long)row_end, j, (unsigned long long)keys[j].len);
    P->p[j] = keys[j].data;
    P->dt[j] = keys[j].dtype;
  }
  return TGP_OK;
}

static int32_t join_mark(tgp_ctx* ctx, const tgp_hash_table* t, const KeyCols& P,
                         uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                         void* stream) {
  if (row_begin == row_end) return TGP_OK;
  const unsigned g = grid_for(row_end - row_begin, 4096);
  hipStream_t s = pick_stream(ctx, stream);
  if (P.n > 1)
    k_probe_mark<true><<<g, 256, 0, s>>>(t->dev(), P, t->bkeys, row_begin, row_end, d_bitmap);
  else
    k_probe_mark<false><<<g, 256, 0, s>>>(t->dev(), P, t->bkeys, row_begin, row_end, d_bitmap);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

static int32_t join_pairs(tgp_ctx* ctx, const tgp_hash_table* t, const KeyCols& P, int32_t left,
                          uint64_t row_begin, uint64_t row_end, uint64_t* d_probe_idx,
                          uint32_t* d_build_payload, uint64_t capacity, uint64_t* h_npairs,
                          void* stream);

int32_t tgp_hash_probe_count(tgp_ctx* ctx, const tgp_hash_table* t, const tgp_column* probe_keys,
                             uint64_t row_begin, uint64_t row_end, uint64_t* d_count,
                             void* stream) {
  TGP_REQUIRE(ctx && t && d_count, "tgp_hash_probe_count: null argument");
  TGP_TRY(require_single_key(t, "tgp_hash_probe_count"));
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_key_column(probe_keys, "tgp_hash_probe_count keys"));
  TGP_REQUIRE(row_begin <= row_end && row_end <= probe_keys->len, "bad probe row range");
  if (row_begin == row_end) return TGP_OK;
  // one resident round (8 workgroups per CU): 0.54 ms for 60M probes against
  // 0.59 at 16 per CU and 0.72 at 4
  const unsigned grid = grid_for((row_end - row_begin + kProbeUnroll - 1) / kProbeUnroll,
                                 (uint64_t)persistent_grid(ctx, 8));
  hipStream_t s = pick_stream(ctx, stream);
  HotDev hd{nullptr, nullptr, nullptr};
  if (want_hot(ctx, t, row_end - row_begin)) {
    void* region = nullptr;
    TGP_TRY(get_scratch(ctx, s, kHotScratchBytes, &region));
    TGP_TRY(build_hot(s, t->dev(), probe_keys->data, probe_keys->dtype, row_begin, row_end,
                      region, &hd));
  }
  k_probe_count<<<grid, 256, hd.hdr ? kHotLdsBytes : 0, s>>>(
      t->dev(), hd, probe_keys->data, probe_keys->dtype, row_begin, row_end,
      reinterpret_cast<unsigned long long*>(d_count));
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_hash_probe_mark(tgp_ctx* ctx, const tgp_hash_table* t, const tgp_column* probe_keys,
                            uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                            void* stream) {
  TGP_REQUIRE(ctx && t && d_bitmap, "tgp_hash_probe_mark: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(require_single_key(t, "tgp_hash_probe_mark"));
  TGP_TRY(check_key_column(probe_keys, "tgp_hash_probe_mark keys"));
  TGP_REQUIRE(row_begin <= row_end && row_end <= probe_keys->len, "bad probe row range");
  KeyCols P;
  TGP_TRY(probe_key_cols(t, probe_keys, 1, row_end, &P));
  return join_mark(ctx, t, P, row_begin, row_end, d_bitmap, stream);
}

int32_t tgp_hash_join_mark(tgp_ctx* ctx, const tgp_hash_table* t, const tgp_column* probe_keys,
                           int32_t nkeys, uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                           void* stream) {
  TGP_REQUIRE(ctx && t && d_bitmap, "tgp_hash_join_mark: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_REQUIRE(row_begin <= row_end, "bad probe row range");
  KeyCols P;
  TGP_TRY(probe_key_cols(t, probe_keys, nkeys, row_end, &P));
  return join_mark(ctx, t, P, row_begin, row_end, d_bitmap, stream);
}

int32_t tgp_hash_probe_pairs_sync(tgp_ctx* ctx, const tgp_hash_table* t,
                                  const tgp_column* probe_keys, uint64_t row_begin,
                                  uint64_t row_end, uint64_t* d_probe_idx,
                                  uint32_t* d_build_payload, uint64_t capacity, uint64_t* h_npairs,
                                  void* stream) {
  TGP_REQUIRE(ctx && t && h_npairs, "tgp_hash_probe_pairs_sync: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(require_single_key(t, "tgp_hash_probe_pairs_sync"));
  TGP_TRY(check_key_column(probe_keys, "tgp_hash_probe_pairs_sync keys"));
  TGP_REQUIRE(row_begin <= row_end && row_end <= probe_keys->len, "bad probe row range");
  *h_npairs = 0;
  KeyCols P;
  TGP_TRY(probe_key_cols(t, probe_keys, 1, row_end, &P));
  return join_pairs(ctx, t, P, 0, row_begin, row_end, d_probe_idx, d_build_payload, capacity,
                    h_npairs, stream);
}

int32_t tgp_hash_join_pairs_sync(tgp_ctx* ctx, const tgp_hash_table* t,
                                 const tgp_column* probe_keys, int32_t nkeys, int32_t join_type,
                                 uint64_t row_begin, uint64_t row_end, uint64_t* d_probe_idx,
                                 uint32_t* d_build_payload, uint64_t capacity, uint64_t* h_npairs,
                                 void* stream) {
  TGP_REQUIRE(ctx && t && h_npairs, "tgp_hash_join_pairs_sync: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_REQUIRE(join_type == TGP_JOIN_INNER || join_type == TGP_JOIN_LEFT,
              "tgp_hash_join_pairs_sync: join type %d (INNER or LEFT)", join_type);
  TGP_REQUIRE(row_begin <= row_end, "bad probe row range");
  *h_npairs = 0;
  KeyCols P;
  TGP_TRY(probe_key_cols(t, probe_keys, nkeys, row_end, &P));
  return join_pairs(ctx, t, P, join_type == TGP_JOIN_LEFT, row_begin, row_end, d_probe_idx,
                    d_build_payload, capacity, h_npairs, stream);
}

static int32_t join_pairs(tgp_ctx* ctx, const tgp_hash_table* t, const KeyCols& P, int32_t left,
                          uint64_t row_begin, uint64_t row_end, uint64_t* d_probe_idx,
                          uint32_t* d_build_payload, uint64_t capacity, uint64_t* h_npairs,
                          void* stream) {
  *h_npairs = 0;
  const bool multi = P.n > 1;
  const uint64_t n = row_end - row_begin;
  if (n == 0) return TGP_OK;
  hipStream_t s = pick_stream(ctx, stream);
  const uint64_t nblocks = (n + kPairBlock - 1) / kPairBlock;
  void* scratch = nullptr;
  TGP_TRY(get_scratch(ctx, s, n * 4 + (nblocks + 1) * 8 + 16, &scratch));
  uint64_t* block_tot = static_cast<uint64_t*>(scratch);
  uint32_t* per_row = reinterpret_cast<uint32_t*>(block_tot + nblocks + 1);
  if (multi)
    k_pairs_count<true><<<(unsigned)nblocks, 256, 0, s>>>(t->dev(), P, t->bkeys, left, row_begin,
                                                          row_end, per_row, block_tot);
  else
    k_pairs_count<false><<<(unsigned)nblocks, 256, 0, s>>>(t->dev(), P, t->bkeys, left, row_begin,
                                                           row_end, per_row, block_tot);
  TGP_HIP(hipGetLastError());
  k_scan_blocks<<<1, 1024, 0, s>>>(block_tot, nblocks);
  TGP_HIP(hipGetLastError());
  uint64_t total = 0;
  TGP_HIP(hipMemcpyAsync(&total, block_tot + nblocks, 8, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  *h_npairs = total;
  if (total > capacity) {
    set_error("join produces %llu pairs, capacity %llu", (unsigned long long)total,
              (unsigned long long)capacity);
    return TGP_ERR_CAPACITY;
  }
  if (total == 0) return TGP_OK;
  TGP_REQUIRE(d_probe_idx && d_build_payload, "tgp_hash_probe_pairs_sync: null output");
  if (multi)
    k_pairs_write<true><<<(unsigned)nblocks, 256, 0, s>>>(
        t->dev(), P, t->bkeys, row_begin, row_end, per_row, block_tot,
        reinterpret_cast<unsigned long long*>(d_probe_idx), d_build_payload);
  else
    k_pairs_write<false><<<(unsigned)nblocks, 256, 0, s>>>(
        t->dev(), P, t->bkeys, row_begin, row_end, per_row, block_tot,
        reinterpret_cast<unsigned long long*>(d_probe_idx), d_build_payload);
  TGP_HIP(hipGetLastError());
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_hash_probe_agg(tgp_ctx* ctx, const tgp_join_agg_query* q, const tgp_group_result* out,
                           void* stream) {
  TGP_REQUIRE(ctx && q && out && q->table, "tgp_hash_probe_agg: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_probe_agg(q, out));
  hipStream_t s = pick_stream(ctx, stream);
  TGP_TRY(fill_u64(s, out->d_count, q->ngroups, 0));
  TGP_TRY(fill_u64(s, out->d_sums, 2 * q->ngroups, 0));
  if (out->d_n_out_of_range) TGP_TRY(fill_u64(s, out->d_n_out_of_range, 1, 0));
  const uint64_t n = q->row_end - q->row_begin;
  if (n == 0) return TGP_OK;
  const ProbeAggDev Q = probe_agg_dev(q, out);
  // 32 workgroups per CU, several rounds of them: the probe rows of a workgroup
  // meet chains of different lengths, and smaller shares even out the CUs' end
  // times (8 per CU: 0.97 ms for 60M probes into 25 groups; 32: 0.76 ms)
  const unsigned grid = grid_for((n + kProbeUnroll - 1) / kProbeUnroll,
                                 (uint64_t)persistent_grid(ctx, 32));
  // LDS tables: per wave with as many copies as fit in 20 KB (8 workgroups of
  // 256 on a CU's 160 KB; 16 copies in 38 KB took 1.0 ms against 0.76), else
  // per wave within 48 KB, else one per workgroup, else the radix-partitioned
  // form (probe-column groups) or global atomics
  constexpr uint64_t kLdsFull = 20 * 1024, kLdsMax = 48 * 1024;
  const uint64_t tab = 3 * 8 * q->ngroups;
  int rshift = -1;
  uint32_t shared_table = 0;
  for (int r = 6; r >= 0 && rshift < 0; --r)
    if ((tab << r) * 4 <= kLdsFull) rshift = r;
  for (int r = 6; r >= 0 && rshift < 0; --r)
    if ((tab << r) * 4 <= kLdsMax) rshift = r;
  if (rshift < 0 && tab <= kLdsMax) { rshift = 0; shared_table = 1; }
  const uint64_t nparts = (q->ngroups + (1u << kPartShift) - 1) >> kPartShift;
  // duplicate build keys: a probe row adds once per match, so the split lo32/hi32
  // words (exact below 2^32 additions each) give way to carry-exact int128 words
  const bool exact = q->table->has_dup != 0;
  const JoinTableDev T = q->table->dev();
  if (rshift >= 0) {
    const uint64_t lds = (tab << rshift) * (shared_table ? 1 : 4);
    if (exact)
      k_probe_agg<true, true><<<grid, 256, lds, s>>>(T, Q, (uint32_t)rshift, shared_table);
    else
      k_probe_agg<true, false><<<grid, 256, lds, s>>>(T, Q, (uint32_t)rshift, shared_table);
  } else if (q->group_by == TGP_JOIN_GROUP_PROBE_COL && nparts < kMaxParts &&
             ctx->tune_join_partition.load() == 0) {
    TGP_REQUIRE(n < (1ull << 32),
                "tgp_hash_probe_agg: a partitioned call takes fewer than 2^32 probe rows");
    // A unique-key table is probed first into one match bit per row, and only
    // matched rows are partitioned (the histogram and the scatter then stream
    // the columns without probing); a table with duplicate keys is probed in
    // the scatter, which adds the rows of several matches directly.
    const bool use_bits = !q->table->has_dup;
    const bool hot = use_bits && want_hot(ctx, q->table, n);
    // scratch: mat[(nparts + 1) * G + 1] | records[n] | match bits | hot-key image
    uint64_t G = (uint64_t)persistent_grid(ctx, 8) < (n + kPartTile - 1) / kPartTile
                     ? (uint64_t)persistent_grid(ctx, 8)
                     : (n + kPartTile - 1) / kPartTile;
    const int64_t gmax = ctx->tune_grid_max.load();
    if (gmax > 0 && G > (uint64_t)gmax) G = (uint64_t)gmax;
    const uint64_t mat_n = (nparts + 1) * G + 1;
    const uint64_t nsb = (mat_n - 1 + kScanChunk - 1) / kScanChunk + 1;  // scan block sums
    const uint64_t head = (8 * (mat_n + nsb) + 255) & ~255ull;
    const uint64_t boff = (head + sizeof(Rec12) * n + 256 + 255) & ~255ull;
    const uint64_t hoff = (boff + (use_bits ? 8 * ((n + 63) / 64 + 4) : 0) + 255) & ~255ull;
    const uint64_t bytes = hoff + (hot ? kHotScratchBytes : 0);
    void* scr = nullptr;
    TGP_TRY(get_scratch(ctx, s, bytes, &scr));
    auto* mat = static_cast<unsigned long long*>(scr);
    auto* recs = reinterpret_cast<Rec12*>(static_cast<uint8_t*>(scr) + head);
    auto* bits = use_bits ? reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(scr) + boff) : nullptr;
    const uint32_t np = (uint32_t)nparts;
    // the aggregation adds to the result (several-match rows arrive first)
    {
      uint64_t* zp[3] = {reinterpret_cast<uint64_t*>(Q.out_count),
                         reinterpret_cast<uint64_t*>(Q.out_sums),
                         reinterpret_cast<uint64_t*>(Q.out_oor)};
      const uint64_t zn[3] = {q->ngroups, 2 * q->ngroups, Q.out_oor ? 1ull : 0ull};
      TGP_TRY(zero_ranges(s, 3, zp, zn));
    }
    if (use_bits) {
      HotDev hd{nullptr, nullptr, nullptr};
      if (hot)
        TGP_TRY(build_hot(s, T, Q.keys, Q.kdt, Q.b, Q.e, static_cast<uint8_t*>(scr) + hoff, &hd));
      const unsigned bgrid = grid_for((n + kProbeUnroll - 1) / kProbeUnroll,
                                      (uint64_t)persistent_grid(ctx, 8));
      const uint64_t blds = hd.hdr ? kHotLdsBytes : 0;
      if (Q.row_limit == ~0u)
        k_probe_bits<false><<<bgrid, 256, blds, s>>>(T, hd, Q.keys, Q.kdt, Q.b, Q.e, Q.row_limit,
                                                     bits);
      else
        k_probe_bits<true><<<bgrid, 256, blds, s>>>(T, hd, Q.keys, Q.kdt, Q.b, Q.e, Q.row_limit,
                                                    bits);
    }
    k_part_hist<<<(unsigned)G, 256, (np + 1) * 4, s>>>(Q, np, bits, mat);
    TGP_TRY(scan_u64(s, reinterpret_cast<uint64_t*>(mat), mat_n - 1,
                     reinterpret_cast<uint64_t*>(mat + mat_n)));
    const uint32_t lds = (np + 1) * 12 + 16;
    if (use_bits)
      k_part_scatter_bits<<<(unsigned)G, 256, lds, s>>>(Q, np, bits, mat, recs);
    else
      k_part_scatter<<<(unsigned)G, 256, lds, s>>>(T, Q, np, mat, recs);
    k_part_agg<<<np + 1, 1024, 0, s>>>(Q, np, G, mat, recs);
  } else {
    TGP_REQUIRE(q->row_end - q->row_begin < (1ull << 32),
                "tgp_hash_probe_agg: a global-table call takes fewer than 2^32 probe rows");
    if (exact) {
      k_probe_agg<false, true><<<grid, 256, 0, s>>>(T, Q, 0, 0);
    } else {
      k_probe_agg<false, false><<<grid, 256, 0, s>>>(T, Q, 0, 0);
      TGP_HIP(hipGetLastError());
      k_split_finalize<<<grid_for(q->ngroups), 256, 0, s>>>(Q.out_sums, q->ngroups);
    }
  }
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_hash_probe_agg_batch(tgp_ctx* ctx, const tgp_join_agg_query* qs, uint32_t n,
                                 const tgp_group_result* outs, void* stream) {
  TGP_REQUIRE(ctx && (n == 0 || (qs && outs)), "tgp_hash_probe_agg_batch: null argument");
  TGP_TRY(check_injected(ctx));
  for (uint32_t i = 0; i < n; ++i) {
    TGP_REQUIRE(qs[i].table, "tgp_hash_probe_agg_batch: query %u has no table", i);
    TGP_TRY(check_probe_agg(&qs[i], &outs[i]));
  }
  hipStream_t s = pick_stream(ctx, stream);
  // queries whose per-wave table fits kBatchWaveWords words run in the batch
  // kernel; the others (more than 213 groups) run alone, after it
  std::vector<ProbeBatchQuery> bq;
  std::vector<ProbeItem> items;
  std::vector<uint32_t> alone;
  bool exact = false;
  for (uint32_t i = 0; i < n; ++i) {
    const tgp_join_agg_query& q = qs[i];
    const uint64_t words = 3 * q.ngroups;
    if (words > kBatchWaveWords) {
      alone.push_back(i);
      continue;
    }
    uint32_t r = 6;
    while ((words << r) > kBatchWaveWords) --r;
    ProbeBatchQuery b{};
    b.T = q.table->dev();
    b.Q = probe_agg_dev(&q, &outs[i]);
    b.rshift = r;
    exact = exact || q.table->has_dup != 0;
    const uint32_t qi = (uint32_t)bq.size();
    bq.push_back(b);
    for (uint64_t r0 = q.row_begin; r0 < q.row_end; r0 += kBatchItemRows)
      items.push_back(ProbeItem{r0, qi, (uint32_t)std::min<uint64_t>(kBatchItemRows,
                                                                     q.row_end - r0)});
  }
  if (!bq.empty()) {
    // scratch (slot 1): counter | queries | items
    const uint64_t qoff = 256, ioff = (qoff + bq.size() * sizeof(ProbeBatchQuery) + 255) & ~255ull;
    const uint64_t bytes = ioff + items.size() * sizeof(ProbeItem) + 16;
    void* scr = nullptr;
    TGP_TRY(get_scratch(ctx, s, bytes, &scr, 1));
    auto* base = static_cast<uint8_t*>(scr);
    auto* counter = reinterpret_cast<unsigned long long*>(base);
    auto* d_q = reinterpret_cast<ProbeBatchQuery*>(base + qoff);
    auto* d_items = reinterpret_cast<ProbeItem*>(base + ioff);
    TGP_HIP(hipMemcpyAsync(d_q, bq.data(), bq.size() * sizeof(ProbeBatchQuery),
                           hipMemcpyHostToDevice, s));
    if (!items.empty())
      TGP_HIP(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(ProbeItem),
                             hipMemcpyHostToDevice, s));
    // the copies read the host vectors before they go (the kernels are not waited for)
    TGP_HIP(hipStreamSynchronize(s));
    TGP_TRY(fill_u64(s, reinterpret_cast<uint64_t*>(counter), 1, 0));
    k_probe_batch_zero<<<(unsigned)bq.size(), 256, 0, s>>>(d_q);
    TGP_HIP(hipGetLastError());
    if (!items.empty()) {
      // 20 KiB of LDS per workgroup: eight per CU
      const uint64_t grid = std::min<uint64_t>(items.size(), (uint64_t)persistent_grid(ctx, 8));
      if (exact)
        k_probe_agg_batch<true><<<(unsigned)grid, 256, 0, s>>>(d_q, d_items, items.size(),
                                                               counter);
      else
        k_probe_agg_batch<false><<<(unsigned)grid, 256, 0, s>>>(d_q, d_items, items.size(),
                                                                counter);
      TGP_HIP(hipGetLastError());
    }
  }
  for (uint32_t i : alone) TGP_TRY(tgp_hash_probe_agg(ctx, &qs[i], &outs[i], s));
  return TGP_OK;
}

int32_t tgp_hash_table_add_bloom(tgp_ctx* ctx, tgp_hash_table* t, uint32_t bits_per_key,
                                 void* stream) {
  TGP_REQUIRE(ctx && t && t->ctx == ctx, "tgp_hash_table_add_bloom: bad argument");
  TGP_TRY(check_injected(ctx));
  TGP_REQUIRE(t->bkeys.n <= 1, "tgp_hash_table_add_bloom: single-key tables only");
  TGP_REQUIRE(bits_per_key >= 1 && bits_per_key <= 64, "bits_per_key %u not in 1..64",
              bits_per_key);
  TGP_REQUIRE(!t->bloom, "tgp_hash_table_add_bloom: the table has a filter already");
  hipStream_t s = pick_stream(ctx, stream);
  uint64_t words = 1;
  while (words * 64 < t->n_rows * (uint64_t)bits_per_key) words <<= 1;
  void* p = nullptr;
  TGP_TRY(hbm_alloc(ctx, words * 8, &p));
  t->bloom = static_cast<unsigned long long*>(p);
  t->bloom_words = words;
  TGP_TRY(fill_u64(s, reinterpret_cast<uint64_t*>(t->bloom), words, 0));
  k_bloom_build<<<grid_for(t->cap), 256, 0, s>>>(static_cast<const Slot*>(t->mem), t->cap, t->bloom,
                                                 words - 1);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

}  // extern "C"
