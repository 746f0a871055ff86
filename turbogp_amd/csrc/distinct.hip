// COUNT(DISTINCT col) over a row range and an optional selection bitmap.
//
// Reference path replaced:
//   exec/vectorized.rs:601-610   count_distinct_masked: exact HashSet of u64 cells
//   query_interpreter/aggregate.rs:1379-1400 and exec/hll.rs:97-128: exact below
//                                100 000 selected cells, else HyperLogLog (p = 14)
//   exec/hll.rs:32-75            HyperLogLog add / estimate over xxh3_64 of the
//                                cell's 8 little-endian bytes
// Cells are the values widened to u64 as the reference stores them (signed
// types sign-extended), so the distinct set and the hashes are the reference's.
//
// HyperLogLog: each workgroup keeps the 16384 registers as 32-bit LDS words
// (LDS atomic max), writes them as bytes to its slice of scratch, and a second
// kernel takes the per-register maximum over the workgroups.  The estimate is
// computed on the host from the 16 KiB of registers, in the reference's order
// of summation, so it is the reference's number bit for bit.
//
// Exact: an open-addressing set of u64 cells in scratch (capacity the power of
// two >= 2x the rows in range, so at most half full and every probe sequence
// ends at an empty slot), inserted with 64-bit CAS; each successful insert is
// one distinct value.  The empty marker value itself is counted by a flag.
#include "tgp_internal.h"

#include <array>
#include <cmath>

namespace tgp {

constexpr uint32_t kHllBits = 14;
constexpr uint32_t kHllRegs = 1u << kHllBits;
constexpr unsigned long long kSetEmpty = ~0ull;
// the interpreter's switch from an exact set to HyperLogLog (aggregate.rs:1387, hll.rs:100)
constexpr uint64_t kDistinctExactBelow = 100000;

// XXH3_64bits(8 little-endian bytes of v), seed 0, default secret: the 4..8-byte
// path of the published XXH3 algorithm (xxHash 0.8).  The constant is
// readLE64(secret + 8) ^ readLE64(secret + 16) of the default secret; that
// path's input word is v with its 32-bit halves swapped.
__host__ __device__ __forceinline__ uint64_t xxh3_u64(uint64_t v) {
  const uint64_t kMx2 = 0x9FB21C651E98DF25ull;
  const uint64_t kFlip = 0x1cad21f72c81017cull ^ 0xdb979083e96dd4deull;
  uint64_t h = ((v << 32) | (v >> 32)) ^ kFlip;
  h ^= ((h << 49) | (h >> 15)) ^ ((h << 24) | (h >> 40));
  h *= kMx2;
  h ^= (h >> 35) + 8;
  h *= kMx2;
  return h ^ (h >> 28);
}

template <int DT>
__device__ __forceinline__ uint64_t cell_at(const uint8_t* col, uint64_t i) {
  if constexpr (DT == TGP_I8) return (uint64_t)(int64_t)reinterpret_cast<const int8_t*>(col)[i];
  else if constexpr (DT == TGP_U8) return col[i];
  else if constexpr (DT == TGP_I16) return (uint64_t)(int64_t)reinterpret_cast<const int16_t*>(col)[i];
  else if constexpr (DT == TGP_U16) return reinterpret_cast<const uint16_t*>(col)[i];
  else if constexpr (DT == TGP_I32) return (uint64_t)(int64_t)reinterpret_cast<const int32_t*>(col)[i];
  else if constexpr (DT == TGP_U32) return reinterpret_cast<const uint32_t*>(col)[i];
  else return reinterpret_cast<const uint64_t*>(col)[i];
}

__device__ __forceinline__ bool selected(const uint64_t* sel, uint64_t i) {
  return !sel || ((sel[i >> 6] >> (i & 63)) & 1);
}

// The register file takes 64 KiB of LDS, so a CU holds two workgroups: they
// are 1024 threads wide, and every thread hashes kHllUnroll independent values
// per step (the hash is a chain of 64-bit multiplies).
constexpr uint32_t kHllBlock = 1024;
constexpr uint32_t kHllUnroll = 4;

// SEL: a selection bitmap is given (its words are read beside the cells).
template <int DT, bool SEL>
__global__ void __launch_bounds__(kHllBlock) k_hll(const uint8_t* __restrict__ col,
                                                   uint64_t row_begin, uint64_t n,
                                                   const uint64_t* __restrict__ sel,
                                                   uint8_t* __restrict__ partial) {
  __shared__ uint32_t reg[kHllRegs];  // 64 KiB
  for (uint32_t i = threadIdx.x; i < kHllRegs; i += kHllBlock) reg[i] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kHllBlock;
  for (uint64_t i0 = blockIdx.x * (uint64_t)kHllBlock + threadIdx.x; i0 < n;
       i0 += kHllUnroll * stride) {
    // all loads first (rows past the end read row i0, which is in range), then
    // the hashes: a load inside each row's branch would wait before the next
    uint64_t c[kHllUnroll];
    bool ok[kHllUnroll];
#pragma unroll
    for (uint32_t k = 0; k < kHllUnroll; ++k) {
      const uint64_t i = i0 + k * stride;
      const uint64_t r = i < n ? i : i0;
      c[k] = cell_at<DT>(col, row_begin + r);
      ok[k] = i < n && (!SEL || selected(sel, r));
    }
#pragma unroll
    for (uint32_t k = 0; k < kHllUnroll; ++k) {
      if (!ok[k]) continue;
      const uint64_t h = xxh3_u64(c[k]);
      const uint64_t w = h << kHllBits;
      const uint32_t rank = (w ? (uint32_t)__clzll((long long)w) : 64u) + 1u;
      atomicMax(&reg[h >> (64 - kHllBits)], rank);
    }
  }
  __syncthreads();
  uint8_t* out = partial + (uint64_t)blockIdx.x * kHllRegs;
  for (uint32_t i = threadIdx.x; i < kHllRegs; i += kHllBlock) out[i] = (uint8_t)reg[i];
}

// Bytewise maximum of two words of four registers each.
__device__ __forceinline__ uint32_t max_u8x4(uint32_t a, uint32_t b) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 32; k += 8) r |= max((a >> k) & 0xffu, (b >> k) & 0xffu) << k;
  return r;
}

// Row-wise maximum of register sets: thread (column c of 16 registers, y) takes
// rows [y * per, (y + 1) * per) of `src` (nrows x 1024 16-byte columns) and writes
// row y of `dst`.  Eight 16-byte loads are in flight per thread.
__global__ void __launch_bounds__(256) k_hll_merge(const uint4* __restrict__ src, uint32_t nrows,
                                                   uint32_t per, uint4* __restrict__ dst) {
  constexpr uint32_t kCols = kHllRegs / 16;
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= kCols) return;
  const uint32_t r0 = blockIdx.y * per;
  const uint32_t r1 = r0 + per < nrows ? r0 + per : nrows;
  uint4 m = make_uint4(0, 0, 0, 0);
  for (uint32_t r = r0; r < r1; r += 8) {
    uint4 x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      x[k] = r + k < r1 ? src[(uint64_t)(r + k) * kCols + c] : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      m.x = max_u8x4(m.x, x[k].x);
      m.y = max_u8x4(m.y, x[k].y);
      m.z = max_u8x4(m.z, x[k].z);
      m.w = max_u8x4(m.w, x[k].w);
    }
  }
  dst[(uint64_t)blockIdx.y * kCols + c] = m;
}

__device__ __forceinline__ uint64_t set_slot(uint64_t v) {  // SplitMix64 finaliser
  v ^= v >> 30;
  v *= 0xbf58476d1ce4e5b9ull;
  v ^= v >> 27;
  v *= 0x94d049bb133111ebull;
  return v ^ (v >> 31);
}

// counters[0]: values inserted (distinct values other than the marker);
// counters[1]: the marker value occurred; counters[2]: a probe run reached
// kMaxProbe slots without an empty one (the set is too small: the caller
// counts again in a set sized by the row count).
constexpr uint32_t kMaxProbe = 4096;
constexpr uint32_t kLdsSetSlots = 4096;  // the workgroup's front set: 32 KiB of LDS
constexpr uint32_t kLdsProbe = 8;        // probe steps in it before going to the device set

template <int DT>
__global__ void __launch_bounds__(256) k_distinct_insert(const uint8_t* __restrict__ col,
                                                         uint64_t row_begin, uint64_t n,
                                                         const uint64_t* __restrict__ sel,
                                                         unsigned long long* __restrict__ table,
                                                         uint64_t mask,
                                                         unsigned long long* counters) {
  // A small set in LDS in front of the device set: a value the workgroup has
  // seen before does not touch the device set (few distinct values otherwise
  // send every row to the same few slots); the lane that puts a value into it
  // inserts the value into the device set.  Once it is half full (many
  // distinct values) the workgroup stops using it.
  __shared__ unsigned long long lset[kLdsSetSlots];
  __shared__ uint32_t lfill;
  for (uint32_t j = threadIdx.x; j < kLdsSetSlots; j += blockDim.x) lset[j] = kSetEmpty;
  if (threadIdx.x == 0) lfill = 0;
  __syncthreads();
  uint32_t inserted = 0, marker = 0, overflow = 0;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  // a run never needs more steps than the set has slots
  const uint32_t max_probe = mask + 1 < kMaxProbe ? (uint32_t)(mask + 1) : kMaxProbe;
  // one row's value: the LDS set, then the device set
  auto add = [&](unsigned long long v) {
    if (v == kSetEmpty) {
      marker = 1;
      return;
    }
    const uint64_t hv = set_slot(v);
    if (lfill < kLdsSetSlots / 2) {
      uint32_t ls = (uint32_t)(hv >> 32) & (kLdsSetSlots - 1);
      for (uint32_t t = 0; t < kLdsProbe; ++t, ls = (ls + 1) & (kLdsSetSlots - 1)) {
        const unsigned long long cur = lset[ls];
        if (cur == v) return;
        if (cur != kSetEmpty) continue;
        const unsigned long long old = atomicCAS(&lset[ls], kSetEmpty, v);
        if (old == kSetEmpty) {
          atomicAdd(&lfill, 1u);
          break;  // this lane inserts v into the device set
        }
        if (old == v) return;
      }
    }
    // a set at most half full ends probe runs at an empty slot within a few
    // steps; a run of max_probe slots means the set is too small and is
    // reported, not continued.  A plain read first spares the atomic for values
    // already present (a stale read only sends the value on to the CAS, which
    // sees the slot as it is).
    uint64_t s = hv & mask;
    uint32_t step = 0;
    for (; step < max_probe; ++step, s = (s + 1) & mask) {
      const unsigned long long seen = __hip_atomic_load(&table[s], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
      if (seen == v) break;
      if (seen != kSetEmpty) continue;
      const unsigned long long old = atomicCAS(&table[s], kSetEmpty, v);
      if (old == kSetEmpty) {
        ++inserted;
        break;
      }
      if (old == v) break;
    }
    overflow |= step == max_probe;
  };
  constexpr int U = 4;
  for (uint64_t i0 = blockIdx.x * 256ull + threadIdx.x; i0 < n; i0 += U * stride) {
    if (overflow) break;  // the count is redone in a larger set: stop early
    // the cells of U rows are loaded before any is added (rows past the end
    // read row i0, which is in range)
    unsigned long long c[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + u * stride;
      const uint64_t r = i < n ? i : i0;
      c[u] = cell_at<DT>(col, row_begin + r);
      ok[u] = i < n && selected(sel, r);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (ok[u]) add(c[u]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    inserted += __shfl_xor(inserted, off);
    marker |= __shfl_xor(marker, off);
    overflow |= __shfl_xor(overflow, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (inserted) atomicAdd(&counters[0], (unsigned long long)inserted);
    if (marker) atomicOr(&counters[1], 1ull);
    if (overflow) atomicOr(&counters[2], 1ull);
  }
}

typedef void (*HllKern)(const uint8_t*, uint64_t, uint64_t, const uint64_t*, uint8_t*);
typedef void (*InsKern)(const uint8_t*, uint64_t, uint64_t, const uint64_t*, unsigned long long*,
                        uint64_t, unsigned long long*);

template <bool SEL>
static HllKern hll_kernel_sel(int32_t dt) {
  switch (dt) {
    case TGP_I8: return k_hll<TGP_I8, SEL>;
    case TGP_U8: return k_hll<TGP_U8, SEL>;
    case TGP_I16: return k_hll<TGP_I16, SEL>;
    case TGP_U16: return k_hll<TGP_U16, SEL>;
    case TGP_I32: return k_hll<TGP_I32, SEL>;
    case TGP_U32: return k_hll<TGP_U32, SEL>;
    case TGP_I64: return k_hll<TGP_I64, SEL>;
    default: return k_hll<TGP_U64, SEL>;
  }
}

static HllKern hll_kernel(int32_t dt, bool sel) {
  return sel ? hll_kernel_sel<true>(dt) : hll_kernel_sel<false>(dt);
}

static InsKern insert_kernel(int32_t dt) {
  switch (dt) {
    case TGP_I8: return k_distinct_insert<TGP_I8>;
    case TGP_U8: return k_distinct_insert<TGP_U8>;
    case TGP_I16: return k_distinct_insert<TGP_I16>;
    case TGP_U16: return k_distinct_insert<TGP_U16>;
    case TGP_I32: return k_distinct_insert<TGP_I32>;
    case TGP_U32: return k_distinct_insert<TGP_U32>;
    case TGP_I64: return k_distinct_insert<TGP_I64>;
    default: return k_distinct_insert<TGP_U64>;
  }
}

static int32_t check_range(const tgp_column* col, uint64_t row_begin, uint64_t row_end,
                           const char* what) {
  TGP_TRY(check_column(col, what));
  TGP_REQUIRE(row_begin <= row_end && row_end <= col->len,
              "%s: rows [%llu, %llu) outside the column's %llu rows", what,
              (unsigned long long)row_begin, (unsigned long long)row_end,
              (unsigned long long)col->len);
  return TGP_OK;
}

// Registers into h_regs (16384 bytes).
static int32_t hll_sketch(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin,
                          uint64_t row_end, const uint64_t* d_sel, uint8_t* h_regs,
                          hipStream_t st) {
  const uint64_t n = row_end - row_begin;
  uint64_t grid = (n + 16 * kHllBlock - 1) / (16 * kHllBlock);
  const uint64_t gmax = (uint64_t)persistent_grid(ctx, 2);  // 64 KiB of LDS per workgroup
  if (grid > gmax) grid = gmax;
  if (grid == 0) grid = 1;
  // partial register sets, then at most 32 merged rows, then the registers
  constexpr uint32_t kPer = 16;
  const uint32_t mid = (uint32_t)((grid + kPer - 1) / kPer);
  void* p = nullptr;
  TGP_TRY(get_scratch(ctx, st, (grid + mid + 1) * kHllRegs, &p));
  uint8_t* partial = static_cast<uint8_t*>(p);
  uint8_t* merged = partial + grid * kHllRegs;
  uint8_t* regs = merged + (uint64_t)mid * kHllRegs;
  hipLaunchKernelGGL(hll_kernel(col->dtype, d_sel != nullptr), dim3((uint32_t)grid), dim3(kHllBlock), 0, st,
                     static_cast<const uint8_t*>(col->data), row_begin, n, d_sel, partial);
  TGP_HIP(hipGetLastError());
  const dim3 cols(kHllRegs / 16 / 256);
  hipLaunchKernelGGL(k_hll_merge, dim3(cols.x, mid), dim3(256), 0, st,
                     reinterpret_cast<const uint4*>(partial), (uint32_t)grid, kPer,
                     reinterpret_cast<uint4*>(merged));
  TGP_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_hll_merge, dim3(cols.x, 1), dim3(256), 0, st,
                     reinterpret_cast<const uint4*>(merged), mid, mid,
                     reinterpret_cast<uint4*>(regs));
  TGP_HIP(hipGetLastError());
  TGP_HIP(hipMemcpyAsync(h_regs, regs, kHllRegs, hipMemcpyDeviceToHost, st));
  TGP_HIP(hipStreamSynchronize(st));
  return TGP_OK;
}

// Inserts the selected cells into a set of `cap` slots; *ok is false when a
// probe run found no empty slot or the set ended more than half full.
static int32_t insert_distinct(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin,
                               uint64_t n, const uint64_t* d_sel, uint64_t cap, uint64_t* out,
                               bool* ok, hipStream_t st) {
  void* p = nullptr;
  TGP_TRY(get_scratch(ctx, st, (cap + 4) * 8, &p));
  auto* counters = static_cast<unsigned long long*>(p);
  auto* table = counters + 4;
  TGP_TRY(fill_u64(st, reinterpret_cast<uint64_t*>(counters), 4, 0));
  TGP_TRY(fill_u64(st, reinterpret_cast<uint64_t*>(table), cap, kSetEmpty));
  uint64_t grid = (n + 1023) / 1024;
  const uint64_t gmax = (uint64_t)persistent_grid(ctx, 8);
  if (grid > gmax) grid = gmax;
  if (grid == 0) grid = 1;
  hipLaunchKernelGGL(insert_kernel(col->dtype), dim3((uint32_t)grid), dim3(256), 0, st,
                     static_cast<const uint8_t*>(col->data), row_begin, n, d_sel, table, cap - 1,
                     counters);
  TGP_HIP(hipGetLastError());
  uint64_t h[3] = {0, 0, 0};
  TGP_HIP(hipMemcpyAsync(h, counters, 24, hipMemcpyDeviceToHost, st));
  TGP_HIP(hipStreamSynchronize(st));
  *out = h[0] + (h[1] ? 1 : 0);
  *ok = h[2] == 0 && 2 * h[0] <= cap;
  return TGP_OK;
}

// Exact COUNT(DISTINCT): from kSizeBySketch rows on, the set is sized at 2.5x
// the HyperLogLog estimate (0.8 % standard error) instead of twice the row
// count, which is smaller to clear and better cached when values repeat (60M
// prices, 10.38M distinct: 2^25 slots instead of 2^27); a set that turns out
// too small is counted again at twice the row count, which always suffices.
constexpr uint64_t kSizeBySketch = 1ull << 20;

static int32_t exact_distinct(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin,
                              uint64_t row_end, const uint64_t* d_sel, uint64_t* out,
                              hipStream_t st) {
  const uint64_t n = row_end - row_begin;
  uint64_t full = 1024;
  while (full < 2 * n) full <<= 1;
  if (n >= kSizeBySketch) {
    uint8_t regs[kHllRegs];
    TGP_TRY(hll_sketch(ctx, col, row_begin, row_end, d_sel, regs, st));
    const uint64_t est = tgp_hll_estimate(regs);
    const uint64_t want = (2 * (est + est / 4 + 1024)) >> ctx->tune_distinct_shrink.load();
    uint64_t cap = 1024;
    while (cap < want && cap < full) cap <<= 1;
    if (cap < full) {
      bool ok = false;
      TGP_TRY(insert_distinct(ctx, col, row_begin, n, d_sel, cap, out, &ok, st));
      if (ok) return TGP_OK;
    }
  }
  bool ok = false;
  TGP_TRY(insert_distinct(ctx, col, row_begin, n, d_sel, full, out, &ok, st));
  TGP_REQUIRE(ok, "tgp_count_distinct: the set of %llu slots overflowed",
              (unsigned long long)full);
  return TGP_OK;
}

}  // namespace tgp

using namespace tgp;

extern "C" {

uint64_t tgp_hll_estimate(const uint8_t* regs) {
  // HyperLogLog::estimate (exec/hll.rs:53-75), summed in register order.  The
  // terms 2^-r are exact doubles, so a table of them gives the same sum as
  // computing each power (16384 pow() calls took 145 us).
  static const auto pow2neg = [] {
    std::array<double, 256> t{};
    for (int r = 0; r < 256; ++r) t[r] = std::ldexp(1.0, -r);
    return t;
  }();
  const double m = (double)kHllRegs;
  const double alpha = 0.7213 / (1.0 + 1.079 / m);
  double sum = 0.0;
  uint64_t zeros = 0;
  for (uint32_t i = 0; i < kHllRegs; ++i) {
    sum += pow2neg[regs[i]];
    zeros += regs[i] == 0;
  }
  const double raw = alpha * m * m / sum;
  const double est = (raw <= 2.5 * m && zeros > 0) ? m * std::log(m / (double)zeros) : raw;
  return (uint64_t)est;
}

int32_t tgp_hll_sketch(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin, uint64_t row_end,
                       const uint64_t* d_sel_bitmap, uint8_t* h_registers, void* stream) {
  TGP_REQUIRE(ctx && h_registers, "tgp_hll_sketch: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_range(col, row_begin, row_end, "tgp_hll_sketch"));
  return hll_sketch(ctx, col, row_begin, row_end, d_sel_bitmap, h_registers,
                    pick_stream(ctx, stream));
}

int32_t tgp_count_distinct(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin,
                           uint64_t row_end, const uint64_t* d_sel_bitmap, int32_t mode,
                           uint64_t* h_count, void* stream) {
  TGP_REQUIRE(ctx && h_count, "tgp_count_distinct: null argument");
  TGP_REQUIRE(mode >= TGP_DISTINCT_AUTO && mode <= TGP_DISTINCT_HLL,
              "tgp_count_distinct: mode %d is not AUTO, EXACT or HLL", mode);
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_range(col, row_begin, row_end, "tgp_count_distinct"));
  hipStream_t st = pick_stream(ctx, stream);
  bool hll = mode == TGP_DISTINCT_HLL;
  if (mode == TGP_DISTINCT_AUTO) {
    uint64_t nsel = row_end - row_begin;
    if (d_sel_bitmap && nsel >= kDistinctExactBelow) {
      void* p = nullptr;
      TGP_TRY(get_scratch(ctx, st, 8, &p));
      TGP_TRY(fill_u64(st, static_cast<uint64_t*>(p), 1, 0));
      TGP_TRY(tgp_bitmap_count_range(ctx, d_sel_bitmap, nsel, 0, nsel,
                                     static_cast<uint64_t*>(p), st));
      TGP_HIP(hipMemcpyAsync(&nsel, p, 8, hipMemcpyDeviceToHost, st));
      TGP_HIP(hipStreamSynchronize(st));
    }
    hll = nsel >= kDistinctExactBelow;
  }
  if (!hll) return exact_distinct(ctx, col, row_begin, row_end, d_sel_bitmap, h_count, st);
  uint8_t regs[kHllRegs];
  TGP_TRY(hll_sketch(ctx, col, row_begin, row_end, d_sel_bitmap, regs, st));
  *h_count = tgp_hll_estimate(regs);
  return TGP_OK;
}

}  // extern "C"
