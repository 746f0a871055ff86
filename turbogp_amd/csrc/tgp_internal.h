// Internal definitions shared by the operator library's translation units.
#pragma once

// The device parts of the library's headers also compile under hipRTC
// (__HIPCC_RTC__: the run-time compiled kernels of jit.hip); the host runtime
// declarations are left out there.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#else
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::size_t size_t;
#endif

#include "../../include/turbogp_gfx950.h"

namespace tgp {

// ---- errors ------------------------------------------------------------------
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define TGP_HIP(call)                                                            \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::tgp::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                       __FILE__, __LINE__);                                      \
      return e_ == hipErrorOutOfMemory ? TGP_ERR_OOM : TGP_ERR_HIP;              \
    }                                                                            \
  } while (0)

#define TGP_REQUIRE(cond, ...)                                                   \
  do {                                                                           \
    if (!(cond)) {                                                               \
      ::tgp::set_error(__VA_ARGS__);                                             \
      return TGP_ERR_INVALID;                                                    \
    }                                                                            \
  } while (0)

#define TGP_TRY(expr)                 \
  do {                                \
    int32_t rc_ = (expr);             \
    if (rc_ != TGP_OK) return rc_;    \
  } while (0)

// Calls f(IntC<I>{}) for I = 0 .. N-1 (compile-time loop: arrays indexed by
// decltype(I)::value stay in registers).
template <int V>
struct IntC {
  static constexpr int value = V;
};
template <int I, int N, typename F>
__device__ __forceinline__ void static_for_impl(F& f) {
  if constexpr (I < N) {
    f(IntC<I>{});
    static_for_impl<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl<0, N>(f);
}

// ---- launch geometry -----------------------------------------------------------
constexpr int kWave = 64;
constexpr int kBlock = 256;            // threads per workgroup for streaming kernels
constexpr int kTileRows = 1024;        // rows staged per tile in the fused kernels

// This thread's wave within its workgroup, as a value the compiler holds in a
// scalar register: loops and branches on it (a wave's slice, its rows' range
// tests) then stay scalar instead of running under exec masks.
__device__ __forceinline__ uint32_t wave_index() {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

}  // namespace tgp

#ifndef __HIPCC_RTC__
namespace tgp {

inline int dtype_width(int32_t dt) {
  switch (dt) {
    case TGP_I8: case TGP_U8: return 1;
    case TGP_I16: case TGP_U16: return 2;
    case TGP_I32: case TGP_U32: return 4;
    case TGP_I64: case TGP_U64: return 8;
    default: return 0;
  }
}
inline bool dtype_signed(int32_t dt) {
  return dt == TGP_I8 || dt == TGP_I16 || dt == TGP_I32 || dt == TGP_I64;
}

// ---- scratch arena: device memory reused by calls on one stream ----------------
// Two regions per stream: slot 0 for a call's own work, slot 1 for calls that
// run other library calls (which use slot 0) while holding buffers of their own.
constexpr int kScratchSlots = 2;
struct Scratch {
  void* ptr[kScratchSlots] = {nullptr, nullptr};
  size_t bytes[kScratchSlots] = {0, 0};
};

}  // namespace tgp

struct tgp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  tgp_device_info info{};
  // HBM tier accounting (memory/tier.rs MemoryTier::Hbm).
  std::mutex mu;
  std::map<void*, uint64_t> allocs;
  uint64_t used = 0, peak = 0, budget = 0;
  // Per-stream scratch (stream-ordered reuse).
  std::map<hipStream_t, tgp::Scratch> scratch;
  std::atomic<int32_t> inject_failures{0};
  std::atomic<int64_t> tune_wg_per_cu{0};
  std::atomic<int64_t> tune_grid_max{0};
  std::atomic<int64_t> tune_no_spec{0};
  std::atomic<int64_t> tune_spec_kernel{0};  // 0: by group count; 1: register fold; 2: replicated
  std::atomic<int64_t> tune_chunk{0};        // 0: static interleave; n: chunks of n slices
  std::atomic<int64_t> tune_slices_per_wave{0};  // fewest slices per wave (0 = 2)
  std::atomic<int64_t> tune_join_partition{0};   // 0: radix-partitioned join GROUP BY; 1: off
  std::atomic<int64_t> tune_join_pack_bits{0};   // multi-column join keys: 0 = 64-bit packing
  std::atomic<int64_t> tune_distinct_shrink{0};  // exact distinct: sketch-sized set / 2^n
  std::atomic<int64_t> tune_group_tag_bits{0};   // hash GROUP BY: 0 = auto; n = verified form, n-bit tags
  std::atomic<int64_t> tune_join_hot{0};         // join hot-key cache: 0 auto, 1 off, 2 on
  std::atomic<int64_t> tune_jit{0};              // run-time compiled shapes: 0 auto, 1 off, 2 on
  std::atomic<int64_t> tune_agg_part{0};         // partitioned wide GROUP BY: 0 auto, 1 off, 2 on
};

namespace tgp {
// Returns the stream to use for a call.
inline hipStream_t pick_stream(tgp_ctx* ctx, void* s) {
  return s ? static_cast<hipStream_t>(s) : ctx->stream;
}
// Consumes one injected failure, if any (host-side failure tests).
int32_t check_injected(tgp_ctx* ctx);
// Device scratch of at least `bytes` for `stream` (grown on demand; not accounted
// against the HBM budget, bounded by the largest call).  `slot` 1 is the outer
// region of composite calls.
int32_t get_scratch(tgp_ctx* ctx, hipStream_t stream, size_t bytes, void** out, int slot = 0);
// Compaction of a bitmap to ascending row indices in two steps on one stream:
// compact_begin counts the set bits (synchronous; the block offsets stay in the
// stream's slot-0 scratch) and compact_finish writes base + i for every set bit
// i (asynchronous).  Nothing may use the stream's slot-0 scratch in between.
int32_t compact_begin(tgp_ctx* ctx, hipStream_t s, const uint64_t* d_bitmap, uint64_t nbits,
                      uint64_t* total);
int32_t compact_finish(tgp_ctx* ctx, hipStream_t s, const uint64_t* d_bitmap, uint64_t nbits,
                       uint64_t base, int32_t index_bytes, void* d_out);
// Sets up, once per device (the device current), the device address of the fixed
// scratch that tgp_bitmap_to_indices_noalloc_sync uses.
int32_t compact_noalloc_init(int32_t device);
// Tier-accounted allocation used by library objects (hash tables).
int32_t hbm_alloc(tgp_ctx* ctx, uint64_t bytes, void** out);
int32_t hbm_free(tgp_ctx* ctx, void* p);
// Device fill helper (kernel based, graph safe).
int32_t fill_u64(hipStream_t s, uint64_t* dst, uint64_t n, uint64_t value);
// Zeroes up to kMaxZeroRanges word ranges with one kernel launch (graph safe);
// null or empty ranges are skipped.
constexpr int kMaxZeroRanges = 4;
int32_t zero_ranges(hipStream_t s, int count, uint64_t* const* ptrs, const uint64_t* n);
// Number of workgroups that fill the device for a persistent streaming kernel.
int persistent_grid(tgp_ctx* ctx, int wg_per_cu);
// Correctly rounded f64 sum of n cells (f64 bits) under an optional bitmap
// (synchronous; *out on return).
int32_t sum_f64_exact(tgp_ctx* ctx, const uint64_t* d_cells, uint64_t n,
                      const uint64_t* d_sel_bitmap, bool neg_zero_start, hipStream_t s,
                      double* out);
// Validates a column descriptor (non-null data, known dtype, 16-byte alignment).
int32_t check_column(const tgp_column* c, const char* what);
// Exclusive prefix sum of n words in place, the total stored at v[n]; bsum is
// scratch of ceil(n / 4096) + 1 words.
int32_t scan_u64(hipStream_t s, uint64_t* v, uint64_t n, uint64_t* bsum);
}  // namespace tgp
#endif  // __HIPCC_RTC__
