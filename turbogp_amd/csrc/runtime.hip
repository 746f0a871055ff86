// Context, device binding, HBM tier accounting, streams, copies and errors.
//
// The HBM tier is the device-resident home of columns and hash tables
// (src/memory/tier.rs:11-29 names MemoryTier::Hbm but the reference has no
// allocator behind it).  Every tgp_hbm_alloc is accounted against a byte
// budget so that a serving loop can admit work by footprint.
#include <cstdarg>
#include <vector>

#include "tgp_internal.h"
#include "jit.h"

namespace tgp {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int32_t check_injected(tgp_ctx* ctx) {
  int32_t n = ctx->inject_failures.load();
  while (n > 0) {
    if (ctx->inject_failures.compare_exchange_weak(n, n - 1)) {
      set_error("injected failure");
      return TGP_ERR_INJECTED;
    }
  }
  return TGP_OK;
}

int32_t check_column(const tgp_column* c, const char* what) {
  TGP_REQUIRE(c != nullptr, "%s: null column", what);
  TGP_REQUIRE(dtype_width(c->dtype) > 0, "%s: unknown dtype %d", what, c->dtype);
  TGP_REQUIRE(c->len == 0 || c->data != nullptr, "%s: null data", what);
  TGP_REQUIRE((reinterpret_cast<uintptr_t>(c->data) & 15) == 0,
              "%s: column data must be 16-byte aligned", what);
  return TGP_OK;
}

int32_t get_scratch(tgp_ctx* ctx, hipStream_t stream, size_t bytes, void** out, int slot) {
  TGP_REQUIRE(slot >= 0 && slot < kScratchSlots, "get_scratch: slot %d", slot);
  std::lock_guard<std::mutex> lk(ctx->mu);
  Scratch& s = ctx->scratch[stream];
  if (s.bytes[slot] < bytes) {
    if (s.ptr[slot]) {
      // The previous buffer may still be read by work queued on this stream.
      TGP_HIP(hipStreamSynchronize(stream));
      TGP_HIP(hipFree(s.ptr[slot]));
      s.ptr[slot] = nullptr;
      s.bytes[slot] = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    TGP_HIP(hipMalloc(&s.ptr[slot], want));
    s.bytes[slot] = want;
  }
  *out = s.ptr[slot];
  return TGP_OK;
}

int32_t hbm_alloc(tgp_ctx* ctx, uint64_t bytes, void** out) {
  TGP_REQUIRE(out != nullptr, "hbm_alloc: null out");
  if (bytes == 0) bytes = 256;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->used + bytes > ctx->budget) {
      set_error("HBM tier budget exceeded: used %llu + %llu > budget %llu",
                (unsigned long long)ctx->used, (unsigned long long)bytes,
                (unsigned long long)ctx->budget);
      return TGP_ERR_BUDGET;
    }
    ctx->used += bytes;  // reserve before the (slow) allocation
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (e != hipSuccess) {
    ctx->used -= bytes;
    set_error("hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
    return TGP_ERR_OOM;
  }
  ctx->allocs[p] = bytes;
  if (ctx->used > ctx->peak) ctx->peak = ctx->used;
  *out = p;
  return TGP_OK;
}

int32_t hbm_free(tgp_ctx* ctx, void* p) {
  if (!p) return TGP_OK;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->allocs.find(p);
    TGP_REQUIRE(it != ctx->allocs.end(), "hbm_free: pointer not owned by this context");
    bytes = it->second;
    ctx->allocs.erase(it);
    ctx->used -= bytes;
  }
  TGP_HIP(hipFree(p));
  return TGP_OK;
}

__global__ void __launch_bounds__(256) k_fill_u64(uint64_t* __restrict__ dst, uint64_t n,
                                                  uint64_t value) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = value;
}

int32_t fill_u64(hipStream_t s, uint64_t* dst, uint64_t n, uint64_t value) {
  if (n == 0) return TGP_OK;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_fill_u64<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(dst, n, value);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

struct ZeroRanges {
  uint64_t* p[kMaxZeroRanges];
  uint64_t n[kMaxZeroRanges];
};

__global__ void __launch_bounds__(256) k_zero_ranges(ZeroRanges r) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
#pragma unroll
  for (int k = 0; k < kMaxZeroRanges; ++k)
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < r.n[k]; i += stride)
      r.p[k][i] = 0;
}

int32_t zero_ranges(hipStream_t s, int count, uint64_t* const* ptrs, const uint64_t* n) {
  TGP_REQUIRE(count >= 0 && count <= kMaxZeroRanges, "zero_ranges: %d ranges", count);
  ZeroRanges r{};
  uint64_t most = 0;
  for (int k = 0; k < count; ++k) {
    if (!ptrs[k] || !n[k]) continue;
    r.p[k] = ptrs[k];
    r.n[k] = n[k];
    most = n[k] > most ? n[k] : most;
  }
  if (most == 0) return TGP_OK;
  uint64_t blocks = (most + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_zero_ranges<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(r);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int persistent_grid(tgp_ctx* ctx, int wg_per_cu) {
  int cus = ctx->info.cu_count > 0 ? ctx->info.cu_count : 256;
  return cus * wg_per_cu;
}

}  // namespace tgp

using namespace tgp;

extern "C" {

const char* tgp_last_error(void) { return g_last_error.c_str(); }
int32_t tgp_abi_version(void) { return TGP_ABI_VERSION; }

int32_t tgp_ctx_create(int32_t device, uint64_t hbm_budget_bytes, tgp_ctx** out) {
  TGP_REQUIRE(out != nullptr, "tgp_ctx_create: null out");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    set_error("no HIP device available (%s)", hipGetErrorString(e));
    return TGP_ERR_NO_DEVICE;
  }
  TGP_REQUIRE(device >= 0 && device < ndev, "device %d out of range [0,%d)", device, ndev);
  TGP_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  TGP_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %d is %s; this library is built for gfx950 only", device,
              prop.gcnArchName);
    return TGP_ERR_NO_DEVICE;
  }
  tgp_ctx* ctx = new tgp_ctx();
  ctx->device = device;
  std::snprintf(ctx->info.name, sizeof(ctx->info.name), "%s", prop.name);
  std::snprintf(ctx->info.arch, sizeof(ctx->info.arch), "%s", prop.gcnArchName);
  ctx->info.cu_count = prop.multiProcessorCount;
  ctx->info.wave_size = prop.warpSize;
  ctx->info.hbm_bytes = (int64_t)prop.totalGlobalMem;
  ctx->info.lds_bytes_per_cu = (int64_t)prop.maxSharedMemoryPerMultiProcessor;
  ctx->info.l2_bytes = prop.l2CacheSize;
  ctx->info.clock_khz = prop.clockRate;
  ctx->info.mem_clock_khz = prop.memoryClockRate;
  ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(0.9 * (double)prop.totalGlobalMem);
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
    delete ctx;
    return TGP_ERR_HIP;
  }
  // the device address of the no-allocation compaction's fixed scratch (once per device)
  const int32_t rc = compact_noalloc_init(device);
  if (rc != TGP_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return TGP_OK;
}

int32_t tgp_ctx_destroy(tgp_ctx* ctx) {
  if (!ctx) return TGP_OK;
  // run-time compiles still in flight finish while every library is alive
  jit_drain();
  (void)hipSetDevice(ctx->device);
  // Streams the caller created may already be destroyed: wait for the device,
  // not for the streams the scratch map remembers.
  (void)hipDeviceSynchronize();
  for (auto& kv : ctx->scratch)
    for (void* p : kv.second.ptr)
      if (p) (void)hipFree(p);
  for (auto& kv : ctx->allocs) (void)hipFree(kv.first);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TGP_OK;
}

int32_t tgp_device_info_get(tgp_ctx* ctx, tgp_device_info* out) {
  TGP_REQUIRE(ctx && out, "tgp_device_info_get: null argument");
  *out = ctx->info;
  return TGP_OK;
}

int32_t tgp_stream_create(tgp_ctx* ctx, void** out_stream) {
  TGP_REQUIRE(ctx && out_stream, "tgp_stream_create: null argument");
  TGP_HIP(hipSetDevice(ctx->device));
  hipStream_t s;
  // Default priority only: streams of mixed priorities are avoided on purpose.
  TGP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out_stream = s;
  return TGP_OK;
}

int32_t tgp_stream_destroy(tgp_ctx* ctx, void* stream) {
  TGP_REQUIRE(ctx && stream, "tgp_stream_destroy: null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TGP_HIP(hipStreamSynchronize(s));
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->scratch.find(s);
    if (it != ctx->scratch.end()) {
      for (void* p : it->second.ptr)
        if (p) (void)hipFree(p);
      ctx->scratch.erase(it);
    }
  }
  TGP_HIP(hipStreamDestroy(s));
  return TGP_OK;
}

int32_t tgp_stream_sync(tgp_ctx* ctx, void* stream) {
  TGP_REQUIRE(ctx, "tgp_stream_sync: null ctx");
  TGP_HIP(hipStreamSynchronize(pick_stream(ctx, stream)));
  return TGP_OK;
}

int32_t tgp_hbm_alloc(tgp_ctx* ctx, uint64_t bytes, void** d_out) {
  TGP_REQUIRE(ctx, "tgp_hbm_alloc: null ctx");
  TGP_TRY(check_injected(ctx));
  return hbm_alloc(ctx, bytes, d_out);
}

int32_t tgp_hbm_free(tgp_ctx* ctx, void* d_ptr) {
  TGP_REQUIRE(ctx, "tgp_hbm_free: null ctx");
  return hbm_free(ctx, d_ptr);
}

int32_t tgp_hbm_usage(tgp_ctx* ctx, uint64_t* used, uint64_t* budget, uint64_t* peak) {
  TGP_REQUIRE(ctx, "tgp_hbm_usage: null ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (used) *used = ctx->used;
  if (budget) *budget = ctx->budget;
  if (peak) *peak = ctx->peak;
  return TGP_OK;
}

int32_t tgp_hbm_set_budget(tgp_ctx* ctx, uint64_t budget_bytes) {
  TGP_REQUIRE(ctx, "tgp_hbm_set_budget: null ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->budget = budget_bytes;
  return TGP_OK;
}

int32_t tgp_memcpy_h2d(tgp_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes,
                       void* stream) {
  TGP_REQUIRE(ctx && (bytes == 0 || (d_dst && h_src)), "tgp_memcpy_h2d: null argument");
  if (bytes == 0) return TGP_OK;
  hipStream_t s = pick_stream(ctx, stream);
  TGP_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_memcpy_d2h(tgp_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes,
                       void* stream) {
  TGP_REQUIRE(ctx && (bytes == 0 || (h_dst && d_src)), "tgp_memcpy_d2h: null argument");
  if (bytes == 0) return TGP_OK;
  hipStream_t s = pick_stream(ctx, stream);
  TGP_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_memcpy_d2d(tgp_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes,
                       void* stream) {
  TGP_REQUIRE(ctx && (bytes == 0 || (d_dst && d_src)), "tgp_memcpy_d2d: null argument");
  if (bytes == 0) return TGP_OK;
  TGP_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, pick_stream(ctx, stream)));
  return TGP_OK;
}

int32_t tgp_fill_u64(tgp_ctx* ctx, uint64_t* d_dst, uint64_t nwords, uint64_t value,
                     void* stream) {
  TGP_REQUIRE(ctx && (nwords == 0 || d_dst), "tgp_fill_u64: null argument");
  return fill_u64(pick_stream(ctx, stream), d_dst, nwords, value);
}

int32_t tgp_tune_set(tgp_ctx* ctx, int32_t key, int64_t value) {
  TGP_REQUIRE(ctx && value >= 0, "tgp_tune_set: bad argument");
  switch (key) {
    case TGP_TUNE_WG_PER_CU: ctx->tune_wg_per_cu.store(value); return TGP_OK;
    case TGP_TUNE_GRID_MAX: ctx->tune_grid_max.store(value); return TGP_OK;
    case TGP_TUNE_SPECIALIZE: ctx->tune_no_spec.store(value ? 0 : 1); return TGP_OK;
    case TGP_TUNE_SPEC_KERNEL:
      TGP_REQUIRE(value <= 2, "TGP_TUNE_SPEC_KERNEL is 0, 1 or 2");
      ctx->tune_spec_kernel.store(value);
      return TGP_OK;
    case TGP_TUNE_CHUNK:
      TGP_REQUIRE(value <= 4096, "TGP_TUNE_CHUNK above 4096 slices");
      ctx->tune_chunk.store(value);
      return TGP_OK;
    case TGP_TUNE_SLICES_PER_WAVE:
      TGP_REQUIRE(value <= 65536, "TGP_TUNE_SLICES_PER_WAVE above 65536");
      ctx->tune_slices_per_wave.store(value);
      return TGP_OK;
    case TGP_TUNE_JOIN_ATOMICS:
      TGP_REQUIRE(value <= 1, "TGP_TUNE_JOIN_ATOMICS is 0 or 1");
      ctx->tune_join_partition.store(value);
      return TGP_OK;
    case TGP_TUNE_JOIN_PACK_BITS:
      TGP_REQUIRE(value <= 63, "TGP_TUNE_JOIN_PACK_BITS is 0..63");
      ctx->tune_join_pack_bits.store(value);
      return TGP_OK;
    case TGP_TUNE_DISTINCT_SHRINK:
      TGP_REQUIRE(value <= 30, "TGP_TUNE_DISTINCT_SHRINK is 0..30");
      ctx->tune_distinct_shrink.store(value);
      return TGP_OK;
    case TGP_TUNE_GROUP_TAG_BITS:
      TGP_REQUIRE(value <= 32, "TGP_TUNE_GROUP_TAG_BITS is 0..32");
      ctx->tune_group_tag_bits.store(value);
      return TGP_OK;
    case TGP_TUNE_JOIN_HOT:
      TGP_REQUIRE(value <= 2, "TGP_TUNE_JOIN_HOT is 0, 1 or 2");
      ctx->tune_join_hot.store(value);
      return TGP_OK;
    case TGP_TUNE_JIT:
      TGP_REQUIRE(value <= 2, "TGP_TUNE_JIT is 0, 1 or 2");
      ctx->tune_jit.store(value);
      return TGP_OK;
    case TGP_TUNE_AGG_PART:
      TGP_REQUIRE(value <= 2, "TGP_TUNE_AGG_PART is 0, 1 or 2");
      ctx->tune_agg_part.store(value);
      return TGP_OK;
  }
  set_error("tgp_tune_set: unknown key %d", key);
  return TGP_ERR_INVALID;
}

int32_t tgp_tune_get(tgp_ctx* ctx, int32_t key, int64_t* value) {
  TGP_REQUIRE(ctx && value, "tgp_tune_get: null argument");
  switch (key) {
    case TGP_TUNE_WG_PER_CU: *value = ctx->tune_wg_per_cu.load(); return TGP_OK;
    case TGP_TUNE_GRID_MAX: *value = ctx->tune_grid_max.load(); return TGP_OK;
    case TGP_TUNE_SPECIALIZE: *value = ctx->tune_no_spec.load() ? 0 : 1; return TGP_OK;
    case TGP_TUNE_SPEC_KERNEL: *value = ctx->tune_spec_kernel.load(); return TGP_OK;
    case TGP_TUNE_CHUNK: *value = ctx->tune_chunk.load(); return TGP_OK;
    case TGP_TUNE_SLICES_PER_WAVE: *value = ctx->tune_slices_per_wave.load(); return TGP_OK;
    case TGP_TUNE_JOIN_ATOMICS: *value = ctx->tune_join_partition.load(); return TGP_OK;
    case TGP_TUNE_JOIN_PACK_BITS: *value = ctx->tune_join_pack_bits.load(); return TGP_OK;
    case TGP_TUNE_DISTINCT_SHRINK: *value = ctx->tune_distinct_shrink.load(); return TGP_OK;
    case TGP_TUNE_GROUP_TAG_BITS: *value = ctx->tune_group_tag_bits.load(); return TGP_OK;
    case TGP_TUNE_JOIN_HOT: *value = ctx->tune_join_hot.load(); return TGP_OK;
    case TGP_TUNE_JIT: *value = ctx->tune_jit.load(); return TGP_OK;
    case TGP_TUNE_AGG_PART: *value = ctx->tune_agg_part.load(); return TGP_OK;
  }
  set_error("tgp_tune_get: unknown key %d", key);
  return TGP_ERR_INVALID;
}

int32_t tgp_event_create(tgp_ctx* ctx, void** out_event) {
  TGP_REQUIRE(ctx && out_event, "tgp_event_create: null argument");
  hipEvent_t e;
  TGP_HIP(hipEventCreate(&e));
  *out_event = e;
  return TGP_OK;
}

int32_t tgp_event_destroy(tgp_ctx* ctx, void* event) {
  TGP_REQUIRE(ctx && event, "tgp_event_destroy: null argument");
  TGP_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return TGP_OK;
}

int32_t tgp_event_record(tgp_ctx* ctx, void* event, void* stream) {
  TGP_REQUIRE(ctx && event, "tgp_event_record: null argument");
  TGP_HIP(hipEventRecord(static_cast<hipEvent_t>(event), pick_stream(ctx, stream)));
  return TGP_OK;
}

int32_t tgp_event_elapsed_ms(tgp_ctx* ctx, void* start, void* end, float* out_ms) {
  TGP_REQUIRE(ctx && start && end && out_ms, "tgp_event_elapsed_ms: null argument");
  TGP_HIP(hipEventSynchronize(static_cast<hipEvent_t>(end)));
  TGP_HIP(hipEventElapsedTime(out_ms, static_cast<hipEvent_t>(start),
                              static_cast<hipEvent_t>(end)));
  return TGP_OK;
}

int32_t tgp_debug_inject_failure(tgp_ctx* ctx, int32_t count) {
  TGP_REQUIRE(ctx && count >= 0, "tgp_debug_inject_failure: bad argument");
  ctx->inject_failures.store(count);
  return TGP_OK;
}

}  // extern "C"
