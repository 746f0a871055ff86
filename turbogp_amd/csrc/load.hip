// Loading the reference's column cells into native-width device columns.
//
// The reference keeps every column as u64 cells (datasource/table.rs:25-30;
// csv.rs:144-161 widens i64 with `v as u64`, and the i32 sidecar is never
// filled, csv.rs:145-149), and a NullBitmap of one bool per row
// (types/null_bitmap.rs).  Here the cells are uploaded in chunks and narrowed
// on the device to the column's declared type (so 8 bytes per row cross PCIe
// once and the resident column has its native width), with every value checked
// to fit; the NULL flags become a device bitmap (bit i = row i is NULL, words
// LSB first as the selection bitmaps).
#include "tgp_internal.h"

namespace tgp {

constexpr uint64_t kLoadChunk = 1ull << 24;  // cells per staged chunk (128 MiB)

// Narrows cells[0, n) into out (first row `row0`); a cell outside the type's
// range lowers *first_bad to its row.
template <int DT>
__global__ void __launch_bounds__(256) k_narrow(const uint64_t* __restrict__ cells, uint64_t n,
                                                uint64_t row0, void* __restrict__ out,
                                                unsigned long long* first_bad) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    const uint64_t c = cells[i];
    const int64_t s = (int64_t)c;
    bool ok;
    if constexpr (DT == TGP_I8) ok = s >= -128 && s <= 127;
    else if constexpr (DT == TGP_U8) ok = c <= 0xffull;
    else if constexpr (DT == TGP_I16) ok = s >= -32768 && s <= 32767;
    else if constexpr (DT == TGP_U16) ok = c <= 0xffffull;
    else if constexpr (DT == TGP_I32) ok = s >= -2147483648ll && s <= 2147483647ll;
    else if constexpr (DT == TGP_U32) ok = c <= 0xffffffffull;
    else ok = true;
    if (!ok) atomicMin(first_bad, (unsigned long long)(row0 + i));
    const uint64_t r = row0 + i;
    if constexpr (DT == TGP_I8 || DT == TGP_U8) static_cast<uint8_t*>(out)[r] = (uint8_t)c;
    else if constexpr (DT == TGP_I16 || DT == TGP_U16) static_cast<uint16_t*>(out)[r] = (uint16_t)c;
    else if constexpr (DT == TGP_I32 || DT == TGP_U32) static_cast<uint32_t*>(out)[r] = (uint32_t)c;
    else static_cast<uint64_t*>(out)[r] = c;
  }
}

template <int DT>
static void launch_narrow(hipStream_t st, const uint64_t* cells, uint64_t n, uint64_t row0,
                          void* out, unsigned long long* bad) {
  uint64_t g = (n + 1023) / 1024;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_narrow<DT>, dim3((uint32_t)(g ? g : 1)), dim3(256), 0, st, cells, n, row0,
                     out, bad);
}

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_column_from_cells(tgp_ctx* ctx, const uint64_t* h_cells, uint64_t n, int32_t dtype,
                              void* d_out, uint64_t* h_first_bad, void* stream) {
  TGP_REQUIRE(ctx && (n == 0 || (h_cells && d_out)), "tgp_column_from_cells: null argument");
  TGP_REQUIRE(dtype_width(dtype) > 0, "tgp_column_from_cells: unknown dtype %d", dtype);
  TGP_TRY(check_injected(ctx));
  if (h_first_bad) *h_first_bad = ~0ull;
  if (n == 0) return TGP_OK;
  hipStream_t st = pick_stream(ctx, stream);
  const uint64_t chunk = n < kLoadChunk ? n : kLoadChunk;
  void* p = nullptr;
  TGP_TRY(get_scratch(ctx, st, 256 + chunk * 8, &p));
  auto* bad = static_cast<unsigned long long*>(p);
  auto* stage = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(p) + 256);
  TGP_TRY(fill_u64(st, reinterpret_cast<uint64_t*>(bad), 1, ~0ull));
  for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
    const uint64_t m = n - r0 < chunk ? n - r0 : chunk;
    // pageable source: the copy has returned once the staging buffer holds the chunk
    TGP_HIP(hipMemcpyAsync(stage, h_cells + r0, m * 8, hipMemcpyHostToDevice, st));
    switch (dtype) {
      case TGP_I8: launch_narrow<TGP_I8>(st, stage, m, r0, d_out, bad); break;
      case TGP_U8: launch_narrow<TGP_U8>(st, stage, m, r0, d_out, bad); break;
      case TGP_I16: launch_narrow<TGP_I16>(st, stage, m, r0, d_out, bad); break;
      case TGP_U16: launch_narrow<TGP_U16>(st, stage, m, r0, d_out, bad); break;
      case TGP_I32: launch_narrow<TGP_I32>(st, stage, m, r0, d_out, bad); break;
      case TGP_U32: launch_narrow<TGP_U32>(st, stage, m, r0, d_out, bad); break;
      case TGP_I64: launch_narrow<TGP_I64>(st, stage, m, r0, d_out, bad); break;
      default: launch_narrow<TGP_U64>(st, stage, m, r0, d_out, bad); break;
    }
    TGP_HIP(hipGetLastError());
    // the next chunk overwrites the staging buffer only after this one is narrowed
    TGP_HIP(hipStreamSynchronize(st));
  }
  uint64_t fb = ~0ull;
  TGP_HIP(hipMemcpyAsync(&fb, bad, 8, hipMemcpyDeviceToHost, st));
  TGP_HIP(hipStreamSynchronize(st));
  if (h_first_bad) *h_first_bad = fb;
  TGP_REQUIRE(fb == ~0ull, "tgp_column_from_cells: cell of row %llu does not fit dtype %d",
              (unsigned long long)fb, dtype);
  return TGP_OK;
}

int32_t tgp_null_bitmap_from_flags(tgp_ctx* ctx, const uint8_t* h_flags, uint64_t n,
                                   uint64_t* d_bitmap, void* stream) {
  TGP_REQUIRE(ctx && (n == 0 || (h_flags && d_bitmap)), "tgp_null_bitmap_from_flags: null argument");
  TGP_TRY(check_injected(ctx));
  if (n == 0) return TGP_OK;
  const uint64_t words = (n + 63) / 64;
  std::vector<uint64_t> w(words, 0);
  for (uint64_t i = 0; i < n; ++i)
    if (h_flags[i]) w[i >> 6] |= 1ull << (i & 63);
  hipStream_t st = pick_stream(ctx, stream);
  TGP_HIP(hipMemcpyAsync(d_bitmap, w.data(), words * 8, hipMemcpyHostToDevice, st));
  TGP_HIP(hipStreamSynchronize(st));
  return TGP_OK;
}

}  // extern "C"
