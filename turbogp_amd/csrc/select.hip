// SELECT columns WHERE <selection> [ORDER BY col [DESC]] [LIMIT k]: the row
// indices of the answer, and a gather of any column at those rows.
//
// Reference path replaced (engine/dispatch.rs):
//   :486-522  SelectStar: selected rows in scan order, first LIMIT
//   :524-618  SelectColumn, :620-733 SelectMulti: selected rows in scan order,
//             stably sorted by the ORDER BY column's u64 cells (`va.cmp(&vb)`,
//             or `vb.cmp(&va)` for DESC), first LIMIT, values gathered as cells
// Selected rows come from the compaction of the selection bitmap
// (tgp_bitmap_to_indices_sync).  ORDER BY sorts (key, row) pairs with a stable
// LSD radix sort (hipCUB): key = the cell for ASC and its complement for DESC,
// so equal keys keep scan order in both directions, as a stable sort_by does.
#include "tgp_internal.h"

#include <hipcub/hipcub.hpp>

namespace tgp {

template <int DT>
__device__ __forceinline__ uint64_t cell_of(const uint8_t* col, uint64_t i) {
  if constexpr (DT == TGP_I8) return (uint64_t)(int64_t)reinterpret_cast<const int8_t*>(col)[i];
  else if constexpr (DT == TGP_U8) return col[i];
  else if constexpr (DT == TGP_I16) return (uint64_t)(int64_t)reinterpret_cast<const int16_t*>(col)[i];
  else if constexpr (DT == TGP_U16) return reinterpret_cast<const uint16_t*>(col)[i];
  else if constexpr (DT == TGP_I32) return (uint64_t)(int64_t)reinterpret_cast<const int32_t*>(col)[i];
  else if constexpr (DT == TGP_U32) return reinterpret_cast<const uint32_t*>(col)[i];
  else return reinterpret_cast<const uint64_t*>(col)[i];
}

__global__ void __launch_bounds__(256) k_iota(uint64_t* __restrict__ out, uint64_t n,
                                              uint64_t base) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) out[i] = base + i;
}

// keys[i] = cell of col at rows[i], complemented for DESC (rows are in range:
// they come from the compaction of a bitmap over the column's rows).
template <int DT>
__global__ void __launch_bounds__(256) k_sort_keys(const uint8_t* __restrict__ col,
                                                   const uint64_t* __restrict__ rows, uint64_t n,
                                                   uint64_t flip, uint64_t* __restrict__ keys) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride)
    keys[i] = cell_of<DT>(col, rows[i]) ^ flip;
}

// out[i] = col[rows[i]] as a u64 cell (CELLS) or at the column's width; a row
// index outside the column writes 0 and raises *bad.
template <int DT, bool CELLS>
__global__ void __launch_bounds__(256) k_gather(const uint8_t* __restrict__ col, uint64_t len,
                                                const uint64_t* __restrict__ rows, uint64_t n,
                                                void* __restrict__ out, unsigned int* bad) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  uint32_t oob = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    const uint64_t r = rows[i];
    const bool ok = r < len;
    oob |= !ok;
    const uint64_t c = ok ? cell_of<DT>(col, r) : 0;
    if constexpr (CELLS) {
      static_cast<uint64_t*>(out)[i] = c;
    } else {
      constexpr int W = (DT == TGP_I8 || DT == TGP_U8) ? 1 : (DT == TGP_I16 || DT == TGP_U16) ? 2
                        : (DT == TGP_I32 || DT == TGP_U32) ? 4 : 8;
      if constexpr (W == 1) static_cast<uint8_t*>(out)[i] = (uint8_t)c;
      else if constexpr (W == 2) static_cast<uint16_t*>(out)[i] = (uint16_t)c;
      else if constexpr (W == 4) static_cast<uint32_t*>(out)[i] = (uint32_t)c;
      else static_cast<uint64_t*>(out)[i] = c;
    }
  }
  if (__ballot(oob) && (threadIdx.x & 63) == 0) atomicOr(bad, 1u);
}

static uint32_t grid_of(uint64_t n) {
  uint64_t g = (n + 1023) / 1024;
  if (g > 8192) g = 8192;
  return (uint32_t)(g ? g : 1);
}

template <int DT>
static void launch_keys(hipStream_t st, const tgp_column* c, const uint64_t* rows, uint64_t n,
                        uint64_t flip, uint64_t* keys) {
  hipLaunchKernelGGL(k_sort_keys<DT>, dim3(grid_of(n)), dim3(256), 0, st,
                     static_cast<const uint8_t*>(c->data), rows, n, flip, keys);
}

template <int DT>
static void launch_gather(hipStream_t st, const tgp_column* c, const uint64_t* rows, uint64_t n,
                          bool cells, void* out, unsigned int* bad) {
  const auto* col = static_cast<const uint8_t*>(c->data);
  if (cells)
    hipLaunchKernelGGL((k_gather<DT, true>), dim3(grid_of(n)), dim3(256), 0, st, col, c->len,
                       rows, n, out, bad);
  else
    hipLaunchKernelGGL((k_gather<DT, false>), dim3(grid_of(n)), dim3(256), 0, st, col, c->len,
                       rows, n, out, bad);
}

#define TGP_BY_DTYPE(dt, F, ...)                 \
  switch (dt) {                                  \
    case TGP_I8: F<TGP_I8>(__VA_ARGS__); break;   \
    case TGP_U8: F<TGP_U8>(__VA_ARGS__); break;   \
    case TGP_I16: F<TGP_I16>(__VA_ARGS__); break; \
    case TGP_U16: F<TGP_U16>(__VA_ARGS__); break; \
    case TGP_I32: F<TGP_I32>(__VA_ARGS__); break; \
    case TGP_U32: F<TGP_U32>(__VA_ARGS__); break; \
    case TGP_I64: F<TGP_I64>(__VA_ARGS__); break; \
    default: F<TGP_U64>(__VA_ARGS__); break;      \
  }

// Device buffers of one call, released on every return path.
struct CallBuffers {
  tgp_ctx* ctx;
  std::vector<void*> p;
  explicit CallBuffers(tgp_ctx* c) : ctx(c) {}
  int32_t get(uint64_t bytes, void** out) {
    TGP_TRY(hbm_alloc(ctx, bytes ? bytes : 8, out));
    p.push_back(*out);
    return TGP_OK;
  }
  ~CallBuffers() {
    for (void* q : p) (void)hbm_free(ctx, q);
  }
};

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_select_rows(tgp_ctx* ctx, uint64_t row_begin, uint64_t row_end,
                        const uint64_t* d_sel_bitmap, const tgp_column* order_col,
                        int32_t descending, uint64_t limit, uint64_t* d_rows, uint64_t capacity,
                        uint64_t* h_count, void* stream) {
  TGP_REQUIRE(ctx && h_count, "tgp_select_rows: null argument");
  TGP_REQUIRE(row_begin <= row_end, "tgp_select_rows: row_begin > row_end");
  TGP_TRY(check_injected(ctx));
  if (order_col) {
    TGP_TRY(check_column(order_col, "tgp_select_rows order column"));
    TGP_REQUIRE(row_end <= order_col->len, "tgp_select_rows: rows up to %llu, order column has %llu",
                (unsigned long long)row_end, (unsigned long long)order_col->len);
  }
  hipStream_t st = pick_stream(ctx, stream);
  *h_count = 0;
  const uint64_t n = row_end - row_begin;
  TGP_REQUIRE(n < (1ull << 31), "tgp_select_rows: %llu rows above the sort's 2^31",
              (unsigned long long)n);
  CallBuffers buf(ctx);
  void* p = nullptr;
  TGP_TRY(buf.get(n * 8, &p));
  uint64_t* rows = static_cast<uint64_t*>(p);
  uint64_t cnt = n;
  if (d_sel_bitmap) {
    if (n) TGP_TRY(tgp_bitmap_to_indices_sync(ctx, d_sel_bitmap, n, row_begin, 8, rows, n, &cnt,
                                              st));
    else cnt = 0;
  } else if (n) {
    hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(256), 0, st, rows, n, row_begin);
    TGP_HIP(hipGetLastError());
  }
  const uint64_t take = limit < cnt ? limit : cnt;
  TGP_REQUIRE(take <= capacity, "tgp_select_rows: %llu rows, capacity %llu",
              (unsigned long long)take, (unsigned long long)capacity);
  if (take == 0) {
    TGP_HIP(hipStreamSynchronize(st));
    return TGP_OK;
  }
  TGP_REQUIRE(d_rows, "tgp_select_rows: null output");
  const uint64_t* first = rows;
  if (order_col && cnt > 1) {
    void* q = nullptr;
    TGP_TRY(buf.get(cnt * 24, &q));
    uint64_t* keys = static_cast<uint64_t*>(q);
    uint64_t* keys_out = keys + cnt;
    uint64_t* rows_out = keys + 2 * cnt;
    TGP_BY_DTYPE(order_col->dtype, launch_keys, st, order_col, rows, cnt,
                 descending ? ~0ull : 0ull, keys);
    TGP_HIP(hipGetLastError());
    size_t temp_bytes = 0;
    TGP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys_out, rows, rows_out,
                                               (int)cnt, 0, 64, st));
    void* temp = nullptr;
    TGP_TRY(buf.get(temp_bytes, &temp));
    TGP_HIP(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_out, rows, rows_out,
                                               (int)cnt, 0, 64, st));
    first = rows_out;
  }
  TGP_HIP(hipMemcpyAsync(d_rows, first, take * 8, hipMemcpyDeviceToDevice, st));
  TGP_HIP(hipStreamSynchronize(st));
  *h_count = take;
  return TGP_OK;
}

int32_t tgp_gather(tgp_ctx* ctx, const tgp_column* col, const uint64_t* d_rows, uint64_t n,
                   int32_t as_cells, void* d_out, void* stream) {
  TGP_REQUIRE(ctx, "tgp_gather: null ctx");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_column(col, "tgp_gather"));
  if (n == 0) return TGP_OK;
  TGP_REQUIRE(d_rows && d_out, "tgp_gather: null rows or output");
  hipStream_t st = pick_stream(ctx, stream);
  void* p = nullptr;
  TGP_TRY(get_scratch(ctx, st, 8, &p));
  auto* bad = static_cast<unsigned int*>(p);
  TGP_TRY(fill_u64(st, static_cast<uint64_t*>(p), 1, 0));
  TGP_BY_DTYPE(col->dtype, launch_gather, st, col, d_rows, n, as_cells != 0, d_out, bad);
  TGP_HIP(hipGetLastError());
  unsigned int h_bad = 0;
  TGP_HIP(hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, st));
  TGP_HIP(hipStreamSynchronize(st));
  TGP_REQUIRE(!h_bad, "tgp_gather: a row index is outside the column's %llu rows",
              (unsigned long long)col->len);
  return TGP_OK;
}

}  // extern "C"
