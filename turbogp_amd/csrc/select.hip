// SELECT columns WHERE <selection> [ORDER BY col [DESC]] [LIMIT k]: the row
// indices of the answer, and a gather of any column at those rows.
//
// Reference path replaced (engine/dispatch.rs):
//   :486-522  SelectStar: selected rows in scan order, first LIMIT
//   :524-618  SelectColumn, :620-733 SelectMulti: selected rows in scan order,
//             stably sorted by the ORDER BY column's u64 cells (`va.cmp(&vb)`,
//             or `vb.cmp(&va)` for DESC), first LIMIT, values gathered as cells
// Selected rows come from the compaction of the selection bitmap.  The order
// is that of (key, position): key = the cell for ASC and its complement for
// DESC, position = the row's place in scan order, so equal keys keep scan order
// in both directions, as a stable sort_by does.  ORDER BY without a small LIMIT
// sorts (key, row) pairs with a stable LSD radix sort (hipCUB).  With LIMIT k
// <= kTopKMax the k first rows are found by a radix select over the 96-bit
// (key, position) value (9 passes of 11-bit digits: a histogram of the rows
// that match the prefix chosen so far, then one workgroup picks the digit that
// holds the k-th row), collected, and sorted in LDS by one workgroup.  The
// positions make every value distinct, so exactly k rows are <= the k-th.
// Buffers live in the stream's outer scratch region: no allocation per call.
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

// keys[i] = cell of col at rows[i] (at base + i without a row list),
// complemented for DESC (rows are in range: they come from the compaction of a
// bitmap over the column's rows).  With `minmax`, the smallest and largest key
// are folded into minmax[0] and minmax[1].
template <int DT>
__global__ void __launch_bounds__(256) k_sort_keys(const uint8_t* __restrict__ col,
                                                   const uint64_t* __restrict__ rows,
                                                   uint64_t base, uint64_t n, uint64_t flip,
                                                   uint64_t* __restrict__ keys,
                                                   unsigned long long* minmax) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  unsigned long long lo = ~0ull, hi = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    const uint64_t k = cell_of<DT>(col, rows ? rows[i] : base + i) ^ flip;
    keys[i] = k;
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  if (!minmax) return;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  // every wave on the same two words: an atomic only where this wave improves
  // on the value read (the minimum only falls and the maximum only rises, so a
  // stale read never skips a needed update)
  if ((threadIdx.x & 63) == 0) {
    if (lo < __atomic_load_n(&minmax[0], __ATOMIC_RELAXED)) atomicMin(&minmax[0], lo);
    if (hi > __atomic_load_n(&minmax[1], __ATOMIC_RELAXED)) atomicMax(&minmax[1], hi);
  }
}

// out[i] = col[rows[i]] as a u64 cell (CELLS) or at the column's width; the
// row index ~0 (no row: an outer join's missing side) writes 0, and any other
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
    oob |= !ok && r != ~0ull;
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
static void launch_keys(hipStream_t st, const tgp_column* c, const uint64_t* rows, uint64_t base,
                        uint64_t n, uint64_t flip, uint64_t* keys, unsigned long long* minmax) {
  hipLaunchKernelGGL(k_sort_keys<DT>, dim3(grid_of(n)), dim3(256), 0, st,
                     static_cast<const uint8_t*>(c->data), rows, base, n, flip, keys, minmax);
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

// ---- top-k: radix select over (key, position) ---------------------------------------
constexpr uint64_t kTopKMax = 4096;  // LIMIT up to this: select + one-workgroup sort
constexpr int kRsBits = 11;
constexpr int kRsBins = 1 << kRsBits;
constexpr int kRsPasses = 9;  // 6 over the 64 key bits, 3 over the 32 position bits

struct RsState {
  unsigned long long prefk, maskk;  // key bits chosen so far
  uint32_t prefp, maskp;            // position bits chosen so far
  unsigned long long remaining;     // rank of the k-th row among those matching the prefix
  unsigned long long taken;         // collect counter
  uint32_t done;                    // every row matching the prefix is among the first k
  uint32_t _pad;
  unsigned long long kmin, kmax;    // smallest and largest key (k_sort_keys)
  uint32_t hist[kRsBins];
};

// A key pass whose digit lies in the leading bits every key shares: its digit
// is that of the smallest key, with no histogram needed.
__device__ __forceinline__ bool rs_shared_digit(const RsState* st, bool on_key, int shift) {
  const unsigned long long diff = st->kmin ^ st->kmax;
  const int common = diff ? __clzll((long long)diff) : 64;
  return on_key && shift >= 64 - common;
}

// Digit (shift, width) of pass p: key bits 63..53, 52..42, 41..31, 30..20,
// 19..9, 8..0, then position bits 31..21, 20..10, 9..0.
__device__ __forceinline__ void rs_digit(int p, int* shift, int* width, bool* on_key) {
  *on_key = p < 6;
  if (p < 6) {
    *shift = p < 5 ? 53 - kRsBits * p : 0;
    *width = p < 5 ? kRsBits : 9;
  } else {
    const int q = p - 6;
    *shift = q < 2 ? 21 - kRsBits * q : 0;
    *width = q < 2 ? kRsBits : 10;
  }
}

__global__ void __launch_bounds__(256) k_rs_init(RsState* st, uint64_t k) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->prefk = st->maskk = 0;
    st->prefp = st->maskp = 0;
    st->remaining = k;
    st->taken = 0;
    st->done = 0;
    st->kmin = ~0ull;
    st->kmax = 0;
  }
  for (int b = threadIdx.x; b < kRsBins; b += blockDim.x) st->hist[b] = 0;
}

__global__ void __launch_bounds__(256) k_rs_hist(const uint64_t* __restrict__ keys, uint64_t n,
                                                 RsState* st, int pass) {
  if (st->done) return;  // decided by an earlier pass
  int shift, width;
  bool on_key;
  rs_digit(pass, &shift, &width, &on_key);
  if (rs_shared_digit(st, on_key, shift)) return;
  __shared__ uint32_t h[kRsBins];
  for (int b = threadIdx.x; b < kRsBins; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const unsigned long long prefk = st->prefk, maskk = st->maskk;
  const uint32_t prefp = st->prefp, maskp = st->maskp;
  const uint64_t dmask = (1ull << width) - 1;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  const uint32_t lane = threadIdx.x & 63;
  constexpr int U = 4;
  for (uint64_t i0 = blockIdx.x * 256ull + threadIdx.x; i0 < n; i0 += U * stride) {
    uint64_t kk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {  // the loads first (rows past the end read row i0)
      const uint64_t i = i0 + u * stride;
      kk[u] = keys[i < n ? i : i0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + u * stride;
      const uint64_t k = kk[u];
      const bool act = i < n && (k & maskk) == prefk && (on_key || ((uint32_t)i & maskp) == prefp);
      const uint32_t d = (uint32_t)((on_key ? (k >> shift) : ((uint32_t)i >> shift)) & dmask);
      // the lanes that share the first active lane's digit add once (high key
      // digits are often the same across a wave)
      const uint64_t active = __ballot(act);
      if (!active) continue;
      const int first = __builtin_ctzll(active);
      const uint32_t d0 = __shfl(d, first);
      const uint64_t same = __ballot(act && d == d0);
      if (lane == (uint32_t)first) atomicAdd(&h[d0], (uint32_t)__popcll(same));
      if (act && d != d0) atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kRsBins; b += blockDim.x)
    if (h[b]) atomicAdd(&st->hist[b], h[b]);
}

// One workgroup: the digit whose bin holds the remaining-th row; clears the bins.
__global__ void __launch_bounds__(256) k_rs_pick(RsState* st, int pass) {
  // `done` is read once for the whole workgroup: this kernel's winning thread
  // may set it, and a wave reading it later must not leave the barriers below
  __shared__ uint32_t skip;
  if (threadIdx.x == 0) skip = st->done;
  __syncthreads();
  if (skip) return;
  {
    int shift, width;
    bool on_key;
    rs_digit(pass, &shift, &width, &on_key);
    if (rs_shared_digit(st, on_key, shift)) {  // every candidate has the smallest key's digit
      if (threadIdx.x == 0) {
        const uint64_t dm = (1ull << width) - 1;
        st->prefk |= ((st->kmin >> shift) & dm) << shift;
        st->maskk |= dm << shift;
      }
      return;
    }
  }
  constexpr int per = kRsBins / 256;
  __shared__ unsigned long long wsum[4];
  uint64_t local[per];
  uint64_t mine = 0;
  for (int j = 0; j < per; ++j) {
    local[j] = st->hist[threadIdx.x * per + j];
    mine += local[j];
  }
  // read before the barrier: the winning thread rewrites it after the barrier
  const uint64_t rem = st->remaining;
  // exclusive scan of the per-thread sums
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  uint64_t incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint64_t cum = incl - mine;
  for (uint32_t w = 0; w < wave; ++w) cum += wsum[w];
  int shift, width;
  bool on_key;
  rs_digit(pass, &shift, &width, &on_key);
  for (int j = 0; j < per; ++j) {
    if (cum < rem && rem <= cum + local[j]) {
      const uint64_t d = (uint64_t)(threadIdx.x * per + j);
      if (on_key) {
        st->prefk |= d << shift;
        st->maskk |= ((1ull << width) - 1) << shift;
      } else {
        st->prefp |= (uint32_t)d << shift;
        st->maskp |= (uint32_t)(((1ull << width) - 1) << shift);
      }
      st->remaining = rem - cum;
      // the bin holds exactly the rows still needed: all of them are taken
      if (local[j] == rem - cum) st->done = 1;
    }
    cum += local[j];
  }
  __syncthreads();
  for (int j = 0; j < per; ++j) st->hist[threadIdx.x * per + j] = 0;
}

// Rows whose (key, position) is <= the k-th: exactly k of them, in any order
// (row = rows[i], or base + i without a row list).
__global__ void __launch_bounds__(256) k_rs_collect(const uint64_t* __restrict__ keys,
                                                    const uint64_t* __restrict__ rows,
                                                    uint64_t base, uint64_t n, RsState* st,
                                                    uint64_t* out_keys, uint64_t* out_rows,
                                                    uint64_t k) {
  // the largest (key, position) taken: the prefix with its undecided bits set
  // (all bits decided after the last pass)
  const unsigned long long tk = st->prefk | ~st->maskk;
  const uint32_t tp = st->maskk == ~0ull ? (st->prefp | ~st->maskp) : 0xffffffffu;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) {
    const uint64_t key = keys[i];
    if (key < tk || (key == tk && (uint32_t)i <= tp)) {
      const unsigned long long at = atomicAdd(&st->taken, 1ull);
      if (at < k) {  // exactly k qualify; the bound keeps a wrong count in range
        out_keys[at] = key;
        out_rows[at] = rows ? rows[i] : base + i;
      }
    }
  }
}

// One workgroup of 1024 sorts m <= kTopKMax (key, row) pairs (rows distinct)
// with a bitonic network in LDS and writes the rows in order.
__global__ void __launch_bounds__(1024) k_topk_sort(const uint64_t* keys, const uint64_t* rows,
                                                    uint64_t m, uint64_t* out_rows) {
  __shared__ unsigned long long sk[kTopKMax], sr[kTopKMax];
  uint32_t M = 1;
  while (M < m) M <<= 1;
  for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
    sk[i] = i < m ? keys[i] : ~0ull;
    sr[i] = i < m ? rows[i] : ~0ull;
  }
  __syncthreads();
  for (uint32_t size = 2; size <= M; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        const uint32_t j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const bool gt = sk[i] > sk[j] || (sk[i] == sk[j] && sr[i] > sr[j]);
          if (gt == up) {
            const unsigned long long tk = sk[i], tr = sr[i];
            sk[i] = sk[j];
            sr[i] = sr[j];
            sk[j] = tk;
            sr[j] = tr;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) out_rows[i] = sr[i];
}

static_assert(sizeof(RsState) % 8 == 0, "state keeps the buffers after it aligned");

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
  uint64_t cnt = n;
  if (d_sel_bitmap) TGP_TRY(compact_begin(ctx, st, d_sel_bitmap, n, &cnt));
  const uint64_t take = limit < cnt ? limit : cnt;
  TGP_REQUIRE(take <= capacity, "tgp_select_rows: %llu rows, capacity %llu",
              (unsigned long long)take, (unsigned long long)capacity);
  if (take == 0) return TGP_OK;
  TGP_REQUIRE(d_rows, "tgp_select_rows: null output");
  const bool sorted = order_col && cnt > 1;
  const bool topk = sorted && take <= kTopKMax && take < cnt;
  // outer scratch: rows[cnt] | keys[cnt] | (full sort) keys_out, rows_out, sort temp
  //                                       | (top-k) state, keys[take], rows[take]
  size_t temp_bytes = 0;
  if (sorted && !topk)
    TGP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, (uint64_t*)nullptr,
                                               (uint64_t*)nullptr, (uint64_t*)nullptr,
                                               (uint64_t*)nullptr, (int)cnt, 0, 64, st));
  const uint64_t extra = !sorted ? 0
                         : topk  ? sizeof(RsState) + 16 * take
                                 : 16 * cnt + ((temp_bytes + 255) & ~(size_t)255);
  void* p = nullptr;
  TGP_TRY(get_scratch(ctx, st, 16 * cnt + extra + 256, &p, 1));
  uint64_t* rows = static_cast<uint64_t*>(p);
  uint64_t* keys = rows + cnt;
  // the row list: compacted from the bitmap, or implicit (row_begin + i) for
  // the top-k path without a bitmap
  const bool implicit_rows = !d_sel_bitmap && topk;
  if (d_sel_bitmap) {
    TGP_TRY(compact_finish(ctx, st, d_sel_bitmap, n, row_begin, 8, rows));
  } else if (!implicit_rows) {
    hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(256), 0, st, rows, n, row_begin);
    TGP_HIP(hipGetLastError());
  }
  const uint64_t* first = rows;
  const uint64_t* row_list = implicit_rows ? nullptr : rows;
  if (sorted) {
    auto* state = reinterpret_cast<RsState*>(keys + cnt);  // top-k only
    if (topk) hipLaunchKernelGGL(k_rs_init, dim3(1), dim3(256), 0, st, state, take);
    TGP_BY_DTYPE(order_col->dtype, launch_keys, st, order_col, row_list, row_begin, cnt,
                 descending ? ~0ull : 0ull, keys, topk ? &state->kmin : nullptr);
    TGP_HIP(hipGetLastError());
    if (topk) {
      auto* tkeys = reinterpret_cast<uint64_t*>(state + 1);
      uint64_t* trows = tkeys + take;
      const uint32_t g = grid_of(cnt);
      for (int pass = 0; pass < kRsPasses; ++pass) {
        hipLaunchKernelGGL(k_rs_hist, dim3(g), dim3(256), 0, st, keys, cnt, state, pass);
        hipLaunchKernelGGL(k_rs_pick, dim3(1), dim3(256), 0, st, state, pass);
      }
      hipLaunchKernelGGL(k_rs_collect, dim3(g), dim3(256), 0, st, keys, row_list, row_begin, cnt,
                         state, tkeys, trows, take);
      hipLaunchKernelGGL(k_topk_sort, dim3(1), dim3(1024), 0, st, tkeys, trows, take, d_rows);
      TGP_HIP(hipGetLastError());
      TGP_HIP(hipStreamSynchronize(st));
      *h_count = take;
      return TGP_OK;
    }
    uint64_t* keys_out = keys + cnt;
    uint64_t* rows_out = keys_out + cnt;
    void* temp = reinterpret_cast<void*>(
        (reinterpret_cast<uintptr_t>(rows_out + cnt) + 255) & ~(uintptr_t)255);
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
