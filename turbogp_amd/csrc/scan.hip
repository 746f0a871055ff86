// Predicate scan -> selection bitmap + count, bitmap algebra, compaction,
// column min/max.
//
// Reference: exec/bitmap.rs (Bitmap :42-273, SetBitIter :278-314, filter_*
// :321-477, filter_u64_avx512_inner :642-694) and query_interpreter/expr.rs
// build_mask (:419-431, AND chain :633-656).  The reference evaluates one
// column pass per predicate leaf plus one AND pass per leaf; here all leaves
// are fused into one pass over HBM, and the bitmap word of 64 rows is a
// single wave64 ballot.
#include "plan.h"

namespace tgp {

// One bit per row: the AND of the predicates.  Lane l of wave w holds row
// row0 + 256 r + 64 w + l, so the wave's ballot for step r is the bitmap word
// of those 64 rows.  When row_begin is not a multiple of 64 the word straddles
// two output words and is OR-ed into both (the output is zeroed first).
__global__ void __launch_bounds__(kBlock) k_scan(DevPlan P) {
  extern __shared__ __align__(16) uint8_t lds[];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = wave_index();
  const uint32_t d = (uint32_t)(P.row_begin & 63);
  const uint64_t nwords = (P.row_end - P.row_begin + 63) / 64;
  unsigned long long cnt = 0;
  for_each_tile(P, lds, [&](const uint8_t* buf, uint64_t row0) {
    const uint32_t sel = eval_select(P, buf, row0);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const unsigned long long word = __ballot((sel >> r) & 1);
      const uint64_t x = row0 + r * kBlock + wave * kWave;  // first row of the word
      if (lane == 0 && x < P.row_end) {
        cnt += __popcll(word);
        if (P.out_bitmap) {
          const uint64_t k = (x - P.base0) >> 6;
          if (d == 0) {
            P.out_bitmap[k] = word;
          } else {
            if (k >= 1 && (word << (64 - d))) atomicOr(&P.out_bitmap[k - 1], word << (64 - d));
            if (k < nwords && (word >> d)) atomicOr(&P.out_bitmap[k], word >> d);
          }
        }
      }
    }
  });
  if (P.out_count && lane == 0 && cnt) atomicAdd(P.out_count, cnt);
}

int32_t run_scan(tgp_ctx* ctx, const DevPlan& P, const PlanInfo& info, hipStream_t s) {
  if (P.out_bitmap && (P.row_begin & 63))
    TGP_TRY(fill_u64(s, reinterpret_cast<uint64_t*>(P.out_bitmap),
                     (P.row_end - P.row_begin + 63) / 64, 0));
  return launch_tiles(ctx, k_scan, P, info.lds_bytes, s);
}

// ---- bitmap algebra -------------------------------------------------------------

__device__ __forceinline__ uint64_t tail_mask(uint64_t nbits, uint64_t word) {
  const uint64_t first = word * 64;
  if (first + 64 <= nbits) return ~0ull;
  const uint64_t valid = nbits > first ? nbits - first : 0;
  return valid ? (~0ull >> (64 - valid)) : 0ull;
}

__global__ void __launch_bounds__(256) k_bitmap_combine(int op, const uint64_t* __restrict__ a,
                                                        const uint64_t* __restrict__ b,
                                                        uint64_t* out, uint64_t nbits) {
  const uint64_t nw = (nbits + 63) / 64;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nw;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = a[i], y = b[i], r;
    if (op == TGP_BIT_AND) r = x & y;
    else if (op == TGP_BIT_OR) r = x | y;
    else r = x & ~y;
    out[i] = r & tail_mask(nbits, i);
  }
}

// A cell widened to 64 bits by its column's signedness (the reference's u64 cells).
__device__ __forceinline__ uint64_t cell64(const void* base, int32_t dtype, uint64_t i) {
  switch (dtype) {
    case TGP_I8: return (uint64_t)(int64_t)static_cast<const int8_t*>(base)[i];
    case TGP_U8: return static_cast<const uint8_t*>(base)[i];
    case TGP_I16: return (uint64_t)(int64_t)static_cast<const int16_t*>(base)[i];
    case TGP_U16: return static_cast<const uint16_t*>(base)[i];
    case TGP_I32: return (uint64_t)(int64_t)static_cast<const int32_t*>(base)[i];
    case TGP_U32: return static_cast<const uint32_t*>(base)[i];
    default: return static_cast<const uint64_t*>(base)[i];
  }
}

template <bool SIGNED>
__device__ __forceinline__ bool compare_cells(int32_t op, uint64_t x, uint64_t y) {
  const bool lt = SIGNED ? (int64_t)x < (int64_t)y : x < y;
  const bool gt = SIGNED ? (int64_t)x > (int64_t)y : x > y;
  switch (op) {
    case TGP_EQ: return x == y;
    case TGP_NE: return x != y;
    case TGP_LT: return lt;
    case TGP_LE: return !gt;
    case TGP_GT: return gt;
    default: return !lt;  // TGP_GE
  }
}

// Column-vs-column narrowing of a selection: wave w handles bitmap word w (64
// rows, one per lane) and keeps bit j only where a[r] op b[r],
// r = row_begin + 64 w + j; bits past n are cleared.
template <bool SIGNED>
__global__ void __launch_bounds__(256) k_compare_columns(const void* a, int32_t adt, const void* b,
                                                         int32_t bdt, int32_t op,
                                                         uint64_t row_begin, uint64_t n,
                                                         uint64_t* __restrict__ bitmap) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwords = (n + 63) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t w = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64; w < nwords;
       w += nwaves) {
    const uint64_t i = w * 64 + lane;
    bool keep = false;
    if (i < n)
      keep = compare_cells<SIGNED>(op, cell64(a, adt, row_begin + i), cell64(b, bdt, row_begin + i));
    const uint64_t bits = __ballot(keep);
    if (lane == 0) bitmap[w] &= bits;
  }
}

__global__ void __launch_bounds__(256) k_bitmap_not(const uint64_t* __restrict__ a, uint64_t* out,
                                                    uint64_t nbits) {
  const uint64_t nw = (nbits + 63) / 64;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nw;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = ~a[i] & tail_mask(nbits, i);
}

__global__ void __launch_bounds__(256) k_bitmap_count(const uint64_t* __restrict__ a,
                                                      uint64_t begin, uint64_t end,
                                                      unsigned long long* out) {
  unsigned long long c = 0;
  const uint64_t w0 = begin / 64, w1 = (end + 63) / 64;
  for (uint64_t i = w0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < w1;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = a[i];
    const uint64_t lo = i * 64, hi = lo + 64;
    if (begin > lo) x &= ~0ull << (begin - lo);
    if (end < hi) x &= (end > lo) ? (~0ull >> (hi - end)) : 0ull;
    c += __popcll(x);
  }
  // wave reduction then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

static unsigned grid_for(uint64_t items) {
  uint64_t g = (items + 255) / 256;
  if (g > 8192) g = 8192;
  return (unsigned)(g ? g : 1);
}

// ---- compaction -------------------------------------------------------------------
// Three passes: per-block popcounts (256 words = 16384 bits per block), an
// exclusive scan of the block counts in one workgroup, then the scatter:
// a wave handles one word at a time, lane l owns bit l, and its output slot
// is the block offset + the word's prefix + popcount(word & lanes below l),
// so 64 consecutive rows are written by 64 consecutive lanes.
constexpr int kCompactWords = 256;

__global__ void __launch_bounds__(256) k_compact_count(const uint64_t* __restrict__ bm,
                                                       uint64_t nbits, uint64_t* block_cnt) {
  __shared__ unsigned long long red[4];
  const uint64_t nw = (nbits + 63) / 64;
  const uint64_t w = blockIdx.x * (uint64_t)kCompactWords + threadIdx.x;
  unsigned long long c = w < nw ? __popcll(bm[w] & tail_mask(nbits, w)) : 0;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Exclusive scan of n block counts in place; total written to block_cnt[n].
__global__ void __launch_bounds__(1024) k_compact_scan(uint64_t* block_cnt, uint64_t n) {
  __shared__ unsigned long long part[1024];
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t b = threadIdx.x * per, e = min(b + per, n);
  unsigned long long s = 0;
  for (uint64_t i = b; i < e; ++i) s += block_cnt[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    unsigned long long v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x ? part[threadIdx.x - 1] : 0;
  for (uint64_t i = b; i < e; ++i) {
    unsigned long long c = block_cnt[i];
    block_cnt[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) block_cnt[n] = part[1023];
}

template <typename IDX>
__global__ void __launch_bounds__(256) k_compact_scatter(const uint64_t* __restrict__ bm,
                                                         uint64_t nbits, uint64_t base,
                                                         const uint64_t* __restrict__ block_off,
                                                         IDX* __restrict__ out) {
  __shared__ uint64_t words[kCompactWords];
  __shared__ unsigned long long pre[kCompactWords];
  const uint64_t nw = (nbits + 63) / 64;
  const uint64_t w0 = blockIdx.x * (uint64_t)kCompactWords;
  const uint64_t w = w0 + threadIdx.x;
  const uint64_t word = w < nw ? (bm[w] & tail_mask(nbits, w)) : 0;
  words[threadIdx.x] = word;
  pre[threadIdx.x] = __popcll(word);
  __syncthreads();
  for (int off = 1; off < kCompactWords; off <<= 1) {
    unsigned long long v = threadIdx.x >= (unsigned)off ? pre[threadIdx.x - off] : 0;
    __syncthreads();
    pre[threadIdx.x] += v;
    __syncthreads();
  }
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t boff = block_off[blockIdx.x];
  for (int i = wave; i < kCompactWords; i += 4) {
    const uint64_t x = words[i];
    if (!x) continue;
    const uint64_t start = boff + (i ? pre[i - 1] : 0);
    if ((x >> lane) & 1) {
      const uint64_t pos = start + __popcll(x & below);
      out[pos] = (IDX)(base + (w0 + i) * 64 + lane);
    }
  }
}

// ---- min / max ------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_minmax(const uint8_t* __restrict__ data, int32_t dtype,
                                                uint64_t n, long long* out) {
  const bool sgn = dtype == TGP_I8 || dtype == TGP_I16 || dtype == TGP_I32 || dtype == TGP_I64;
  long long mn = sgn ? LLONG_MAX : (long long)~0ull, mx = sgn ? LLONG_MIN : 0;
  unsigned long long umn = ~0ull, umx = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    long long v;
    switch (dtype) {
      case TGP_I8: v = ((const int8_t*)data)[i]; break;
      case TGP_U8: v = data[i]; break;
      case TGP_I16: v = ((const int16_t*)data)[i]; break;
      case TGP_U16: v = ((const uint16_t*)data)[i]; break;
      case TGP_I32: v = ((const int32_t*)data)[i]; break;
      case TGP_U32: v = ((const uint32_t*)data)[i]; break;
      default: v = ((const long long*)data)[i]; break;
    }
    if (sgn) { mn = v < mn ? v : mn; mx = v > mx ? v : mx; }
    else {
      const unsigned long long u = (unsigned long long)v;
      umn = u < umn ? u : umn; umx = u > umx ? u : umx;
    }
  }
  if (sgn) { atomicMin(&out[0], mn); atomicMax(&out[1], mx); }
  else {
    atomicMin((unsigned long long*)&out[0], umn);
    atomicMax((unsigned long long*)&out[1], umx);
  }
}

}  // namespace tgp

using namespace tgp;

namespace tgp {

// Blocks of the compaction passes over nbits bits; the passes keep nblocks + 1
// offset words (every word is written before it is read: nothing to clear).
static uint64_t compact_blocks(uint64_t nbits) {
  return ((nbits + 63) / 64 + kCompactWords - 1) / kCompactWords;
}

// compact_begin / compact_finish with the block offsets in `bc`.
static int32_t compact_count_into(hipStream_t s, const uint64_t* d_bitmap, uint64_t nbits,
                                  uint64_t* bc, uint64_t* total) {
  *total = 0;
  if (nbits == 0) return TGP_OK;
  const uint64_t nblocks = compact_blocks(nbits);
  k_compact_count<<<(unsigned)nblocks, 256, 0, s>>>(d_bitmap, nbits, bc);
  TGP_HIP(hipGetLastError());
  k_compact_scan<<<1, 1024, 0, s>>>(bc, nblocks);
  TGP_HIP(hipGetLastError());
  TGP_HIP(hipMemcpyAsync(total, bc + nblocks, 8, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

static int32_t compact_scatter_from(hipStream_t s, const uint64_t* d_bitmap, uint64_t nbits,
                                    uint64_t base, int32_t index_bytes, void* d_out,
                                    const uint64_t* bc) {
  if (nbits == 0) return TGP_OK;
  const uint64_t nblocks = compact_blocks(nbits);
  if (index_bytes == 8)
    k_compact_scatter<unsigned long long><<<(unsigned)nblocks, 256, 0, s>>>(
        d_bitmap, nbits, base, bc, static_cast<unsigned long long*>(d_out));
  else
    k_compact_scatter<uint32_t><<<(unsigned)nblocks, 256, 0, s>>>(d_bitmap, nbits, base, bc,
                                                                  static_cast<uint32_t*>(d_out));
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

// ---- compaction without allocation -----------------------------------------------------
// The block offsets of tgp_bitmap_to_indices_noalloc_sync live in this file-scope
// device array (512 KiB per GPU, held until the process ends) instead of the
// stream's scratch, which is allocated on first use.  Host code never passes the
// array's name to a kernel: its device address, per device, comes from
// hipGetSymbolAddress at context creation (compact_noalloc_init).
constexpr uint64_t kNoAllocMaxBlocks = 65536;  // 2^30 bits of 16384-bit blocks
constexpr uint64_t kNoAllocMaxBits = kNoAllocMaxBlocks * kCompactWords * 64;
static __device__ __align__(256) uint64_t scratch[kNoAllocMaxBlocks + 1];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static uint64_t* g_noalloc_ptr[kMaxDevices] = {};

int32_t compact_noalloc_init(int32_t device) {
  TGP_REQUIRE(device >= 0 && device < kMaxDevices, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return TGP_OK;
  int cur = -1;
  TGP_HIP(hipGetDevice(&cur));
  TGP_REQUIRE(cur == device, "compact_noalloc_init: device %d is not current (%d)", device, cur);
  void* p = nullptr;
  TGP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  TGP_REQUIRE(p != nullptr, "hipGetSymbolAddress gave no address for the compaction scratch");
  g_noalloc_ptr[device] = static_cast<uint64_t*>(p);
  return TGP_OK;
}

int32_t compact_begin(tgp_ctx* ctx, hipStream_t s, const uint64_t* d_bitmap, uint64_t nbits,
                      uint64_t* total) {
  *total = 0;
  if (nbits == 0) return TGP_OK;
  const uint64_t nw = (nbits + 63) / 64;
  const uint64_t nblocks = (nw + kCompactWords - 1) / kCompactWords;
  void* scratch = nullptr;
  TGP_TRY(get_scratch(ctx, s, (nblocks + 1) * 8, &scratch));
  uint64_t* bc = static_cast<uint64_t*>(scratch);
  return compact_count_into(s, d_bitmap, nbits, bc, total);
}

int32_t compact_finish(tgp_ctx* ctx, hipStream_t s, const uint64_t* d_bitmap, uint64_t nbits,
                       uint64_t base, int32_t index_bytes, void* d_out) {
  if (nbits == 0) return TGP_OK;
  void* scratch = nullptr;  // the block offsets compact_begin left (same size: same buffer)
  TGP_TRY(get_scratch(ctx, s, (compact_blocks(nbits) + 1) * 8, &scratch));
  return compact_scatter_from(s, d_bitmap, nbits, base, index_bytes, d_out,
                              static_cast<const uint64_t*>(scratch));
}

}  // namespace tgp

extern "C" {

uint64_t tgp_bitmap_words(uint64_t nbits) { return (nbits + 63) / 64; }

int32_t tgp_scan_filter(tgp_ctx* ctx, const tgp_column* cols, int32_t ncols, const tgp_pred* preds,
                        int32_t npreds, uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                        uint64_t* d_count, void* stream) {
  TGP_REQUIRE(ctx, "tgp_scan_filter: null ctx");
  TGP_TRY(check_injected(ctx));
  DevPlan P;
  PlanInfo info;
  if (ncols > 0) TGP_TRY(plan_scan(cols, ncols, preds, npreds, row_begin, row_end, &P, &info));
  else TGP_REQUIRE(npreds == 0, "predicates need columns");
  P.out_bitmap = reinterpret_cast<unsigned long long*>(d_bitmap);
  P.out_count = reinterpret_cast<unsigned long long*>(d_count);
  if (ncols == 0) {
    // No column to stage: a one-byte dummy slot would still need memory, so
    // an empty predicate list over no columns selects every row directly.
    hipStream_t s = pick_stream(ctx, stream);
    const uint64_t n = row_end - row_begin;
    if (d_bitmap) {
      TGP_TRY(fill_u64(s, d_bitmap, tgp_bitmap_words(n), ~0ull));
      if (n % 64) {
        const uint64_t last = ~0ull >> (64 - n % 64);
        TGP_HIP(hipMemcpyAsync(d_bitmap + n / 64, &last, 8, hipMemcpyHostToDevice, s));
        TGP_HIP(hipStreamSynchronize(s));
      }
    }
    if (d_count) {
      uint64_t cur = 0;
      TGP_HIP(hipMemcpyAsync(&cur, d_count, 8, hipMemcpyDeviceToHost, s));
      TGP_HIP(hipStreamSynchronize(s));
      cur += n;
      TGP_HIP(hipMemcpyAsync(d_count, &cur, 8, hipMemcpyHostToDevice, s));
      TGP_HIP(hipStreamSynchronize(s));
    }
    return TGP_OK;
  }
  return run_scan(ctx, P, info, pick_stream(ctx, stream));
}

int32_t tgp_bitmap_combine(tgp_ctx* ctx, int32_t op, const uint64_t* d_a, const uint64_t* d_b,
                           uint64_t* d_out, uint64_t nbits, void* stream) {
  TGP_REQUIRE(ctx && (nbits == 0 || (d_a && d_b && d_out)), "tgp_bitmap_combine: null argument");
  TGP_REQUIRE(op >= TGP_BIT_AND && op <= TGP_BIT_ANDNOT, "tgp_bitmap_combine: bad op %d", op);
  TGP_TRY(check_injected(ctx));
  if (nbits == 0) return TGP_OK;
  k_bitmap_combine<<<grid_for((nbits + 63) / 64), 256, 0, pick_stream(ctx, stream)>>>(
      op, d_a, d_b, d_out, nbits);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_bitmap_not(tgp_ctx* ctx, const uint64_t* d_a, uint64_t* d_out, uint64_t nbits,
                       void* stream) {
  TGP_REQUIRE(ctx && (nbits == 0 || (d_a && d_out)), "tgp_bitmap_not: null argument");
  TGP_TRY(check_injected(ctx));
  if (nbits == 0) return TGP_OK;
  k_bitmap_not<<<grid_for((nbits + 63) / 64), 256, 0, pick_stream(ctx, stream)>>>(d_a, d_out,
                                                                                 nbits);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_scan_compare_columns(tgp_ctx* ctx, const tgp_column* a, int32_t op,
                                 const tgp_column* b, int32_t is_signed, uint64_t row_begin,
                                 uint64_t row_end, uint64_t* d_bitmap, void* stream) {
  TGP_REQUIRE(ctx, "tgp_scan_compare_columns: null ctx");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_column(a, "tgp_scan_compare_columns left column"));
  TGP_TRY(check_column(b, "tgp_scan_compare_columns right column"));
  TGP_REQUIRE(op >= TGP_EQ && op <= TGP_GE, "tgp_scan_compare_columns: op %d", op);
  TGP_REQUIRE(row_begin <= row_end && row_end <= a->len && row_end <= b->len,
              "tgp_scan_compare_columns: rows [%llu, %llu) outside the columns",
              (unsigned long long)row_begin, (unsigned long long)row_end);
  const uint64_t n = row_end - row_begin;
  if (n == 0) return TGP_OK;
  TGP_REQUIRE(d_bitmap, "tgp_scan_compare_columns: null bitmap");
  const unsigned g = grid_for((n + 63) / 64 * 64);
  hipStream_t st = pick_stream(ctx, stream);
  if (is_signed)
    k_compare_columns<true><<<g, 256, 0, st>>>(a->data, a->dtype, b->data, b->dtype, op,
                                               row_begin, n, d_bitmap);
  else
    k_compare_columns<false><<<g, 256, 0, st>>>(a->data, a->dtype, b->data, b->dtype, op,
                                                row_begin, n, d_bitmap);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_bitmap_count_range(tgp_ctx* ctx, const uint64_t* d_a, uint64_t nbits, uint64_t begin,
                               uint64_t end, uint64_t* d_count, void* stream) {
  TGP_REQUIRE(ctx && d_count, "tgp_bitmap_count_range: null argument");
  TGP_TRY(check_injected(ctx));
  if (end > nbits) end = nbits;
  if (begin >= end) return TGP_OK;  // count_ones_range returns 0 (bitmap.rs:249-251)
  TGP_REQUIRE(d_a, "tgp_bitmap_count_range: null bitmap");
  const uint64_t words = (end + 63) / 64 - begin / 64;
  k_bitmap_count<<<grid_for(words), 256, 0, pick_stream(ctx, stream)>>>(
      d_a, begin, end, reinterpret_cast<unsigned long long*>(d_count));
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_bitmap_to_indices_sync(tgp_ctx* ctx, const uint64_t* d_bitmap, uint64_t nbits,
                                   uint64_t base, int32_t index_bytes, void* d_out,
                                   uint64_t out_capacity, uint64_t* h_count, void* stream) {
  TGP_REQUIRE(ctx && h_count, "tgp_bitmap_to_indices_sync: null argument");
  TGP_REQUIRE(index_bytes == 4 || index_bytes == 8, "index_bytes must be 4 or 8");
  TGP_TRY(check_injected(ctx));
  *h_count = 0;
  if (nbits == 0) return TGP_OK;
  TGP_REQUIRE(d_bitmap, "tgp_bitmap_to_indices_sync: null bitmap");
  if (index_bytes == 4)
    TGP_REQUIRE(base + nbits <= (1ull << 32), "u32 indices cannot hold rows up to %llu",
                (unsigned long long)(base + nbits));
  hipStream_t s = pick_stream(ctx, stream);
  uint64_t total = 0;
  TGP_TRY(compact_begin(ctx, s, d_bitmap, nbits, &total));
  *h_count = total;
  if (total > out_capacity) {
    set_error("compaction needs %llu entries, capacity %llu", (unsigned long long)total,
              (unsigned long long)out_capacity);
    return TGP_ERR_CAPACITY;
  }
  if (total == 0) return TGP_OK;
  TGP_REQUIRE(d_out, "tgp_bitmap_to_indices_sync: null output");
  TGP_TRY(compact_finish(ctx, s, d_bitmap, nbits, base, index_bytes, d_out));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_bitmap_to_indices_noalloc_sync(tgp_ctx* ctx, const uint64_t* d_bitmap,
                                           uint64_t nbits, uint64_t base, int32_t index_bytes,
                                           void* d_out, uint64_t out_capacity, uint64_t* h_count,
                                           void* stream) {
  TGP_REQUIRE(ctx && h_count, "tgp_bitmap_to_indices_noalloc_sync: null argument");
  TGP_REQUIRE(index_bytes == 4 || index_bytes == 8, "index_bytes must be 4 or 8");
  TGP_TRY(check_injected(ctx));
  *h_count = 0;
  TGP_REQUIRE(nbits <= kNoAllocMaxBits,
              "tgp_bitmap_to_indices_noalloc_sync: %llu bits > the maximum %llu (2^30) its "
              "fixed scratch covers; use tgp_bitmap_to_indices_sync",
              (unsigned long long)nbits, (unsigned long long)kNoAllocMaxBits);
  if (nbits == 0) return TGP_OK;
  TGP_REQUIRE(d_bitmap, "tgp_bitmap_to_indices_noalloc_sync: null bitmap");
  if (index_bytes == 4)
    TGP_REQUIRE(base + nbits <= (1ull << 32), "u32 indices cannot hold rows up to %llu",
                (unsigned long long)(base + nbits));
  uint64_t* bc = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    bc = ctx->device >= 0 && ctx->device < kMaxDevices ? g_noalloc_ptr[ctx->device] : nullptr;
  }
  TGP_REQUIRE(bc != nullptr, "tgp_bitmap_to_indices_noalloc_sync: scratch of device %d not set up",
              ctx->device);
  hipStream_t s = pick_stream(ctx, stream);
  uint64_t total = 0;
  TGP_TRY(compact_count_into(s, d_bitmap, nbits, bc, &total));
  *h_count = total;
  if (total > out_capacity) {
    set_error("compaction needs %llu entries, capacity %llu", (unsigned long long)total,
              (unsigned long long)out_capacity);
    return TGP_ERR_CAPACITY;
  }
  if (total == 0) return TGP_OK;
  TGP_REQUIRE(d_out, "tgp_bitmap_to_indices_noalloc_sync: null output");
  TGP_TRY(compact_scatter_from(s, d_bitmap, nbits, base, index_bytes, d_out, bc));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_column_minmax_sync(tgp_ctx* ctx, const tgp_column* col, int64_t* h_min,
                               int64_t* h_max) {
  TGP_REQUIRE(ctx && col && h_min && h_max, "tgp_column_minmax_sync: null argument");
  TGP_TRY(check_column(col, "tgp_column_minmax_sync"));
  TGP_REQUIRE(col->len > 0, "tgp_column_minmax_sync: empty column");
  hipStream_t s = ctx->stream;
  void* scratch = nullptr;
  TGP_TRY(get_scratch(ctx, s, 16, &scratch));
  const bool sgn = dtype_signed(col->dtype);
  long long init[2] = {sgn ? LLONG_MAX : (long long)~0ull, sgn ? LLONG_MIN : 0};
  TGP_HIP(hipMemcpyAsync(scratch, init, 16, hipMemcpyHostToDevice, s));
  k_minmax<<<grid_for(col->len), 256, 0, s>>>(static_cast<const uint8_t*>(col->data), col->dtype,
                                              col->len, static_cast<long long*>(scratch));
  TGP_HIP(hipGetLastError());
  long long r[2];
  TGP_HIP(hipMemcpyAsync(r, scratch, 16, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  *h_min = r[0];
  *h_max = r[1];
  return TGP_OK;
}

}  // extern "C"
