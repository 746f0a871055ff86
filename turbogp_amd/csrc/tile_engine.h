// LDS-staged tile engine shared by the scan and the fused aggregation kernels.
//
// Staging.  A workgroup of 256 threads (4 waves) walks tiles of 1024 rows.  Each
// referenced column ("slot") of a tile is width KiB at its native width, i.e.
// `width` groups of 64 16-byte chunks.  A wave moves one such group with a
// single gfx950 LDS-DMA instruction (global_load_lds_dwordx4): lane l reads
// 16 bytes from HBM and the hardware writes them to LDS at base + 16 l, so the
// data never occupies VGPRs.  Tiles are double-buffered: the DMA of tile t+1 is
// in flight while tile t is evaluated from LDS.
//
// Tiles start at a multiple of 64 rows (the row range's first row rounded
// down), so every chunk is 16-byte aligned, every 64-row wave slice of a tile is
// one 64-bit word of the selection bitmap (bit i <-> row i, LSB first;
// exec/bitmap.rs:36-45), and rows outside [row_begin, row_end) are masked.
//
// Evaluation.  Row j of a tile belongs to thread j % 256 (R = 4 rows per
// thread).  The query is a small per-slot program prepared on the host: for
// each slot, the predicates, group-key parts and aggregate factors that read
// it.  The kernel walks slots and their ops in wave-uniform loops, so all plan
// metadata is loaded once per tile into scalar registers, and the per-row work
// is straight-line vector code on the 4 rows.
#pragma once

#include "tgp_internal.h"

namespace tgp {

constexpr int kRows = kTileRows / kBlock;  // rows per thread per tile

// Where a slot's tile segment sits in a stage buffer and how to widen it.
struct DevCol {
  uint32_t lds_off;
  int32_t width;      // 1, 2, 4 or 8
  int32_t is_signed;  // sign-extend when widening
  int32_t _pad;
};

struct DevSlot {
  const uint8_t* data;
  uint64_t limit_bytes;  // bytes readable from data (column length * width)
  int32_t width;
  uint32_t lds_off;
  uint64_t _pad;
};

struct DevPred {
  DevCol c;
  int32_t cmp;        // tgp_cmp
  int32_t is_signed;  // compare as int64 (else uint64)
  int64_t lit;
};

struct DevPart {
  DevCol c;
  int32_t divisor;    // 1 => none
  uint32_t magic;     // round-up multiplier for the divisor
  int32_t shift;      // l - 1
  int32_t _pad;
  int64_t base, card, stride;
};

struct DevFactor {
  DevCol c;
  int64_t fa, fb;     // factor = fa + fb * value
};

struct DevAgg {
  int32_t kind;       // tgp_agg_kind
  int32_t nf;
  int64_t coef;
  DevFactor f[TGP_MAX_FACTORS];
};

// Everything a kernel needs about one query (kernel argument, < 4 KiB).
struct DevPlan {
  uint64_t row_begin, row_end;
  uint64_t base0;        // row_begin rounded down to a multiple of 64
  uint64_t ntiles;
  DevSlot slots[TGP_MAX_COLS];
  DevPred preds[TGP_MAX_PREDS];
  DevPart parts[TGP_MAX_GROUP_PARTS];
  DevAgg aggs[TGP_MAX_AGGS];
  int32_t nslots, npreds, nparts, naggs;
  uint32_t stage_bytes;  // LDS bytes of one stage buffer (sum of widths x 1 KiB)
  uint32_t minmax_kinds; // 3 bits per aggregate: its kind if MIN/MAX, else 0
  uint32_t lds_rshift;   // LDS table: 2^lds_rshift copies of every word
  uint32_t _pad1;
  uint64_t ngroups;
  const uint64_t* sel_bitmap;
  // outputs
  unsigned long long* out_count;
  unsigned long long* out_sums;
  unsigned long long* out_oor;
  unsigned long long* out_bitmap;  // scan only
};

// ---- staging ----------------------------------------------------------------------

// Issues the LDS-DMA of tile `t` into `buf`.  Wave w moves chunk groups w, w+4, ...
// A group that straddles the end of a column is copied byte by byte (only the
// final tile of a column can have one).
__device__ __forceinline__ void stage_issue(const DevPlan& P, uint64_t t, uint8_t* buf) {
  const uint32_t lane = threadIdx.x & 63, wave = wave_index();
  const uint64_t row0 = P.base0 + t * kTileRows;
  uint32_t g = 0;  // group index across slots, uniform
  for (int s = 0; s < P.nslots; ++s) {
    const DevSlot& S = P.slots[s];
    const uint32_t ngroups = (uint32_t)S.width;  // width KiB per tile
    const uint8_t* src = S.data + row0 * (uint64_t)S.width;
    const uint64_t left = S.limit_bytes > row0 * (uint64_t)S.width
                              ? S.limit_bytes - row0 * (uint64_t)S.width : 0;
    for (uint32_t k = 0; k < ngroups; ++k, ++g) {
      if ((g & 3) != wave) continue;
      const uint64_t off = (uint64_t)k * 1024;
      if (off >= left) continue;  // past the column: rows are masked, LDS left as is
      uint8_t* dst = buf + S.lds_off + (uint32_t)off;
      if (off + 1024 <= left) {
        __builtin_amdgcn_global_load_lds(
            reinterpret_cast<const void*>(src + off + lane * 16), dst, 16, 0, 0);
      } else {
        for (uint32_t b = lane; b < 1024; b += 64)
          dst[b] = off + b < left ? src[off + b] : 0;
      }
    }
  }
}

// Waits for this wave's outstanding DMA; a barrier must follow before other
// waves read the data.
__device__ __forceinline__ void stage_wait() { __builtin_amdgcn_s_waitcnt(0); }

// ---- row evaluation -----------------------------------------------------------------

// Loads a slot's value for this thread's R rows (rows j = tid + 256 r).
__device__ __forceinline__ void load_col(const DevCol& c, const uint8_t* buf, int64_t (&x)[kRows]) {
  const uint8_t* base = buf + c.lds_off;
  const uint32_t tid = threadIdx.x;
  switch (c.width) {
    case 1:
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const uint32_t v = base[tid + r * kBlock];
        x[r] = c.is_signed ? (int64_t)(int8_t)v : (int64_t)v;
      }
      break;
    case 2:
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const uint32_t v = reinterpret_cast<const uint16_t*>(base)[tid + r * kBlock];
        x[r] = c.is_signed ? (int64_t)(int16_t)v : (int64_t)v;
      }
      break;
    case 4:
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const uint32_t v = reinterpret_cast<const uint32_t*>(base)[tid + r * kBlock];
        x[r] = c.is_signed ? (int64_t)(int32_t)v : (int64_t)v;
      }
      break;
    default:
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        x[r] = reinterpret_cast<const int64_t*>(base)[tid + r * kBlock];
      break;
  }
}

// bitmap.rs filter_* semantics on the 64-bit widened value.
__device__ __forceinline__ bool cmp_value(int32_t op, bool sgn, int64_t v, int64_t lit) {
  if (sgn) {
    switch (op) {
      case TGP_EQ: return v == lit;
      case TGP_NE: return v != lit;
      case TGP_LT: return v < lit;
      case TGP_LE: return v <= lit;
      case TGP_GT: return v > lit;
      default: return v >= lit;
    }
  }
  const uint64_t u = (uint64_t)v, l = (uint64_t)lit;
  switch (op) {
    case TGP_EQ: return u == l;
    case TGP_NE: return u != l;
    case TGP_LT: return u < l;
    case TGP_LE: return u <= l;
    case TGP_GT: return u > l;
    default: return u >= l;
  }
}

// Selection bits (bit r = row r of this thread) of a staged tile: row range,
// the optional input bitmap, then the AND of the predicates.
__device__ __forceinline__ uint32_t eval_select(const DevPlan& P, const uint8_t* buf,
                                                uint64_t row0) {
  uint32_t sel = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const uint64_t row = row0 + threadIdx.x + r * kBlock;
    if (row >= P.row_begin && row < P.row_end) sel |= 1u << r;
  }
  if (P.sel_bitmap) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint64_t row = row0 + threadIdx.x + r * kBlock;
      if ((sel >> r) & 1) {
        const uint64_t i = row - P.row_begin;
        if (!((P.sel_bitmap[i >> 6] >> (i & 63)) & 1)) sel &= ~(1u << r);
      }
    }
  }
  for (int p = 0; p < P.npreds; ++p) {
    const DevPred& pr = P.preds[p];
    int64_t x[kRows];
    load_col(pr.c, buf, x);
    const bool sgn = pr.is_signed != 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      if (!cmp_value(pr.cmp, sgn, x[r], pr.lit)) sel &= ~(1u << r);
  }
  return sel;
}

// Dense group ids of the thread's rows (-1: a key part is out of range).
__device__ __forceinline__ void eval_groups(const DevPlan& P, const uint8_t* buf,
                                            int64_t (&gid)[kRows]) {
#pragma unroll
  for (int r = 0; r < kRows; ++r) gid[r] = 0;
  for (int k = 0; k < P.nparts; ++k) {
    const DevPart& D = P.parts[k];
    int64_t x[kRows];
    load_col(D.c, buf, x);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      int64_t v = x[r];
      if (D.divisor > 1) {
        if (v >= 0 && v <= 0xffffffffll) {
          const uint32_t n = (uint32_t)v;
          const uint32_t q = __umulhi(D.magic, n);
          v = (int64_t)((q + ((n - q) >> 1)) >> D.shift);
        } else {
          v = v / D.divisor;
        }
      }
      const int64_t part = v - D.base;
      gid[r] = (part < 0 || part >= D.card || gid[r] < 0) ? -1 : gid[r] + part * D.stride;
    }
  }
}

// Per-row value of aggregate A for the thread's rows, in wrapping 64-bit
// arithmetic (expr.rs binop wrapping_add/sub/mul).
__device__ __forceinline__ void eval_agg(const DevAgg& A, const uint8_t* buf,
                                         uint64_t (&v)[kRows]) {
  if (A.kind == TGP_AGG_COUNT_NONZERO) {
    int64_t x[kRows];
    load_col(A.f[0].c, buf, x);
#pragma unroll
    for (int r = 0; r < kRows; ++r) v[r] = x[r] != 0 ? 1 : 0;
    return;
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) v[r] = (uint64_t)A.coef;
  for (int k = 0; k < A.nf; ++k) {
    const DevFactor& F = A.f[k];
    int64_t x[kRows];
    load_col(F.c, buf, x);
    const bool assign = k == 0 && A.coef == 1;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      uint64_t f;
      if (F.fb == 1) f = (uint64_t)F.fa + (uint64_t)x[r];
      else if (F.fb == -1) f = (uint64_t)F.fa - (uint64_t)x[r];
      else f = (uint64_t)F.fa + (uint64_t)F.fb * (uint64_t)x[r];
      v[r] = assign ? f : v[r] * f;
    }
  }
}

// MIN / MAX run as an unsigned maximum of an order-preserving code with
// identity 0 (the value a cleared result holds): signed values flip the sign
// bit, and MIN complements the code.
__host__ __device__ constexpr bool is_minmax(int32_t kind) {
  return kind >= TGP_AGG_MIN && kind <= TGP_AGG_MAX_U64;
}
__host__ __device__ __forceinline__ uint64_t mm_encode(int32_t kind, uint64_t v) {
  const uint64_t k = (kind == TGP_AGG_MIN || kind == TGP_AGG_MAX) ? v ^ (1ull << 63) : v;
  return (kind == TGP_AGG_MIN || kind == TGP_AGG_MIN_U64) ? ~k : k;
}
__host__ __device__ __forceinline__ uint64_t mm_decode(int32_t kind, uint64_t c) {
  const uint64_t k = (kind == TGP_AGG_MIN || kind == TGP_AGG_MIN_U64) ? ~c : c;
  return (kind == TGP_AGG_MIN || kind == TGP_AGG_MAX) ? k ^ (1ull << 63) : k;
}

// Split sums: a value v is added as its unsigned low 32 bits into one 64-bit word
// and its signed high 32 bits into another, with non-returning atomics.  Each word
// stays exact for fewer than 2^32 additions, and
//   sum(v) = sum(lo32) + 2^32 * sum(hi32)
// is rebuilt as int128 words (lo, hi) by split_to_i128.
__device__ __forceinline__ void split_add(unsigned long long* p, int64_t v) {
  atomicAdd(p, (unsigned long long)(uint32_t)v);
  const int64_t h = v >> 32;
  if (h) atomicAdd(p + 1, (unsigned long long)h);
}

__host__ __device__ __forceinline__ void split_to_i128(uint64_t slo, int64_t shi, uint64_t* lo,
                                                       int64_t* hi) {
  const uint64_t add = (uint64_t)shi << 32;
  const uint64_t l = slo + add;
  *hi = (shi >> 32) + (int64_t)(l < slo);
  *lo = l;
}

// Exact 128-bit accumulation: value = hi * 2^64 + lo.
__device__ __forceinline__ void add128(uint64_t& lo, int32_t& hi, int64_t v) {
  const uint64_t s = lo + (uint64_t)v;
  hi += (int32_t)(s < lo) - (int32_t)(v < 0);
  lo = s;
}

// Adds (lo, hi) to an int128 stored as two words with 64-bit atomics.  The
// carries out of the low word are counted exactly whatever the order of the
// concurrent additions, because each addition crosses at most one multiple of
// 2^64 and sees the running sum just before it.
__device__ __forceinline__ void atomic_add128(unsigned long long* p, uint64_t lo, int64_t hi) {
  if (lo == 0 && hi == 0) return;
  const uint64_t old = atomicAdd(p, (unsigned long long)lo);
  const int64_t h = hi + (int64_t)((old + lo) < old);
  if (h != 0) atomicAdd(p + 1, (unsigned long long)h);
}

// Adds one signed 64-bit value to an int128 in memory (LDS or global).
__device__ __forceinline__ void atomic_add128_i64(unsigned long long* p, int64_t v) {
  const unsigned long long old = atomicAdd(p, (unsigned long long)v);
  const long long h = (long long)((old + (uint64_t)v) < old) - (long long)(v < 0);
  if (h) atomicAdd(p + 1, (unsigned long long)h);
}

// The persistent tile loop with double-buffered LDS-DMA staging.  `body` is
// called with (stage buffer, first row of the tile) for each tile this
// workgroup owns.
template <typename Body>
__device__ __forceinline__ void for_each_tile(const DevPlan& P, uint8_t* lds, Body&& body) {
  uint64_t t = blockIdx.x;
  if (t >= P.ntiles) return;
  stage_issue(P, t, lds);
  for (uint32_t i = 0; t < P.ntiles; t += gridDim.x, ++i) {
    stage_wait();
    __syncthreads();  // tile t is in LDS; every wave is done with the other buffer
    const uint64_t tn = t + gridDim.x;
    uint8_t* cur = lds + (i & 1) * P.stage_bytes;
    uint8_t* nxt = lds + ((i + 1) & 1) * P.stage_bytes;
    if (tn < P.ntiles) stage_issue(P, tn, nxt);
    body(cur, P.base0 + t * kTileRows);
  }
  __syncthreads();
}

}  // namespace tgp
