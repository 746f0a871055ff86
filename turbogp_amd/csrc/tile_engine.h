// LDS-staged tile engine shared by the scan and the fused aggregation kernels.
//
// A workgroup of 256 threads walks tiles of kTileRows (1024) rows of the
// query's row range.  For each tile every referenced column ("slot") is
// copied HBM -> VGPR -> LDS at its native width with 16-byte loads; the wave
// reads 64 consecutive 16-byte chunks, so each load instruction touches one
// contiguous KiB.  Rows are then evaluated from LDS, row j of the tile by
// thread j % 256, so that the 64 lanes of a wave hold 64 consecutive rows and
// a wave-wide ballot of a predicate is exactly one 64-bit word of the
// selection bitmap (bit i <-> row i, LSB first; exec/bitmap.rs:36-45).
//
// The loads of tile t+1 are issued into registers before tile t is evaluated
// from LDS, so HBM traffic overlaps the evaluation.  Row ranges may start at any
// row: each slot's tile segment starts at the 16-byte boundary at or below its
// first byte ("phase" bytes early), which keeps every load aligned.
#pragma once

#include "tgp_internal.h"

namespace tgp {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct DevSlot {
  const uint8_t* data;
  uint64_t limit_bytes;  // bytes readable from data (column length * width)
  int32_t width;         // 1, 2, 4 or 8
  int32_t dtype;
  uint32_t lds_off;      // byte offset of the slot's segment in the stage area
  uint32_t chunk_pre;    // prefix count of 16-byte chunks before this slot
};

struct DevPred {
  int32_t slot, op, is_signed, _pad;
  int64_t lit;
};

struct DevPart {
  int32_t slot;
  int32_t divisor;   // 1 => no division
  uint32_t magic;    // 32-bit round-up multiplier for divisor (see host planner)
  int32_t shift;     // l - 1
  int64_t base;
  int64_t card;
  int64_t stride;
};

struct DevAgg {
  int32_t kind;      // tgp_agg_kind
  int32_t nf;
  int32_t fslot[TGP_MAX_FACTORS];
  int32_t _pad;
  int64_t coef;
  int64_t fa[TGP_MAX_FACTORS];
  int64_t fb[TGP_MAX_FACTORS];
};

// Everything a kernel needs about one query; passed by value (kernarg) or,
// for the batched executor, read from a device array.
struct DevPlan {
  uint64_t row_begin, row_end;
  uint64_t ntiles;
  DevSlot slots[TGP_MAX_COLS];
  DevPred preds[TGP_MAX_PREDS];
  DevPart parts[TGP_MAX_GROUP_PARTS];
  DevAgg aggs[TGP_MAX_AGGS];
  int32_t nslots, npreds, nparts, naggs;
  uint32_t nchunks;          // total 16-byte chunks per tile (all slots)
  uint32_t stage_bytes;      // LDS bytes of the stage area
  uint64_t ngroups;
  const uint64_t* sel_bitmap;
  // outputs
  unsigned long long* out_count;
  unsigned long long* out_sums;
  unsigned long long* out_oor;
  unsigned long long* out_bitmap;  // scan only
};

// ---- staging -------------------------------------------------------------------

// Chunk c (0 <= c < P.nchunks) belongs to the slot whose prefix is the largest
// one <= c.  Slots are few (<= 8), so a linear search is cheapest.
__device__ __forceinline__ int slot_of_chunk(const DevPlan& P, uint32_t c) {
  int s = 0;
#pragma unroll
  for (int i = 1; i < TGP_MAX_COLS; ++i)
    if (i < P.nslots && c >= P.slots[i].chunk_pre) s = i;
  return s;
}

// Loads this thread's share of tile `tile_row0` into registers.  `ldsaddr[i]`
// receives the LDS byte address the chunk goes to (0xffffffff => nothing).
template <int NST>
__device__ __forceinline__ void stage_load(const DevPlan& P, uint64_t tile_row0, uint64_t tile_rows,
                                           u32x4 (&reg)[NST], uint32_t (&ldsaddr)[NST]) {
#pragma unroll
  for (int i = 0; i < NST; ++i) {
    const uint32_t c = threadIdx.x + i * kBlock;
    ldsaddr[i] = 0xffffffffu;
    reg[i] = u32x4{0u, 0u, 0u, 0u};
    if (c < P.nchunks) {
      const int s = slot_of_chunk(P, c);
      const DevSlot& S = P.slots[s];
      const uint32_t cc = c - S.chunk_pre;
      const uint64_t first = tile_row0 * (uint64_t)S.width;
      const uint64_t g0 = first & ~(uint64_t)15;
      const uint32_t phase = (uint32_t)(first - g0);
      const uint64_t need = (uint64_t)phase + tile_rows * (uint64_t)S.width;  // bytes of this tile
      const uint64_t off = (uint64_t)cc * 16;
      if (off < need) {
        const uint64_t gb = g0 + off;
        ldsaddr[i] = S.lds_off + (uint32_t)off;
        if (gb + 16 <= S.limit_bytes) {
          reg[i] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(S.data + gb));
        } else if (gb < S.limit_bytes) {
          // Last chunk of the column: copy the valid bytes only.
          uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
          for (int b = 0; b < 16; ++b)
            if (gb + b < S.limit_bytes) w[b >> 2] |= (uint32_t)S.data[gb + b] << ((b & 3) * 8);
          reg[i] = u32x4{w[0], w[1], w[2], w[3]};
        }
      }
    }
  }
}

template <int NST>
__device__ __forceinline__ void stage_store(uint8_t* lds, const u32x4 (&reg)[NST],
                                            const uint32_t (&ldsaddr)[NST]) {
#pragma unroll
  for (int i = 0; i < NST; ++i)
    if (ldsaddr[i] != 0xffffffffu) *reinterpret_cast<u32x4*>(lds + ldsaddr[i]) = reg[i];
}

// Value of slot s at tile row j, widened to 64 bits by the slot's dtype.
__device__ __forceinline__ int64_t lds_value(const DevPlan& P, const uint8_t* lds, int s,
                                             uint64_t tile_row0, uint32_t j) {
  const DevSlot& S = P.slots[s];
  const uint32_t phase = (uint32_t)((tile_row0 * (uint64_t)S.width) & 15);
  const uint8_t* p = lds + S.lds_off + phase + j * (uint32_t)S.width;
  switch (S.dtype) {
    case TGP_I8: return (int64_t)*reinterpret_cast<const int8_t*>(p);
    case TGP_U8: return (int64_t)*p;
    case TGP_I16: return (int64_t)*reinterpret_cast<const int16_t*>(p);
    case TGP_U16: return (int64_t)*reinterpret_cast<const uint16_t*>(p);
    case TGP_I32: return (int64_t)*reinterpret_cast<const int32_t*>(p);
    case TGP_U32: return (int64_t)*reinterpret_cast<const uint32_t*>(p);
    default: return *reinterpret_cast<const int64_t*>(p);
  }
}

// bitmap.rs filter_* semantics on the 64-bit widened value.
__device__ __forceinline__ bool eval_pred(const DevPred& p, int64_t v) {
  if (p.is_signed) {
    const int64_t l = p.lit;
    switch (p.op) {
      case TGP_EQ: return v == l;
      case TGP_NE: return v != l;
      case TGP_LT: return v < l;
      case TGP_LE: return v <= l;
      case TGP_GT: return v > l;
      default: return v >= l;
    }
  } else {
    const uint64_t u = (uint64_t)v, l = (uint64_t)p.lit;
    switch (p.op) {
      case TGP_EQ: return u == l;
      case TGP_NE: return u != l;
      case TGP_LT: return u < l;
      case TGP_LE: return u <= l;
      case TGP_GT: return u > l;
      default: return u >= l;
    }
  }
}

__device__ __forceinline__ bool row_selected(const DevPlan& P, const uint8_t* lds,
                                             uint64_t tile_row0, uint32_t j) {
  bool sel = true;
  for (int p = 0; p < P.npreds; ++p) {
    const DevPred& pr = P.preds[p];
    sel = sel && eval_pred(pr, lds_value(P, lds, pr.slot, tile_row0, j));
  }
  if (P.sel_bitmap) {
    const uint64_t i = tile_row0 + j - P.row_begin;
    sel = sel && ((P.sel_bitmap[i >> 6] >> (i & 63)) & 1);
  }
  return sel;
}

// Dense group id; returns -1 when a key part is outside [0, card).
__device__ __forceinline__ int64_t group_id(const DevPlan& P, const uint8_t* lds,
                                            uint64_t tile_row0, uint32_t j) {
  int64_t gid = 0;
  for (int k = 0; k < P.nparts; ++k) {
    const DevPart& D = P.parts[k];
    int64_t v = lds_value(P, lds, D.slot, tile_row0, j);
    if (D.divisor > 1) {
      if (v >= 0 && v <= 0xffffffffll) {
        const uint32_t n = (uint32_t)v;
        const uint32_t t = __umulhi(D.magic, n);
        v = (int64_t)((t + ((n - t) >> 1)) >> D.shift);
      } else {
        v = v / D.divisor;  // reference semantics need v >= 0 here; kept total
      }
    }
    const int64_t part = v - D.base;
    if (part < 0 || part >= D.card) return -1;
    gid += part * D.stride;
  }
  return gid;
}

// Per-row value of aggregate a, wrapping 64-bit arithmetic (expr.rs binop).
__device__ __forceinline__ int64_t agg_value(const DevPlan& P, const DevAgg& A, const uint8_t* lds,
                                             uint64_t tile_row0, uint32_t j) {
  if (A.kind == TGP_AGG_COUNT_NONZERO)
    return lds_value(P, lds, A.fslot[0], tile_row0, j) != 0 ? 1 : 0;
  uint64_t v = (uint64_t)A.coef;
  for (int k = 0; k < A.nf; ++k) {
    const uint64_t x = (uint64_t)lds_value(P, lds, A.fslot[k], tile_row0, j);
    uint64_t f;
    if (A.fb[k] == 1) f = (uint64_t)A.fa[k] + x;
    else if (A.fb[k] == -1) f = (uint64_t)A.fa[k] - x;
    else f = (uint64_t)A.fa[k] + (uint64_t)A.fb[k] * x;
    v = (k == 0 && A.coef == 1) ? f : v * f;
  }
  return (int64_t)v;
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

}  // namespace tgp
