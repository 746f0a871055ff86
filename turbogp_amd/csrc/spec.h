// Shape-specialised scan-aggregate kernels: shapes, arguments, device helpers.
//
// The generic tile engine (tile_engine.h) interprets any query.  For the query
// shapes that carry the workloads (the Q1 family: date filter + low-cardinality
// GROUP BY + price/discount/tax sums; the Q6 selective scan-sum) this file
// compiles the query's structure into the kernel: the column dtypes, which
// column each predicate / key part / factor reads, the factor forms.  Literals,
// bounds, key bases and coefficients stay runtime arguments.  The planner
// derives a Shape from any tgp_agg_query and launches the matching kernel, or
// falls back to the generic engine.
//
// Kernel design (HBM-bound streaming, no LDS staging):
//   - a wave owns slices of 256 rows; lane l holds rows 4l..4l+3 of a slice, so
//     each column is one contiguous 16-byte-multiple vector load per lane and
//     a wave's loads cover one contiguous span (i64: 2 KiB, u8: 256 B);
//   - predicates are closed integer ranges computed exactly on the host
//     (compared in 32 bits for narrow columns), branch free;
//   - group ids index a wave-private histogram in LDS updated with 64-bit LDS
//     atomics (one atomic per row per aggregate, independent of the group
//     count); the histogram is folded into lane-owned exact int128
//     accumulators after every slice, so its 64-bit words cannot overflow
//     (at most 256 values of magnitude < 2^55; larger values take an exact
//     slow path for that slice);
//   - larger group counts use a workgroup LDS table with carry-exact atomics.
#pragma once

#ifndef __HIPCC_RTC__
#include <cstring>
#include <vector>
#endif

#include "plan.h"

namespace tgp {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kSR = 4;                 // rows per lane
constexpr int kSliceRows = kWave * kSR;  // 256 rows per wave slice
constexpr int kMaxSpecPreds = 4;
constexpr int kHistWords = 64;         // per-wave histogram words (groups * (1 + naggs))

// Factor forms: value = fa + fb * x with fb = 1 (kFPlus), -1 (kFMinus); kFIdent
// is x itself (fa = 0, fb = 1); kFGeneral multiplies.
enum FactorForm : int { kFGeneral = 0, kFPlus = 1, kFMinus = 2, kFIdent = 3 };

struct Shape {
  int nslots;
  int dt[TGP_MAX_COLS];
  int npreds;
  int pslot[kMaxSpecPreds];
  int pneg[kMaxSpecPreds];  // 1: row passes when the value is OUTSIDE [lo, hi]
  int nparts;
  int gslot[TGP_MAX_GROUP_PARTS];
  int gdiv[TGP_MAX_GROUP_PARTS];  // 1: the key part divides by a runtime divisor
  int naggs;
  int kind[TGP_MAX_AGGS];
  int coef1[TGP_MAX_AGGS];  // 1: coefficient is 1
  int nf[TGP_MAX_AGGS];
  int fslot[TGP_MAX_AGGS][TGP_MAX_FACTORS];
  int fform[TGP_MAX_AGGS][TGP_MAX_FACTORS];
  int table;                // 0: wave-private histogram; 1: workgroup LDS table
  int has_sel;              // 1: an input selection bitmap is ANDed in
};

struct SpecArgs {
  uint64_t row_begin, row_end, base0, nslices;
  const uint8_t* col[TGP_MAX_COLS];
  uint64_t col_len[TGP_MAX_COLS];
  int64_t plo[kMaxSpecPreds], phi[kMaxSpecPreds];
  int64_t gbase[TGP_MAX_GROUP_PARTS], gcard[TGP_MAX_GROUP_PARTS], gstride[TGP_MAX_GROUP_PARTS];
  uint32_t gmagic[TGP_MAX_GROUP_PARTS];
  int32_t gshift[TGP_MAX_GROUP_PARTS];
  int32_t gdivisor[TGP_MAX_GROUP_PARTS];
  int32_t _pad;
  int64_t coef[TGP_MAX_AGGS];
  int64_t fa[TGP_MAX_AGGS][TGP_MAX_FACTORS];
  int64_t fb[TGP_MAX_AGGS][TGP_MAX_FACTORS];
  uint64_t ngroups;
  uint32_t words;  // groups * (1 + naggs)
  uint32_t _pad2;
  const uint64_t* sel_bitmap;
  unsigned long long* out_count;
  unsigned long long* out_sums;
  unsigned long long* out_oor;
};

// ---- device helpers -------------------------------------------------------------------

constexpr int width_of(int dt) {
  return (dt == TGP_I8 || dt == TGP_U8) ? 1 : (dt == TGP_I16 || dt == TGP_U16) ? 2
         : (dt == TGP_I32 || dt == TGP_U32) ? 4 : 8;
}

// A vector load, streaming (non-temporal: the line is not kept in the caches)
// or cached (for data other waves read again soon).
template <bool NT, typename T>
__device__ __forceinline__ T vload(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// The lane's 4 consecutive values of a column, widened to int64 (u64 values
// keep their bit pattern).  `fast`: the whole slice lies inside the column.
template <int DT, bool NT = true>
__device__ __forceinline__ void load4(const uint8_t* col, uint64_t len, uint64_t row, bool fast,
                                      int64_t (&x)[kSR]) {
  constexpr int W = width_of(DT);
  if (fast) {
    const uint8_t* p = col + row * W;
    if constexpr (W == 8) {
      const u32x4 a = vload<NT>(reinterpret_cast<const u32x4*>(p));
      const u32x4 b = vload<NT>(reinterpret_cast<const u32x4*>(p) + 1);
      x[0] = (int64_t)(((uint64_t)a.y << 32) | a.x);
      x[1] = (int64_t)(((uint64_t)a.w << 32) | a.z);
      x[2] = (int64_t)(((uint64_t)b.y << 32) | b.x);
      x[3] = (int64_t)(((uint64_t)b.w << 32) | b.z);
    } else if constexpr (W == 4) {
      const u32x4 a = vload<NT>(reinterpret_cast<const u32x4*>(p));
      const uint32_t v[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        x[r] = DT == TGP_I32 ? (int64_t)(int32_t)v[r] : (int64_t)v[r];
    } else if constexpr (W == 2) {
      typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
      const u32x2 a = vload<NT>(reinterpret_cast<const u32x2*>(p));
      const uint32_t v[4] = {a.x & 0xffffu, a.x >> 16, a.y & 0xffffu, a.y >> 16};
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        x[r] = DT == TGP_I16 ? (int64_t)(int16_t)v[r] : (int64_t)v[r];
    } else {
      const uint32_t a = vload<NT>(reinterpret_cast<const uint32_t*>(p));
#pragma unroll
      for (int r = 0; r < kSR; ++r) {
        const uint32_t b = (a >> (8 * r)) & 0xffu;
        x[r] = DT == TGP_I8 ? (int64_t)(int8_t)b : (int64_t)b;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      const uint64_t i = row + r;
      int64_t v = 0;
      if (i < len) {
        if constexpr (DT == TGP_I8) v = ((const int8_t*)col)[i];
        else if constexpr (DT == TGP_U8) v = col[i];
        else if constexpr (DT == TGP_I16) v = ((const int16_t*)col)[i];
        else if constexpr (DT == TGP_U16) v = ((const uint16_t*)col)[i];
        else if constexpr (DT == TGP_I32) v = ((const int32_t*)col)[i];
        else if constexpr (DT == TGP_U32) v = ((const uint32_t*)col)[i];
        else v = ((const int64_t*)col)[i];
      }
      x[r] = v;
    }
  }
}

// x in [lo, hi] in the column's value domain (bounds prepared on the host).
template <int DT>
__device__ __forceinline__ bool in_range(int64_t x, int64_t lo, int64_t hi) {
  if constexpr (DT == TGP_U64) {
    const uint64_t u = (uint64_t)x;
    return u >= (uint64_t)lo && u <= (uint64_t)hi;
  } else if constexpr (DT == TGP_U32) {
    const uint32_t u = (uint32_t)x;
    return u >= (uint32_t)lo && u <= (uint32_t)hi;
  } else if constexpr (width_of(DT) <= 4) {
    const int32_t v = (int32_t)x;
    return v >= (int32_t)lo && v <= (int32_t)hi;
  } else {
    return x >= lo && x <= hi;
  }
}

template <Shape S>
struct SliceVals {
  int64_t x[TGP_MAX_COLS][kSR];
};

template <Shape S, int I, bool NT = true>
__device__ __forceinline__ void load_all(const SpecArgs& A, uint64_t row, bool fast,
                                         SliceVals<S>& v) {
  if constexpr (I < S.nslots) {
    load4<S.dt[I], NT>(A.col[I], A.col_len[I], row, fast, v.x[I]);
    load_all<S, I + 1, NT>(A, row, fast, v);
  }
}

// Evaluates one slice for this lane: selection bits, group ids, row values.
template <Shape S>
__device__ __forceinline__ uint32_t eval_slice(const SpecArgs& A, uint64_t row,
                                               const SliceVals<S>& v, int32_t (&gid)[kSR],
                                               uint32_t& oor) {
  // rows of the query's range; a slice wholly inside it (all but the first and
  // last slice of a query) skips the per-row compares
  uint32_t sel = (1u << kSR) - 1;
  const uint64_t slice0 = row - (threadIdx.x & 63) * kSR;
  if (slice0 < A.row_begin || slice0 + kSliceRows > A.row_end) {
    sel = 0;
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if (row + r >= A.row_begin && row + r < A.row_end) sel |= 1u << r;
  }
  if constexpr (S.has_sel) {
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      if (!((sel >> r) & 1)) continue;
      const uint64_t i = row + r - A.row_begin;
      if (!((A.sel_bitmap[i >> 6] >> (i & 63)) & 1)) sel &= ~(1u << r);
    }
  }
  static_for<S.npreds>([&](auto P) {
    constexpr int p = decltype(P)::value;
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      const bool in = in_range<S.dt[S.pslot[p]]>(v.x[S.pslot[p]][r], A.plo[p], A.phi[p]);
      if (in == (S.pneg[p] != 0)) sel &= ~(1u << r);
    }
  });
  // group id: sum of part * stride over the key parts, branch free; a part
  // outside [0, card) makes the row out of range.  ngroups <= 2^20, so parts and
  // strides fit the 24-bit multiplier.
  uint32_t ok = (1u << kSR) - 1;
#pragma unroll
  for (int r = 0; r < kSR; ++r) gid[r] = 0;
  static_for<S.nparts>([&](auto K) {
    constexpr int k = decltype(K)::value;
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      int64_t x = v.x[S.gslot[k]][r];
      if constexpr (S.gdiv[k]) {
        if (x >= 0 && x <= 0xffffffffll) {
          const uint32_t n = (uint32_t)x;
          const uint32_t q = __umulhi(A.gmagic[k], n);
          x = (int64_t)((q + ((n - q) >> 1)) >> A.gshift[k]);
        } else {
          x = x / A.gdivisor[k];
        }
      }
      const uint64_t part = (uint64_t)(x - A.gbase[k]);
      if (part >= (uint64_t)A.gcard[k]) ok &= ~(1u << r);
      gid[r] += (int32_t)__umul24((uint32_t)part, (uint32_t)A.gstride[k]);
    }
  });
  if constexpr (S.nparts > 0) {
    const uint32_t bad = sel & ~ok;
    if (bad) {
      oor += __popc(bad);
      sel &= ok;
    }
#pragma unroll
    for (int r = 0; r < kSR; ++r) gid[r] = ((ok >> r) & 1) ? gid[r] : 0;
  }
  return sel;
}

template <Shape S, int a>
__device__ __forceinline__ void agg_values(const SpecArgs& A, const SliceVals<S>& v,
                                           uint64_t (&out)[kSR]) {
  if constexpr (S.kind[a] == TGP_AGG_COUNT_NONZERO) {
#pragma unroll
    for (int r = 0; r < kSR; ++r) out[r] = v.x[S.fslot[a][0]][r] != 0;
  } else {
#pragma unroll
    for (int r = 0; r < kSR; ++r) {
      uint64_t p = (uint64_t)A.coef[a];
      static_for<S.nf[a]>([&](auto K) {
        constexpr int k = decltype(K)::value;
        const uint64_t x = (uint64_t)v.x[S.fslot[a][k]][r];
        uint64_t f;
        if constexpr (S.fform[a][k] == kFIdent) f = x;
        else if constexpr (S.fform[a][k] == kFPlus) f = (uint64_t)A.fa[a][k] + x;
        else if constexpr (S.fform[a][k] == kFMinus) f = (uint64_t)A.fa[a][k] - x;
        else f = (uint64_t)A.fa[a][k] + (uint64_t)A.fb[a][k] * x;
        if constexpr (k == 0 && S.coef1[a] != 0) p = f;
        else p = p * f;
      });
      out[r] = p;
    }
  }
}

// |v| >= 2^55 for some lane/row: the slice's histogram could overflow.
__device__ __forceinline__ bool big_value(uint64_t v) {
  const int64_t s = (int64_t)v;
  const uint64_t m = (uint64_t)(s ^ (s >> 63));
  return (m >> 55) != 0;
}

// MIN / MAX aggregates of shape s as a plan's minmax_kinds (3 bits per aggregate:
// its kind, 0 for the sums).  MIN / MAX accumulate an order code (mm_encode) with
// an unsigned maximum whose identity is 0, and the result is decoded afterwards.
constexpr uint32_t shape_minmax_kinds(const Shape& s) {
  uint32_t k = 0;
  for (int a = 0; a < s.naggs; ++a)
    if (is_minmax(s.kind[a])) k |= (uint32_t)s.kind[a] << (3 * a);
  return k;
}

template <Shape S, int a>
__device__ __forceinline__ void accumulate_aggs(const SpecArgs& A, const SliceVals<S>& v,
                                                uint32_t sel, const int32_t (&gid)[kSR],
                                                unsigned long long* hist, bool& big) {
  if constexpr (a < S.naggs && is_minmax(S.kind[a])) {
    uint64_t val[kSR];
    agg_values<S, a>(A, v, val);
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if ((sel >> r) & 1) {
        const unsigned long long c = mm_encode(S.kind[a], val[r]);
        if constexpr (S.table == 0) atomicMax(&hist[gid[r] * (1 + S.naggs) + 1 + a], c);
        else atomicMax(&hist[A.ngroups + ((uint64_t)gid[r] * S.naggs + a) * 2], c);
      }
    accumulate_aggs<S, a + 1>(A, v, sel, gid, hist, big);
  } else if constexpr (a < S.naggs) {
    uint64_t val[kSR];
    agg_values<S, a>(A, v, val);
    bool b = false;
#pragma unroll
    for (int r = 0; r < kSR; ++r) b |= ((sel >> r) & 1) && big_value(val[r]);
    if (__ballot(b)) {
      // exact slow path for this slice and aggregate: straight to the result
      big = true;
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1)
          atomic_add128_i64(&A.out_sums[((uint64_t)gid[r] * S.naggs + a) * 2], (int64_t)val[r]);
    } else {
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1) {
          if constexpr (S.table == 0)
            atomicAdd(&hist[gid[r] * (1 + S.naggs) + 1 + a], (unsigned long long)val[r]);
          else
            atomic_add128_i64(&hist[A.ngroups + ((uint64_t)gid[r] * S.naggs + a) * 2],
                              (int64_t)val[r]);
        }
    }
    accumulate_aggs<S, a + 1>(A, v, sel, gid, hist, big);
  }
}

// ---- the compiled shapes ------------------------------------------------------------

constexpr Shape shape_q1(int group /*0 ab, 1 a, 2 b, 3 month, 4 none*/, int has_sel) {
  Shape s{};
  s.nslots = 7;
  const int dts[7] = {TGP_I32, TGP_I64, TGP_I64, TGP_I64, TGP_I64, TGP_U8, TGP_U8};
  for (int i = 0; i < 7; ++i) s.dt[i] = dts[i];
  s.npreds = 1;
  s.pslot[0] = 0;
  if (group == 0) { s.nparts = 2; s.gslot[0] = 5; s.gslot[1] = 6; }
  if (group == 1) { s.nparts = 1; s.gslot[0] = 5; }
  if (group == 2) { s.nparts = 1; s.gslot[0] = 6; }
  if (group == 3) { s.nparts = 1; s.gslot[0] = 0; s.gdiv[0] = 1; }
  s.naggs = 4;
  for (int a = 0; a < 4; ++a) { s.kind[a] = TGP_AGG_SUM; s.coef1[a] = 1; }
  s.nf[0] = 1; s.fslot[0][0] = 1; s.fform[0][0] = kFIdent;
  s.nf[1] = 1; s.fslot[1][0] = 2; s.fform[1][0] = kFIdent;
  s.nf[2] = 2; s.fslot[2][0] = 2; s.fform[2][0] = kFIdent; s.fslot[2][1] = 3; s.fform[2][1] = kFMinus;
  s.nf[3] = 3; s.fslot[3][0] = 2; s.fform[3][0] = kFIdent; s.fslot[3][1] = 3; s.fform[3][1] = kFMinus;
  s.fslot[3][2] = 4; s.fform[3][2] = kFPlus;
  s.table = group == 3 ? 1 : 0;
  s.has_sel = has_sel;
  return s;
}

constexpr Shape shape_q6() {
  Shape s{};
  s.nslots = 4;
  const int dts[4] = {TGP_I32, TGP_I64, TGP_I64, TGP_I64};
  for (int i = 0; i < 4; ++i) s.dt[i] = dts[i];
  s.npreds = 3;
  s.pslot[0] = 0; s.pslot[1] = 1; s.pslot[2] = 2;
  s.naggs = 1;
  s.kind[0] = TGP_AGG_SUM;
  s.coef1[0] = 1;
  s.nf[0] = 2; s.fslot[0][0] = 3; s.fform[0][0] = kFIdent; s.fslot[0][1] = 1; s.fform[0][1] = kFIdent;
  return s;
}

inline constexpr Shape kShapes[] = {
    shape_q1(0, 0), shape_q1(1, 0), shape_q1(2, 0), shape_q1(3, 0), shape_q1(4, 0), shape_q6(),
};
inline constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

#ifndef __HIPCC_RTC__
inline bool shape_eq(const Shape& a, const Shape& b) {
  return std::memcmp(&a, &b, sizeof(Shape)) == 0;
}

// Builds the shape and arguments of q; false when q has no specialised form.
bool extract_shape(const tgp_agg_query* q, Shape* s, SpecArgs* A);
#endif

}  // namespace tgp
