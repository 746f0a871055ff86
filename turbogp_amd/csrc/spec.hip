// Shape-specialised scan-aggregate kernels.
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
#include <type_traits>
#include <utility>
#include <vector>

#include "plan.h"

namespace tgp {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Calls f(std::integral_constant<int, I>{}) for I = 0 .. N-1 (compile-time loop).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

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

// The lane's 4 consecutive values of a column, widened to int64 (u64 values
// keep their bit pattern).  `fast`: the whole slice lies inside the column.
template <int DT>
__device__ __forceinline__ void load4(const uint8_t* col, uint64_t len, uint64_t row, bool fast,
                                      int64_t (&x)[kSR]) {
  constexpr int W = width_of(DT);
  if (fast) {
    const uint8_t* p = col + row * W;
    if constexpr (W == 8) {
      const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
      const u32x4 b = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p) + 1);
      x[0] = (int64_t)(((uint64_t)a.y << 32) | a.x);
      x[1] = (int64_t)(((uint64_t)a.w << 32) | a.z);
      x[2] = (int64_t)(((uint64_t)b.y << 32) | b.x);
      x[3] = (int64_t)(((uint64_t)b.w << 32) | b.z);
    } else if constexpr (W == 4) {
      const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
      const uint32_t v[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        x[r] = DT == TGP_I32 ? (int64_t)(int32_t)v[r] : (int64_t)v[r];
    } else if constexpr (W == 2) {
      typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
      const u32x2 a = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
      const uint32_t v[4] = {a.x & 0xffffu, a.x >> 16, a.y & 0xffffu, a.y >> 16};
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        x[r] = DT == TGP_I16 ? (int64_t)(int16_t)v[r] : (int64_t)v[r];
    } else {
      const uint32_t a = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
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

template <Shape S, int I>
__device__ __forceinline__ void load_all(const SpecArgs& A, uint64_t row, bool fast,
                                         SliceVals<S>& v) {
  if constexpr (I < S.nslots) {
    load4<S.dt[I]>(A.col[I], A.col_len[I], row, fast, v.x[I]);
    load_all<S, I + 1>(A, row, fast, v);
  }
}

// Evaluates one slice for this lane: selection bits, group ids, row values.
template <Shape S>
__device__ __forceinline__ uint32_t eval_slice(const SpecArgs& A, uint64_t row,
                                               const SliceVals<S>& v, int32_t (&gid)[kSR],
                                               uint32_t& oor) {
  uint32_t sel = 0;
#pragma unroll
  for (int r = 0; r < kSR; ++r)
    if (row + r >= A.row_begin && row + r < A.row_end) sel |= 1u << r;
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
      const int64_t part = x - A.gbase[k];
      if (part < 0 || part >= A.gcard[k]) gid[r] = -1;
      else if (gid[r] >= 0) gid[r] += (int32_t)(part * A.gstride[k]);
    }
  });
  if constexpr (S.nparts > 0) {
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if (((sel >> r) & 1) && gid[r] < 0) {
        ++oor;
        sel &= ~(1u << r);
      }
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

template <Shape S, int a>
__device__ __forceinline__ void accumulate_aggs(const SpecArgs& A, const SliceVals<S>& v,
                                                uint32_t sel, const int32_t (&gid)[kSR],
                                                unsigned long long* hist, bool& big) {
  if constexpr (a < S.naggs) {
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

template <Shape S>
__global__ void __launch_bounds__(kBlock) k_spec(SpecArgs A) {
  extern __shared__ __align__(16) unsigned long long lmem[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // table 0: per-wave histograms of kHistWords words; table 1: one workgroup table
  // (counts[ngroups], then (lo, hi) sums) followed by nothing else.
  unsigned long long* hist = S.table == 0 ? lmem + wave * kHistWords : lmem;
  if constexpr (S.table == 0) {
    hist[lane] = 0;
  } else {
    const uint64_t words = A.ngroups * (1 + 2 * (uint64_t)S.naggs);
    for (uint64_t i = threadIdx.x; i < words; i += kBlock) lmem[i] = 0;
    __syncthreads();
  }
  // lane-owned exact accumulator of histogram word `lane` (table 0)
  uint64_t acc_lo = 0;
  int64_t acc_hi = 0;
  uint32_t oor = 0;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t gw = (uint64_t)blockIdx.x * (kBlock / kWave) + wave;
  uint64_t min_len = A.col_len[0];
#pragma unroll
  for (int s = 1; s < S.nslots; ++s) min_len = A.col_len[s] < min_len ? A.col_len[s] : min_len;
  for (uint64_t k = gw; k < A.nslices; k += nwaves) {
    const uint64_t slice0 = A.base0 + k * kSliceRows;
    const uint64_t row = slice0 + lane * kSR;
    const bool fast = slice0 + kSliceRows <= min_len;
    SliceVals<S> v;
    load_all<S, 0>(A, row, fast, v);
    int32_t gid[kSR];
    const uint32_t sel = eval_slice<S>(A, row, v, gid, oor);
    if (!__ballot(sel != 0)) continue;
#pragma unroll
    for (int r = 0; r < kSR; ++r)
      if ((sel >> r) & 1) {
        if constexpr (S.table == 0) atomicAdd(&hist[gid[r] * (1 + S.naggs)], 1ull);
        else atomicAdd(&hist[gid[r]], 1ull);
      }
    bool big = false;
    accumulate_aggs<S, 0>(A, v, sel, gid, hist, big);
    if constexpr (S.table == 0) {
      // fold the slice's histogram into the lane-owned int128 accumulators
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      const unsigned long long w = hist[lane];
      hist[lane] = 0;
      const uint64_t s = acc_lo + w;
      // count words are non-negative; sum words are signed 64-bit partials
      const bool is_count = (lane % (1 + S.naggs)) == 0;
      acc_hi += (int64_t)(s < acc_lo) - ((!is_count && (int64_t)w < 0) ? 1 : 0);
      acc_lo = s;
      __builtin_amdgcn_wave_barrier();
    }
  }
  // ---- flush to the result ----
  if constexpr (S.table == 0) {
    // workgroup reduction of the 4 waves' lane accumulators, then device atomics
    __syncthreads();
    unsigned long long* red = lmem + 4 * kHistWords;  // [wave][lane] (lo, hi)
    red[(wave * 64 + lane) * 2] = acc_lo;
    red[(wave * 64 + lane) * 2 + 1] = (unsigned long long)acc_hi;
    __syncthreads();
    if (threadIdx.x < A.words) {
      uint64_t lo = 0;
      int64_t hi = 0;
      for (int w = 0; w < kBlock / kWave; ++w) {
        const uint64_t l2 = red[(w * 64 + threadIdx.x) * 2];
        const uint64_t s = lo + l2;
        hi += (int64_t)red[(w * 64 + threadIdx.x) * 2 + 1] + (int64_t)(s < lo);
        lo = s;
      }
      const uint32_t g = threadIdx.x / (1 + S.naggs), a = threadIdx.x % (1 + S.naggs);
      if (a == 0) {
        if (lo) atomicAdd(&A.out_count[g], (unsigned long long)lo);
      } else {
        atomic_add128(&A.out_sums[((uint64_t)g * S.naggs + (a - 1)) * 2], lo, hi);
      }
    }
  } else {
    __syncthreads();
    for (uint64_t g = threadIdx.x; g < A.ngroups; g += kBlock) {
      const unsigned long long c = lmem[g];
      if (!c) continue;
      atomicAdd(&A.out_count[g], c);
      for (int a = 0; a < S.naggs; ++a) {
        const uint64_t i = A.ngroups + (g * S.naggs + a) * 2;
        atomic_add128(&A.out_sums[(g * S.naggs + a) * 2], lmem[i], (int64_t)lmem[i + 1]);
      }
    }
  }
  if (oor && A.out_oor) atomicAdd(A.out_oor, (unsigned long long)oor);
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

constexpr Shape kShapes[] = {
    shape_q1(0, 0), shape_q1(1, 0), shape_q1(2, 0), shape_q1(3, 0), shape_q1(4, 0), shape_q6(),
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

bool shape_eq(const Shape& a, const Shape& b) { return std::memcmp(&a, &b, sizeof(Shape)) == 0; }

template <int I>
int32_t launch_shape(tgp_ctx* ctx, const Shape& s, const SpecArgs& A, hipStream_t st,
                     bool* launched) {
  if constexpr (I < kNumShapes) {
    if (shape_eq(s, kShapes[I])) {
      constexpr Shape S = kShapes[I];
      auto kern = k_spec<S>;
      uint32_t lds = S.table == 0 ? (4 * kHistWords + 4 * 64 * 2) * 8
                                  : (uint32_t)(A.ngroups * (1 + 2 * S.naggs) * 8);
      int per_cu = 0;
      TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, lds));
      if (per_cu < 1) per_cu = 1;
      const int64_t want = ctx->tune_wg_per_cu.load();
      if (want > 0 && want < per_cu) per_cu = (int)want;
      uint64_t grid = (uint64_t)persistent_grid(ctx, per_cu);
      const uint64_t need = (A.nslices + 3) / 4;
      if (grid > need) grid = need;
      const int64_t gmax = ctx->tune_grid_max.load();
      if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
      if (grid == 0) grid = 1;
      kern<<<dim3((unsigned)grid), dim3(kBlock), lds, st>>>(A);
      TGP_HIP(hipGetLastError());
      *launched = true;
      return TGP_OK;
    }
    return launch_shape<I + 1>(ctx, s, A, st, launched);
  } else {
    return TGP_OK;
  }
}

// ---- host: shape extraction ----------------------------------------------------------

typedef __int128 i128;

// A set of values as at most two closed intervals, or "all but one point".
struct ValueSet {
  int n;           // number of intervals (0 = empty)
  i128 lo[2], hi[2];
  bool not_point;  // the set is everything except lo[0] (n unused)
};

static void dtype_range(int dt, i128* mn, i128* mx) {
  switch (dt) {
    case TGP_I8: *mn = -128; *mx = 127; return;
    case TGP_U8: *mn = 0; *mx = 255; return;
    case TGP_I16: *mn = -32768; *mx = 32767; return;
    case TGP_U16: *mn = 0; *mx = 65535; return;
    case TGP_I32: *mn = -(i128)2147483648ll; *mx = 2147483647; return;
    case TGP_U32: *mn = 0; *mx = 4294967295ll; return;
    case TGP_I64: *mn = -((i128)1 << 63); *mx = ((i128)1 << 63) - 1; return;
    default: *mn = 0; *mx = ((i128)1 << 64) - 1; return;  // U64: value domain = bit pattern
  }
}

// {x in the dtype's value range : x op lit}, where the value is compared as
// int64 (is_signed) or as its uint64 bit pattern.  Values of signed columns are
// sign-extended; u64 values are their own bit pattern.
static ValueSet pred_set(int dt, int op, int is_signed, int64_t lit) {
  i128 mn, mx;
  dtype_range(dt, &mn, &mx);
  const i128 two64 = (i128)1 << 64, two63 = (i128)1 << 63;
  const i128 L = is_signed ? (i128)lit : (i128)(uint64_t)lit;
  const i128 clo = is_signed ? -two63 : 0, chi = is_signed ? two63 - 1 : two64 - 1;
  i128 a = clo, b = chi;
  switch (op) {
    case TGP_EQ: case TGP_NE: a = L; b = L; break;
    case TGP_LT: b = L - 1; break;
    case TGP_LE: b = L; break;
    case TGP_GT: a = L + 1; break;
    default: a = L; break;
  }
  // preimage of the compare-domain interval [a, b] in the value domain
  i128 lo[2] = {0, 0}, hi[2] = {-1, -1};
  if (dt == TGP_U64) {
    if (!is_signed) { lo[0] = a; hi[0] = b; }
    else {  // v < 2^63 compares as v, v >= 2^63 as v - 2^64
      lo[0] = a > 0 ? a : 0; hi[0] = b < two63 - 1 ? b : two63 - 1;
      lo[1] = (a + two64) > two63 ? (a + two64) : two63;
      hi[1] = (b + two64) < two64 - 1 ? (b + two64) : two64 - 1;
    }
  } else if (is_signed) {
    lo[0] = a; hi[0] = b;
  } else {  // v >= 0 compares as v, v < 0 as v + 2^64
    lo[0] = a > 0 ? a : 0; hi[0] = b;
    lo[1] = a - two64; hi[1] = b - two64 < -1 ? b - two64 : -1;
  }
  ValueSet r{};
  for (int k = 0; k < 2; ++k) {
    i128 l = lo[k] < mn ? mn : lo[k], h = hi[k] > mx ? mx : hi[k];
    if (l <= h) { r.lo[r.n] = l; r.hi[r.n] = h; ++r.n; }
  }
  if (op == TGP_NE) {
    ValueSet ne{};
    if (r.n == 0) { ne.n = 1; ne.lo[0] = mn; ne.hi[0] = mx; }   // nothing equal: all pass
    else { ne.not_point = true; ne.lo[0] = r.lo[0]; }          // EQ set is one point
    return ne;
  }
  return r;
}

// Intersection of two interval sets of at most two intervals each; false when
// the result needs more than two intervals.
static bool intersect(const ValueSet& x, const ValueSet& y, ValueSet* out) {
  ValueSet r{};
  for (int i = 0; i < x.n; ++i)
    for (int j = 0; j < y.n; ++j) {
      const i128 l = x.lo[i] > y.lo[j] ? x.lo[i] : y.lo[j];
      const i128 h = x.hi[i] < y.hi[j] ? x.hi[i] : y.hi[j];
      if (l > h) continue;
      if (r.n == 2) return false;
      r.lo[r.n] = l;
      r.hi[r.n] = h;
      ++r.n;
    }
  *out = r;
  return true;
}

static int form_of(int64_t fa, int64_t fb) {
  if (fb == 1) return fa == 0 ? kFIdent : kFPlus;
  if (fb == -1) return kFMinus;
  return kFGeneral;
}

// Builds the shape and arguments of q; false when q has no specialised form.
static bool extract_shape(const tgp_agg_query* q, Shape* s, SpecArgs* A) {
  std::memset(s, 0, sizeof(*s));
  std::memset(A, 0, sizeof(*A));
  s->nslots = q->ncols;
  for (int i = 0; i < q->ncols; ++i) {
    s->dt[i] = q->cols[i].dtype;
    A->col[i] = static_cast<const uint8_t*>(q->cols[i].data);
    A->col_len[i] = q->cols[i].len;
  }
  // predicates: one merged range per column (ascending column), then the
  // single-point exclusions
  bool empty = false;
  std::vector<std::pair<int, i128>> excl;
  for (int c = 0; c < q->ncols; ++c) {
    i128 mn, mx;
    dtype_range(q->cols[c].dtype, &mn, &mx);
    ValueSet acc{};
    acc.n = 1;
    acc.lo[0] = mn;
    acc.hi[0] = mx;
    for (int p = 0; p < q->npreds; ++p) {
      if (q->preds[p].col != c) continue;
      const ValueSet v =
          pred_set(q->cols[c].dtype, q->preds[p].op, q->preds[p].is_signed, q->preds[p].lit);
      if (v.not_point) {
        excl.push_back({c, v.lo[0]});
        continue;
      }
      if (!intersect(acc, v, &acc)) return false;
    }
    if (acc.n == 0) {
      empty = true;
    } else if (acc.n == 2) {
      return false;  // not one range
    } else if (!(acc.lo[0] == mn && acc.hi[0] == mx)) {
      if (s->npreds >= kMaxSpecPreds) return false;
      const int p = s->npreds++;
      s->pslot[p] = c;
      A->plo[p] = (int64_t)(uint64_t)acc.lo[0];  // u64 columns: bit pattern
      A->phi[p] = (int64_t)(uint64_t)acc.hi[0];
    }
  }
  for (auto& e : excl) {
    if (s->npreds >= kMaxSpecPreds) return false;
    const int p = s->npreds++;
    s->pslot[p] = e.first;
    s->pneg[p] = 1;
    A->plo[p] = A->phi[p] = (int64_t)(uint64_t)e.second;
  }
  if (empty) {  // a predicate selects nothing: one never-true range on column 0
    std::memset(s->pslot, 0, sizeof(s->pslot));
    std::memset(s->pneg, 0, sizeof(s->pneg));
    s->npreds = 1;
    A->plo[0] = 1;
    A->phi[0] = 0;
  }
  // narrow compares use 32-bit bounds: clamp to the dtype range (already done) and
  // keep never-true ranges never-true after the truncation
  for (int p = 0; p < s->npreds; ++p) {
    const int w = width_of(s->dt[s->pslot[p]]);
    if (w <= 4 && !s->pneg[p] && A->plo[p] > A->phi[p]) { A->plo[p] = 1; A->phi[p] = 0; }
  }
  uint64_t ng = 1;
  s->nparts = q->nparts;
  for (int k = 0; k < q->nparts; ++k) {
    const tgp_group_part& g = q->parts[k];
    s->gslot[k] = g.col;
    s->gdiv[k] = g.divisor > 1;
    A->gdivisor[k] = g.divisor;
    if (g.divisor > 1) divisor_magic((uint32_t)g.divisor, &A->gmagic[k], &A->gshift[k]);
    A->gbase[k] = g.base;
    A->gcard[k] = g.card;
    A->gstride[k] = (int64_t)ng;
    ng *= (uint64_t)g.card;
  }
  if (ng > (1u << 20)) return false;
  A->ngroups = ng;
  s->naggs = q->naggs;
  for (int a = 0; a < q->naggs; ++a) {
    const tgp_agg& G = q->aggs[a];
    s->kind[a] = G.kind;
    if (G.kind == TGP_AGG_COUNT_NONZERO) {
      s->nf[a] = 1;
      s->fslot[a][0] = G.fcol[0];
      continue;
    }
    s->coef1[a] = G.coef == 1;
    A->coef[a] = G.coef;
    s->nf[a] = G.nfactors;
    for (int k = 0; k < G.nfactors; ++k) {
      s->fslot[a][k] = G.fcol[k];
      s->fform[a][k] = form_of(G.fa[k], G.fb[k]);
      A->fa[a][k] = G.fa[k];
      A->fb[a][k] = G.fb[k];
    }
  }
  A->words = (uint32_t)(ng * (1 + (uint64_t)q->naggs));
  s->table = A->words <= (uint32_t)kHistWords ? 0 : 1;
  if (s->table == 1 && ng * (1 + 2 * (uint64_t)q->naggs) * 8 > kLdsTableBytes) return false;
  s->has_sel = q->d_sel_bitmap != nullptr;
  A->sel_bitmap = q->d_sel_bitmap;
  A->row_begin = q->row_begin;
  A->row_end = q->row_end;
  A->base0 = q->row_begin & ~(uint64_t)(kSliceRows - 1);
  A->nslices = q->row_end > q->row_begin
                   ? (q->row_end - A->base0 + kSliceRows - 1) / kSliceRows : 0;
  return true;
}

// Index of the compiled shape that runs q, or -1.
int32_t spec_shape_index(const tgp_agg_query* q) {
  Shape s;
  SpecArgs A;
  if (!extract_shape(q, &s, &A)) return -1;
  for (int i = 0; i < kNumShapes; ++i)
    if (shape_eq(s, kShapes[i])) return i;
  return -1;
}

// Launches the specialised kernel for q when one is compiled; *launched tells.
// The result buffers must already be zeroed.
int32_t try_spec(tgp_ctx* ctx, const tgp_agg_query* q, const tgp_group_result* out,
                 hipStream_t st, bool* launched) {
  *launched = false;
  if (ctx->tune_no_spec.load()) return TGP_OK;
  Shape s;
  SpecArgs A;
  if (!extract_shape(q, &s, &A)) return TGP_OK;
  A.out_count = reinterpret_cast<unsigned long long*>(out->d_count);
  A.out_sums = reinterpret_cast<unsigned long long*>(out->d_sums);
  A.out_oor = reinterpret_cast<unsigned long long*>(out->d_n_out_of_range);
  if (A.nslices == 0) {
    *launched = true;
    return TGP_OK;
  }
  return launch_shape<0>(ctx, s, A, st, launched);
}

}  // namespace tgp
