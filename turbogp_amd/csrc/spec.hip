// Shape-specialised scan-aggregate kernel (one query per launch) and the
// planner's shape extraction.  Design notes are in spec.h.
#include "wave_agg.h"

namespace tgp {

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


// Single-query kernel on the wave-level blocks of wave_agg.h.  With chunk == 0
// wave w of the grid takes slices w, w + nwaves, ... (static interleave: all
// waves sweep the table together, no atomics); with chunk > 0 the waves take
// chunks of `chunk` slices from a device counter (a slow wave takes fewer).  A
// wave accumulates in its replicated LDS histogram and flushes into the result
// every 2^20 rows and at the end.  The LDS region of a wave is words << rshift
// words (dynamic shared memory of 4 such regions per workgroup).
template <Shape S>
__device__ __forceinline__ void wave_rows(const SpecArgs& A, uint64_t s, uint64_t min_len,
                                          unsigned long long* reg, uint32_t rshift,
                                          uint32_t& oor, uint64_t& since) {
  if (since + kSliceRows > kFlushRows) {
    wave_flush(A.out_count, A.out_sums, A.words, S.naggs, reg, rshift);
    since = 0;
  }
  SliceVals<S> v;
  wave_load<S>(A, s, min_len, v);
  wave_slice<S>(A, s, v, reg, rshift, oor);
  since += kSliceRows;
}

template <Shape S>
__global__ void __launch_bounds__(kBlock) k_wave(SpecArgs A, uint32_t rshift, uint32_t chunk,
                                                 unsigned long long* counter) {
  extern __shared__ __align__(16) unsigned long long lreg[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t region = A.words << rshift;
  unsigned long long* reg = lreg + wave * region;
  for (uint32_t w = lane; w < region; w += 64) reg[w] = 0;
  __builtin_amdgcn_wave_barrier();
  uint64_t min_len = A.col_len[0];
#pragma unroll
  for (int s = 1; s < S.nslots; ++s) min_len = A.col_len[s] < min_len ? A.col_len[s] : min_len;
  uint32_t oor = 0;
  uint64_t since = 0;  // rows added to the region since its last flush
  if (chunk == 0) {
    const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
    for (uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + wave; s < A.nslices; s += nwaves)
      wave_rows<S>(A, s, min_len, reg, rshift, oor, since);
  } else {
    uint64_t k = grab_chunk(counter);
    while (k * chunk < A.nslices) {
      const uint64_t c0 = k * chunk;
      const uint64_t c1 = c0 + chunk < A.nslices ? c0 + chunk : A.nslices;
      const uint64_t next = grab_chunk(counter);  // in flight while this chunk runs
      for (uint64_t s = c0; s < c1; ++s) wave_rows<S>(A, s, min_len, reg, rshift, oor, since);
      k = next;
    }
  }
  // workgroup flush: every wave folds its copies, then one thread per word adds
  // the 4 waves' totals (as int128) into the result
  wave_fold_copies(reg, A.words, rshift);
  __syncthreads();
  constexpr uint32_t W1 = 1 + S.naggs;
  for (uint32_t w = threadIdx.x; w < A.words; w += kBlock) {
    const bool is_count = w % W1 == 0;
    uint64_t lo = 0;
    int64_t hi = 0;
    for (int v = 0; v < kBlock / kWave; ++v) {
      const uint64_t x = lreg[v * region + w];
      const uint64_t t = lo + x;
      hi += (int64_t)(t < lo) - ((!is_count && (int64_t)x < 0) ? 1 : 0);
      lo = t;
    }
    if (lo || hi) add_result_word(A.out_count, A.out_sums, w, S.naggs, lo, hi);
  }
  if (oor && A.out_oor) atomicAdd(A.out_oor, (unsigned long long)oor);
}

// Queries of at least this many slices (4M rows) take the replicated form, and
// queries of more histogram words than kHistWords from kWaveManySlices (256K rows).
constexpr uint64_t kWaveMinSlices = 16384;
constexpr uint64_t kWaveManySlices = 1024;

// Fewest slices a wave of a single-query kernel gets before the grid shrinks
// (each workgroup ends with a flush of device atomics, so small queries run on
// fewer, longer waves).
static uint64_t min_slices_per_wave(tgp_ctx* ctx) {
  const int64_t v = ctx->tune_slices_per_wave.load();
  return v > 0 ? (uint64_t)v : 2;
}

// Zeroes the result (and the chunk counter when given) in one launch.
static int32_t zero_result(hipStream_t st, const SpecArgs& A, int naggs,
                           unsigned long long* counter) {
  uint64_t* p[4] = {reinterpret_cast<uint64_t*>(A.out_count),
                    reinterpret_cast<uint64_t*>(A.out_sums),
                    reinterpret_cast<uint64_t*>(A.out_oor), reinterpret_cast<uint64_t*>(counter)};
  const uint64_t n[4] = {A.ngroups, A.ngroups * (uint64_t)naggs * 2, 1, 1};
  return zero_ranges(st, 4, p, n);
}

template <int I>
int32_t launch_shape(tgp_ctx* ctx, const Shape& s, const SpecArgs& A, hipStream_t st,
                     bool* launched) {
  if constexpr (I < kNumShapes) {
    if (shape_eq(s, kShapes[I])) {
      constexpr Shape S = kShapes[I];
      // Kernel form (measured over 20K..60M rows): the replicated-histogram waves
      // for large queries of several groups, and from 256K rows up for queries
      // whose histogram outgrows a wave's registers (the register fold then
      // goes through a workgroup table with carry-exact LDS atomics); the
      // register fold otherwise (less LDS to set up and flush, which small
      // queries feel).
      const int rshift = wave_rshift(A.words);
      const int64_t form = ctx->tune_spec_kernel.load();
      const bool large = A.nslices >= kWaveMinSlices;
      const bool many = A.words > (uint32_t)kHistWords && A.nslices >= kWaveManySlices;
      const bool wave = rshift >= 0 &&
                        (form == 2 || (form == 0 && A.ngroups >= 2 && (large || many)));
      int per_cu = 0;
      uint32_t lds = 0;
      if (wave) {
        lds = (uint32_t)((A.words << rshift) * (kBlock / kWave) * 8);
        TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_wave<S>, kBlock, lds));
      } else {
        lds = S.table == 0 ? (4 * kHistWords + 4 * 64 * 2) * 8
                           : (uint32_t)(A.ngroups * (1 + 2 * S.naggs) * 8);
        TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spec<S>, kBlock, lds));
      }
      if (per_cu < 1) per_cu = 1;
      const int64_t want = ctx->tune_wg_per_cu.load();
      if (want > 0 && want < per_cu) per_cu = (int)want;
      // Large scans stream best from few concurrent waves as long as those hide
      // the per-row work: measured on 60M rows, one group and 2..4 groups run
      // fastest at 2 workgroups per CU, more groups at 3.
      if (want == 0 && large && (wave || A.ngroups == 1)) {
        const int cap = A.ngroups <= 4 ? 2 : 3;
        if (per_cu > cap) per_cu = cap;
      }
      uint64_t grid = (uint64_t)persistent_grid(ctx, per_cu);
      const int64_t gmax = ctx->tune_grid_max.load();
      if (wave) {
        // 0: static interleave (default); > 0: slices per dynamically taken chunk
        const uint64_t chunk = (uint64_t)ctx->tune_chunk.load();
        // every workgroup ends with one device atomic per histogram word, all on
        // the same addresses: queries of many words run on fewer, longer waves
        uint64_t spw = min_slices_per_wave(ctx);
        if (ctx->tune_slices_per_wave.load() == 0 && A.words / 64 > spw) spw = A.words / 64;
        const uint64_t per_wg = (kBlock / kWave) * (chunk ? chunk : spw);
        const uint64_t need = (A.nslices + per_wg - 1) / per_wg;
        if (grid > need) grid = need;
        if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
        if (grid == 0) grid = 1;
        void* counter = nullptr;
        TGP_TRY(get_scratch(ctx, st, 256, &counter));
        auto* cnt = static_cast<unsigned long long*>(counter);
        TGP_TRY(zero_result(st, A, S.naggs, cnt));
        k_wave<S><<<dim3((unsigned)grid), dim3(kBlock), lds, st>>>(A, (uint32_t)rshift,
                                                                  (uint32_t)chunk, cnt);
      } else {
        const uint64_t per_wg = (kBlock / kWave) * min_slices_per_wave(ctx);
        const uint64_t need = (A.nslices + per_wg - 1) / per_wg;
        if (grid > need) grid = need;
        if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
        if (grid == 0) grid = 1;
        TGP_TRY(zero_result(st, A, S.naggs, nullptr));
        k_spec<S><<<dim3((unsigned)grid), dim3(kBlock), lds, st>>>(A);
      }
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
bool extract_shape(const tgp_agg_query* q, Shape* s, SpecArgs* A) {
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
    if (G.kind != TGP_AGG_SUM && G.kind != TGP_AGG_COUNT_NONZERO) return false;  // MIN/MAX
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

// Launches the specialised kernel for q when one is compiled (after zeroing the
// result buffers in the same stream); *launched tells.
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
    TGP_TRY(zero_result(st, A, q->naggs, nullptr));
    *launched = true;
    return TGP_OK;
  }
  return launch_shape<0>(ctx, s, A, st, launched);
}

}  // namespace tgp
