// Shape-specialised scan-aggregate kernel (one query per launch) and the
// planner's shape extraction.  Design notes are in spec.h.
#include "jit.h"

namespace tgp {

template <Shape S>
__global__ void __launch_bounds__(kBlock) k_spec(SpecArgs A) {
  spec_body<S>(A);
}

template <Shape S>
__global__ void __launch_bounds__(kBlock) k_wave(SpecArgs A, uint32_t rshift, uint32_t chunk,
                                                 unsigned long long* counter) {
  wave_body<S>(A, rshift, chunk, counter);
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

// The kernels that run one shape: ahead-of-time template instances, or the
// run-time compiled module functions of the same bodies (jit.hip).

// Chooses the kernel form and grid of shape S for the query's arguments and
// launches it (after zeroing the result in the same stream).
int32_t launch_shape(tgp_ctx* ctx, const Shape& S, const ShapeKernels& K, SpecArgs A,
                     hipStream_t st, unsigned long long* counter, bool zeroed) {
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
  } else {
    lds = S.table == 0 ? (4 * kHistWords + 4 * 64 * 2) * 8
                       : (uint32_t)(A.ngroups * (1 + 2 * S.naggs) * 8);
  }
  TGP_TRY(K.occupancy(wave, lds, &per_cu));
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
    uint32_t chunk = (uint32_t)ctx->tune_chunk.load();
    // every workgroup ends with one device atomic per histogram word, all on
    // the same addresses: queries of many words run on fewer, longer waves
    uint64_t spw = min_slices_per_wave(ctx);
    if (ctx->tune_slices_per_wave.load() == 0 && A.words / 64 > spw) spw = A.words / 64;
    const uint64_t per_wg = (kBlock / kWave) * (chunk ? chunk : spw);
    const uint64_t need = (A.nslices + per_wg - 1) / per_wg;
    if (grid > need) grid = need;
    if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
    if (grid == 0) grid = 1;
    unsigned long long* cnt = counter;
    if (!cnt) {
      void* scr = nullptr;
      TGP_TRY(get_scratch(ctx, st, 256, &scr));
      cnt = static_cast<unsigned long long*>(scr);
    }
    if (!zeroed) TGP_TRY(zero_result(st, A, S.naggs, cnt));
    else if (chunk) TGP_TRY(fill_u64(st, reinterpret_cast<uint64_t*>(cnt), 1, 0));
    uint32_t rs = (uint32_t)rshift;
    void* args[] = {&A, &rs, &chunk, &cnt};
    TGP_TRY(K.launch(true, grid, lds, st, args));
  } else {
    const uint64_t per_wg = (kBlock / kWave) * min_slices_per_wave(ctx);
    const uint64_t need = (A.nslices + per_wg - 1) / per_wg;
    if (grid > need) grid = need;
    if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
    if (grid == 0) grid = 1;
    if (!zeroed) TGP_TRY(zero_result(st, A, S.naggs, nullptr));
    void* args[] = {&A};
    TGP_TRY(K.launch(false, grid, lds, st, args));
  }
  // MIN / MAX accumulate order codes: decoded into values
  return finalize_minmax(st, A.out_count, A.out_sums, A.ngroups, S.naggs, shape_minmax_kinds(S));
}

// The ahead-of-time kernels of s, if s is one of kShapes.
template <int I>
bool aot_kernels(const Shape& s, ShapeKernels* K) {
  if constexpr (I < kNumShapes) {
    if (shape_eq(s, kShapes[I])) {
      constexpr Shape S = kShapes[I];
      K->spec = reinterpret_cast<const void*>(&k_spec<S>);
      K->wave = reinterpret_cast<const void*>(&k_wave<S>);
      return true;
    }
    return aot_kernels<I + 1>(s, K);
  } else {
    return false;
  }
}

// Queries of a shape without ahead-of-time kernels, TGP_TUNE_JIT 0 (default):
// from kJitMinRows rows the shape is compiled on a background thread and the
// query runs on the generic engine until the compile is done; 1: never
// compiled; 2: compiled before the query runs, whatever its size.
constexpr uint64_t kJitMinRows = 1ull << 20;

int32_t find_shape_kernels(const Shape& s, bool wait, ShapeKernels* K, bool* ok) {
  *ok = false;
  *K = ShapeKernels{};
  if (aot_kernels<0>(s, K)) {
    *ok = true;
    return TGP_OK;
  }
  bool ready = false;
  if (jit_shape_kernels(s, wait, &K->jit, &ready) != TGP_OK) return TGP_OK;  // failed: none
  *ok = ready;
  return TGP_OK;
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
    if (G.kind != TGP_AGG_SUM && G.kind != TGP_AGG_COUNT_NONZERO && !is_minmax(G.kind))
      return false;
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
  return TGP_SPEC_SHAPE_RUNTIME;
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
  ShapeKernels K;
  if (!aot_kernels<0>(s, &K)) {
    const int64_t mode = ctx->tune_jit.load();
    if (mode == 1 || (mode == 0 && A.row_end - A.row_begin < kJitMinRows)) return TGP_OK;
    // a shape still compiling, or that failed to compile, runs on the generic
    // engine (a failure stays in the compile cache and tgp_jit_info's count)
    bool ready = false;
    if (jit_shape_kernels(s, mode == 2, &K.jit, &ready) != TGP_OK || !ready) return TGP_OK;
  }
  TGP_TRY(launch_shape(ctx, s, K, A, st, nullptr, false));
  *launched = true;
  return TGP_OK;
}

}  // namespace tgp

extern "C" int32_t tgp_jit_info(uint64_t* compiled, uint64_t* failed, uint64_t* launches,
                                double* compile_ms) {
  TGP_REQUIRE(compiled && failed && launches && compile_ms, "tgp_jit_info: null argument");
  const tgp::JitStats st = tgp::jit_stats();
  *compiled = st.compiled;
  *failed = st.failed;
  *launches = st.launches;
  *compile_ms = st.compile_ms;
  return TGP_OK;
}
