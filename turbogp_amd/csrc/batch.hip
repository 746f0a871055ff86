// Batched executor: thousands of filter + GROUP BY queries in one persistent
// kernel launch, captured in a hipGraph.
//
// Every query of a batch is planned into its specialised shape (spec.h) and cut
// into 256-row slices.  The slices of all queries form one flattened list; wave
// w of the persistent grid processes the w-th equal share of that list, i.e. a
// contiguous run of slices that crosses query boundaries.  A wave accumulates
// into a wave-private LDS region (counts and 64-bit partial sums per group) and
// flushes it into the query's exact int128 result with device atomics when it
// moves to the next query, when 2^20 rows have been added since the last flush
// (the partials stay exact for |value| < 2^43; larger values take an exact slow
// path per slice), and at the end.  Load balance is by construction: every
// wave gets the same number of slices, whatever the sizes of the queries.
//
// The one-launch-per-query form (tgp_group_agg per query) is kept as the
// baseline mode.  The reference has no batched executor; its per-query entry is
// QueryEngine::execute (engine/mod.rs:1487).
#include "spec.h"

namespace tgp {

constexpr int kBatchWords = 512;                 // per-wave region (groups * (1 + naggs))
constexpr uint64_t kFlushRows = 1ull << 20;      // rows between flushes
constexpr int kBatchBigBits = 43;                // |value| < 2^43 on the fast path

struct BatchQuery {
  SpecArgs a;           // arguments; a.out_* point into the batch's result arena
  uint64_t slice0;      // first flattened slice of the query
  int32_t shape;        // index into kShapes
  int32_t words;        // groups * (1 + naggs)
  int32_t naggs;
  int32_t _pad;
};

__device__ __forceinline__ bool big_batch(uint64_t v) {
  const int64_t s = (int64_t)v;
  return (((uint64_t)(s ^ (s >> 63))) >> kBatchBigBits) != 0;
}

// One slice of query Q (shape S) into the wave-private region `reg`.
template <Shape S>
__device__ __forceinline__ void batch_slice(const SpecArgs& A, uint64_t slice,
                                            unsigned long long* reg, uint32_t& oor) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t slice0 = A.base0 + slice * kSliceRows;
  const uint64_t row = slice0 + lane * kSR;
  uint64_t min_len = A.col_len[0];
#pragma unroll
  for (int s = 1; s < S.nslots; ++s) min_len = A.col_len[s] < min_len ? A.col_len[s] : min_len;
  const bool fast = slice0 + kSliceRows <= min_len;
  SliceVals<S> v;
  load_all<S, 0>(A, row, fast, v);
  int32_t gid[kSR];
  const uint32_t sel = eval_slice<S>(A, row, v, gid, oor);
  if (!__ballot(sel != 0)) return;
  constexpr int W1 = 1 + S.naggs;
#pragma unroll
  for (int r = 0; r < kSR; ++r)
    if ((sel >> r) & 1) atomicAdd(&reg[gid[r] * W1], 1ull);
  static_for<S.naggs>([&](auto AI) {
    constexpr int a = decltype(AI)::value;
    uint64_t val[kSR];
    agg_values<S, a>(A, v, val);
    bool b = false;
#pragma unroll
    for (int r = 0; r < kSR; ++r) b |= ((sel >> r) & 1) && big_batch(val[r]);
    if (__ballot(b)) {
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1)
          atomic_add128_i64(&A.out_sums[((uint64_t)gid[r] * S.naggs + a) * 2], (int64_t)val[r]);
    } else {
#pragma unroll
      for (int r = 0; r < kSR; ++r)
        if ((sel >> r) & 1) atomicAdd(&reg[gid[r] * W1 + 1 + a], (unsigned long long)val[r]);
    }
  });
}

// Adds the wave region into query Q's result and clears it.
__device__ __forceinline__ void batch_flush(const BatchQuery& Q, unsigned long long* reg,
                                            int naggs) {
  const uint32_t lane = threadIdx.x & 63;
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const uint32_t W1 = 1 + naggs;
  for (uint32_t w = lane; w < (uint32_t)Q.words; w += 64) {
    const unsigned long long x = reg[w];
    if (!x) continue;
    reg[w] = 0;
    const uint32_t g = w / W1, k = w % W1;
    if (k == 0) atomicAdd(&Q.a.out_count[g], x);
    else atomic_add128(&Q.a.out_sums[((uint64_t)g * naggs + (k - 1)) * 2], x,
                       (int64_t)x < 0 ? -1 : 0);
  }
  __builtin_amdgcn_wave_barrier();
}

template <int I>
__device__ __forceinline__ void dispatch_slice(int shape, const SpecArgs& A, uint64_t slice,
                                               unsigned long long* reg, uint32_t& oor) {
  if constexpr (I < kNumShapes) {
    if constexpr (kShapes[I].nslots > 0 && kShapes[I].has_sel == 0) {
      if (shape == I) {
        batch_slice<kShapes[I]>(A, slice, reg, oor);
        return;
      }
    }
    dispatch_slice<I + 1>(shape, A, slice, reg, oor);
  }
}

__global__ void __launch_bounds__(kBlock) k_batch(const BatchQuery* __restrict__ qs, uint32_t nq,
                                                  uint64_t total_slices) {
  __shared__ unsigned long long lreg[(kBlock / kWave) * kBatchWords];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long* reg = lreg + wave * kBatchWords;
  for (int w = lane; w < kBatchWords; w += 64) reg[w] = 0;
  __builtin_amdgcn_wave_barrier();
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t gw = (uint64_t)blockIdx.x * (kBlock / kWave) + wave;
  const uint64_t s0 = total_slices * gw / nwaves, s1 = total_slices * (gw + 1) / nwaves;
  if (s0 >= s1) return;
  // first query whose slices include s0: the last q with qs[q].slice0 <= s0
  uint32_t lo = 0, hi = nq;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (qs[mid].slice0 <= s0) lo = mid; else hi = mid;
  }
  uint32_t q = __builtin_amdgcn_readfirstlane(lo);
  uint32_t oor = 0;
  uint64_t rows_since_flush = 0;
  uint64_t q_end = q + 1 < nq ? qs[q + 1].slice0 : total_slices;
  for (uint64_t s = s0; s < s1; ++s) {
    while (s >= q_end) {  // next query (skips queries without slices)
      if (rows_since_flush) batch_flush(qs[q], reg, qs[q].naggs);
      if (oor) { atomicAdd(qs[q].a.out_oor, (unsigned long long)oor); oor = 0; }
      rows_since_flush = 0;
      ++q;
      q_end = q + 1 < nq ? qs[q + 1].slice0 : total_slices;
    }
    const BatchQuery& Q = qs[q];
    dispatch_slice<0>(Q.shape, Q.a, s - Q.slice0, reg, oor);
    rows_since_flush += kSliceRows;
    if (rows_since_flush >= kFlushRows) {
      batch_flush(Q, reg, Q.naggs);
      rows_since_flush = 0;
    }
  }
  if (rows_since_flush) batch_flush(qs[q], reg, qs[q].naggs);
  if (oor) atomicAdd(qs[q].a.out_oor, (unsigned long long)oor);
}

}  // namespace tgp

using namespace tgp;

struct tgp_batch {
  tgp_ctx* ctx = nullptr;
  uint32_t max_queries = 0, n = 0;
  // host copies
  std::vector<tgp_agg_query> queries;
  std::vector<BatchQuery> plan;
  std::vector<uint64_t> count_off, sums_off, oor_off, ngroups;
  std::vector<int32_t> fallback;  // queries without a batch shape (run alone)
  uint64_t result_words = 0, total_slices = 0;
  // device
  BatchQuery* d_plan = nullptr;
  uint64_t* d_results = nullptr;
  uint64_t plan_cap = 0, results_cap = 0;
  // tuning / graph
  uint32_t wg_per_cu = 0;  // 0 = occupancy
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  hipStream_t graph_stream = nullptr;
};

static void drop_graph(tgp_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
  b->graph = nullptr;
  b->graph_def = nullptr;
}

static int32_t batch_launch(tgp_batch* b, hipStream_t s) {
  TGP_TRY(fill_u64(s, b->d_results, b->result_words, 0));
  if (b->total_slices) {
    int per_cu = 0;
    TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, kBlock, 0));
    if (per_cu < 1) per_cu = 1;
    if (b->wg_per_cu && (int)b->wg_per_cu < per_cu) per_cu = (int)b->wg_per_cu;
    uint64_t grid = (uint64_t)persistent_grid(b->ctx, per_cu);
    const uint64_t need = (b->total_slices + 3) / 4;
    if (grid > need) grid = need;
    const int64_t gmax = b->ctx->tune_grid_max.load();
    if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
    if (grid == 0) grid = 1;
    k_batch<<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(b->d_plan, (uint32_t)b->plan.size(),
                                                        b->total_slices);
    TGP_HIP(hipGetLastError());
  }
  return TGP_OK;
}

extern "C" {

int32_t tgp_batch_create(tgp_ctx* ctx, uint32_t max_queries, tgp_batch** out) {
  TGP_REQUIRE(ctx && out && max_queries > 0, "tgp_batch_create: bad argument");
  auto* b = new tgp_batch();
  b->ctx = ctx;
  b->max_queries = max_queries;
  *out = b;
  return TGP_OK;
}

int32_t tgp_batch_destroy(tgp_batch* b) {
  if (!b) return TGP_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  drop_graph(b);
  if (b->d_plan) hbm_free(b->ctx, b->d_plan);
  if (b->d_results) hbm_free(b->ctx, b->d_results);
  delete b;
  return TGP_OK;
}

int32_t tgp_batch_set_queries(tgp_batch* b, const tgp_agg_query* qs, uint32_t n, void* stream) {
  TGP_REQUIRE(b && (n == 0 || qs), "tgp_batch_set_queries: null argument");
  TGP_REQUIRE(n <= b->max_queries, "batch of %u queries above its capacity %u", n,
              b->max_queries);
  TGP_TRY(check_injected(b->ctx));
  hipStream_t s = pick_stream(b->ctx, stream);
  drop_graph(b);
  b->n = n;
  b->queries.assign(qs, qs + n);
  b->plan.clear();
  b->fallback.clear();
  b->count_off.assign(n, 0);
  b->sums_off.assign(n, 0);
  b->oor_off.assign(n, 0);
  b->ngroups.assign(n, 0);
  // result arena layout per query: count[ngroups], sums[ngroups*naggs*2], oor
  uint64_t words = 0;
  for (uint32_t i = 0; i < n; ++i) {
    DevPlan P;
    PlanInfo info;
    TGP_TRY(plan_agg(&qs[i], &P, &info));  // validation (same rules as tgp_group_agg)
    b->ngroups[i] = P.ngroups;
    b->count_off[i] = words;
    words += P.ngroups;
    b->sums_off[i] = words;
    words += P.ngroups * qs[i].naggs * 2;
    b->oor_off[i] = words;
    words += 1;
  }
  b->result_words = words;
  if (b->results_cap < words) {
    if (b->d_results) TGP_TRY(hbm_free(b->ctx, b->d_results));
    b->d_results = nullptr;
    void* p = nullptr;
    TGP_TRY(hbm_alloc(b->ctx, words * 8, &p));
    b->d_results = static_cast<uint64_t*>(p);
    b->results_cap = words;
  }
  uint64_t slices = 0;
  for (uint32_t i = 0; i < n; ++i) {
    Shape sh;
    BatchQuery bq{};
    int idx = -1;
    if (extract_shape(&qs[i], &sh, &bq.a)) {
      for (int k = 0; k < kNumShapes; ++k)
        if (shape_eq(sh, kShapes[k])) idx = k;
    }
    if (idx < 0 || sh.has_sel || bq.a.words > (uint32_t)kBatchWords) {
      b->fallback.push_back((int32_t)i);
      continue;
    }
    bq.shape = idx;
    bq.words = (int32_t)bq.a.words;
    bq.naggs = qs[i].naggs;
    bq.slice0 = slices;
    bq.a.out_count = reinterpret_cast<unsigned long long*>(b->d_results + b->count_off[i]);
    bq.a.out_sums = reinterpret_cast<unsigned long long*>(b->d_results + b->sums_off[i]);
    bq.a.out_oor = reinterpret_cast<unsigned long long*>(b->d_results + b->oor_off[i]);
    slices += bq.a.nslices;
    b->plan.push_back(bq);
  }
  b->total_slices = slices;
  const uint64_t bytes = b->plan.size() * sizeof(BatchQuery);
  if (b->plan_cap < b->plan.size()) {
    if (b->d_plan) TGP_TRY(hbm_free(b->ctx, b->d_plan));
    b->d_plan = nullptr;
    void* p = nullptr;
    TGP_TRY(hbm_alloc(b->ctx, bytes ? bytes : 64, &p));
    b->d_plan = static_cast<BatchQuery*>(p);
    b->plan_cap = b->plan.size();
  }
  if (bytes) {
    TGP_HIP(hipMemcpyAsync(b->d_plan, b->plan.data(), bytes, hipMemcpyHostToDevice, s));
    TGP_HIP(hipStreamSynchronize(s));
  }
  return TGP_OK;
}

int32_t tgp_batch_set_tuning(tgp_batch* b, uint32_t wg_per_cu, int32_t use_graph) {
  TGP_REQUIRE(b, "tgp_batch_set_tuning: null batch");
  drop_graph(b);
  b->wg_per_cu = wg_per_cu;
  b->use_graph = use_graph != 0;
  return TGP_OK;
}

int32_t tgp_batch_run(tgp_batch* b, int32_t mode, void* stream) {
  TGP_REQUIRE(b, "tgp_batch_run: null batch");
  TGP_REQUIRE(mode == TGP_BATCH_PERSISTENT || mode == TGP_BATCH_PER_QUERY, "bad mode %d", mode);
  TGP_TRY(check_injected(b->ctx));
  hipStream_t s = pick_stream(b->ctx, stream);
  if (mode == TGP_BATCH_PER_QUERY) {
    // the one-launch-per-query form: tgp_group_agg for every query
    for (uint32_t i = 0; i < b->n; ++i) {
      tgp_group_result r{b->d_results + b->count_off[i], b->d_results + b->sums_off[i],
                         b->d_results + b->oor_off[i]};
      TGP_TRY(tgp_group_agg(b->ctx, &b->queries[i], &r, s));
    }
    return TGP_OK;
  }
  if (b->use_graph) {
    if (!b->graph || b->graph_stream != s) {
      drop_graph(b);
      TGP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int32_t rc = batch_launch(b, s);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      TGP_TRY(rc);
      if (e != hipSuccess) {
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return TGP_ERR_HIP;
      }
      b->graph_def = g;
      TGP_HIP(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
      b->graph_stream = s;
    }
    TGP_HIP(hipGraphLaunch(b->graph, s));
  } else {
    TGP_TRY(batch_launch(b, s));
  }
  // queries without a batch shape run alone, into the same arena
  for (int32_t i : b->fallback) {
    tgp_group_result r{b->d_results + b->count_off[i], b->d_results + b->sums_off[i],
                       b->d_results + b->oor_off[i]};
    TGP_TRY(tgp_group_agg(b->ctx, &b->queries[i], &r, s));
  }
  return TGP_OK;
}

int32_t tgp_batch_run_timed(tgp_batch* b, int32_t mode, void* stream, float* h_query_ms,
                            float* h_total_ms) {
  TGP_REQUIRE(b && h_total_ms, "tgp_batch_run_timed: null argument");
  hipStream_t s = pick_stream(b->ctx, stream);
  std::vector<hipEvent_t> ev;
  const uint32_t nev = mode == TGP_BATCH_PER_QUERY ? b->n + 1 : 2;
  ev.resize(nev);
  for (auto& e : ev) TGP_HIP(hipEventCreate(&e));
  int32_t rc = TGP_OK;
  TGP_HIP(hipEventRecord(ev[0], s));
  if (mode == TGP_BATCH_PER_QUERY) {
    for (uint32_t i = 0; i < b->n && rc == TGP_OK; ++i) {
      tgp_group_result r{b->d_results + b->count_off[i], b->d_results + b->sums_off[i],
                         b->d_results + b->oor_off[i]};
      rc = tgp_group_agg(b->ctx, &b->queries[i], &r, s);
      if (rc == TGP_OK) TGP_HIP(hipEventRecord(ev[i + 1], s));
    }
  } else {
    rc = tgp_batch_run(b, mode, s);
    if (rc == TGP_OK) TGP_HIP(hipEventRecord(ev[1], s));
  }
  if (rc == TGP_OK) {
    TGP_HIP(hipEventSynchronize(ev[nev - 1]));
    float total = 0;
    TGP_HIP(hipEventElapsedTime(&total, ev[0], ev[nev - 1]));
    *h_total_ms = total;
    if (h_query_ms)
      for (uint32_t i = 0; i < b->n; ++i) {
        if (mode == TGP_BATCH_PER_QUERY)
          TGP_HIP(hipEventElapsedTime(&h_query_ms[i], ev[0], ev[i + 1]));
        else
          h_query_ms[i] = total;  // every query of a persistent batch completes at its end
      }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int32_t tgp_batch_info(tgp_batch* b, uint64_t* result_words, uint64_t* total_slices,
                       uint32_t* n_batched, uint32_t* n_fallback) {
  TGP_REQUIRE(b, "tgp_batch_info: null batch");
  if (result_words) *result_words = b->result_words;
  if (total_slices) *total_slices = b->total_slices;
  if (n_batched) *n_batched = (uint32_t)b->plan.size();
  if (n_fallback) *n_fallback = (uint32_t)b->fallback.size();
  return TGP_OK;
}

int32_t tgp_batch_result_layout(tgp_batch* b, uint32_t i, uint64_t* count_off, uint64_t* sums_off,
                                uint64_t* oor_off, uint64_t* ngroups) {
  TGP_REQUIRE(b && i < b->n, "tgp_batch_result_layout: bad query index");
  if (count_off) *count_off = b->count_off[i];
  if (sums_off) *sums_off = b->sums_off[i];
  if (oor_off) *oor_off = b->oor_off[i];
  if (ngroups) *ngroups = b->ngroups[i];
  return TGP_OK;
}

int32_t tgp_batch_results_sync(tgp_batch* b, uint64_t* h_words, uint64_t nwords, void* stream) {
  TGP_REQUIRE(b && (nwords == 0 || h_words), "tgp_batch_results_sync: null argument");
  TGP_REQUIRE(nwords <= b->result_words, "asked for %llu result words, arena holds %llu",
              (unsigned long long)nwords, (unsigned long long)b->result_words);
  hipStream_t s = pick_stream(b->ctx, stream);
  if (nwords) TGP_HIP(hipMemcpyAsync(h_words, b->d_results, nwords * 8, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  return TGP_OK;
}

int32_t tgp_batch_autotune(tgp_batch* b, void* stream, uint32_t* best_wg_per_cu, float* best_ms) {
  TGP_REQUIRE(b, "tgp_batch_autotune: null batch");
  hipStream_t s = pick_stream(b->ctx, stream);
  hipEvent_t e0, e1;
  TGP_HIP(hipEventCreate(&e0));
  TGP_HIP(hipEventCreate(&e1));
  const uint32_t cand[] = {1, 2, 3, 4, 6, 8};
  float best = 1e30f;
  uint32_t pick = 0;
  const bool graph = b->use_graph;
  for (uint32_t c : cand) {
    b->wg_per_cu = c;
    drop_graph(b);
    TGP_TRY(batch_launch(b, s));  // warm
    float ms_min = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      TGP_HIP(hipEventRecord(e0, s));
      TGP_TRY(batch_launch(b, s));
      TGP_HIP(hipEventRecord(e1, s));
      TGP_HIP(hipEventSynchronize(e1));
      float ms = 0;
      TGP_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms_min = ms < ms_min ? ms : ms_min;
    }
    if (ms_min < best) {
      best = ms_min;
      pick = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->wg_per_cu = pick;
  b->use_graph = graph;
  drop_graph(b);
  if (best_wg_per_cu) *best_wg_per_cu = pick;
  if (best_ms) *best_ms = best;
  return TGP_OK;
}

}  // extern "C"
