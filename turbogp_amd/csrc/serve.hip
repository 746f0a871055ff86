// Continuous-batching query server over device-resident columns.
//
// Requests (filter + GROUP BY aggregates, and hash-join aggregates) are
// submitted from any thread at any time.  Two worker threads run them:
//
//   - the aggregate worker keeps up to `batch_slots` batches in flight, each on
//     its own stream.  Whenever a slot is free it takes every waiting aggregate
//     request (up to `max_batch_queries`) into one persistent-kernel batch
//     (tgp_batch_*), so requests that arrive while others run join the next
//     batch and batches overlap on the device.  Requests whose row range is at
//     least `heavy_rows` (full-table scans) are batched on a slot of their own
//     (slot 0, in the shared scan with 128-slice chunks: each slice is read
//     once for all of them), so they do not hold back the small requests.
//   - `join_slots` join workers (a thread and a stream each) admit a join when
//     the hash table it will build fits the HBM budget (the context's tier
//     budget, optionally capped by `hash_budget_bytes`), build it and run the
//     fused probe->aggregate; a join waits while the budget is held by others,
//     and is rejected if it can never fit.
//
// With `per_query_launch` the aggregate worker instead runs each request as its
// own tgp_group_agg launch (one request per slot): the one-launch-per-query
// serving form that batching is measured against.
//
// Completion copies each request's exact result words to host memory and
// stamps it.  Latency is measured from the request's arrival time (given by the
// submitter, so an open-loop load generator that falls behind does not hide
// queueing time).  The reference serves one query at a time behind a
// connection-pool semaphore (server/pool.rs, engine/mod.rs:522-545
// route_and_execute); it has no batching.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tgp_internal.h"

using namespace tgp;

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  uint64_t id = 0;
  tgp_serve_request req{};
  uint64_t ngroups = 0;
  int32_t naggs = 0;
  uint64_t rows = 0;   // rows in the aggregate's range / probe range
  uint64_t words = 0;  // result words: count[ngroups] + sums[ngroups*naggs*2] + oor
  int64_t t_arrive = 0, t_start = 0, t_done = 0;
  int32_t status = TGP_OK;
  bool done = false;
  bool deferred = false;  // waited at least once for HBM budget
  std::vector<uint64_t> result;
  std::string error;      // tgp_last_error() text of a failed request
};

struct AggSlot {
  tgp_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  std::vector<Request*> reqs;
  bool busy = false;
  bool heavy = false;
  bool single = false;          // running its requests as tgp_group_agg launches
  uint64_t* h_arena = nullptr;  // pinned copy of the batch's (or single launches') results
  uint64_t h_cap = 0;
  uint64_t* d_res = nullptr;    // single launches: the requests' result words, one after another
  uint64_t d_cap = 0;
};

// A resident hash table over a prefix of one build side (key and payload
// columns), shared by every join on that build side: a join over the first nb
// build rows probes it with the rows from nb on excluded (build_rows).  A join
// that needs a longer prefix builds a replacement; the outgrown table is kept
// until no join in flight probes it.
constexpr size_t kMaxProbeBatch = 1024;  // joins per probe batch

struct SharedTable {
  const void* keys = nullptr;
  const void* pay = nullptr;
  int32_t kdt = 0, pdt = 0;
  tgp_hash_table* table = nullptr;  // ready table (nullptr until the first build ends)
  uint64_t rows = 0;                // build rows it holds
  uint64_t building = 0;            // rows of the build in progress (0: none)
  uint32_t refs = 0;                // joins in flight probing `table`
};

struct JoinSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  Request* req = nullptr;
  SharedTable* shared = nullptr;    // the shared table this join probes
  tgp_hash_table* shared_table = nullptr;  // ... the instance it holds a reference to
  SharedTable* builds = nullptr;    // the shared table this slot is (re)building
  uint64_t build_len = 0;           // ... over this many build rows
  tgp_hash_table* table = nullptr;
  uint64_t table_bytes = 0;
  void* pool = nullptr;       // the slot's table memory, reused join after join
  uint64_t pool_bytes = 0;    // grows to the largest table the slot has built
  uint64_t* d_res = nullptr;  // count[ng] | sums[ng*2] | oor (per request of a probe batch)
  uint64_t d_cap = 0;
  bool busy = false;
  // a probe batch: joins whose shared tables are ready, probed in one launch, each
  // holding a reference to the table instance it probes
  std::vector<Request*> batch;
  std::vector<std::pair<SharedTable*, tgp_hash_table*>> holds;
  std::vector<uint64_t> h_res;
};

}  // namespace

struct tgp_server {
  tgp_ctx* ctx = nullptr;
  tgp_serve_config cfg{};
  std::mutex mu;
  // New work wakes only the workers that can take it: one submit used to wake
  // every worker thread, which then contended with the submitter for `mu`.
  std::condition_variable agg_cv;   // aggregate requests / stop
  std::condition_variable join_cv;  // join requests / stop
  std::condition_variable done_cv;  // completions
  std::deque<Request*> agg_q, heavy_q, join_q;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread agg_thread;
  std::vector<std::thread> join_threads;
  std::vector<AggSlot> agg_slots;
  std::vector<JoinSlot> join_slots;
  uint64_t hash_in_use = 0;  // bytes of the tables of the joins in flight
  uint64_t pool_held = 0;    // bytes of all join slots' table pools
  bool heavy_slot = false;  // slot 0 serves the heavy queue
  tgp_serve_stats stats{};
  std::vector<std::unique_ptr<SharedTable>> shared;  // share_build_tables: per build side
  // outgrown shared tables still probed by joins in flight: (table, refs)
  std::vector<std::pair<tgp_hash_table*, uint32_t>> retired;
};

namespace {

int32_t result_shape(const tgp_serve_request& r, uint64_t* ngroups, int32_t* naggs,
                     uint64_t* rows) {
  if (r.kind == TGP_SERVE_AGG) {
    TGP_TRY(tgp_group_agg_ngroups(&r.agg, ngroups));
    *naggs = r.agg.naggs;
    *rows = r.agg.row_end > r.agg.row_begin ? r.agg.row_end - r.agg.row_begin : 0;
    return TGP_OK;
  }
  if (r.kind == TGP_SERVE_JOIN_AGG) {
    const tgp_join_agg_query& j = r.join.probe;
    TGP_REQUIRE(j.group_by == TGP_JOIN_GROUP_NONE || j.ngroups > 0,
                "join request: ngroups must be > 0");
    TGP_REQUIRE(j.ngroups <= (1ull << 26), "join request: %llu groups above 2^26",
                (unsigned long long)j.ngroups);
    TGP_REQUIRE(r.join.build_keys.data && r.join.build_keys.len > 0,
                "join request: empty build side");
    *ngroups = j.group_by == TGP_JOIN_GROUP_NONE ? 1 : j.ngroups;
    *naggs = 1;
    *rows = j.row_end > j.row_begin ? j.row_end - j.row_begin : 0;
    return TGP_OK;
  }
  set_error("serve request: unknown kind %d", r.kind);
  return TGP_ERR_INVALID;
}

void finish(tgp_server* s, Request* r, int32_t status) {
  r->status = status;
  if (status != TGP_OK) r->error = tgp_last_error();
  r->t_done = now_ns();
  r->done = true;
  s->stats.completed += 1;
  if (status != TGP_OK) s->stats.failed += 1;
}

// ---- aggregate worker -------------------------------------------------------------

// Grows the slot's pinned result copy to `words` (by replacement: it is idle).
int32_t ensure_host_arena(AggSlot& sl, uint64_t words) {
  if (words <= sl.h_cap) return TGP_OK;
  if (sl.h_arena) (void)hipHostFree(sl.h_arena);
  sl.h_arena = nullptr;
  sl.h_cap = 0;
  if (hipHostMalloc((void**)&sl.h_arena, words * 8) != hipSuccess) {
    set_error("serve: hipHostMalloc of %llu bytes failed", (unsigned long long)words * 8);
    return TGP_ERR_OOM;
  }
  sl.h_cap = words;
  return TGP_OK;
}

// The slot's requests ran as single-query launches, their results one after
// another in d_res: one copy out, then each request takes its words.
int32_t harvest_single(tgp_server* s, AggSlot& sl) {
  uint64_t total = 0;
  for (Request* r : sl.reqs) total += r->words;
  int32_t rc = ensure_host_arena(sl, total);
  if (rc == TGP_OK && (hipMemcpyAsync(sl.h_arena, sl.d_res, total * 8, hipMemcpyDeviceToHost,
                                      sl.stream) != hipSuccess ||
                       hipStreamSynchronize(sl.stream) != hipSuccess)) {
    set_error("serve: result copy failed");
    rc = TGP_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  uint64_t off = 0;
  for (Request* r : sl.reqs) {
    if (rc == TGP_OK) r->result.assign(sl.h_arena + off, sl.h_arena + off + r->words);
    off += r->words;
    finish(s, r, rc);
  }
  sl.reqs.clear();
  sl.busy = false;
  if (sl.heavy) s->stats.heavy_in_flight -= 1;
  s->done_cv.notify_all();
  return rc;
}

// Runs the requests as single-query launches (tgp_group_agg) one after another
// on the slot's stream.
int32_t launch_single(tgp_server* s, AggSlot& sl, const std::vector<Request*>& take) {
  uint64_t total = 0;
  for (Request* r : take) total += r->words;
  if (sl.d_cap < total) {
    // grows by doubling: a free synchronises the device
    const uint64_t cap = total > 2 * sl.d_cap ? total : 2 * sl.d_cap;
    if (sl.d_res) TGP_TRY(hbm_free(s->ctx, sl.d_res));
    sl.d_res = nullptr;
    sl.d_cap = 0;
    void* p = nullptr;
    TGP_TRY(hbm_alloc(s->ctx, cap * 8, &p));
    sl.d_res = static_cast<uint64_t*>(p);
    sl.d_cap = cap;
  }
  uint64_t off = 0;
  for (Request* r : take) {
    uint64_t* base = sl.d_res + off;
    const tgp_group_result out{base, base + r->ngroups,
                               base + r->ngroups + r->ngroups * (uint64_t)r->naggs * 2};
    TGP_TRY(tgp_group_agg(s->ctx, &r->req.agg, &out, sl.stream));
    off += r->words;
  }
  TGP_HIP(hipEventRecord(sl.done_ev, sl.stream));
  return TGP_OK;
}

int32_t harvest_batch(tgp_server* s, AggSlot& sl);

// Copies a finished batch's (or single launch's) results out and completes its
// requests.
int32_t harvest_agg(tgp_server* s, AggSlot& sl) {
  const int64_t t0 = now_ns();
  const int32_t rc = sl.single ? harvest_single(s, sl) : harvest_batch(s, sl);
  std::lock_guard<std::mutex> lk(s->mu);
  s->stats.harvest_ns += (uint64_t)(now_ns() - t0);
  return rc;
}

int32_t harvest_batch(tgp_server* s, AggSlot& sl) {
  uint64_t words = 0;
  int32_t rc = tgp_batch_info(sl.batch, &words, nullptr, nullptr, nullptr);
  if (rc == TGP_OK) rc = ensure_host_arena(sl, words);
  if (rc == TGP_OK) rc = tgp_batch_results_sync(sl.batch, sl.h_arena, words, sl.stream);
  std::lock_guard<std::mutex> lk(s->mu);
  for (uint32_t i = 0; i < sl.reqs.size(); ++i) {
    Request* r = sl.reqs[i];
    uint64_t off = 0;
    if (rc == TGP_OK) rc = tgp_batch_result_layout(sl.batch, i, &off, nullptr, nullptr, nullptr);
    if (rc == TGP_OK) r->result.assign(sl.h_arena + off, sl.h_arena + off + r->words);
    finish(s, r, rc);
  }
  sl.reqs.clear();
  sl.busy = false;
  if (sl.heavy) s->stats.heavy_in_flight -= 1;
  s->done_cv.notify_all();
  return rc;
}

constexpr size_t kSingleMax = 8;

// Launches the given requests as one batch on the slot.  Up to kSingleMax
// requests run as single-query launches instead (tgp_group_agg, the same
// results): a persistent batch of a few small queries costs more in planning
// and launch than their own small kernels (measured at half the one-launch
// capacity: batched p50 0.30 ms against 0.11 ms one launch each).
void launch_agg(tgp_server* s, AggSlot& sl, std::vector<Request*>&& take, bool heavy) {
  std::vector<tgp_agg_query> qs(take.size());
  for (size_t i = 0; i < take.size(); ++i) qs[i] = take[i]->req.agg;
  const int64_t t0 = now_ns();
  int32_t rc = TGP_OK;
  sl.single = s->cfg.per_query_launch != 0 || (!heavy && take.size() <= kSingleMax);
  if (sl.single) {
    rc = launch_single(s, sl, take);
  } else {
    rc = tgp_batch_set_queries(sl.batch, qs.data(), (uint32_t)qs.size(), sl.stream);
    if (rc == TGP_OK) rc = tgp_batch_run(sl.batch, TGP_BATCH_PERSISTENT, sl.stream);
    if (rc == TGP_OK && hipEventRecord(sl.done_ev, sl.stream) != hipSuccess) {
      set_error("serve: hipEventRecord failed");
      rc = TGP_ERR_HIP;
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->stats.launch_ns += (uint64_t)(now_ns() - t0);
  if (rc != TGP_OK) {
    for (Request* r : take) finish(s, r, rc);
    if (heavy) s->stats.heavy_in_flight -= 1;
    s->done_cv.notify_all();
    return;
  }
  for (Request* r : take) r->t_start = t0;
  sl.reqs = std::move(take);
  sl.busy = true;
  sl.heavy = heavy;
  s->stats.batches += 1;
  s->stats.batched_requests += sl.reqs.size();
  if (sl.reqs.size() > s->stats.max_batch) s->stats.max_batch = sl.reqs.size();
}

// Requests in flight on the other aggregate slots below which a free slot does
// not wait to fill a batch.
constexpr size_t kLightLoad = 16;

void agg_worker(tgp_server* s) {
  (void)hipSetDevice(s->ctx->device);
  const uint32_t maxq = s->cfg.per_query_launch ? 1 : s->cfg.max_batch_queries;
  for (;;) {
    bool progressed = false;
    for (AggSlot& sl : s->agg_slots) {
      if (sl.busy && hipEventQuery(sl.done_ev) == hipSuccess) {
        (void)harvest_agg(s, sl);
        progressed = true;
      }
    }
    // launch: a free slot takes the waiting requests (heavy ones on slot 0)
    for (size_t si = 0; si < s->agg_slots.size(); ++si) {
      AggSlot& sl = s->agg_slots[si];
      if (sl.busy) continue;
      std::vector<Request*> take;
      bool heavy = false;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        const bool heavy_slot = s->heavy_slot && si == 0;
        if (heavy_slot) {
          const uint32_t hmax = s->cfg.per_query_launch ? 1 : maxq;
          while (!s->heavy_q.empty() && take.size() < hmax) {
            take.push_back(s->heavy_q.front());
            s->heavy_q.pop_front();
          }
          if (!take.empty()) {
            heavy = true;
            s->stats.heavy_in_flight += 1;
          }
        } else if (!s->agg_q.empty()) {
          // continuous batching with a bounded wait: while other batches run,
          // gather a larger batch (a quarter of the maximum, or the oldest
          // request's wait reaching batch_wait_us) instead of launching a
          // handful of requests; a device that is idle or runs at most
          // kLightLoad requests (light load: the wait would only add latency)
          // launches at once
          size_t others_inflight = 0;
          for (size_t j = 0; j < s->agg_slots.size(); ++j)
            if (j != si && !(s->heavy_slot && j == 0) && s->agg_slots[j].busy)
              others_inflight += s->agg_slots[j].reqs.size();
          const int64_t waited = now_ns() - s->agg_q.front()->t_arrive;
          const bool go = others_inflight <= kLightLoad || s->agg_q.size() >= (maxq + 3) / 4 ||
                          waited >= (int64_t)s->cfg.batch_wait_us * 1000;
          while (go && !s->agg_q.empty() && take.size() < maxq) {
            take.push_back(s->agg_q.front());
            s->agg_q.pop_front();
          }
        }
      }
      if (take.empty()) continue;
      launch_agg(s, sl, std::move(take), heavy);
      progressed = true;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    bool busy = false;
    for (AggSlot& sl : s->agg_slots) busy |= sl.busy;
    if (s->stop && !busy && s->agg_q.empty() && s->heavy_q.empty()) return;
    if (!progressed) {
      // poll in-flight batches every few microseconds; sleep when idle
      if (busy) s->agg_cv.wait_for(lk, std::chrono::microseconds(20));
      else s->agg_cv.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

// ---- join worker --------------------------------------------------------------------

uint64_t hash_budget(tgp_server* s) {
  uint64_t used = 0, budget = 0, peak = 0;
  (void)tgp_hbm_usage(s->ctx, &used, &budget, &peak);
  // room the tier has now, plus the table pools (all of it is table memory)
  uint64_t room = budget > used ? budget - used : 0;
  room += s->pool_held;
  if (s->cfg.hash_budget_bytes && s->cfg.hash_budget_bytes < room) room = s->cfg.hash_budget_bytes;
  return room;
}

// Drops one join's reference to the shared table instance it probed; returns an
// outgrown instance that no join probes any more (to be destroyed outside the lock).
tgp_hash_table* drop_ref(tgp_server* s, SharedTable* st, tgp_hash_table* t) {  // holds s->mu
  if (t == st->table) {
    st->refs -= 1;
    return nullptr;
  }
  for (size_t i = 0; i < s->retired.size(); ++i)  // outgrown while the join probed it
    if (s->retired[i].first == t && --s->retired[i].second == 0) {
      s->retired.erase(s->retired.begin() + (long)i);
      return t;
    }
  return nullptr;
}

// A probe batch has ended: one copy of all its results, split per request.
int32_t harvest_probe_batch(tgp_server* s, JoinSlot& sl) {
  uint64_t total = 0;
  for (Request* r : sl.batch) total += r->words;
  sl.h_res.resize(total);
  int32_t rc = TGP_OK;
  if (hipMemcpyAsync(sl.h_res.data(), sl.d_res, total * 8, hipMemcpyDeviceToHost,
                     sl.stream) != hipSuccess ||
      hipStreamSynchronize(sl.stream) != hipSuccess) {
    set_error("serve: join result copy failed");
    rc = TGP_ERR_HIP;
  }
  std::vector<tgp_hash_table*> free_now;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    uint64_t off = 0;
    for (size_t i = 0; i < sl.batch.size(); ++i) {
      Request* r = sl.batch[i];
      r->result.assign(sl.h_res.begin() + (long)off, sl.h_res.begin() + (long)(off + r->words));
      off += r->words;
      if (tgp_hash_table* t = drop_ref(s, sl.holds[i].first, sl.holds[i].second))
        free_now.push_back(t);
      finish(s, r, rc);
    }
    sl.batch.clear();
    sl.holds.clear();
    sl.busy = false;
    s->done_cv.notify_all();
  }
  for (tgp_hash_table* t : free_now) (void)tgp_hash_table_destroy(s->ctx, t);
  return rc;
}

// Probes every join of the slot's batch in one tgp_hash_probe_agg_batch launch.
int32_t start_probe_batch(tgp_server* s, JoinSlot& sl) {
  uint64_t total = 0;
  const int64_t t0 = now_ns();
  for (Request* r : sl.batch) {
    r->t_start = t0;
    total += r->words;
  }
  if (sl.d_cap < total) {
    if (sl.d_res) TGP_TRY(hbm_free(s->ctx, sl.d_res));
    sl.d_res = nullptr;
    sl.d_cap = 0;
    void* p = nullptr;
    const uint64_t want = std::max<uint64_t>(total, 2 * sl.d_cap);
    TGP_TRY(hbm_alloc(s->ctx, want * 8, &p));
    sl.d_res = static_cast<uint64_t*>(p);
    sl.d_cap = want;
  }
  std::vector<tgp_join_agg_query> qs(sl.batch.size());
  std::vector<tgp_group_result> outs(sl.batch.size());
  uint64_t off = 0;
  for (size_t i = 0; i < sl.batch.size(); ++i) {
    const Request* r = sl.batch[i];
    const tgp_serve_join& J = r->req.join;
    qs[i] = J.probe;
    qs[i].table = sl.holds[i].second;
    qs[i].build_rows = (uint32_t)J.build_keys.len;
    qs[i].ngroups = r->ngroups;
    uint64_t* base = sl.d_res + off;
    outs[i] = tgp_group_result{base, base + r->ngroups, base + r->ngroups * 3};
    off += r->words;
  }
  TGP_TRY(tgp_hash_probe_agg_batch(s->ctx, qs.data(), (uint32_t)qs.size(), outs.data(),
                                   sl.stream));
  TGP_HIP(hipEventRecord(sl.done_ev, sl.stream));
  return TGP_OK;
}

int32_t harvest_join(tgp_server* s, JoinSlot& sl) {
  if (!sl.batch.empty()) return harvest_probe_batch(s, sl);
  Request* r = sl.req;
  r->result.resize(r->words);
  int32_t rc = TGP_OK;
  if (hipMemcpyAsync(r->result.data(), sl.d_res, r->words * 8, hipMemcpyDeviceToHost,
                     sl.stream) != hipSuccess ||
      hipStreamSynchronize(sl.stream) != hipSuccess) {
    set_error("serve: join result copy failed");
    rc = TGP_ERR_HIP;
  }
  // the table lives in the slot's pool: destroying it frees no device memory
  if (sl.table) {
    const int32_t rd = tgp_hash_table_destroy(s->ctx, sl.table);
    if (rc == TGP_OK) rc = rd;
  }
  tgp_hash_table* free_now = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->hash_in_use -= sl.table_bytes;
    if (sl.shared) free_now = drop_ref(s, sl.shared, sl.shared_table);
    sl.shared = nullptr;
    sl.shared_table = nullptr;
    sl.table = nullptr;
    sl.table_bytes = 0;
    sl.req = nullptr;
    sl.busy = false;
    finish(s, r, rc);
    s->done_cv.notify_all();
  }
  if (free_now) (void)tgp_hash_table_destroy(s->ctx, free_now);
  return rc;
}

SharedTable* find_shared(tgp_server* s, const tgp_serve_join& J) {  // caller holds s->mu
  for (auto& t : s->shared)
    if (t->keys == J.build_keys.data && t->pay == J.build_payload.data &&
        t->kdt == J.build_keys.dtype && (!J.build_payload.data || t->pdt == J.build_payload.dtype))
      return t.get();
  return nullptr;
}

void drop_pool(tgp_server* s, JoinSlot& sl) {  // caller holds s->mu
  if (!sl.pool) return;
  (void)hbm_free(s->ctx, sl.pool);
  s->pool_held -= sl.pool_bytes;
  sl.pool = nullptr;
  sl.pool_bytes = 0;
}

// A slot's pool grows to the largest table it has to hold.  Freeing and
// allocating device memory synchronises the device, so it happens only when a
// join is larger than any before it on this slot.  When the tier is short, the
// pools of idle slots are given back first.
int32_t grow_pool(tgp_server* s, JoinSlot& sl, uint64_t need) {
  std::lock_guard<std::mutex> lk(s->mu);
  drop_pool(s, sl);
  void* p = nullptr;
  int32_t rc = hbm_alloc(s->ctx, need, &p);
  if (rc == TGP_ERR_BUDGET) {
    for (JoinSlot& o : s->join_slots)
      if (&o != &sl && !o.busy) drop_pool(s, o);
    rc = hbm_alloc(s->ctx, need, &p);
  }
  if (rc != TGP_OK) return rc;
  sl.pool = p;
  sl.pool_bytes = need;
  s->pool_held += need;
  return TGP_OK;
}

int32_t start_join(tgp_server* s, JoinSlot& sl, Request* r, uint64_t need) {
  const tgp_serve_join& J = r->req.join;
  r->t_start = now_ns();
  // result buffer: count | sums | oor
  if (sl.d_cap < r->words) {
    if (sl.d_res) TGP_TRY(hbm_free(s->ctx, sl.d_res));
    sl.d_res = nullptr;
    sl.d_cap = 0;
    void* p = nullptr;
    TGP_TRY(hbm_alloc(s->ctx, r->words * 8, &p));
    sl.d_res = static_cast<uint64_t*>(p);
    sl.d_cap = r->words;
  }
  const tgp_column* pay = J.build_payload.data ? &J.build_payload : nullptr;
  tgp_join_agg_query q = J.probe;
  if (sl.builds) {
    // (re)build the build side's shared table over this join's prefix; the tier
    // accounts it from now on, so the reservation taken at admission is released
    tgp_hash_table* t = nullptr;
    tgp_column bk = J.build_keys, bp{};
    bk.len = sl.build_len;
    if (pay) {
      bp = *pay;
      bp.len = sl.build_len;
    }
    const int32_t rc = tgp_hash_build(s->ctx, &bk, pay ? &bp : nullptr, &t, sl.stream);
    std::lock_guard<std::mutex> lk(s->mu);
    s->hash_in_use -= sl.table_bytes;
    sl.table_bytes = 0;
    SharedTable* st = sl.builds;
    sl.builds = nullptr;
    st->building = 0;
    if (rc != TGP_OK) return rc;
    s->stats.table_builds += 1;
    if (st->table) {
      if (st->refs) s->retired.push_back({st->table, st->refs});
      else (void)tgp_hash_table_destroy(s->ctx, st->table);
    }
    st->table = t;
    st->rows = sl.build_len;
    st->refs = 0;
    sl.shared = st;
  }
  if (sl.shared) {
    std::lock_guard<std::mutex> lk(s->mu);
    sl.shared->refs += 1;
    sl.shared_table = sl.shared->table;
    q.table = sl.shared_table;
    q.build_rows = (uint32_t)J.build_keys.len;
  } else {
    if (sl.pool_bytes < need) TGP_TRY(grow_pool(s, sl, need));
    tgp_hash_table* t = nullptr;
    TGP_TRY(tgp_hash_build_in(s->ctx, &J.build_keys, pay, sl.pool, sl.pool_bytes, &t, sl.stream));
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stats.table_builds += 1;
    }
    sl.table = t;
    sl.table_bytes = need;
    q.table = t;
  }
  q.ngroups = r->ngroups;
  tgp_group_result out{sl.d_res, sl.d_res + r->ngroups, sl.d_res + r->ngroups * 3};
  TGP_TRY(tgp_hash_probe_agg(s->ctx, &q, &out, sl.stream));
  TGP_HIP(hipEventRecord(sl.done_ev, sl.stream));
  sl.req = r;
  sl.busy = true;
  return TGP_OK;
}

// One join worker per slot: hash builds are synchronous, so each worker has
// its own thread and the builds of different joins overlap.
void join_worker(tgp_server* s, size_t slot) {
  (void)hipSetDevice(s->ctx->device);
  JoinSlot& sl = s->join_slots[slot];
  for (;;) {
    bool progressed = false;
    if (sl.busy && hipEventQuery(sl.done_ev) == hipSuccess) {
      (void)harvest_join(s, sl);
      progressed = true;
    }
    Request* r = nullptr;
    bool new_batch = false;
    uint64_t need = 0;
    if (!sl.busy && s->cfg.share_build_tables) {
      // a join on a build side whose resident table holds its prefix probes it now;
      // one that needs a (longer) table builds it when the budget allows, in arrival
      // order; joins waiting for a build in progress let later ones pass
      std::lock_guard<std::mutex> lk(s->mu);
      // every queued join whose resident table holds its prefix: one probe batch (the
      // one-launch-per-query form: the first such join, alone)
      for (auto it = s->join_q.begin();
           it != s->join_q.end() && sl.batch.size() < kMaxProbeBatch;) {
        SharedTable* st = find_shared(s, (*it)->req.join);
        if (st && st->table && st->rows >= (*it)->req.join.build_keys.len) {
          s->stats.admitted_joins += 1;
          if (s->cfg.per_query_launch) {
            sl.shared = st;
            r = *it;
            s->join_q.erase(it);
            sl.busy = true;
            break;
          }
          sl.batch.push_back(*it);
          sl.holds.push_back({st, st->table});
          st->refs += 1;
          it = s->join_q.erase(it);
        } else {
          ++it;
        }
      }
      if (!sl.batch.empty()) {
        sl.busy = true;
        new_batch = true;
      }
      for (auto it = s->join_q.begin(); !sl.busy && it != s->join_q.end();) {
        Request* q = *it;
        const tgp_serve_join& J = q->req.join;
        SharedTable* st = find_shared(s, J);
        if (st && st->building) {
          ++it;
          continue;
        }
        // the (re)build covers the longest prefix of this build side that a queued
        // join asks for, so the joins behind this one need no further build
        uint64_t len = J.build_keys.len;
        for (Request* o : s->join_q) {
          const tgp_serve_join& K = o->req.join;
          if (K.build_keys.data == J.build_keys.data && K.build_keys.dtype == J.build_keys.dtype &&
              K.build_payload.data == J.build_payload.data &&
              (!J.build_payload.data || K.build_payload.dtype == J.build_payload.dtype))
            len = std::max<uint64_t>(len, K.build_keys.len);
        }
        need = tgp_hash_table_bytes_for(J.build_keys.len);
        uint64_t used = 0, budget = 0, peak = 0;
        (void)tgp_hbm_usage(s->ctx, &used, &budget, &peak);
        const uint64_t cap = s->cfg.hash_budget_bytes ? s->cfg.hash_budget_bytes : budget;
        if (tgp_hash_table_bytes_for(len) + s->hash_in_use <= hash_budget(s) &&
            tgp_hash_table_bytes_for(len) <= cap)
          need = tgp_hash_table_bytes_for(len);
        else
          len = J.build_keys.len;
        if (need > cap) {
          it = s->join_q.erase(it);
          s->stats.rejected += 1;
          set_error("serve: join needs %llu bytes of hash table, budget %llu",
                    (unsigned long long)need, (unsigned long long)cap);
          finish(s, q, TGP_ERR_BUDGET);
          s->done_cv.notify_all();
          progressed = true;
          continue;
        }
        if (s->hash_in_use + need > hash_budget(s)) {
          if (!q->deferred) {
            q->deferred = true;
            s->stats.deferred += 1;
          }
          break;
        }
        if (!st) {
          s->shared.push_back(std::make_unique<SharedTable>());
          st = s->shared.back().get();
          st->keys = J.build_keys.data;
          st->pay = J.build_payload.data;
          st->kdt = J.build_keys.dtype;
          st->pdt = J.build_payload.dtype;
        }
        st->building = len;
        sl.build_len = len;
        s->join_q.erase(it);
        s->hash_in_use += need;
        s->stats.admitted_joins += 1;
        sl.busy = true;
        sl.builds = st;
        sl.table_bytes = need;  // reserved until the build ends
        r = q;
        break;
      }
    } else if (!sl.busy) {
      std::lock_guard<std::mutex> lk(s->mu);
      while (!s->join_q.empty()) {
        Request* front = s->join_q.front();
        need = tgp_hash_table_bytes_for(front->req.join.build_keys.len);
        const uint64_t room = hash_budget(s);
        uint64_t used = 0, budget = 0, peak = 0;
        (void)tgp_hbm_usage(s->ctx, &used, &budget, &peak);
        const uint64_t cap = s->cfg.hash_budget_bytes ? s->cfg.hash_budget_bytes : budget;
        if (need > cap) {
          // can never fit: reject
          s->join_q.pop_front();
          s->stats.rejected += 1;
          set_error("serve: join needs %llu bytes of hash table, budget %llu",
                    (unsigned long long)need, (unsigned long long)cap);
          finish(s, front, TGP_ERR_BUDGET);
          s->done_cv.notify_all();
          progressed = true;
          continue;
        }
        if (s->hash_in_use + need > room) {
          // wait for admitted tables to be released
          if (!front->deferred) {
            front->deferred = true;
            s->stats.deferred += 1;
          }
          break;
        }
        s->join_q.pop_front();
        s->hash_in_use += need;
        s->stats.admitted_joins += 1;
        sl.busy = true;  // claimed; start_join fills the slot
        sl.table_bytes = need;
        r = front;
        break;
      }
    }
    if (new_batch) {
      const int32_t rc = start_probe_batch(s, sl);
      if (rc != TGP_OK) {
        std::vector<tgp_hash_table*> free_now;
        {
          std::lock_guard<std::mutex> lk(s->mu);
          for (size_t i = 0; i < sl.batch.size(); ++i) {
            if (tgp_hash_table* t = drop_ref(s, sl.holds[i].first, sl.holds[i].second))
              free_now.push_back(t);
            finish(s, sl.batch[i], rc);
          }
          sl.batch.clear();
          sl.holds.clear();
          sl.busy = false;
          s->done_cv.notify_all();
        }
        for (tgp_hash_table* t : free_now) (void)tgp_hash_table_destroy(s->ctx, t);
      }
      progressed = true;
    }
    if (r) {
      const int32_t rc = start_join(s, sl, r, need);
      if (rc != TGP_OK) {
        if (sl.table) (void)tgp_hash_table_destroy(s->ctx, sl.table);
        std::lock_guard<std::mutex> lk(s->mu);
        bool others = false;
        for (JoinSlot& o : s->join_slots) others |= o.busy && &o != &sl;
        if (sl.builds) sl.builds->building = 0;  // failed before its build started
        if (sl.shared && sl.shared_table) {      // failed after taking a reference
          if (sl.shared_table == sl.shared->table) sl.shared->refs -= 1;
          else
            for (auto& rt : s->retired)
              if (rt.first == sl.shared_table) rt.second -= 1;  // freed with the server
        }
        sl.builds = nullptr;
        sl.shared = nullptr;
        sl.shared_table = nullptr;
        s->hash_in_use -= sl.table_bytes;
        sl.table = nullptr;
        sl.table_bytes = 0;
        sl.req = nullptr;
        sl.busy = false;
        if (rc == TGP_ERR_BUDGET && others) {
          // the tier filled up meanwhile: wait for the joins in flight to release it
          s->join_q.push_front(r);
          s->stats.admitted_joins -= 1;
          if (!r->deferred) {
            r->deferred = true;
            s->stats.deferred += 1;
          }
        } else {
          finish(s, r, rc);
          s->done_cv.notify_all();
        }
      }
      progressed = true;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stop && !sl.busy && s->join_q.empty()) return;
    if (!progressed) {
      if (sl.busy) s->join_cv.wait_for(lk, std::chrono::microseconds(20));
      else s->join_cv.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

void release(tgp_server* s) {
  for (AggSlot& sl : s->agg_slots) {
    if (sl.batch) (void)tgp_batch_destroy(sl.batch);
    if (sl.done_ev) (void)hipEventDestroy(sl.done_ev);
    if (sl.stream) (void)tgp_stream_destroy(s->ctx, sl.stream);  // frees its scratch too
    if (sl.h_arena) (void)hipHostFree(sl.h_arena);
    if (sl.d_res) (void)hbm_free(s->ctx, sl.d_res);
  }
  for (auto& st : s->shared)
    if (st->table) (void)tgp_hash_table_destroy(s->ctx, st->table);
  s->shared.clear();
  for (auto& rt : s->retired) (void)tgp_hash_table_destroy(s->ctx, rt.first);
  s->retired.clear();
  for (JoinSlot& sl : s->join_slots) {
    if (sl.table) (void)tgp_hash_table_destroy(s->ctx, sl.table);
    drop_pool(s, sl);
    if (sl.d_res) (void)hbm_free(s->ctx, sl.d_res);
    if (sl.done_ev) (void)hipEventDestroy(sl.done_ev);
    if (sl.stream) (void)tgp_stream_destroy(s->ctx, sl.stream);  // frees its scratch too
  }
}

}  // namespace

extern "C" {

int64_t tgp_serve_now_ns(void) { return now_ns(); }

int32_t tgp_serve_create(tgp_ctx* ctx, const tgp_serve_config* cfg, tgp_server** out) {
  TGP_REQUIRE(ctx && cfg && out, "tgp_serve_create: null argument");
  TGP_REQUIRE(cfg->max_batch_queries > 0 && cfg->max_batch_queries <= (1u << 16),
              "max_batch_queries %u not in 1..65536", cfg->max_batch_queries);
  TGP_REQUIRE(cfg->batch_slots >= 1 && cfg->batch_slots <= 8, "batch_slots %u not in 1..8",
              cfg->batch_slots);
  TGP_REQUIRE(cfg->join_slots >= 1 && cfg->join_slots <= 8, "join_slots %u not in 1..8",
              cfg->join_slots);
  TGP_TRY(check_injected(ctx));
  *out = nullptr;
  auto s = std::make_unique<tgp_server>();
  s->ctx = ctx;
  s->cfg = *cfg;
  s->agg_slots.resize(cfg->batch_slots);
  s->join_slots.resize(cfg->join_slots);
  int32_t rc = TGP_OK;
  s->heavy_slot = cfg->heavy_rows > 0 && cfg->batch_slots >= 2;
  for (size_t si = 0; si < s->agg_slots.size(); ++si) {
    AggSlot& sl = s->agg_slots[si];
    if (rc == TGP_OK) rc = tgp_batch_create(ctx, cfg->max_batch_queries, &sl.batch);
    if (rc == TGP_OK) {
      // the heavy slot's batches are many scans over the same large ranges: the
      // shared scan (order 3) reads each slice once for all of them, in 128-slice chunks
      const bool heavy = s->heavy_slot && si == 0;
      const tgp_batch_tuning t{cfg->wg_per_cu, heavy ? 128u : cfg->chunk_slices,
                               heavy ? 3 : cfg->order, heavy ? 1 : cfg->cached_loads, 0, 1};
      rc = tgp_batch_set_tuning(sl.batch, &t);
    }
    if (rc == TGP_OK && (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreateWithFlags(&sl.done_ev, hipEventDisableTiming) !=
                             hipSuccess)) {
      set_error("tgp_serve_create: stream/event creation failed");
      rc = TGP_ERR_HIP;
    }
  }
  for (JoinSlot& sl : s->join_slots) {
    if (rc == TGP_OK && (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreateWithFlags(&sl.done_ev, hipEventDisableTiming) !=
                             hipSuccess)) {
      set_error("tgp_serve_create: stream/event creation failed");
      rc = TGP_ERR_HIP;
    }
  }
  if (rc != TGP_OK) {
    release(s.get());
    return rc;
  }
  tgp_server* raw = s.release();
  raw->agg_thread = std::thread(agg_worker, raw);
  for (size_t j = 0; j < raw->join_slots.size(); ++j)
    raw->join_threads.emplace_back(join_worker, raw, j);
  *out = raw;
  return TGP_OK;
}

int32_t tgp_serve_destroy(tgp_server* s) {
  if (!s) return TGP_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->agg_cv.notify_all();
  s->join_cv.notify_all();
  if (s->agg_thread.joinable()) s->agg_thread.join();
  for (std::thread& t : s->join_threads)
    if (t.joinable()) t.join();
  release(s);
  delete s;
  return TGP_OK;
}

int32_t tgp_serve_submit(tgp_server* s, const tgp_serve_request* req, int64_t arrival_ns,
                         uint64_t* out_id) {
  TGP_REQUIRE(s && req && out_id, "tgp_serve_submit: null argument");
  const int64_t t_in = now_ns();
  auto r = std::make_unique<Request>();
  r->req = *req;
  TGP_TRY(result_shape(*req, &r->ngroups, &r->naggs, &r->rows));
  r->words = r->ngroups + r->ngroups * (uint64_t)r->naggs * 2 + 1;
  if (req->kind == TGP_SERVE_AGG) {
    // validation now, so a bad request fails its submit, not a whole batch
    uint64_t ng = 0;
    TGP_TRY(tgp_group_agg_ngroups(&req->agg, &ng));
    int32_t strategy = 0, shape = 0;
    TGP_TRY(tgp_group_agg_explain(&req->agg, &strategy, &shape));
  }
  r->t_arrive = arrival_ns ? arrival_ns : now_ns();
  bool is_join = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    TGP_REQUIRE(!s->stop, "tgp_serve_submit: server is stopping");
    r->id = s->next_id++;
    *out_id = r->id;
    Request* raw = r.get();
    s->reqs.emplace(raw->id, std::move(r));
    s->stats.submitted += 1;
    s->stats.submit_ns += (uint64_t)(now_ns() - t_in);
    is_join = raw->req.kind == TGP_SERVE_JOIN_AGG;
    if (is_join) s->join_q.push_back(raw);
    else if (s->heavy_slot && raw->rows >= s->cfg.heavy_rows) s->heavy_q.push_back(raw);
    else s->agg_q.push_back(raw);
  }
  // any idle join worker can take a join; the one aggregate worker takes the rest
  if (is_join) s->join_cv.notify_all();
  else s->agg_cv.notify_one();
  return TGP_OK;
}

int32_t tgp_serve_wait(tgp_server* s, uint64_t id, int64_t timeout_ns, int32_t* out_done) {
  TGP_REQUIRE(s && out_done, "tgp_serve_wait: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  TGP_REQUIRE(it != s->reqs.end(), "tgp_serve_wait: unknown request %llu",
              (unsigned long long)id);
  Request* r = it->second.get();
  if (timeout_ns < 0) s->done_cv.wait(lk, [&] { return r->done; });
  else s->done_cv.wait_for(lk, std::chrono::nanoseconds(timeout_ns), [&] { return r->done; });
  *out_done = r->done ? 1 : 0;
  return TGP_OK;
}

int32_t tgp_serve_result(tgp_server* s, uint64_t id, uint64_t* h_words, uint64_t nwords,
                         tgp_serve_timing* timing) {
  TGP_REQUIRE(s, "tgp_serve_result: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  TGP_REQUIRE(it != s->reqs.end(), "tgp_serve_result: unknown request %llu",
              (unsigned long long)id);
  Request* r = it->second.get();
  TGP_REQUIRE(r->done, "tgp_serve_result: request %llu not finished", (unsigned long long)id);
  if (timing) {
    timing->arrival_ns = r->t_arrive;
    timing->start_ns = r->t_start;
    timing->done_ns = r->t_done;
    timing->status = r->status;
    timing->deferred = r->deferred ? 1 : 0;
    timing->result_words = r->words;
  }
  if (r->status != TGP_OK) {
    set_error("request %llu failed: %s", (unsigned long long)id, r->error.c_str());
    return r->status;
  }
  TGP_REQUIRE(nwords == 0 || h_words, "tgp_serve_result: null buffer");
  TGP_REQUIRE(nwords == 0 || nwords >= r->words, "tgp_serve_result: %llu words needed",
              (unsigned long long)r->words);
  if (nwords) std::memcpy(h_words, r->result.data(), r->words * 8);
  return TGP_OK;
}

int32_t tgp_serve_release(tgp_server* s, uint64_t id) {
  TGP_REQUIRE(s, "tgp_serve_release: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  TGP_REQUIRE(it != s->reqs.end() && it->second->done, "tgp_serve_release: request %llu "
              "unknown or not finished", (unsigned long long)id);
  s->reqs.erase(it);
  return TGP_OK;
}

int32_t tgp_serve_replay(tgp_server* s, const tgp_serve_request* reqs,
                         const int64_t* arrival_offset_ns, uint32_t n, uint64_t* out_ids,
                         int64_t* out_start_ns) {
  TGP_REQUIRE(s && (n == 0 || (reqs && arrival_offset_ns && out_ids)),
              "tgp_serve_replay: null argument");
  for (uint32_t i = 1; i < n; ++i)
    TGP_REQUIRE(arrival_offset_ns[i] >= arrival_offset_ns[i - 1],
                "tgp_serve_replay: arrival offsets must be non-decreasing");
  const int64_t start = now_ns() + 1000000;  // 1 ms to set up
  if (out_start_ns) *out_start_ns = start;
  int32_t rc = TGP_OK;
  std::string err;
  // the open-loop submitter: request i enters at start + offset[i] whatever the
  // state of the earlier ones, and carries that time as its arrival
  std::thread submitter([&] {
    for (uint32_t i = 0; i < n && rc == TGP_OK; ++i) {
      const int64_t at = start + arrival_offset_ns[i];
      const int64_t t = now_ns();
      if (at - t > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(at - t - 100000));
      while (now_ns() < at) {
      }
      rc = tgp_serve_submit(s, &reqs[i], at, &out_ids[i]);
      if (rc != TGP_OK) err = tgp_last_error();
    }
  });
  submitter.join();
  if (rc != TGP_OK) {
    set_error("tgp_serve_replay: %s", err.c_str());
    return rc;
  }
  for (uint32_t i = 0; i < n; ++i) {
    int32_t done = 0;
    TGP_TRY(tgp_serve_wait(s, out_ids[i], -1, &done));
  }
  return TGP_OK;
}

int32_t tgp_serve_get_stats(tgp_server* s, tgp_serve_stats* out) {
  TGP_REQUIRE(s && out, "tgp_serve_get_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return TGP_OK;
}

}  // extern "C"
