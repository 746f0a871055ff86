// Host-side query planner: validates a tgp_agg_query (or a scan's predicate
// list) and compiles it into the tile engine's per-slot program (DevPlan).
#pragma once

#include "tile_engine.h"

namespace tgp {

enum AggStrategy {
  kAggReg1 = 0,   // one group: per-thread registers
  kAggReg8 = 1,   // <= 8 groups, <= 4 aggregates: per-thread registers
  kAggLds = 2,    // per-workgroup LDS table + 64-bit LDS atomics
  kAggGlobal = 3,       // global table of split sums, 64-bit device atomics per row
  kAggGlobalExact = 4,  // global table of int128 sums (2^32 rows or more)
  kAggPart = 5          // rows partitioned by group range, one LDS table per partition
};

struct PlanInfo {
  uint32_t lds_bytes = 0;  // dynamic LDS of the kernel (2 stage buffers + table)
  int strategy = kAggReg1;
  // kAggPart geometry, set whenever the groups can be partitioned (nparts > 0):
  // 2^part_shift groups per partition
  uint32_t part_shift = 0, nparts = 0;
};

// Radix-partitioned aggregation (kAggPart): the LDS table of one partition, the
// most partitions (the scatter keeps 12 bytes of LDS per partition), and the
// smallest row range that takes it by default (below it the global table's
// device atomics cost less than the extra passes).
constexpr uint32_t kPartTableBytes = 64 * 1024;
constexpr uint32_t kPartMax = 4096;
constexpr uint64_t kPartMinRows = 1ull << 18;

// A work item of the batch's generic path: tiles [t0, t1) of plan q, aggregated in
// an LDS table of split sums (mode 1), straight into the result as split sums
// (mode 0, fewer than 2^32 rows) or as int128 words (mode 2).
struct GenericItem {
  uint32_t q, t0, t1, mode;
};

// The batch's queries whose groups exceed an LDS table (kAggPart's passes, for all
// of them at once): per query, its plans, partition geometry, its items of
// kGenericTiles tiles (histogram and scatter) and where its counts and records live.
struct PartQuery {
  uint32_t plan;      // full plan: index in the generic plans
  uint32_t hplan;     // histogram plan (predicate and key columns only)
  uint32_t shift, nparts;
  uint32_t nitems, naggs;
  uint64_t mat_base;  // counts: mat[mat_base + p * nitems + j] (item j, partition p)
  uint64_t grp_base;  // records (local index l): grp[grp_base + l],
  uint64_t val_base;  //   val[val_base + l * naggs + a]
};
struct PartItem {  // tiles [t0, t1) of query pq, its item j
  uint32_t pq, j, t0, t1;
};
struct PartChunk {  // records [c * kPartChunk, (c + 1) * kPartChunk) of query pq
  uint32_t pq, c;
};
constexpr uint64_t kPartChunk = 1ull << 15;  // records per aggregation work item

#ifndef __HIPCC_RTC__
// Turns MIN / MAX codes (kinds: 3 bits per aggregate) of a result of ngroups x naggs
// into values (0 for an empty group); the sums are left as they are.
int32_t finalize_minmax(hipStream_t s, unsigned long long* cnt, unsigned long long* sums,
                        uint64_t ngroups, int32_t naggs, uint32_t kinds);
// The histogram plan of q for kAggPart: only its predicate and key columns.
int32_t plan_part_hist(const tgp_agg_query* q, DevPlan* H);
// Queues the passes of the batch's partitioned queries: histogram items, the scan of
// mat (mat_n words, the total at mat[mat_n]; bsum ceil(mat_n / 4096) + 1 words),
// scatter items, aggregation chunks.  counters: 3 words, zero.
int32_t launch_batch_part(tgp_ctx* ctx, const DevPlan* d_plans, const DevPlan* d_hplans,
                          const PartQuery* d_pq, const PartItem* d_items, uint64_t nitems,
                          const PartChunk* d_chunks, uint64_t nchunks, uint32_t hstage_max,
                          uint32_t stage_max, uint32_t nparts_max, uint32_t table_max,
                          unsigned long long* counters, unsigned long long* mat, uint64_t mat_n,
                          uint64_t* bsum, uint16_t* grp, uint64_t* val, hipStream_t s);
#endif

// Queues the generic items (one persistent launch; 2 stage buffers of stage_max
// bytes + an LDS table of table_max bytes per workgroup) and the MIN/MAX
// finalisation of the plans in d_minmax (minmax_words: the largest groups x aggs).
#ifndef __HIPCC_RTC__
int32_t launch_batch_generic(tgp_ctx* ctx, const DevPlan* d_plans, const GenericItem* d_items,
                             uint64_t nitems, unsigned long long* counter, uint32_t stage_max,
                             uint32_t table_max, const uint32_t* d_minmax, uint32_t nminmax,
                             uint64_t minmax_words, hipStream_t s);
#endif

// LDS table limit of the kAggLds strategy, and of its copies.
constexpr uint32_t kLdsTableBytes = 48 * 1024;
constexpr uint32_t kLdsCopyBytes = 16 * 1024;
// LDS of the register strategies' final workgroup reduction.
constexpr uint32_t kRegReduceBytes = 4 * 8 * (1 + 2 * TGP_MAX_AGGS) * 8;

// Fills slots and tile geometry of P for columns `cols` and the row range.
int32_t plan_slots(const tgp_column* cols, int32_t ncols, uint64_t row_begin, uint64_t row_end,
                   DevPlan* P, PlanInfo* info);
// Predicate-only program (scans).
int32_t plan_scan(const tgp_column* cols, int32_t ncols, const tgp_pred* preds, int32_t npreds,
                  uint64_t row_begin, uint64_t row_end, DevPlan* P, PlanInfo* info);
// Full aggregation program (predicates, group parts, aggregates, strategy).
int32_t plan_agg(const tgp_agg_query* q, DevPlan* P, PlanInfo* info);
// Round-up multiplier for unsigned 32-bit division by d >= 2.
void divisor_magic(uint32_t d, uint32_t* magic, int32_t* shift);

#ifndef __HIPCC_RTC__
// Workgroups per CU for the persistent kernels: as many as LDS and registers
// allow, at most 4.
template <typename K>
inline int32_t launch_tiles(tgp_ctx* ctx, K kern, const DevPlan& P, uint32_t lds_bytes,
                            hipStream_t s) {
  if (P.ntiles == 0) return TGP_OK;
  if (lds_bytes > 65536)
    TGP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_bytes));
  int per_cu = 0;
  TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, lds_bytes));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 4) per_cu = 4;
  const int64_t want = ctx->tune_wg_per_cu.load();
  if (want > 0 && want < per_cu) per_cu = (int)want;
  uint64_t grid = (uint64_t)persistent_grid(ctx, per_cu);
  const int64_t gmax = ctx->tune_grid_max.load();
  if (gmax > 0 && grid > (uint64_t)gmax) grid = (uint64_t)gmax;
  if (grid > P.ntiles) grid = P.ntiles;
  kern<<<dim3((unsigned)grid), dim3(kBlock), lds_bytes, s>>>(P);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}
#endif  // __HIPCC_RTC__

}  // namespace tgp
