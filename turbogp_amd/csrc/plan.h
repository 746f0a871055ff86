// Host-side query planner: validates a tgp_agg_query (or a scan's predicate
// list) and lays it out as a DevPlan for the tile engine.
#pragma once

#include "tile_engine.h"

namespace tgp {

enum AggStrategy {
  kAggReg1 = 0,   // one group: per-thread registers
  kAggReg8 = 1,   // <= 8 groups, <= 4 aggregates: per-thread registers
  kAggLds = 2,    // per-workgroup LDS table + 64-bit LDS atomics
  kAggGlobal = 3  // global table, 64-bit device atomics per row
};

struct PlanInfo {
  int nst = 0;              // staging rounds (16-byte chunks per thread per tile)
  uint32_t lds_bytes = 0;   // dynamic LDS of the kernel (stage + table)
  int strategy = kAggReg1;
};

// Largest staging-round count the kernels are instantiated for.
constexpr int kMaxNst = 17;
// LDS table limit of the kAggLds strategy.
constexpr uint32_t kLdsTableBytes = 48 * 1024;

int nst_bucket(int nst);  // smallest instantiated round count >= nst (0 => too big)

// Fills the slot, chunk and tile fields of P for columns `cols`.
int32_t plan_slots(const tgp_column* cols, int32_t ncols, uint64_t row_begin, uint64_t row_end,
                   DevPlan* P, PlanInfo* info);
int32_t plan_preds(const tgp_pred* preds, int32_t npreds, int32_t ncols, DevPlan* P);
// Full aggregation plan (slots, predicates, group parts, aggregates, strategy).
int32_t plan_agg(const tgp_agg_query* q, DevPlan* P, PlanInfo* info);
// Round-up multiplier for unsigned 32-bit division by d >= 2.
void divisor_magic(uint32_t d, uint32_t* magic, int32_t* shift);

}  // namespace tgp
