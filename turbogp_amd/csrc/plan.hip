// Host-side planner for the tile engine: compiles a query into per-slot ops.
#include <vector>

#include "plan.h"

namespace tgp {

// Unsigned division by d (2 <= d < 2^31) as q = (t + ((n - t) >> 1)) >> (l - 1),
// t = mulhi(m, n), l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, section 4).  Exact for every 32-bit n.
void divisor_magic(uint32_t d, uint32_t* magic, int32_t* shift) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  unsigned __int128 num = ((unsigned __int128)1 << 32) * (((unsigned __int128)1 << l) - d);
  *magic = (uint32_t)(num / d + 1);
  *shift = l - 1;
}

int32_t plan_slots(const tgp_column* cols, int32_t ncols, uint64_t row_begin, uint64_t row_end,
                   DevPlan* P, PlanInfo* info) {
  TGP_REQUIRE(ncols >= 1 && ncols <= TGP_MAX_COLS, "ncols %d outside [1,%d]", ncols,
              TGP_MAX_COLS);
  TGP_REQUIRE(cols != nullptr, "null column array");
  TGP_REQUIRE(row_begin <= row_end, "row_begin %llu > row_end %llu",
              (unsigned long long)row_begin, (unsigned long long)row_end);
  P->row_begin = row_begin;
  P->row_end = row_end;
  P->base0 = row_begin & ~63ull;
  P->ntiles = row_end > row_begin ? (row_end - P->base0 + kTileRows - 1) / kTileRows : 0;
  P->nslots = ncols;
  uint32_t off = 0;
  for (int s = 0; s < ncols; ++s) {
    TGP_TRY(check_column(&cols[s], "column"));
    TGP_REQUIRE(cols[s].len >= row_end, "column %d has %llu rows < row_end %llu", s,
                (unsigned long long)cols[s].len, (unsigned long long)row_end);
    DevSlot& S = P->slots[s];
    S.data = static_cast<const uint8_t*>(cols[s].data);
    S.width = dtype_width(cols[s].dtype);
    S.limit_bytes = cols[s].len * (uint64_t)S.width;
    S.lds_off = off;
    off += (uint32_t)(kTileRows * S.width);
  }
  P->stage_bytes = off;
  info->lds_bytes = 2 * off;
  return TGP_OK;
}

namespace {
DevCol col_of(const tgp_column* cols, const DevPlan* P, int s) {
  DevCol c{};
  c.lds_off = P->slots[s].lds_off;
  c.width = P->slots[s].width;
  c.is_signed = dtype_signed(cols[s].dtype) ? 1 : 0;
  return c;
}

int32_t plan_preds(const tgp_column* cols, const tgp_pred* preds, int32_t npreds, int32_t ncols,
                   DevPlan* P) {
  TGP_REQUIRE(npreds >= 0 && npreds <= TGP_MAX_PREDS, "npreds %d outside [0,%d]", npreds,
              TGP_MAX_PREDS);
  TGP_REQUIRE(npreds == 0 || preds != nullptr, "null predicate array");
  P->npreds = npreds;
  for (int p = 0; p < npreds; ++p) {
    TGP_REQUIRE(preds[p].col >= 0 && preds[p].col < ncols, "predicate %d: column %d out of range",
                p, preds[p].col);
    TGP_REQUIRE(preds[p].op >= TGP_EQ && preds[p].op <= TGP_GE, "predicate %d: bad op %d", p,
                preds[p].op);
    DevPred& d = P->preds[p];
    d.c = col_of(cols, P, preds[p].col);
    d.cmp = preds[p].op;
    d.is_signed = preds[p].is_signed ? 1 : 0;
    d.lit = preds[p].lit;
  }
  return TGP_OK;
}
}  // namespace

int32_t plan_scan(const tgp_column* cols, int32_t ncols, const tgp_pred* preds, int32_t npreds,
                  uint64_t row_begin, uint64_t row_end, DevPlan* P, PlanInfo* info) {
  std::memset(P, 0, sizeof(*P));
  TGP_TRY(plan_slots(cols, ncols, row_begin, row_end, P, info));
  TGP_TRY(plan_preds(cols, preds, npreds, ncols, P));
  P->ngroups = 1;
  return TGP_OK;
}

int32_t plan_agg(const tgp_agg_query* q, DevPlan* P, PlanInfo* info) {
  TGP_REQUIRE(q != nullptr, "null query");
  std::memset(P, 0, sizeof(*P));
  TGP_TRY(plan_slots(q->cols, q->ncols, q->row_begin, q->row_end, P, info));
  TGP_TRY(plan_preds(q->cols, q->preds, q->npreds, q->ncols, P));
  TGP_REQUIRE(q->nparts >= 0 && q->nparts <= TGP_MAX_GROUP_PARTS, "nparts %d out of range",
              q->nparts);
  TGP_REQUIRE(q->naggs >= 0 && q->naggs <= TGP_MAX_AGGS, "naggs %d out of range", q->naggs);
  P->nparts = q->nparts;
  P->naggs = q->naggs;
  uint64_t ngroups = 1;
  for (int k = 0; k < q->nparts; ++k) {
    const tgp_group_part& g = q->parts[k];
    TGP_REQUIRE(g.col >= 0 && g.col < q->ncols, "group part %d: column %d out of range", k, g.col);
    TGP_REQUIRE(g.divisor >= 1, "group part %d: divisor %d < 1", k, g.divisor);
    TGP_REQUIRE(g.card >= 1 && g.card <= (1ll << 31), "group part %d: card %lld out of range", k,
                (long long)g.card);
    DevPart& D = P->parts[k];
    D.c = col_of(q->cols, P, g.col);
    D.divisor = g.divisor;
    D.magic = 0;
    D.shift = 0;
    if (g.divisor > 1) divisor_magic((uint32_t)g.divisor, &D.magic, &D.shift);
    D.base = g.base;
    D.card = g.card;
    D.stride = (int64_t)ngroups;
    ngroups *= (uint64_t)g.card;
    TGP_REQUIRE(ngroups <= (1ull << 31), "group key space %llu too large",
                (unsigned long long)ngroups);
  }
  P->ngroups = ngroups;
  for (int a = 0; a < q->naggs; ++a) {
    const tgp_agg& A = q->aggs[a];
    DevAgg& D = P->aggs[a];
    TGP_REQUIRE(A.kind >= TGP_AGG_SUM && A.kind <= TGP_AGG_MAX_U64, "agg %d: bad kind %d", a,
                A.kind);
    TGP_REQUIRE(A.nfactors >= 0 && A.nfactors <= TGP_MAX_FACTORS, "agg %d: nfactors %d", a,
                A.nfactors);
    TGP_REQUIRE(A.kind != TGP_AGG_COUNT_NONZERO || A.nfactors >= 1,
                "agg %d: COUNT(col) needs its column in fcol[0]", a);
    D.kind = A.kind;
    D.nf = A.kind == TGP_AGG_COUNT_NONZERO ? 1 : A.nfactors;
    if (is_minmax(A.kind)) P->minmax_kinds |= (uint32_t)A.kind << (3 * a);
    D.coef = A.coef;
    for (int k = 0; k < D.nf; ++k) {
      TGP_REQUIRE(A.fcol[k] >= 0 && A.fcol[k] < q->ncols, "agg %d factor %d: column %d", a, k,
                  A.fcol[k]);
      D.f[k].c = col_of(q->cols, P, A.fcol[k]);
      D.f[k].fa = A.fa[k];
      D.f[k].fb = A.fb[k];
    }
  }
  P->sel_bitmap = q->d_sel_bitmap;
  const uint64_t table_bytes = ngroups * (1 + 2 * (uint64_t)q->naggs) * 8;
  // copies of the LDS table's words while they fit kLdsCopyBytes (up to 32: the
  // lanes of one atomic then rarely meet on a word)
  uint32_t rs = 0;
  while (rs < 5 && (table_bytes << (rs + 1)) <= kLdsCopyBytes) ++rs;
  P->lds_rshift = rs;
  if (ngroups == 1)
    info->strategy = kAggReg1;
  else if (ngroups <= 8 && q->naggs <= 4)  // faster than an LDS table with copies
    info->strategy = kAggReg8;
  else if (q->row_end - q->row_begin >= (1ull << 32))  // split sums need fewer rows
    info->strategy = kAggGlobalExact;
  else if (table_bytes <= kLdsTableBytes)
    info->strategy = kAggLds;
  else
    info->strategy = kAggGlobal;
  if (info->strategy == kAggGlobal) {
    // partitions of 2^s groups whose table (count, two words per aggregate) fits
    // kPartTableBytes; a record keeps the local group in 16 bits
    const uint64_t wbytes = (1 + 2 * (uint64_t)q->naggs) * 8;
    uint32_t sh = 0;
    while (sh < 16 && (wbytes << (sh + 1)) <= kPartTableBytes) ++sh;
    const uint64_t np = (ngroups + (1ull << sh) - 1) >> sh;
    if (np <= kPartMax) {
      info->part_shift = sh;
      info->nparts = (uint32_t)np;
      if (q->row_end - q->row_begin >= kPartMinRows) info->strategy = kAggPart;
    }
  }
  if (info->strategy == kAggLds) info->lds_bytes += (uint32_t)(table_bytes << rs);
  if ((info->strategy == kAggReg1 || info->strategy == kAggReg8) &&
      info->lds_bytes < kRegReduceBytes)
    info->lds_bytes = kRegReduceBytes;
  return TGP_OK;
}

}  // namespace tgp
