/*
 * turbogp_gfx950.h -- C ABI of the MI355X (gfx950) operator library.
 *
 * This is the drop-in boundary for turboGP's data-parallel operators:
 * predicate scan -> selection bitmap -> compaction, grouped aggregation with
 * exact integer sums, and hash-join build/probe.  Every entry point is plain
 * C: integers, plain pointers and sizes, opaque handles.  No torch or HIP
 * types appear in a signature (a stream is passed as `void*`, NULL = the
 * context's own stream).
 *
 * Conventions
 *   - Every function returns int32_t: TGP_OK (0) or a negative TGP_ERR_*.
 *     Nothing aborts; tgp_last_error() gives a message for the calling thread.
 *   - Buffers named d_* are DEVICE pointers (allocate them with tgp_hbm_alloc,
 *     which accounts them in the context's HBM tier).  Buffers named h_* are
 *     host pointers.
 *   - Operator calls are asynchronous on `stream` unless the name ends in
 *     _sync.  Calls that overlap on one context and one stream are ordered;
 *     different streams may run concurrently (results of two calls never share
 *     writable memory unless the caller passes the same output buffer).
 *   - Integer semantics follow the reference engine: a row value is computed
 *     in wrapping 64-bit two's complement (query_interpreter/expr.rs binop,
 *     :2037-2104); sums are accumulated EXACTLY in 128 bits (the reference
 *     sums expressions in f64, query_interpreter/aggregate.rs:1485-1792; the
 *     exact sum is the value the f64 approximates).
 *
 * Reference interfaces replaced (paths relative to the turboGP checkout):
 *   src/kernel/mod.rs:47-74     Operator                 -> TGP_OP_*
 *   src/kernel/mod.rs:84-92     PredicateOp              -> tgp_kernel_params.pred*_op
 *   src/kernel/mod.rs:96-148    KernelParams/KernelResult-> tgp_kernel_params/_result
 *   src/kernel/mod.rs:158-177   trait Kernel::execute    -> tgp_kernel_execute
 *   src/kernel/mod.rs:179-242   KernelTable select/list  -> tgp_kernel_table_*
 *   src/memory/tier.rs:11-29    MemoryTier (Hbm)         -> TGP_TIER_*, tgp_hbm_*
 *   src/datasource/table.rs:25  Table column cells       -> tgp_column
 *   src/exec/bitmap.rs:42-314   Bitmap + SetBitIter      -> tgp_scan_filter, tgp_bitmap_*
 *   src/exec/bitmap.rs:321-477  filter_{eq..ge}_{u64,i64}-> tgp_pred
 *   query_interpreter/expr.rs:419-1577  build_mask (AND/OR/BETWEEN) -> tgp_scan_filter + tgp_bitmap_combine
 *   query_interpreter/aggregate.rs:507-697, accum_grouped.rs:55-409 -> tgp_group_agg
 *   accum_grouped.rs:198-335, engine/dispatch.rs:1125-1140, :1377-1385 (GROUP BY any
 *       key tuple)                                                  -> tgp_hash_group_agg_sync
 *   kernel/aggregate.rs:19-147 SumF64*, exec/vectorized.rs:128-136, :525-533 -> tgp_sum_f64_sync
 *   kernel/aggregate.rs:149-185 CountDistinctScalar   -> tgp_kernel_execute (AggregateCountDistinct)
 *   src/exec/join_hash_table.rs:38-241, kernel/hash.rs:84-225      -> tgp_hash_build / tgp_hash_probe_*
 *   query_interpreter/join.rs:544-939 hash_join_with_keys (+ A9 agg)-> tgp_hash_probe_agg
 *   src/exec/join.rs:56-199 hash_join / semi_join / anti_join       -> tgp_hash_probe_pairs / _mark
 */
#ifndef TURBOGP_GFX950_H
#define TURBOGP_GFX950_H

#ifndef __HIPCC_RTC__ /* hipRTC (run-time compiled kernels) predefines these types */
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define TGP_ABI_VERSION 1

/* ---- status codes -------------------------------------------------------- */
#define TGP_OK 0
#define TGP_ERR_INVALID (-1)     /* bad argument (null, size, range, alignment) */
#define TGP_ERR_HIP (-2)         /* a HIP runtime call failed */
#define TGP_ERR_OOM (-3)         /* device allocation failed */
#define TGP_ERR_BUDGET (-4)      /* HBM-tier budget would be exceeded */
#define TGP_ERR_UNSUPPORTED (-5) /* shape outside the compiled limits */
#define TGP_ERR_NO_DEVICE (-6)   /* no usable gfx950 device */
#define TGP_ERR_CAPACITY (-7)    /* hash table or output buffer too small */
#define TGP_ERR_CANCELLED (-8)   /* request cancelled before it ran */
#define TGP_ERR_INJECTED (-9)    /* failure injected by tgp_debug_inject_failure */

/* ---- element types and comparisons --------------------------------------- */
enum tgp_dtype {
  TGP_I8 = 0, TGP_U8 = 1, TGP_I16 = 2, TGP_U16 = 3,
  TGP_I32 = 4, TGP_U32 = 5, TGP_I64 = 6, TGP_U64 = 7
};
/* bitmap.rs filter_{eq,ne,lt,le,gt,ge}_*: EQ/NE bitwise, the rest ordered. */
enum tgp_cmp { TGP_EQ = 0, TGP_NE = 1, TGP_LT = 2, TGP_LE = 3, TGP_GT = 4, TGP_GE = 5 };
enum tgp_bitop { TGP_BIT_AND = 0, TGP_BIT_OR = 1, TGP_BIT_ANDNOT = 2 };

/* Compile-time limits of the fused kernels. */
#define TGP_MAX_COLS 8   /* distinct columns referenced by one query */
#define TGP_MAX_PREDS 8  /* AND-ed predicate leaves */
#define TGP_MAX_AGGS 6   /* aggregates besides the implicit COUNT(*) */
#define TGP_MAX_FACTORS 3
#define TGP_MAX_GROUP_PARTS 2

/* A device-resident column at its native width (row 0 at `data`). */
typedef struct tgp_column {
  const void* data;  /* device pointer */
  uint64_t len;      /* rows */
  int32_t dtype;     /* enum tgp_dtype */
  int32_t _pad;
} tgp_column;

/* One predicate leaf: cols[col] <op> lit.  Values are widened to 64 bits
 * (sign- or zero-extended by dtype) and compared as int64 when is_signed
 * (bitmap.rs filter_*_i64, INT columns) or as uint64 otherwise
 * (bitmap.rs filter_*_u64, DATE columns; expr.rs:1482-1508). */
typedef struct tgp_pred {
  int32_t col;
  int32_t op;        /* enum tgp_cmp */
  int32_t is_signed;
  int32_t _pad;
  int64_t lit;
} tgp_pred;

/* One part of a composite dense group key:
 *   part = (v / divisor) - base,  0 <= part < card   (divisor >= 1; v >= 0
 *   when divisor > 1).  group id = part0 + card0 * part1.
 * A row whose part falls outside [0, card) is counted in no group and is
 * reported through tgp_group_result.n_out_of_range. */
typedef struct tgp_group_part {
  int32_t col;
  int32_t divisor;
  int64_t base;
  int64_t card;
} tgp_group_part;

enum tgp_agg_kind {
  TGP_AGG_SUM = 0,           /* SUM(coef * prod_k (fa[k] + fb[k]*cols[fcol[k]])) */
  TGP_AGG_COUNT_NONZERO = 1, /* COUNT(col): rows with cols[fcol[0]] != 0
                                (accum_grouped.rs:243-247 counts non-zero cells) */
  /* MIN / MAX of the same per-row value as SUM (wrapping 64-bit; for MIN(col):
   * nfactors 1, coef 1, fa 0, fb 1).  The result's sum words hold the value:
   * low word = the 64 bits, high word = their sign extension (signed kinds) or 0
   * (unsigned kinds), so the int128 reading is the value; 0 for a group with no
   * rows. */
  TGP_AGG_MIN = 2,     /* signed int64 order: minmax_vec on INT (aggregate.rs:1822-1836) */
  TGP_AGG_MAX = 3,
  TGP_AGG_MIN_U64 = 4, /* raw u64 order: MinCol/MaxCol (accum_grouped.rs:269-280, :320-329) */
  TGP_AGG_MAX_U64 = 5
};

typedef struct tgp_agg {
  int32_t kind;      /* enum tgp_agg_kind */
  int32_t nfactors;  /* 0..TGP_MAX_FACTORS (0 => SUM(coef)) */
  int32_t fcol[TGP_MAX_FACTORS];
  int32_t _pad;
  int64_t coef;
  int64_t fa[TGP_MAX_FACTORS];
  int64_t fb[TGP_MAX_FACTORS];
} tgp_agg;

/* A filter + GROUP BY + aggregate query over rows [row_begin, row_end). */
typedef struct tgp_agg_query {
  uint64_t row_begin, row_end;
  tgp_column cols[TGP_MAX_COLS];
  int32_t ncols, npreds, nparts, naggs;
  tgp_pred preds[TGP_MAX_PREDS];
  tgp_group_part parts[TGP_MAX_GROUP_PARTS];
  tgp_agg aggs[TGP_MAX_AGGS];
  const uint64_t* d_sel_bitmap; /* optional extra AND mask (bit i <-> row_begin+i), or NULL */
} tgp_agg_query;

/* Dense result of tgp_group_agg: for g < ngroups,
 *   count[g]                 COUNT(*) of rows in group g (0 => group absent)
 *   sums[(g*naggs + a)*2+0]  low 64 bits of the exact int128 sum of agg a
 *   sums[(g*naggs + a)*2+1]  high 64 bits (two's complement)
 * Device buffers sized ngroups and ngroups*naggs*2 words. */
typedef struct tgp_group_result {
  uint64_t* d_count;
  uint64_t* d_sums;
  uint64_t* d_n_out_of_range; /* one word, may be NULL */
} tgp_group_result;

/* ---- context, device, streams --------------------------------------------- */
typedef struct tgp_ctx tgp_ctx;

typedef struct tgp_device_info {
  char name[64];
  char arch[32];          /* e.g. "gfx950" */
  int32_t cu_count;
  int32_t wave_size;
  int64_t hbm_bytes;
  int64_t lds_bytes_per_cu;
  int64_t l2_bytes;
  int32_t clock_khz;
  int32_t mem_clock_khz;
} tgp_device_info;

/* Mirrors KernelTable::new + detect_cpu (kernel/mod.rs:191, cpu.rs:67):
 * binds `device`, checks it is gfx950, creates the context stream and an HBM
 * tier with a byte budget (0 => 90% of the device's HBM). */
int32_t tgp_ctx_create(int32_t device, uint64_t hbm_budget_bytes, tgp_ctx** out);
int32_t tgp_ctx_destroy(tgp_ctx* ctx);
int32_t tgp_device_info_get(tgp_ctx* ctx, tgp_device_info* out);
const char* tgp_last_error(void);
int32_t tgp_abi_version(void);
/* Every stream the library runs work on gets device scratch of its own (kept
 * for reuse).  tgp_stream_destroy waits for the stream, releases that scratch
 * and destroys the stream; it also accepts streams made outside the library.
 * Scratch of streams destroyed elsewhere is released by tgp_ctx_destroy. */
int32_t tgp_stream_create(tgp_ctx* ctx, void** out_stream);
int32_t tgp_stream_destroy(tgp_ctx* ctx, void* stream);
int32_t tgp_stream_sync(tgp_ctx* ctx, void* stream);

/* Device-side timing on a stream (hipEvents); used by benchmarks. */
int32_t tgp_event_create(tgp_ctx* ctx, void** out_event);
int32_t tgp_event_destroy(tgp_ctx* ctx, void* event);
int32_t tgp_event_record(tgp_ctx* ctx, void* event, void* stream);
/* Milliseconds between two recorded events (waits for `end`). */
int32_t tgp_event_elapsed_ms(tgp_ctx* ctx, void* start, void* end, float* out_ms);

/* Launch tunables (used by the autotuner; 0 = automatic). */
enum tgp_tune_key {
  TGP_TUNE_WG_PER_CU = 0, /* workgroups per CU of the persistent scan/aggregate kernels */
  TGP_TUNE_GRID_MAX = 1,  /* upper bound on their grid size */
  TGP_TUNE_SPECIALIZE = 2, /* 1 (default): use shape-specialised kernels when a query
                              matches one; 0: always the generic engine */
  TGP_TUNE_SPEC_KERNEL = 3, /* specialised kernel form: 1: a shared per-wave histogram
                               folded into registers after every slice; 2: waves with
                               replicated LDS histograms (where the groups fit); 0
                               (default): 2 for queries of 2..4 groups and >= 4M rows,
                               else 1 */
  TGP_TUNE_CHUNK = 4,       /* form 2's schedule: 0 (default) static interleave of slices;
                               n > 0: chunks of n slices (<= 4096) taken from a counter */
  TGP_TUNE_SLICES_PER_WAVE = 5, /* fewest 256-row slices per wave of the specialised
                                   kernels before their grid shrinks (0 = default 2) */
  TGP_TUNE_JOIN_ATOMICS = 6,    /* join GROUP BY a probe column with more groups than LDS
                                   holds: 0 (default) radix-partitioned, 1 global atomics */
  TGP_TUNE_JOIN_PACK_BITS = 7,  /* multi-column join tables built from now on keep only the
                                   low n bits (1..63) of the packed key, so that distinct
                                   keys collide and the full-key verification is exercised
                                   (tests); 0 (default): all 64 bits */
  TGP_TUNE_DISTINCT_SHRINK = 8, /* exact COUNT(DISTINCT) over >= 2^20 rows sizes its set from
                                   the HyperLogLog estimate; n > 0 divides that size by 2^n
                                   (0..30), so that the set overflows and the count is redone
                                   in a set sized by the row count (tests); 0 (default) */
  TGP_TUNE_GROUP_TAG_BITS = 9,  /* hash GROUP BY over several key columns: n in 1..32 makes
                                   every such query take the hashed form with full-key
                                   verification and keeps only n bits of each slot's hash tag,
                                   so that distinct tuples collide (tests); 0 (default): exact
                                   packing when the key ranges fit 63 bits */
  TGP_TUNE_JOIN_HOT = 10,       /* hot-key cache of join probes (a sampled image of the most
                                   frequent probe keys' entries, looked up in LDS before the
                                   table): 0 (default) for unique-key single-column tables of
                                   >= 8 MiB probed by >= 4M rows; 1 off; 2 for every probe of a
                                   unique-key single-column table (tests) */
  TGP_TUNE_JIT = 11,            /* queries with a shape-specialised form (tgp_group_agg_explain
                                   spec_shape >= 0) but no kernel compiled ahead of time: 0
                                   (default) compile that form at run time (hipRTC, about a
                                   second per shape, cached for the process) for queries of
                                   >= 2^20 rows, run smaller ones on the generic engine; 1
                                   never compile; 2 always (tests) */
  TGP_TUNE_AGG_PART = 12        /* generic engine, groups too many for an LDS table (strategy
                                   3 or 5 of tgp_group_agg_explain): 0 (default) partition the
                                   rows by group range and aggregate each partition in LDS for
                                   >= 2^18 rows; 1 never (one global table); 2 at any row
                                   count (tests).  Never while the stream is being captured */
};
int32_t tgp_tune_set(tgp_ctx* ctx, int32_t key, int64_t value);
int32_t tgp_tune_get(tgp_ctx* ctx, int32_t key, int64_t* value);

/* ---- HBM tier (memory/tier.rs MemoryTier::Hbm: resident columns) ------------ */
int32_t tgp_hbm_alloc(tgp_ctx* ctx, uint64_t bytes, void** d_out);
int32_t tgp_hbm_free(tgp_ctx* ctx, void* d_ptr);
int32_t tgp_hbm_usage(tgp_ctx* ctx, uint64_t* used, uint64_t* budget, uint64_t* peak);
int32_t tgp_hbm_set_budget(tgp_ctx* ctx, uint64_t budget_bytes);
int32_t tgp_memcpy_h2d(tgp_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes, void* stream);
int32_t tgp_memcpy_d2h(tgp_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes, void* stream);
int32_t tgp_memcpy_d2d(tgp_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes, void* stream);
/* Fills with a kernel (graph-capture safe; no memset node). */
int32_t tgp_fill_u64(tgp_ctx* ctx, uint64_t* d_dst, uint64_t nwords, uint64_t value, void* stream);
/* Column min/max (as int64 if signed dtype, else as uint64 bits), one pass. */
int32_t tgp_column_minmax_sync(tgp_ctx* ctx, const tgp_column* col, int64_t* h_min, int64_t* h_max);

/* ---- loading: reference cells -> native-width device columns ------------------ */
/* The reference's u64 cells of one column (datasource/table.rs:25-30; i64 as
 * `v as u64`, csv.rs:144-161) become a resident column of type `dtype` at d_out
 * (n * width bytes, e.g. from tgp_hbm_alloc): uploaded in chunks and narrowed on
 * the device.  A cell outside the type's range (signed types: the cell read as
 * i64) is an error; *h_first_bad receives its row (UINT64_MAX when all fit). */
int32_t tgp_column_from_cells(tgp_ctx* ctx, const uint64_t* h_cells, uint64_t n, int32_t dtype,
                              void* d_out, uint64_t* h_first_bad, void* stream);
/* NullBitmap (types/null_bitmap.rs: one bool per row, true = NULL) as a device
 * bitmap of tgp_bitmap_words(n) words at d_bitmap: bit i set = row i is NULL. */
int32_t tgp_null_bitmap_from_flags(tgp_ctx* ctx, const uint8_t* h_flags, uint64_t n,
                                   uint64_t* d_bitmap, void* stream);

/* ---- predicate scan, selection bitmap, compaction ------------------------- */
/* Bitmap layout (bitmap.rs:36-45): bit i of the selection is bit (i & 63) of
 * 64-bit word i >> 6, i.e. LSB-first bytes on a little-endian host; the words
 * past the last row are written with zero tail bits (bitmap.rs:57-67).
 * Capacity of a bitmap for n rows: tgp_bitmap_words(n) words. */
uint64_t tgp_bitmap_words(uint64_t nbits);

/* AND of preds over rows [row_begin, row_end): writes the bitmap (bit i <->
 * row row_begin+i) when d_bitmap != NULL and adds the number of selected rows
 * to *d_count (device u64, caller zeroes it) when d_count != NULL.
 * Replaces filter_* + Bitmap::and_inplace chains (expr.rs:633-656) and the
 * kernel-table scans ScanEq/ScanRange/ScanMultiPredicate (kernel/scan.rs). */
int32_t tgp_scan_filter(tgp_ctx* ctx, const tgp_column* cols, int32_t ncols,
                        const tgp_pred* preds, int32_t npreds,
                        uint64_t row_begin, uint64_t row_end,
                        uint64_t* d_bitmap, uint64_t* d_count, void* stream);

/* Column-vs-column comparison narrowing a selection: bit i of d_bitmap (row
 * row_begin+i) stays set only where a[row] op b[row], cells compared as signed
 * 64-bit values when is_signed (INT / DATE columns) and as u64 otherwise (EQ
 * and NE agree either way); bits past the range are cleared.  d_bitmap holds
 * the selection to narrow (all ones: no selection yet).  Replaces the Col-op-Col
 * fast path of eval_comparison_vec (query_interpreter/expr.rs:1382-1451). */
int32_t tgp_scan_compare_columns(tgp_ctx* ctx, const tgp_column* a, int32_t op,
                                 const tgp_column* b, int32_t is_signed, uint64_t row_begin,
                                 uint64_t row_end, uint64_t* d_bitmap, void* stream);

/* f64 column (cells = f64 bits, as the reference stores FLOAT columns) vs an f64
 * literal -> a fresh bitmap (bit i <-> row row_begin+i; d_count, if non-null,
 * receives the number of set bits, added to its current value).  op TGP_EQ with
 * eq_epsilon = 0 is IEEE equality (filter_eq_f64's AVX-512 path, _CMP_EQ_OQ:
 * -0.0 == 0.0, NaN never equal); eq_epsilon = 1 is |cell - lit| <= 1e-6 *
 * max(|lit|, 1) (filter_eq_f64_epsilon, what `WHERE f = lit` runs); TGP_NE is
 * NOT of the epsilon equality (filter_ne_f64); TGP_LT/LE/GT/GE are ordered IEEE
 * compares (false for NaN).  Replaces exec/bitmap.rs:479-580 and :696-753. */
int32_t tgp_scan_filter_f64(tgp_ctx* ctx, const tgp_column* col, int32_t op, int32_t eq_epsilon,
                            double lit, uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                            uint64_t* d_count, void* stream);

/* ---- expression predicates (the interpreter's per-row eval fallback) ------------------
 * A WHERE term that is not a column-vs-literal comparison (arithmetic on columns,
 * AND / OR / NOT trees, BETWEEN, IN, CASE) is evaluated per row by the reference
 * (query_interpreter/expr.rs:1453-1463: eval of both sides, binop, truthy) with
 * Value2 semantics (expr.rs:1626-1800, :2037-2141; types.rs:351-368):
 *   - a column reads as Int (INT: the cell as int64), Float (FLOAT: the f64 bits)
 *     or Date (DATE: the low 32 bits as int32);
 *   - + - * of two Ints wrap in int64; otherwise (a Float or a Date operand) the
 *     f64 result; / is always f64 (NULL when the divisor is 0); % of two Ints is the
 *     truncated int64 remainder, else f64 fmod (NULL for 0);
 *   - = <> < <= > >= compare as f64 (so Ints above 2^53 compare rounded); NULL
 *     makes = < <= > >= false and <> true (<> is NOT of =);
 *   - AND / OR / NOT on truthiness (Int != 0, Float != 0; NULL and Date false);
 *     unary minus wraps Ints, negates Floats, gives NULL otherwise;
 *   - BETWEEN lo AND hi is lo <= v AND v <= hi; IN is any = ; CASE takes the first
 *     truthy WHEN (TGP_EX_IF nests them), NULL without ELSE.
 * The expression arrives as a postfix program of at most TGP_MAX_EXPR_OPS ops over
 * a stack at most TGP_MAX_EXPR_DEPTH deep (checked); bit i of d_bitmap (row
 * row_begin+i) stays set only where the program's final value is truthy, and bits
 * past the range are cleared. */
#define TGP_MAX_EXPR_OPS 64
#define TGP_MAX_EXPR_DEPTH 8
enum tgp_expr_code {
  TGP_EX_COL = 0,    /* push column `arg` read as imm's kind (TGP_EX_INT/FLOAT/DATE) */
  TGP_EX_INT = 1,    /* push Int(imm) */
  TGP_EX_FLOAT = 2,  /* push Float(f64 bits in imm) */
  TGP_EX_DATE = 3,   /* push Date(int32 imm) */
  TGP_EX_NULL = 4,   /* push NULL */
  TGP_EX_ADD = 5, TGP_EX_SUB = 6, TGP_EX_MUL = 7, TGP_EX_DIV = 8, TGP_EX_MOD = 9,
  TGP_EX_EQ = 10, TGP_EX_NE = 11, TGP_EX_LT = 12, TGP_EX_LE = 13, TGP_EX_GT = 14,
  TGP_EX_GE = 15,
  TGP_EX_AND = 16, TGP_EX_OR = 17, TGP_EX_NOT = 18, TGP_EX_NEG = 19,
  TGP_EX_BETWEEN = 20,      /* v lo hi -> lo <= v AND v <= hi */
  TGP_EX_NOT_BETWEEN = 21,
  TGP_EX_IN = 22,           /* v x1 .. xn (n = arg >= 1) -> any v = xk */
  TGP_EX_NOT_IN = 23,
  TGP_EX_IF = 24            /* cond then else -> then if cond is truthy, else else */
};
typedef struct tgp_expr_op {
  int32_t code; /* enum tgp_expr_code */
  int32_t arg;  /* column index (TGP_EX_COL), list length (TGP_EX_IN / NOT_IN) */
  int64_t imm;  /* literal bits, or the column kind of TGP_EX_COL */
} tgp_expr_op;
int32_t tgp_scan_expr(tgp_ctx* ctx, const tgp_column* cols, int32_t ncols,
                      const tgp_expr_op* prog, int32_t nops, uint64_t row_begin,
                      uint64_t row_end, uint64_t* d_bitmap, void* stream);

/* Bitmap algebra over nbits bits (bitmap.rs and/or/not/and_inplace):
 * out = a op b (out may alias a); NOT clears tail bits (bitmap.rs:121-136). */
int32_t tgp_bitmap_combine(tgp_ctx* ctx, int32_t op, const uint64_t* d_a, const uint64_t* d_b,
                           uint64_t* d_out, uint64_t nbits, void* stream);
int32_t tgp_bitmap_not(tgp_ctx* ctx, const uint64_t* d_a, uint64_t* d_out, uint64_t nbits,
                       void* stream);
/* count_ones_range(begin, end) (bitmap.rs:248-273), added to *d_count. */
int32_t tgp_bitmap_count_range(tgp_ctx* ctx, const uint64_t* d_a, uint64_t nbits,
                               uint64_t begin, uint64_t end, uint64_t* d_count, void* stream);

/* Compaction (SetBitIter, bitmap.rs:278-314): ascending indices of the set
 * bits, each plus `base`, as u64 (index_bytes 8, like Vec<usize>) or u32
 * (index_bytes 4).  d_out needs room for popcount entries; *h_count receives
 * the number written.  Synchronous. */
int32_t tgp_bitmap_to_indices_sync(tgp_ctx* ctx, const uint64_t* d_bitmap, uint64_t nbits,
                                   uint64_t base, int32_t index_bytes, void* d_out,
                                   uint64_t out_capacity, uint64_t* h_count, void* stream);

/* tgp_bitmap_to_indices_sync without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under graph capture of its
 * neighbours, tight serving loops, harnesses that forbid hipMalloc).  The block
 * offsets live in a fixed 512 KiB device array per GPU (set up at context
 * creation, held until the process ends), so nbits is at most 2^30
 * (TGP_ERR_INVALID beyond).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, threads or contexts) race and must be
 * serialised by the caller.  Same results as tgp_bitmap_to_indices_sync. */
int32_t tgp_bitmap_to_indices_noalloc_sync(tgp_ctx* ctx, const uint64_t* d_bitmap,
                                           uint64_t nbits, uint64_t base, int32_t index_bytes,
                                           void* d_out, uint64_t out_capacity, uint64_t* h_count,
                                           void* stream);

/* ---- fused filter + GROUP BY + exact aggregation ---------------------------- */
/* One pass over the referenced columns: preds AND-ed (plus the optional
 * bitmap), dense composite group key, COUNT(*) and up to TGP_MAX_AGGS sums,
 * all exact.  ngroups = prod(card) (1 when nparts == 0).  The result buffers
 * are overwritten.  Replaces execute_grouped -> try_accumulator_grouped /
 * HashMap grouping + sum_vec (aggregate.rs:507-697, :1485-1792;
 * accum_grouped.rs:55-409) and the scalar aggregates (aggregate.rs:1055-1289). */
int32_t tgp_group_agg(tgp_ctx* ctx, const tgp_agg_query* q, const tgp_group_result* out,
                      void* stream);
/* Plan of a query without running it (host only): the accumulation strategy
 * of the generic engine (0 registers/1 group, 1 registers/<=8 groups, 2 LDS
 * table, 3 global table, 4 global table of int128 words for ranges of 2^32 rows or
 * more, 5 rows partitioned by group range into LDS tables) and the index of the shape-specialised kernel compiled ahead of time that
 * would run it, or TGP_SPEC_SHAPE_RUNTIME when the query has a specialised form
 * that is compiled at run time (see TGP_TUNE_JIT), or -1 (none: the generic
 * engine runs it). */
#define TGP_SPEC_SHAPE_RUNTIME 1000
int32_t tgp_group_agg_explain(const tgp_agg_query* q, int32_t* strategy, int32_t* spec_shape);
/* Counters of the run-time compiled specialised kernels of this process: shapes
 * compiled, compiles that failed (those queries run on the generic engine),
 * launches, and the milliseconds spent compiling.  No counterpart in the
 * reference (its kernels are monomorphised by rustc ahead of time). */
int32_t tgp_jit_info(uint64_t* compiled, uint64_t* failed, uint64_t* launches, double* compile_ms);
/* Number of groups the query's key space spans (prod of card). */
int32_t tgp_group_agg_ngroups(const tgp_agg_query* q, uint64_t* out_ngroups);

/* ---- GROUP BY arbitrary integer key tuples (device hash table) ------------------ */
/* The groups are the distinct key tuples of the selected rows, for any key values
 * (sparse int64 ids, several columns, NULL keys): what the reference's accumulator
 * path keys by a combined hash in per-chunk FxHashMaps merged per key
 * (query_interpreter/accum_grouped.rs:198-335, aggregate.rs:582-627) and the
 * dispatch path groups in an FxHashMap (engine/dispatch.rs:1125-1140 one key,
 * :1377-1385 several).  Here groups are the true key tuples (no two tuples merge)
 * and a NULL key is a value of its own: all rows whose key k is NULL share one
 * group for that part.  Keys are compared as the reference's u64 cells (narrow
 * columns widened by their signedness).
 *
 * `agg` gives the columns, the filter (preds AND the optional selection bitmap),
 * the row range and the aggregates exactly as for tgp_group_agg; agg.nparts must
 * be 0.  key_col[k] indexes agg.cols; d_key_nulls[k] (optional) is a NULL bitmap
 * of key k: bit i set = the key of row row_begin + i is NULL. */
#define TGP_MAX_GROUP_KEYS 4
typedef struct tgp_hash_group_query {
  tgp_agg_query agg;
  int32_t nkeys;                          /* 1..TGP_MAX_GROUP_KEYS */
  int32_t key_col[TGP_MAX_GROUP_KEYS];
  int32_t _pad[3];
  const uint64_t* d_key_nulls[TGP_MAX_GROUP_KEYS];
} tgp_hash_group_query;
/* Output, one row per group for up to `capacity` groups (device buffers), in no
 * particular order:
 *   d_keys[g * nkeys + k]   key k of group g as its u64 cell (0 when NULL)
 *   d_key_nulls[g]          bit k set = key k of group g is NULL (may be NULL)
 *   d_count[g], d_sums[(g * naggs + a) * 2 + {0,1}]  as tgp_group_result */
typedef struct tgp_hash_group_result {
  uint64_t* d_keys;
  uint32_t* d_key_nulls;
  uint64_t* d_count;
  uint64_t* d_sums;
  uint64_t capacity;
} tgp_hash_group_result;
/* Synchronous: *h_ngroups = number of groups written.  More than `capacity`
 * groups: TGP_ERR_CAPACITY with *h_ngroups = the number of groups when the table
 * held them all, else a lower bound above capacity (call again with at least
 * that capacity).  Device memory for the table: about 16 * capacity * (1 + 2 *
 * naggs) bytes of the stream's scratch. */
int32_t tgp_hash_group_agg_sync(tgp_ctx* ctx, const tgp_hash_group_query* q,
                                const tgp_hash_group_result* out, uint64_t* h_ngroups,
                                void* stream);

/* ---- batched executor (many queries, one persistent launch) --------------- */
/* A batch holds up to max_queries tgp_agg_query descriptors.  PERSISTENT mode
 * runs every query whose shape has a specialised kernel in ONE persistent
 * kernel launch (captured once in a hipGraph and replayed); the others run one
 * launch each after it.  PER_QUERY mode runs tgp_group_agg for each query (the
 * one-launch-per-query baseline).  Results of query i live in the batch's
 * device arena at the offsets tgp_batch_result_layout gives, in the layout of
 * tgp_group_result (count words, then (lo, hi) sum words, then the
 * out-of-range count).  Not thread safe per batch. */
typedef struct tgp_batch tgp_batch;
enum tgp_batch_mode { TGP_BATCH_PERSISTENT = 0, TGP_BATCH_PER_QUERY = 1 };
int32_t tgp_batch_create(tgp_ctx* ctx, uint32_t max_queries, tgp_batch** out);
int32_t tgp_batch_destroy(tgp_batch* b);
/* Plans and uploads the queries (synchronous); invalidates the captured graph. */
int32_t tgp_batch_set_queries(tgp_batch* b, const tgp_agg_query* qs, uint32_t n, void* stream);
/* Launch settings of the persistent kernel.  wg_per_cu: workgroups per CU (0 =
 * as many as fit); chunk_slices: 256-row slices per dynamically scheduled unit
 * (0 = default 32); order: 0 = units in query order (the slices of query 0,
 * then of query 1, ...), 1 = units sorted by table position, so the queries
 * that cover a table region read it at about the same time, 2 = as 1 with
 * units cut at absolute table chunks and chunk k served by XCD k % 8, so those
 * queries also share one L2, 3 = shared scan: the queries that read the same
 * columns with the same aggregate forms (the Q1 layout) share one read of each
 * table chunk (chunk_slices <= 4096 per chunk): a workgroup loads each slice of
 * the chunk once and applies every overlapping query to it; the other queries
 * run as in 2; cached_loads: 1 = keep column lines in the caches (pays off with
 * orders 1 and 2), 0 = streaming loads;
 * use_graph: replay a captured hipGraph (0 = plain launches); items_per_grab:
 * work units a wave takes per counter atomic in orders 1 and 2 (0 = 1; at
 * most 64). */
typedef struct tgp_batch_tuning {
  uint32_t wg_per_cu;
  uint32_t chunk_slices;
  int32_t order;
  int32_t cached_loads;
  int32_t use_graph;
  uint32_t items_per_grab;
} tgp_batch_tuning;
int32_t tgp_batch_set_tuning(tgp_batch* b, const tgp_batch_tuning* t);
int32_t tgp_batch_get_tuning(tgp_batch* b, tgp_batch_tuning* t);
/* The batch's autotuner: times the persistent launch on the current queries
 * over schedules (order, cached loads), chunk sizes and workgroups per CU, keeps
 * the fastest settings and returns them. */
int32_t tgp_batch_autotune(tgp_batch* b, void* stream, tgp_batch_tuning* best, float* best_ms);
int32_t tgp_batch_run(tgp_batch* b, int32_t mode, void* stream);
/* tgp_batch_run, synchronous, with device timing: *h_total_ms for the whole
 * batch and, when h_query_ms != NULL, each query's completion time measured
 * from the start of the batch (PERSISTENT: all equal the total). */
int32_t tgp_batch_run_timed(tgp_batch* b, int32_t mode, void* stream, float* h_query_ms,
                            float* h_total_ms);
int32_t tgp_batch_info(tgp_batch* b, uint64_t* result_words, uint64_t* total_slices,
                       uint32_t* n_batched, uint32_t* n_fallback);
int32_t tgp_batch_result_layout(tgp_batch* b, uint32_t i, uint64_t* count_off, uint64_t* sums_off,
                                uint64_t* oor_off, uint64_t* ngroups);
/* Copies the first nwords words of the result arena to the host (synchronous). */
int32_t tgp_batch_results_sync(tgp_batch* b, uint64_t* h_words, uint64_t nwords, void* stream);

/* ---- SUM over f64 cells (FLOAT / DECIMAL columns) ------------------------------ */
/* SUM of the doubles whose bits are d_cells[i], i < n, selected by the optional
 * bitmap (bit i <-> cell i; NULL = all), into *h_sum: the exact sum rounded once
 * to the nearest double (ties to even), whatever the order of the cells.
 * Replaces the AggregateSumF64 kernels (kernel/aggregate.rs:19-147) and the
 * dispatch path's masked f64 sums (exec/vectorized.rs:128-136, :525-533;
 * engine/dispatch.rs:400-404), which add in sequence (or lane by lane) and round
 * at every step; where their partial sums are exact the two agree.  NaN, +Inf and
 * -Inf propagate as in IEEE addition.  A zero total is -0.0 only when
 * neg_zero_start and every selected cell is -0.0 (or none is selected), as
 * Iterator::sum over f64 starts from -0.0; the dispatch sums start from +0.0
 * (pass 0).  Synchronous. */
int32_t tgp_sum_f64_sync(tgp_ctx* ctx, const uint64_t* d_cells, uint64_t n,
                         const uint64_t* d_sel_bitmap, int32_t neg_zero_start, double* h_sum,
                         void* stream);

/* ---- COUNT(DISTINCT col) ------------------------------------------------------ */
/* Distinct u64 cells (values widened as the reference stores them, signed types
 * sign-extended) among rows [row_begin, row_end) of `col` selected by the
 * optional bitmap (bit i <-> row_begin + i, NULL = all).
 *   EXACT: the count of count_distinct_masked (exec/vectorized.rs:601-610);
 *   HLL:   HyperLogLog, 16384 registers over xxh3_64 of each cell's 8 LE bytes,
 *          the estimate of exec/hll.rs:53-75;
 *   AUTO:  as COUNT(DISTINCT col) in query_interpreter/aggregate.rs:1379-1400:
 *          exact below 100 000 selected rows, else HyperLogLog.
 * Synchronous: *h_count is set on return. */
enum tgp_distinct_mode { TGP_DISTINCT_AUTO = 0, TGP_DISTINCT_EXACT = 1, TGP_DISTINCT_HLL = 2 };
int32_t tgp_count_distinct(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin,
                           uint64_t row_end, const uint64_t* d_sel_bitmap, int32_t mode,
                           uint64_t* h_count, void* stream);
/* The 16384 HyperLogLog registers (HyperLogLog::add, exec/hll.rs:40-50) of the
 * same selection, into host memory; mergeable by a bytewise max (HyperLogLog::merge). */
int32_t tgp_hll_sketch(tgp_ctx* ctx, const tgp_column* col, uint64_t row_begin, uint64_t row_end,
                       const uint64_t* d_sel_bitmap, uint8_t* h_registers, void* stream);
/* HyperLogLog::estimate (exec/hll.rs:53-75) of 16384 host registers. */
uint64_t tgp_hll_estimate(const uint8_t* h_registers);

/* ---- SELECT rows: compaction, ORDER BY, LIMIT, gather -------------------------- */
/* The rows a SELECT returns (engine/dispatch.rs:486-733): the rows of
 * [row_begin, row_end) selected by the bitmap (bit i <-> row_begin + i; NULL =
 * all), in scan order, or stably sorted by the u64 cells of order_col
 * (ascending, or descending when `descending`; equal cells keep scan order),
 * the first `limit` of them, as absolute row indices into d_rows (capacity
 * entries).  *h_count = the number written.  Synchronous. */
int32_t tgp_select_rows(tgp_ctx* ctx, uint64_t row_begin, uint64_t row_end,
                        const uint64_t* d_sel_bitmap, const tgp_column* order_col,
                        int32_t descending, uint64_t limit, uint64_t* d_rows, uint64_t capacity,
                        uint64_t* h_count, void* stream);
/* d_out[i] = col[d_rows[i]] for i < n: as u64 cells (ResultColumn values,
 * signed types sign-extended) when as_cells, else at the column's width.  A row
 * index of TGP_GATHER_NO_ROW gives 0: the missing side of an outer join's row,
 * which hash_join fills with 0 cells (exec/join.rs:118-124, :141-152).  Any
 * other row outside the column is an error (TGP_ERR_INVALID).  Synchronous. */
#define TGP_GATHER_NO_ROW 0xFFFFFFFFFFFFFFFFull
int32_t tgp_gather(tgp_ctx* ctx, const tgp_column* col, const uint64_t* d_rows, uint64_t n,
                   int32_t as_cells, void* d_out, void* stream);

/* ---- result encoding of exact int128 values (host) ------------------------------ */
/* ResultColumn cells are u64 (engine/result.rs:37-53) and pgwire prints a cell as
 * the decimal text of the u64 (server/async_pgwire.rs:881-910): an int128 sum
 * travels instead as NUMERIC text (type OID 1700) in ResultColumn.string_values.
 * tgp_int128_to_numeric writes "[-]digits[.digits]" of (lo + 2^64 hi) * 10^-scale;
 * tgp_int128_avg_to_numeric writes the same value divided by `count`, with
 * frac_digits (>= scale) fraction digits, rounded half away from zero.  buf
 * receives the text and a terminating NUL; *out_len its length (also when buf
 * is too small, which returns TGP_ERR_INVALID). */
int32_t tgp_int128_to_numeric(uint64_t lo, uint64_t hi, int32_t scale, char* buf,
                              uint64_t buf_size, uint64_t* out_len);
int32_t tgp_int128_avg_to_numeric(uint64_t lo, uint64_t hi, uint64_t count, int32_t scale,
                                  int32_t frac_digits, char* buf, uint64_t buf_size,
                                  uint64_t* out_len);

/* ---- hash join --------------------------------------------------------------- */
typedef struct tgp_hash_table tgp_hash_table;

/* Build an open-addressing table (64-bit CAS on the key slot) over the
 * build rows: key column (int64/uint64/int32 widened) and an optional 32-bit
 * payload column stored inline (NULL => the build row index is the payload).
 * Duplicate keys keep one slot each (multiset, like JoinHashTable chains,
 * join_hash_table.rs:127-198).  Capacity = next_pow2(2*n) slots of 16 bytes,
 * accounted in the HBM tier.  Synchronous (returns once the table is built). */
int32_t tgp_hash_build(tgp_ctx* ctx, const tgp_column* keys, const tgp_column* payload,
                       tgp_hash_table** out, void* stream);
/* As tgp_hash_build, with the slots in caller memory (16-byte aligned, at least
 * tgp_hash_table_bytes_for(keys->len) bytes; e.g. a buffer a server reuses for
 * one join after another): no device allocation and no device-wide
 * synchronisation of a free when the table is destroyed, which leaves the
 * caller's memory alone.  Build keys equal to INT64_MIN (rare) still take a
 * small allocation of their own. */
int32_t tgp_hash_build_in(tgp_ctx* ctx, const tgp_column* keys, const tgp_column* payload,
                          void* d_mem, uint64_t mem_bytes, tgp_hash_table** out, void* stream);
int32_t tgp_hash_table_destroy(tgp_ctx* ctx, tgp_hash_table* t);
/* Bytes of HBM the table holds (admission control). */
int32_t tgp_hash_table_footprint(const tgp_hash_table* t, uint64_t* out_bytes);
/* Bytes a table over n build rows would take, before building it. */
uint64_t tgp_hash_table_bytes_for(uint64_t n_build_rows);
int32_t tgp_hash_table_info(const tgp_hash_table* t, uint64_t* n_rows, uint64_t* capacity,
                            int32_t* has_duplicates);

/* Total number of (probe row, build row) matches over rows [row_begin,row_end)
 * of the probe keys, added to *d_count (HashProbe, kernel/hash.rs:160-187). */
int32_t tgp_hash_probe_count(tgp_ctx* ctx, const tgp_hash_table* t, const tgp_column* probe_keys,
                             uint64_t row_begin, uint64_t row_end, uint64_t* d_count, void* stream);

/* Semi/anti join marks: bit i set iff probe row row_begin+i has >= 1 match
 * (exec/join.rs semi_join); anti = tgp_bitmap_not of it. */
int32_t tgp_hash_probe_mark(tgp_ctx* ctx, const tgp_hash_table* t, const tgp_column* probe_keys,
                            uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap, void* stream);

/* Inner-join pairs (probe row, build payload) in ascending probe-row order
 * (exec/join.rs hash_join INNER; query_interpreter/join.rs:643-794).  Within
 * one probe row, duplicate matches appear in ascending payload order (build-row
 * order when the table was built without a payload column).
 * Synchronous; *h_npairs receives the number of pairs (capacity exceeded =>
 * TGP_ERR_CAPACITY with *h_npairs set to the number needed). */
int32_t tgp_hash_probe_pairs_sync(tgp_ctx* ctx, const tgp_hash_table* t,
                                  const tgp_column* probe_keys, uint64_t row_begin,
                                  uint64_t row_end, uint64_t* d_probe_idx, uint32_t* d_build_payload,
                                  uint64_t capacity, uint64_t* h_npairs, void* stream);

/* ---- multi-column keys and LEFT joins (exec/join.rs:51-160) ----
 * A table over 2..TGP_MAX_JOIN_KEYS key columns is keyed by a 64-bit packing of
 * each row's cells, which is not collision-free; every candidate match is then
 * checked against all key columns, as hash_join does after probe_all
 * (exec/join.rs:51-66 pack_key, :102-111).  The table keeps the build key
 * column descriptors: those columns must stay allocated while it is used.
 * Cells are compared as 64-bit values (narrow columns widened by their
 * signedness), the reference's u64 cells.  Count and aggregate probes take
 * single-key tables only. */
#define TGP_MAX_JOIN_KEYS 4
#define TGP_JOIN_NO_MATCH 0xFFFFFFFFu /* LEFT join pair of a probe row without a match */
enum tgp_join_type {
  TGP_JOIN_INNER = 0, /* JoinType::Inner */
  TGP_JOIN_LEFT = 1   /* JoinType::Left: a probe row without a match gives one pair
                         (row, TGP_JOIN_NO_MATCH); the reference fills the build
                         columns of that row with 0 */
};
int32_t tgp_hash_build_multi(tgp_ctx* ctx, const tgp_column* keys, int32_t nkeys,
                             const tgp_column* payload, tgp_hash_table** out, void* stream);
/* Pairs of a join over nkeys probe key columns (nkeys = the table's key count)
 * in ascending probe-row order, matches of one probe row in ascending payload
 * order; join_type TGP_JOIN_INNER or TGP_JOIN_LEFT.  Otherwise as
 * tgp_hash_probe_pairs_sync (exec/join.rs:68-160 hash_join). */
int32_t tgp_hash_join_pairs_sync(tgp_ctx* ctx, const tgp_hash_table* t,
                                 const tgp_column* probe_keys, int32_t nkeys, int32_t join_type,
                                 uint64_t row_begin, uint64_t row_end, uint64_t* d_probe_idx,
                                 uint32_t* d_build_payload, uint64_t capacity, uint64_t* h_npairs,
                                 void* stream);
/* The non-equi conjuncts of a join's ON clause, applied to each match as the
 * interpreter's hash_join does (query_interpreter/join.rs:1146-1195 with
 * eval_non_equi_match :1233-1300): the npairs pairs of
 * tgp_hash_join_pairs_sync over probe rows [row_begin, row_end) (its build
 * payloads being build row indices) are kept when the program (tgp_scan_expr's
 * codes and semantics), whose column c < nleft is left_cols[c] at the probe
 * row and whose column c >= nleft is right_cols[c - nleft] at the build row, is
 * truthy.  Under TGP_JOIN_LEFT a probe row whose matches all fail keeps one pair
 * (row, TGP_JOIN_NO_MATCH) in its place, and (row, TGP_JOIN_NO_MATCH) pairs stay.
 * Order is kept.  *h_nout <= npairs pairs are written to d_out_probe_idx /
 * d_out_build_row (not the input arrays).  A pair naming a row outside
 * [row_begin, row_end) or beyond the right columns fails the call.
 * nleft + nright <= TGP_MAX_COLS.  Synchronous. */
int32_t tgp_join_residual_sync(tgp_ctx* ctx, const tgp_column* left_cols, int32_t nleft,
                               const tgp_column* right_cols, int32_t nright,
                               const tgp_expr_op* prog, int32_t nops, int32_t join_type,
                               uint64_t row_begin, uint64_t row_end,
                               const uint64_t* d_probe_idx, const uint32_t* d_build_row,
                               uint64_t npairs, uint64_t* d_out_probe_idx,
                               uint32_t* d_out_build_row, uint64_t* h_nout, void* stream);
/* Semi-join marks over nkeys probe key columns (exec/join.rs:162-199
 * semi_join / anti_join: anti = tgp_bitmap_not of the marks). */
int32_t tgp_hash_join_mark(tgp_ctx* ctx, const tgp_hash_table* t, const tgp_column* probe_keys,
                           int32_t nkeys, uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                           void* stream);

enum tgp_join_group_by {
  TGP_JOIN_GROUP_BUILD_PAYLOAD = 0, /* GROUP BY the inline build payload */
  TGP_JOIN_GROUP_PROBE_COL = 1,     /* GROUP BY a probe-side column */
  TGP_JOIN_GROUP_NONE = 2           /* one group */
};

/* Fused probe -> GROUP BY -> SUM(value), COUNT(*) over the inner-join rows,
 * without materialising them.  Group keys must be dense: key - key_base in
 * [0, ngroups) (others are counted in *d_n_out_of_range).  Output layout as
 * tgp_group_result with naggs = 1.  Replaces hash_join_with_keys +
 * try_accumulator_grouped (join.rs:544-939, accum_grouped.rs:198-335). */
typedef struct tgp_join_agg_query {
  const tgp_hash_table* table;
  tgp_column probe_keys;
  tgp_column value;       /* summed column (int64 / int32 ...) */
  tgp_column group_col;   /* used when group_by == TGP_JOIN_GROUP_PROBE_COL */
  int32_t group_by;
  uint32_t build_rows;    /* 0: every build row matches; n > 0: only build rows below n
                             (a join with the first n rows of a table built over more of
                             the same build side, e.g. one kept resident for many joins) */
  int64_t key_base;
  uint64_t ngroups;
  uint64_t row_begin, row_end;
} tgp_join_agg_query;
int32_t tgp_hash_probe_agg(tgp_ctx* ctx, const tgp_join_agg_query* q, const tgp_group_result* out,
                           void* stream);
/* Adds a blocked bloom filter to a single-key table (the reference's join pre-filter,
 * exec/bloom_filter.rs:62-330: "definitely absent" keys skip the table): every stored
 * key sets 4 bits of one 64-bit word, the filter has bits_per_key bits per build row
 * (rounded up to a power of two of words).  Every later probe of the table tests a
 * key's word first (one load the L2 holds) and probes the table only when all 4 bits
 * are set.  No false negatives: every result equals the unfiltered one.  Pays off
 * when most probe keys miss the build side. */
int32_t tgp_hash_table_add_bloom(tgp_ctx* ctx, tgp_hash_table* t, uint32_t bits_per_key,
                                 void* stream);
/* n probe aggregations in one launch, with the results of n tgp_hash_probe_agg
 * calls, bit for bit (the queries may probe different tables, row ranges and build
 * prefixes).  For the serving loop's join requests, which the reference runs one
 * at a time (route_and_execute, engine/mod.rs:522-545, each a hash_join_with_keys,
 * join.rs:544-939): work items of 16K probe rows of any query go to persistent
 * workgroups.  A query of more than 213 groups runs alone within the call.
 * Returns once the work is queued (it waits for the stream's earlier work and its
 * argument copies, not for the probes). */
int32_t tgp_hash_probe_agg_batch(tgp_ctx* ctx, const tgp_join_agg_query* qs, uint32_t n,
                                 const tgp_group_result* outs, void* stream);

/* ---- serving: continuous batching over resident columns -------------------- */
/* The reference serves one query at a time per pooled connection
 * (route_and_execute, engine/mod.rs:522-545; server/pool.rs).  A tgp_server
 * instead admits requests continuously: waiting aggregate requests form the next
 * persistent-kernel batch whenever one of `batch_slots` batches finishes (batches
 * overlap on the device, one stream each), and join requests are admitted when
 * their hash table fits the HBM budget (deferred while it does not, rejected
 * with TGP_ERR_BUDGET when it never can).  Results are exact and copied to host
 * memory per request; latency runs from the request's arrival time. */
typedef struct tgp_server tgp_server;
enum tgp_serve_kind { TGP_SERVE_AGG = 0, TGP_SERVE_JOIN_AGG = 1 };
typedef struct tgp_serve_join {
  tgp_column build_keys;       /* build side: all `len` rows are inserted */
  tgp_column build_payload;    /* data NULL => the build row index is the payload */
  tgp_join_agg_query probe;    /* `table` is filled in by the server */
} tgp_serve_join;
typedef struct tgp_serve_request {
  int32_t kind; /* tgp_serve_kind */
  int32_t _pad;
  tgp_agg_query agg;           /* kind TGP_SERVE_AGG */
  tgp_serve_join join;         /* kind TGP_SERVE_JOIN_AGG */
} tgp_serve_request;
typedef struct tgp_serve_config {
  uint32_t max_batch_queries;  /* requests per batch (a take of at most 8 runs as that
                                  many single-query launches, with the same results) */
  uint32_t batch_slots;        /* batches in flight (1..8) */
  uint32_t join_slots;         /* joins in flight (1..8) */
  uint32_t batch_wait_us;      /* while the other slots hold more than 16 requests, a free
                                  slot waits for max_batch_queries / 4 requests or until the
                                  oldest has waited this long (0: launch whatever is queued) */
  uint64_t heavy_rows;         /* ranges of >= this many rows run alone, one at a time (0 = off) */
  uint64_t hash_budget_bytes;  /* cap on hash-table HBM (0 = the tier's budget) */
  uint32_t wg_per_cu, chunk_slices; /* batch tuning (tgp_batch_tuning) */
  int32_t order, cached_loads;
  int32_t per_query_launch;    /* 1: no batching, one tgp_group_agg launch per aggregate
                                  request and one tgp_hash_probe_agg launch per join */
  int32_t share_build_tables;  /* 1: one resident hash table per build side (key and payload
                                  columns) serves every join on a prefix of it, probed with
                                  the request's prefix bound (build_rows); it grows (a new
                                  build, over the longest prefix a queued join asks for) when
                                  a request needs a longer prefix and is held in the HBM tier
                                  until the server ends.  The joins whose table is ready are
                                  probed together, one tgp_hash_probe_agg_batch launch per
                                  take.  0: every join builds its own table, admitted by its
                                  footprint */
} tgp_serve_config;
typedef struct tgp_serve_timing {
  int64_t arrival_ns, start_ns, done_ns; /* tgp_serve_now_ns() clock */
  int32_t status, deferred;              /* deferred: waited for HBM budget */
  uint64_t result_words;
} tgp_serve_timing;
typedef struct tgp_serve_stats {
  uint64_t submitted, completed, failed, rejected, deferred, admitted_joins;
  uint64_t batches, batched_requests, max_batch, heavy_in_flight;
  /* host time (ns, summed over calls): inside tgp_serve_submit, launching aggregate
   * work (batch planning and upload, or the single launches), and completing it
   * (result copies) */
  uint64_t submit_ns, launch_ns, harvest_ns;
  uint64_t table_builds;       /* hash tables built (with share_build_tables: per build side
                                  and growth, not per join) */
} tgp_serve_stats;
int64_t tgp_serve_now_ns(void);
int32_t tgp_serve_create(tgp_ctx* ctx, const tgp_serve_config* cfg, tgp_server** out);
/* Finishes the requests already submitted, then stops the workers. */
int32_t tgp_serve_destroy(tgp_server* s);
/* Queues one request; arrival_ns 0 = now.  Thread safe. */
int32_t tgp_serve_submit(tgp_server* s, const tgp_serve_request* req, int64_t arrival_ns,
                         uint64_t* out_id);
/* Waits until request `id` is finished (timeout_ns < 0: no limit). */
int32_t tgp_serve_wait(tgp_server* s, uint64_t id, int64_t timeout_ns, int32_t* out_done);
/* Result words of a finished request: count[ngroups] | sums[ngroups*naggs*2] | oor
 * (tgp_group_result layout, naggs = 1 for joins); returns the request's status. */
int32_t tgp_serve_result(tgp_server* s, uint64_t id, uint64_t* h_words, uint64_t nwords,
                         tgp_serve_timing* timing);
/* Forgets a finished request. */
int32_t tgp_serve_release(tgp_server* s, uint64_t id);
/* Open-loop replay: request i is submitted at start + arrival_offset_ns[i]
 * (non-decreasing) from a submitter thread, with that time as its arrival;
 * returns when all are finished. */
int32_t tgp_serve_replay(tgp_server* s, const tgp_serve_request* reqs,
                         const int64_t* arrival_offset_ns, uint32_t n, uint64_t* out_ids,
                         int64_t* out_start_ns);
int32_t tgp_serve_get_stats(tgp_server* s, tgp_serve_stats* out);

/* ---- synthetic data (counter-based, seeded) ------------------------------- */
/* Row i of column `column_id`: u = mix64(seed ^ (column_id << 40) ^ i)
 * (SplitMix64 finaliser), value = lo + mulhi64(u, hi - lo + 1).  Written at
 * the column's native width.  Deterministic for any row range. */
int32_t tgp_synth_uniform(tgp_ctx* ctx, tgp_column* col, uint64_t seed, uint64_t column_id,
                          int64_t lo, int64_t hi, uint64_t row_begin, uint64_t row_end,
                          void* stream);
/* out[i] = table[idx(i)], idx(i) = number of entries of the ascending table
 * cdf[0..n_cdf) that are <= unit(u_i), unit(u) = (u >> 11) * 2^-53, u_i the
 * counter hash above (cdf[n_cdf-1] must be 1.0).  Used for the Zipf
 * foreign-key column: cdf is the cumulative rank distribution and table maps
 * a rank to a build key. */
int32_t tgp_synth_lookup_cdf(tgp_ctx* ctx, tgp_column* out, uint64_t seed, uint64_t column_id,
                             const double* d_cdf, uint64_t n_cdf, const int64_t* d_table,
                             uint64_t row_begin, uint64_t row_end, void* stream);

/* ---- Kernel trait / KernelTable mirror (kernel/mod.rs:47-242) ------------- */
enum tgp_operator {           /* kernel/mod.rs:47-74, same order */
  TGP_OP_SCAN_EQ_U64 = 0, TGP_OP_SCAN_RANGE_U64 = 1, TGP_OP_SCAN_MULTI_PREDICATE = 2,
  TGP_OP_HASH_BUILD = 3, TGP_OP_HASH_PROBE = 4, TGP_OP_AGGREGATE_SUM_F64 = 5,
  TGP_OP_AGGREGATE_COUNT_DISTINCT = 6, TGP_OP_SIMILARITY_HAMMING = 7, TGP_OP_LEAPFROG_JOIN = 8
};
enum tgp_pred_op { TGP_PRED_EQ = 0, TGP_PRED_GT = 1, TGP_PRED_LT = 2 }; /* mod.rs:84-92 */
enum tgp_tier {               /* memory/tier.rs:11-29, same order */
  TGP_TIER_L1L2 = 0, TGP_TIER_L3 = 1, TGP_TIER_DDR5 = 2, TGP_TIER_HBM = 3, TGP_TIER_CXL = 4,
  TGP_TIER_NVME = 5, TGP_TIER_NVMEOF = 6, TGP_TIER_NETWORK = 7
};
#define TGP_TARGET_GFX950 100 /* device target beside CpuTarget (cpu.rs:30-41) */

/* repr(C) image of KernelParams (mod.rs:96-119). */
typedef struct tgp_kernel_params {
  uint64_t target_u64, low_u64, high_u64;
  uint32_t max_distance, _pad0;
  uint64_t cell_count;
  uint64_t target2_u64, target3_u64;
  int32_t pred1_op, pred2_op, pred3_op;
  uint8_t predicate_count, _pad1[3];
} tgp_kernel_params;
/* repr(C) image of KernelResult (mod.rs:141-148). */
typedef struct tgp_kernel_result {
  uint64_t count;
  double sum;
  uint64_t mask;
} tgp_kernel_result;

typedef struct tgp_kernel_table tgp_kernel_table;
typedef struct tgp_kernel tgp_kernel;

int32_t tgp_kernel_table_new(tgp_ctx* ctx, tgp_kernel_table** out);
int32_t tgp_kernel_table_destroy(tgp_kernel_table* t);
/* select(op, tier): exact (op, gfx950, tier) -> (op, gfx950, Hbm) -> any with
 * op (mod.rs:209-221).  *out = NULL when no kernel implements op. */
int32_t tgp_kernel_table_select(tgp_kernel_table* t, int32_t op, int32_t tier, tgp_kernel** out);
int32_t tgp_kernel_table_len(tgp_kernel_table* t, int32_t* out_n);
/* list(): i-th entry's operator, target, tier and name. */
int32_t tgp_kernel_table_entry(tgp_kernel_table* t, int32_t i, int32_t* op, int32_t* target,
                               int32_t* tier, const char** name);
int32_t tgp_kernel_info(const tgp_kernel* k, int32_t* op, int32_t* target, int32_t* tier,
                        const char** name);
/* Kernel::execute(input, output, &params) -> KernelResult (mod.rs:171-176).
 * `input`/`output` are host or device pointers with the reference's layouts:
 *   scans:      input = cell_count u64 cells; result.count (+ mask of first 64)
 *   HashBuild:  input = cell_count u64 keys; writes a tgp_hash_table* to output
 *   HashProbe:  input = [tgp_hash_table* (8 B)][cell_count u64 keys]; result.count
 *   AggregateSumF64: input = cell_count f64 bits; result.sum (the exact sum
 *               rounded once, tgp_sum_f64_sync with the -0.0 start of
 *               Iterator::sum) and result.count = cell_count (aggregate.rs:19-147)
 *   AggregateCountDistinct: input = cell_count u64 cells; result.count = the
 *               number of distinct cells (aggregate.rs:149-185)
 * Host inputs are staged through HBM (PCIe); tier HBM kernels expect device
 * pointers.  Errors return a status; `result` is then zeroed. */
int32_t tgp_kernel_execute(const tgp_kernel* k, const void* input, void* output,
                           const tgp_kernel_params* params, tgp_kernel_result* result);

/* ---- test hook -------------------------------------------------------------- */
/* Make the next `count` operator calls on ctx fail with TGP_ERR_INJECTED
 * before touching the device (failure-path tests on the host). */
int32_t tgp_debug_inject_failure(tgp_ctx* ctx, int32_t count);

#ifdef __cplusplus
}
#endif
#endif /* TURBOGP_GFX950_H */
