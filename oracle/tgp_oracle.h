/*
 * tgp_oracle.h -- CPU oracle for the turboGP scan / aggregate / join path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * engine's CPU algorithms (turboGP, Rust), used as the checker by tests/,
 * by __graft_entry__.smoke() and as the CPU baseline leg of bench.py.  The
 * product library never links, loads or calls it.
 *
 * Reference files restated (paths relative to the turboGP checkout):
 *   src/exec/bitmap.rs            Bitmap, filter_*, and/or/not, SetBitIter
 *   src/kernel/scan.rs            ScanEq / ScanRange / ScanMultiPredicate (scalar)
 *   src/engine/query_interpreter/expr.rs      build_mask AND chain, per-row eval, binop
 *   src/engine/query_interpreter/aggregate.rs execute_grouped, sum_vec
 *   src/engine/query_interpreter/accum_grouped.rs  combined key hash
 *   src/exec/join_hash_table.rs   JoinHashTable (CRC32 hash, tagged chains)
 *   src/engine/query_interpreter/join.rs      hash_join_with_keys
 *
 * The structs below have the same layout as the product's C ABI structs so
 * the same query description can be handed to both; the fields here hold
 * HOST pointers.
 */
#ifndef TGP_ORACLE_H
#define TGP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { const void* data; uint64_t len; int32_t dtype; int32_t _pad; } orc_column;
typedef struct { int32_t col, op, is_signed, _pad; int64_t lit; } orc_pred;
typedef struct { int32_t col, divisor; int64_t base, card; } orc_part;
typedef struct {
  int32_t kind, nfactors, fcol[3], _pad;
  int64_t coef, fa[3], fb[3];
} orc_agg;
typedef struct {
  uint64_t row_begin, row_end;
  orc_column cols[8];
  int32_t ncols, npreds, nparts, naggs;
  orc_pred preds[8];
  orc_part parts[2];
  orc_agg aggs[6];
  const uint64_t* sel_bitmap; /* host bitmap words, bit i <-> row_begin+i, or NULL */
} orc_query;

/* generator */
uint64_t orc_mix64(uint64_t z);
void orc_synth_uniform(void* out, int32_t dtype, uint64_t seed, uint64_t column_id, int64_t lo,
                       int64_t hi, uint64_t row_begin, uint64_t row_end);
void orc_synth_cdf(void* out, int32_t dtype, uint64_t seed, uint64_t column_id, const double* cdf,
                   uint64_t n_cdf, const int64_t* table, uint64_t row_begin, uint64_t row_end);

/* 512 when the AVX-512 filter path runs on this host, else 0; orc_set_simd(0)
 * forces the scalar path (tests compare both), orc_set_simd(512) restores it. */
int32_t orc_simd_level(void);
void orc_set_simd(int32_t level);

/* bitmap.rs */
void orc_widen_cells(const void* col, int32_t dtype, uint64_t n, uint64_t* cells);
void orc_filter_cells(const uint64_t* cells, uint64_t n, int32_t op, int32_t is_signed, int64_t lit,
                      uint8_t* bits);
void orc_bitmap_and_inplace(uint8_t* a, const uint8_t* b, uint64_t nbytes);
void orc_bitmap_or_inplace(uint8_t* a, const uint8_t* b, uint64_t nbytes);
void orc_bitmap_not(const uint8_t* a, uint8_t* out, uint64_t nbits);
uint64_t orc_bitmap_count_range(const uint8_t* bits, uint64_t nbits, uint64_t start, uint64_t end);
uint64_t orc_set_bits(const uint8_t* bits, uint64_t nbits, uint64_t* out);

/* kernel/scan.rs scalar kernels (count + first-64 mask) */
uint64_t orc_scan_multi(const uint64_t* cells, uint64_t n, int32_t npred, const int32_t* ops,
                        const uint64_t* targets, uint64_t* mask);

/* filter -> bitmap over a query's predicates (one pass per leaf + AND) */
int32_t orc_query_filter(const orc_query* q, uint8_t* bits /* (n+7)/8 bytes */);

/* execute_grouped.  Aggregate kinds: 0 SUM, 1 COUNT(col) (non-zero cells),
 * 2/3 MIN/MAX in signed order, 4/5 MIN/MAX in raw u64 order (the value in the
 * low word, its sign extension / 0 in the high word).
 * mode 0: exact int128 sums (out_sums lo,hi pairs);
 * mode 1: the reference's f64 sums (out_f64).  Both fill out_count.
 * ngroups = prod(card).  nthreads <= 0 => all cores. */
int32_t orc_group_agg(const orc_query* q, int32_t mode, int32_t nthreads, uint64_t* out_count,
                      uint64_t* out_sums, double* out_f64, uint64_t* out_oor);

/* GROUP BY the key tuple (accum_grouped.rs:198-335, dispatch.rs:1377-1385) with a
 * tuple-keyed map instead of the combined hash; NULL parts (key_nulls[k]: bitmap
 * words, bit i <-> row row_begin + i, or NULL) are values of their own.  Returns
 * the group count G; when G <= capacity writes the groups in ascending tuple order
 * (signed keys as int64, others as u64, NULL last): out_keys[g*nkeys+k] (u64 cells,
 * 0 for NULL), out_nulls[g] (bit k = key k NULL), out_count[g], out_sums as
 * orc_group_agg mode 0. */
int64_t orc_group_by_keys(const orc_query* q, int32_t nkeys, const int32_t* key_cols,
                          const uint64_t* const* key_nulls, uint64_t capacity, uint64_t* out_keys,
                          uint32_t* out_nulls, uint64_t* out_count, uint64_t* out_sums);

/* join_hash_table.rs + join.rs + accumulator grouping.
 * group_by: 0 build payload, 1 probe column, 2 none.
 * Values are summed exactly (int128 as lo,hi pairs).  Rows whose group key -
 * key_base is outside [0, ngroups) count in *out_oor. */
int32_t orc_join_agg(const int64_t* build_keys, const int32_t* build_payload, uint64_t n_build,
                     const int64_t* probe_keys, const int64_t* probe_vals,
                     const int32_t* probe_group, uint64_t row_begin, uint64_t row_end,
                     int32_t group_by, int64_t key_base, uint64_t ngroups, int32_t nthreads,
                     uint64_t* out_count, uint64_t* out_sums, uint64_t* out_oor);
/* probe_all pairs in probe order: returns the number of pairs; writes at most
 * capacity (probe row, build row) pairs, build rows of one probe row in
 * ascending order. */
uint64_t orc_join_pairs(const int64_t* build_keys, uint64_t n_build, const int64_t* probe_keys,
                        uint64_t n_probe, uint64_t* out_probe, uint32_t* out_build,
                        uint64_t capacity);

/* JoinHashTable::hash (CRC32 + multiply) for the distribution test. */
uint64_t orc_join_hash(uint64_t key);

/* COUNT(DISTINCT col): xxh3_64 of a u64 cell's 8 LE bytes (xxhash-rust 0.8,
 * exec/hll.rs:34), HyperLogLog p = 14 registers and estimate (exec/hll.rs),
 * and the count (mode 0: exact below 100 000 selected cells, else HLL; 1:
 * exact; 2: HLL).  sel: bitmap over the n elements or NULL. */
uint64_t orc_xxh3_u64(uint64_t v);
void orc_hll_registers(const void* col, int32_t dtype, uint64_t n, const uint64_t* sel,
                       uint8_t* regs);
uint64_t orc_hll_estimate(const uint8_t* regs);
uint64_t orc_count_distinct(const void* col, int32_t dtype, uint64_t n, const uint64_t* sel,
                            int32_t mode);

/* f64 filters (exec/bitmap.rs:479-580): op 0 EQ (eps: filter_eq_f64_epsilon),
 * 1 NE, 2 LT, 3 LE, 4 GT, 5 GE; cells are f64 bits; bits: (n+7)/8 bytes. */
void orc_filter_f64(const uint64_t* cells, uint64_t n, int32_t op, int32_t eps, double val,
                    uint8_t* bits);

/* Expression predicates (query_interpreter/expr.rs eval / binop): the postfix
 * program of tgp_scan_expr; clears the bits of rows whose value is not truthy. */
typedef struct { int32_t code, arg; int64_t imm; } orc_expr_op;
int32_t orc_scan_expr(const orc_column* cols, int32_t ncols, const orc_expr_op* prog,
                      int32_t nops, uint64_t row_begin, uint64_t row_end, uint8_t* bits);
int64_t orc_join_residual(const orc_column* cols, int32_t ncols, int32_t nleft,
                          const orc_expr_op* prog, int32_t nops, int32_t left,
                          const uint64_t* probe, const uint32_t* build, uint64_t npairs,
                          uint32_t no_row, uint64_t* out_probe, uint32_t* out_build);

#ifdef __cplusplus
}
#endif
#endif
