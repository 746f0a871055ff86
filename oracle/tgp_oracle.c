/*
 * tgp_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see tgp_oracle.h).
 *
 * Every function restates a reference algorithm and cites the file:line it
 * follows.  Two deliberate departures, both documented where they occur:
 *   - sums are also kept as exact int128 (mode 0); mode 1 reproduces the
 *     reference's f64 sums;
 *   - group keys are dense parts (v / divisor - base), with integer division;
 *     the reference's INT '/' yields a float (expr.rs:2037-2090).
 */
#include "tgp_oracle.h"

#include <math.h>
#include <immintrin.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#if defined(__SSE4_2__)
#include <nmmintrin.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ---- generator ---------------------------------------------------------------
 * Counter-based: row i of column c is mix64(seed ^ (c << 40) ^ i) mapped to
 * [lo, hi] by a 64x64->128 multiply-high.  mix64 is the SplitMix64 finaliser. */
uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static int dt_width(int32_t dt) {
  switch (dt) {
    case 0: case 1: return 1;
    case 2: case 3: return 2;
    case 4: case 5: return 4;
    default: return 8;
  }
}

static void put_native(void* out, int32_t dt, uint64_t i, int64_t v) {
  switch (dt_width(dt)) {
    case 1: ((uint8_t*)out)[i] = (uint8_t)v; break;
    case 2: ((uint16_t*)out)[i] = (uint16_t)v; break;
    case 4: ((uint32_t*)out)[i] = (uint32_t)v; break;
    default: ((uint64_t*)out)[i] = (uint64_t)v; break;
  }
}

/* Value of a native column as the reference's u64 cell holds it:
 * datasource/csv.rs:144-161 stores integers as `v as u64` (sign-extended). */
static int64_t get_native(const void* col, int32_t dt, uint64_t i) {
  switch (dt) {
    case 0: return ((const int8_t*)col)[i];
    case 1: return ((const uint8_t*)col)[i];
    case 2: return ((const int16_t*)col)[i];
    case 3: return ((const uint16_t*)col)[i];
    case 4: return ((const int32_t*)col)[i];
    case 5: return ((const uint32_t*)col)[i];
    default: return ((const int64_t*)col)[i];
  }
}

void orc_synth_uniform(void* out, int32_t dtype, uint64_t seed, uint64_t column_id, int64_t lo,
                       int64_t hi, uint64_t row_begin, uint64_t row_end) {
  const uint64_t range = (uint64_t)hi - (uint64_t)lo + 1;
  const uint64_t key = seed ^ (column_id << 40);
#pragma omp parallel for schedule(static)
  for (uint64_t i = row_begin; i < row_end; ++i) {
    const uint64_t u = orc_mix64(key ^ i);
    const int64_t v = range ? lo + (int64_t)(((u128)u * range) >> 64) : (int64_t)u;
    put_native(out, dtype, i - row_begin, v);
  }
}

void orc_synth_cdf(void* out, int32_t dtype, uint64_t seed, uint64_t column_id, const double* cdf,
                   uint64_t n_cdf, const int64_t* table, uint64_t row_begin, uint64_t row_end) {
  const uint64_t key = seed ^ (column_id << 40);
#pragma omp parallel for schedule(static)
  for (uint64_t i = row_begin; i < row_end; ++i) {
    const uint64_t u = orc_mix64(key ^ i);
    const double x = (double)(u >> 11) * 0x1.0p-53;
    uint64_t lo = 0, hi = n_cdf;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (cdf[mid] <= x) lo = mid + 1; else hi = mid;
    }
    if (lo >= n_cdf) lo = n_cdf - 1;
    put_native(out, dtype, i - row_begin, table[lo]);
  }
}

/* ---- bitmap.rs ------------------------------------------------------------------ */
void orc_widen_cells(const void* col, int32_t dtype, uint64_t n, uint64_t* cells) {
  for (uint64_t i = 0; i < n; ++i) cells[i] = (uint64_t)get_native(col, dtype, i);
}

/* filter_u64_avx512_inner (bitmap.rs:642-694): four 8-lane compares per step into
 * four mask bytes, then one 8-lane step at a time, then the scalar tail.  Chosen at
 * run time when the host has AVX-512F/BW, as the reference's filter_* pick their
 * AVX-512 path with is_x86_feature_detected (the library is built in one place
 * and run on another host, so the dispatch is at run time, not -march=native). */
#define ORC_FILTER_LOOP(CMP, PRED)                                                   \
  do {                                                                               \
    while (i + 32 <= n) {                                                            \
      const __m512i v0 = _mm512_loadu_si512((const void*)(cells + i));               \
      const __m512i v1 = _mm512_loadu_si512((const void*)(cells + i + 8));           \
      const __m512i v2 = _mm512_loadu_si512((const void*)(cells + i + 16));          \
      const __m512i v3 = _mm512_loadu_si512((const void*)(cells + i + 24));          \
      bits[b] = (uint8_t)CMP(v0, t, PRED);                                           \
      bits[b + 1] = (uint8_t)CMP(v1, t, PRED);                                       \
      bits[b + 2] = (uint8_t)CMP(v2, t, PRED);                                       \
      bits[b + 3] = (uint8_t)CMP(v3, t, PRED);                                       \
      i += 32;                                                                       \
      b += 4;                                                                        \
    }                                                                                \
    while (i + 8 <= n) {                                                             \
      bits[b++] = (uint8_t)CMP(_mm512_loadu_si512((const void*)(cells + i)), t, PRED); \
      i += 8;                                                                        \
    }                                                                                \
  } while (0)

__attribute__((target("avx512f,avx512bw"))) static uint64_t filter_cells_avx512(
    const uint64_t* cells, uint64_t n, int32_t op, int32_t is_signed, int64_t lit, uint8_t* bits) {
  const __m512i t = _mm512_set1_epi64(lit);
  uint64_t i = 0, b = 0;
  if (is_signed) {
    switch (op) {
      case 0: ORC_FILTER_LOOP(_mm512_cmp_epi64_mask, _MM_CMPINT_EQ); break;
      case 1: ORC_FILTER_LOOP(_mm512_cmp_epi64_mask, _MM_CMPINT_NE); break;
      case 2: ORC_FILTER_LOOP(_mm512_cmp_epi64_mask, _MM_CMPINT_LT); break;
      case 3: ORC_FILTER_LOOP(_mm512_cmp_epi64_mask, _MM_CMPINT_LE); break;
      case 4: ORC_FILTER_LOOP(_mm512_cmp_epi64_mask, _MM_CMPINT_NLE); break;
      default: ORC_FILTER_LOOP(_mm512_cmp_epi64_mask, _MM_CMPINT_NLT); break;
    }
  } else {
    switch (op) {
      case 0: ORC_FILTER_LOOP(_mm512_cmp_epu64_mask, _MM_CMPINT_EQ); break;
      case 1: ORC_FILTER_LOOP(_mm512_cmp_epu64_mask, _MM_CMPINT_NE); break;
      case 2: ORC_FILTER_LOOP(_mm512_cmp_epu64_mask, _MM_CMPINT_LT); break;
      case 3: ORC_FILTER_LOOP(_mm512_cmp_epu64_mask, _MM_CMPINT_LE); break;
      case 4: ORC_FILTER_LOOP(_mm512_cmp_epu64_mask, _MM_CMPINT_NLE); break;
      default: ORC_FILTER_LOOP(_mm512_cmp_epu64_mask, _MM_CMPINT_NLT); break;
    }
  }
  return i; /* rows done; the caller finishes the tail */
}

static int g_use_avx512 = -1; /* -1: not probed yet; orc_set_simd overrides */

int32_t orc_simd_level(void) {
  if (g_use_avx512 < 0) {
    __builtin_cpu_init();
    g_use_avx512 = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
  }
  return g_use_avx512 ? 512 : 0;
}

void orc_set_simd(int32_t level) {
  g_use_avx512 = -1;
  if (level < 512) g_use_avx512 = 0;
  else (void)orc_simd_level(); /* only where the host has it */
}

/* filter_{eq,ne,lt,le,gt,ge}_{u64,i64}, bitmap.rs:321-477 (scalar) and :642-694
 * (AVX-512): bit i = cmp(cells[i], lit), LSB-first bytes (bitmap.rs:36-45). */
void orc_filter_cells(const uint64_t* cells, uint64_t n, int32_t op, int32_t is_signed, int64_t lit,
                      uint8_t* bits) {
  memset(bits, 0, (n + 7) / 8);
  uint64_t start = 0;
  if (orc_simd_level() >= 512) start = filter_cells_avx512(cells, n, op, is_signed, lit, bits);
  for (uint64_t i = start; i < n; ++i) {
    int hit;
    if (is_signed) {
      const int64_t c = (int64_t)cells[i];
      switch (op) {
        case 0: hit = c == lit; break;
        case 1: hit = c != lit; break;
        case 2: hit = c < lit; break;
        case 3: hit = c <= lit; break;
        case 4: hit = c > lit; break;
        default: hit = c >= lit; break;
      }
    } else {
      const uint64_t c = cells[i], l = (uint64_t)lit;
      switch (op) {
        case 0: hit = c == l; break;
        case 1: hit = c != l; break;
        case 2: hit = c < l; break;
        case 3: hit = c <= l; break;
        case 4: hit = c > l; break;
        default: hit = c >= l; break;
      }
    }
    if (hit) bits[i >> 3] |= (uint8_t)(1u << (i & 7));
  }
}

/* Bitmap::and_inplace / or_inplace (bitmap.rs:162-190), scalar path. */
void orc_bitmap_and_inplace(uint8_t* a, const uint8_t* b, uint64_t nbytes) {
  for (uint64_t i = 0; i < nbytes; ++i) a[i] &= b[i];
}
void orc_bitmap_or_inplace(uint8_t* a, const uint8_t* b, uint64_t nbytes) {
  for (uint64_t i = 0; i < nbytes; ++i) a[i] |= b[i];
}

/* Bitmap::not (bitmap.rs:121-136): complement, unused high bits of the last
 * byte cleared. */
void orc_bitmap_not(const uint8_t* a, uint8_t* out, uint64_t nbits) {
  const uint64_t nb = (nbits + 7) / 8;
  for (uint64_t i = 0; i < nb; ++i) out[i] = (uint8_t)~a[i];
  const unsigned unused = (unsigned)((8 - (nbits & 7)) & 7);
  if (unused && nb) out[nb - 1] &= (uint8_t)((1u << (8 - unused)) - 1);
}

static int bit_get(const uint8_t* bits, uint64_t i) { return (bits[i >> 3] >> (i & 7)) & 1; }

/* Bitmap::count_ones_range (bitmap.rs:248-273): set bits in [start, end). */
uint64_t orc_bitmap_count_range(const uint8_t* bits, uint64_t nbits, uint64_t start, uint64_t end) {
  if (start >= end || start >= nbits) return 0;
  if (end > nbits) end = nbits;
  uint64_t c = 0;
  for (uint64_t i = start; i < end; ++i) c += bit_get(bits, i);
  return c;
}

/* SetBitIter::next (bitmap.rs:278-314): ascending indices of set bits,
 * skipping zero bytes. */
uint64_t orc_set_bits(const uint8_t* bits, uint64_t nbits, uint64_t* out) {
  uint64_t pos = 0, n = 0;
  while (pos < nbits) {
    const uint64_t byte_idx = pos >> 3;
    const uint8_t byte = bits[byte_idx];
    const uint8_t shifted = (uint8_t)(byte >> (pos & 7));
    if (byte == 0 || shifted == 0) {
      pos = (byte_idx + 1) << 3;
      continue;
    }
    const uint64_t r = pos + (uint64_t)__builtin_ctz(shifted);
    if (r >= nbits) break;
    out[n++] = r;
    pos = r + 1;
  }
  return n;
}

/* ScanMultiPredicateScalar (kernel/scan.rs:126-160): count of cells matching
 * all of the first min(npred,3) predicates {0 Eq, 1 Gt, 2 Lt} (unsigned);
 * *mask = hits of the first 64 cells (ScanEqScalar, scan.rs:44-58). */
uint64_t orc_scan_multi(const uint64_t* cells, uint64_t n, int32_t npred, const int32_t* ops,
                        const uint64_t* targets, uint64_t* mask) {
  if (npred > 3) npred = 3;
  uint64_t count = 0, m = 0;
  for (uint64_t i = 0; i < n; ++i) {
    int hit = 1;
    for (int p = 0; p < npred; ++p) {
      const uint64_t c = cells[i], t = targets[p];
      hit &= ops[p] == 0 ? c == t : ops[p] == 1 ? c > t : c < t;
    }
    count += (uint64_t)hit;
    if (i < 64) m |= (uint64_t)hit << i;
  }
  if (mask) *mask = m;
  return count;
}

/* ---- predicate mask (expr.rs:633-656 AND chain over filter_* leaves) -------------- */
int32_t orc_query_filter(const orc_query* q, uint8_t* bits) {
  const uint64_t n = q->row_end - q->row_begin;
  const uint64_t nb = (n + 7) / 8;
  if (q->npreds == 0) {
    memset(bits, 0xFF, nb);
    const unsigned unused = (unsigned)((8 - (n & 7)) & 7);
    if (unused && nb) bits[nb - 1] >>= unused; /* Bitmap::all_ones, bitmap.rs:57-67 */
  }
  uint64_t* cells = (uint64_t*)malloc((n ? n : 1) * 8);
  uint8_t* leaf = (uint8_t*)malloc(nb ? nb : 1);
  if (!cells || !leaf) { free(cells); free(leaf); return -3; }
  for (int p = 0; p < q->npreds; ++p) {
    const orc_column* c = &q->cols[q->preds[p].col];
    const int w = dt_width(c->dtype);
    /* u64 cells (the reference's column layout) are filtered where they lie;
     * narrower columns are widened to cells first */
    const uint64_t* src = (const uint64_t*)((const uint8_t*)c->data + q->row_begin * w);
    if (w != 8) {
      orc_widen_cells((const uint8_t*)c->data + q->row_begin * w, c->dtype, n, cells);
      src = cells;
    }
    orc_filter_cells(src, n, q->preds[p].op, q->preds[p].is_signed, q->preds[p].lit,
                     p == 0 ? bits : leaf);
    if (p > 0) orc_bitmap_and_inplace(bits, leaf, nb);
  }
  if (q->sel_bitmap) orc_bitmap_and_inplace(bits, (const uint8_t*)q->sel_bitmap, nb);
  free(cells);
  free(leaf);
  return 0;
}

/* ---- grouping (aggregate.rs:547-697: per-chunk FxHashMap<keyhash, Vec<usize>>,
 * merged in chunk order, then sum_vec per group) ---------------------------------- */
#define CHUNK 65536
#define KEY_MULT 0x517cc1b727220a95ull /* accum_grouped.rs:217, aggregate.rs:599 */

typedef struct {
  uint64_t hash;
  int64_t part[2];
  uint64_t* rows;
  uint64_t n, cap;
  int used;
} grp_t;

typedef struct {
  grp_t* slots;
  uint64_t cap, n;
} gmap_t;

static void gmap_init(gmap_t* m, uint64_t cap) {
  uint64_t c = 16;
  while (c < cap * 2) c <<= 1;
  m->slots = (grp_t*)calloc(c, sizeof(grp_t));
  m->cap = c;
  m->n = 0;
}

static grp_t* gmap_get(gmap_t* m, uint64_t hash, const int64_t* part);

static void gmap_grow(gmap_t* m) {
  gmap_t nm;
  gmap_init(&nm, m->cap);
  for (uint64_t i = 0; i < m->cap; ++i)
    if (m->slots[i].used) {
      grp_t* g = gmap_get(&nm, m->slots[i].hash, m->slots[i].part);
      g->rows = m->slots[i].rows;
      g->n = m->slots[i].n;
      g->cap = m->slots[i].cap;
    }
  free(m->slots);
  *m = nm;
}

static grp_t* gmap_get(gmap_t* m, uint64_t hash, const int64_t* part) {
  if ((m->n + 1) * 2 > m->cap) gmap_grow(m);
  uint64_t i = orc_mix64(hash) & (m->cap - 1);
  for (;;) {
    grp_t* g = &m->slots[i];
    if (!g->used) {
      g->used = 1;
      g->hash = hash;
      g->part[0] = part[0];
      g->part[1] = part[1];
      m->n++;
      return g;
    }
    if (g->hash == hash) return g; /* groups are identified by the combined hash (F6) */
    i = (i + 1) & (m->cap - 1);
  }
}

static void grp_push(grp_t* g, uint64_t row) {
  if (g->n == g->cap) {
    g->cap = g->cap ? g->cap * 2 : 64;
    g->rows = (uint64_t*)realloc(g->rows, g->cap * 8);
  }
  g->rows[g->n++] = row;
}

static void gmap_free(gmap_t* m, int free_rows) {
  if (free_rows)
    for (uint64_t i = 0; i < m->cap; ++i)
      if (m->slots[i].used) free(m->slots[i].rows);
  free(m->slots);
}

/* Per-row value of an aggregate, wrapping 64-bit (expr.rs binop :2066-2077). */
static int64_t agg_row_value(const orc_query* q, const orc_agg* a, uint64_t row) {
  if (a->kind == 1) {
    const orc_column* c = &q->cols[a->fcol[0]];
    return get_native(c->data, c->dtype, row) != 0;
  }
  uint64_t v = (uint64_t)a->coef;
  for (int k = 0; k < a->nfactors; ++k) {
    const orc_column* c = &q->cols[a->fcol[k]];
    const uint64_t x = (uint64_t)get_native(c->data, c->dtype, row);
    const uint64_t f = (uint64_t)a->fa[k] + (uint64_t)a->fb[k] * x;
    v = (k == 0 && a->coef == 1) ? f : v * f;
  }
  return (int64_t)v;
}

/* Is the aggregate a plain SUM(col) (sum_vec Col on INT: exact wrapping i64,
 * aggregate.rs:1498-1504)? */
static int agg_is_plain_col(const orc_agg* a) {
  return a->kind == 0 && a->nfactors == 1 && a->coef == 1 && a->fa[0] == 0 && a->fb[0] == 1;
}

int32_t orc_group_agg(const orc_query* q, int32_t mode, int32_t nthreads, uint64_t* out_count,
                      uint64_t* out_sums, double* out_f64, uint64_t* out_oor) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const uint64_t n = q->row_end - q->row_begin;
  uint64_t ngroups = 1, stride[2] = {1, 1};
  for (int k = 0; k < q->nparts; ++k) {
    stride[k] = ngroups;
    ngroups *= (uint64_t)q->parts[k].card;
  }
  memset(out_count, 0, ngroups * 8);
  if (out_sums) memset(out_sums, 0, ngroups * q->naggs * 16);
  if (out_f64) memset(out_f64, 0, ngroups * q->naggs * sizeof(double));
  uint64_t oor = 0;

  uint8_t* bits = (uint8_t*)malloc((n + 7) / 8 + 1);
  uint64_t* idx = (uint64_t*)malloc((n ? n : 1) * 8);
  if (!bits || !idx) { free(bits); free(idx); return -3; }
  orc_query_filter(q, bits);
  const uint64_t nsel = orc_set_bits(bits, n, idx); /* aggregate.rs:551 iter_set_bits */
  free(bits);
  for (uint64_t i = 0; i < nsel; ++i) idx[i] += q->row_begin;

  const uint64_t nchunks = (nsel + CHUNK - 1) / CHUNK;
  gmap_t* maps = (gmap_t*)calloc(nchunks ? nchunks : 1, sizeof(gmap_t));
  uint64_t* chunk_oor = (uint64_t*)calloc(nchunks ? nchunks : 1, 8);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
  for (uint64_t c = 0; c < nchunks; ++c) {
    gmap_init(&maps[c], 64);
    const uint64_t b = c * CHUNK, e = b + CHUNK < nsel ? b + CHUNK : nsel;
    for (uint64_t i = b; i < e; ++i) {
      const uint64_t row = idx[i];
      int64_t part[2] = {0, 0};
      uint64_t h = 0;
      int bad = 0;
      for (int k = 0; k < q->nparts; ++k) {
        const orc_part* P = &q->parts[k];
        const orc_column* col = &q->cols[P->col];
        int64_t v = get_native(col->data, col->dtype, row);
        if (P->divisor > 1) v = v / P->divisor;
        part[k] = v - P->base;
        if (part[k] < 0 || part[k] >= P->card) bad = 1;
        h = h * KEY_MULT + (uint64_t)v; /* combined key hash, accum_grouped.rs:213-218 */
      }
      if (bad) { chunk_oor[c]++; continue; }
      grp_push(gmap_get(&maps[c], h, part), row);
    }
  }
  /* merge in chunk order: each group's rows stay ascending (aggregate.rs:612-627) */
  gmap_t global;
  gmap_init(&global, 64);
  for (uint64_t c = 0; c < nchunks; ++c) {
    oor += chunk_oor[c];
    for (uint64_t i = 0; i < maps[c].cap; ++i) {
      grp_t* s = &maps[c].slots[i];
      if (!s->used) continue;
      grp_t* g = gmap_get(&global, s->hash, s->part);
      for (uint64_t r = 0; r < s->n; ++r) grp_push(g, s->rows[r]);
    }
    gmap_free(&maps[c], 1);
  }
  free(maps);
  free(chunk_oor);
  free(idx);

  /* per group, per aggregate (aggregate.rs:1291 eval_agg_expr -> :1485 sum_vec) */
  for (uint64_t i = 0; i < global.cap; ++i) {
    grp_t* g = &global.slots[i];
    if (!g->used) continue;
    uint64_t gid = 0;
    for (int k = 0; k < q->nparts; ++k) gid += (uint64_t)g->part[k] * stride[k];
    out_count[gid] = g->n; /* COUNT(*) = indices.len(), aggregate.rs:1372-1373 */
    for (int a = 0; a < q->naggs; ++a) {
      const orc_agg* A = &q->aggs[a];
      if (A->kind >= 2 && A->kind <= 5) {
        /* MIN/MAX: signed (minmax_vec on INT, aggregate.rs:1822-1836) or raw u64
         * (accum_grouped.rs:269-280: MinCol starts at u64::MAX, MaxCol at 0) */
        const int sgn = A->kind == 2 || A->kind == 3, is_min = A->kind == 2 || A->kind == 4;
        uint64_t best = 0;
        for (uint64_t r = 0; r < g->n; ++r) {
          const uint64_t v = (uint64_t)agg_row_value(q, A, g->rows[r]);
          int better;
          if (r == 0) better = 1;
          else if (sgn) better = is_min ? (int64_t)v < (int64_t)best : (int64_t)v > (int64_t)best;
          else better = is_min ? v < best : v > best;
          if (better) best = v;
        }
        if (mode == 0) {
          out_sums[(gid * q->naggs + a) * 2] = best;
          out_sums[(gid * q->naggs + a) * 2 + 1] = sgn && (int64_t)best < 0 ? ~0ull : 0ull;
        } else {
          out_f64[gid * q->naggs + a] = sgn ? (double)(int64_t)best : (double)best;
        }
      } else if (mode == 0) {
        i128 s = 0;
        for (uint64_t r = 0; r < g->n; ++r) s += (i128)agg_row_value(q, A, g->rows[r]);
        out_sums[(gid * q->naggs + a) * 2] = (uint64_t)(u128)s;
        out_sums[(gid * q->naggs + a) * 2 + 1] = (uint64_t)((u128)s >> 64);
      } else {
        if (A->kind == 1 || agg_is_plain_col(A)) {
          int64_t s = 0; /* exact wrapping i64 (aggregate.rs:1498-1504) */
          for (uint64_t r = 0; r < g->n; ++r)
            s = (int64_t)((uint64_t)s + (uint64_t)agg_row_value(q, A, g->rows[r]));
          out_f64[gid * q->naggs + a] = (double)s;
        } else {
          double s = 0.0; /* per-row value as f64, summed in row order (:1605-1683) */
          for (uint64_t r = 0; r < g->n; ++r) s += (double)agg_row_value(q, A, g->rows[r]);
          out_f64[gid * q->naggs + a] = s;
        }
      }
    }
  }
  gmap_free(&global, 1);
  if (out_oor) *out_oor = oor;
  return 0;
}

/* ---- GROUP BY the key tuple itself -------------------------------------------------
 * accum_grouped.rs:198-335 groups the selected rows per 65 536-row chunk in an
 * FxHashMap keyed by the combined hash h = h * 0x517cc1b727220a95 + cell of the key
 * cells (:213-218) and merges the chunk maps per key (:291-335); dispatch.rs:1377-1385
 * does the same for a multi-key GROUP BY.  Here a group is the key TUPLE (so two
 * tuples with one combined hash stay apart) and a NULL key part is a value of its
 * own.  One map over all selected rows in row order; COUNT(*), exact int128 sums and
 * MIN/MAX as orc_group_agg mode 0 computes them.  Groups are written in ascending
 * tuple order: key k compared as int64 when key_signed[k], else as u64, NULL after
 * every value. */
typedef struct {
  uint64_t key[4];
  uint32_t nulls;
  int used;
  uint64_t count;
  i128 sum[6];
  uint64_t best[6];
} tgrp_t;
typedef struct { tgrp_t* s; uint64_t cap, n; } tmap_t;

static uint64_t tuple_hash(const uint64_t* key, int nk, uint32_t nulls) {
  uint64_t h = 0x6a09e667f3bcc908ull ^ nulls;
  for (int k = 0; k < nk; ++k) h = orc_mix64(h ^ key[k]);
  return h;
}

static tgrp_t* tmap_get(tmap_t* m, const uint64_t* key, int nk, uint32_t nulls);
static void tmap_grow(tmap_t* m, int nk) {
  tmap_t nm = {(tgrp_t*)calloc(m->cap * 2, sizeof(tgrp_t)), m->cap * 2, 0};
  for (uint64_t i = 0; i < m->cap; ++i)
    if (m->s[i].used) {
      tgrp_t* g = tmap_get(&nm, m->s[i].key, nk, m->s[i].nulls);
      *g = m->s[i];
    }
  free(m->s);
  *m = nm;
}
static tgrp_t* tmap_get(tmap_t* m, const uint64_t* key, int nk, uint32_t nulls) {
  if ((m->n + 1) * 2 > m->cap) tmap_grow(m, nk);
  uint64_t i = tuple_hash(key, nk, nulls) & (m->cap - 1);
  for (;;) {
    tgrp_t* g = &m->s[i];
    if (!g->used) {
      memset(g, 0, sizeof(*g));
      g->used = 1;
      memcpy(g->key, key, sizeof(uint64_t) * (size_t)nk);
      g->nulls = nulls;
      m->n++;
      return g;
    }
    if (g->nulls == nulls && !memcmp(g->key, key, sizeof(uint64_t) * (size_t)nk)) return g;
    i = (i + 1) & (m->cap - 1);
  }
}

static int g_sort_nk;
static const int32_t* g_sort_signed;
static int cmp_tuple(const void* pa, const void* pb) {
  const tgrp_t* a = (const tgrp_t*)pa;
  const tgrp_t* b = (const tgrp_t*)pb;
  for (int k = 0; k < g_sort_nk; ++k) {
    const int na = (a->nulls >> k) & 1, nb = (b->nulls >> k) & 1;
    if (na != nb) return na - nb;
    if (na) continue;
    if (a->key[k] == b->key[k]) continue;
    if (g_sort_signed[k]) return (int64_t)a->key[k] < (int64_t)b->key[k] ? -1 : 1;
    return a->key[k] < b->key[k] ? -1 : 1;
  }
  return 0;
}

int64_t orc_group_by_keys(const orc_query* q, int32_t nkeys, const int32_t* key_cols,
                          const uint64_t* const* key_nulls, uint64_t capacity, uint64_t* out_keys,
                          uint32_t* out_nulls, uint64_t* out_count, uint64_t* out_sums) {
  if (nkeys < 1 || nkeys > 4) return -1;
  const uint64_t n = q->row_end - q->row_begin;
  uint8_t* bits = (uint8_t*)malloc((n + 7) / 8 + 1);
  uint64_t* idx = (uint64_t*)malloc((n ? n : 1) * 8);
  if (!bits || !idx) { free(bits); free(idx); return -3; }
  orc_query_filter(q, bits);
  const uint64_t nsel = orc_set_bits(bits, n, idx);
  free(bits);
  tmap_t m = {(tgrp_t*)calloc(64, sizeof(tgrp_t)), 64, 0};
  int32_t sgn[4] = {0, 0, 0, 0};
  for (int k = 0; k < nkeys; ++k) {
    const int32_t dt = q->cols[key_cols[k]].dtype;
    sgn[k] = dt == 0 || dt == 2 || dt == 4 || dt == 6;
  }
  for (uint64_t i = 0; i < nsel; ++i) {
    const uint64_t rel = idx[i], row = q->row_begin + rel;
    uint64_t key[4] = {0, 0, 0, 0};
    uint32_t nulls = 0;
    for (int k = 0; k < nkeys; ++k) {
      const uint64_t* nb = key_nulls ? key_nulls[k] : NULL;
      if (nb && ((nb[rel >> 6] >> (rel & 63)) & 1)) { nulls |= 1u << k; continue; }
      const orc_column* c = &q->cols[key_cols[k]];
      key[k] = (uint64_t)get_native(c->data, c->dtype, row);
    }
    tgrp_t* g = tmap_get(&m, key, nkeys, nulls);
    g->count++;
    for (int a = 0; a < q->naggs; ++a) {
      const orc_agg* A = &q->aggs[a];
      const uint64_t v = (uint64_t)agg_row_value(q, A, row);
      if (A->kind >= 2 && A->kind <= 5) {
        const int sg = A->kind == 2 || A->kind == 3, is_min = A->kind == 2 || A->kind == 4;
        int better;
        if (g->count == 1) better = 1;
        else if (sg) better = is_min ? (int64_t)v < (int64_t)g->best[a]
                                     : (int64_t)v > (int64_t)g->best[a];
        else better = is_min ? v < g->best[a] : v > g->best[a];
        if (better) g->best[a] = v;
      } else {
        g->sum[a] += (i128)(int64_t)v;
      }
    }
  }
  free(idx);
  tgrp_t* list = (tgrp_t*)malloc((m.n ? m.n : 1) * sizeof(tgrp_t));
  uint64_t G = 0;
  for (uint64_t i = 0; i < m.cap; ++i)
    if (m.s[i].used) list[G++] = m.s[i];
  free(m.s);
  g_sort_nk = nkeys;
  g_sort_signed = sgn;
  qsort(list, G, sizeof(tgrp_t), cmp_tuple);
  if (G <= capacity) {
    for (uint64_t g = 0; g < G; ++g) {
      for (int k = 0; k < nkeys; ++k) out_keys[g * nkeys + k] = list[g].key[k];
      if (out_nulls) out_nulls[g] = list[g].nulls;
      out_count[g] = list[g].count;
      for (int a = 0; a < q->naggs; ++a) {
        const orc_agg* A = &q->aggs[a];
        uint64_t lo, hi;
        if (A->kind >= 2 && A->kind <= 5) {
          lo = list[g].best[a];
          hi = (A->kind == 2 || A->kind == 3) && (int64_t)lo < 0 ? ~0ull : 0ull;
        } else {
          lo = (uint64_t)(u128)list[g].sum[a];
          hi = (uint64_t)((u128)list[g].sum[a] >> 64);
        }
        out_sums[(g * q->naggs + a) * 2] = lo;
        out_sums[(g * q->naggs + a) * 2 + 1] = hi;
      }
    }
  }
  free(list);
  return (int64_t)G;
}

/* ---- JoinHashTable (exec/join_hash_table.rs) ------------------------------------ */
static uint32_t crc32c_u64(uint64_t v) {
#if defined(__SSE4_2__)
  return (uint32_t)_mm_crc32_u64(0, v);
#else
  /* CRC-32C (Castagnoli, reflected polynomial 0x82F63B78), bitwise */
  uint32_t crc = 0;
  for (int i = 0; i < 8; ++i) {
    crc ^= (uint32_t)(v & 0xff);
    v >>= 8;
    for (int b = 0; b < 8; ++b) crc = (crc >> 1) ^ (0x82F63B78u & (0u - (crc & 1)));
  }
  return crc;
#endif
}

/* JoinHashTable::hash (join_hash_table.rs:89-104) */
uint64_t orc_join_hash(uint64_t key) { return (uint64_t)crc32c_u64(key) * 0x8648DBDB00000001ull; }

typedef struct { uint64_t key; uint32_t row_idx, next; } jentry_t;
typedef struct {
  uint64_t* dir;
  jentry_t* entries;
  uint64_t len, dir_size;
  unsigned shift;
} jht_t;

static uint16_t bloom_tag(uint64_t h) { return (uint16_t)((h >> 16) & 0xffff) | 1; } /* :112-114 */

/* JoinHashTable::new (:76-85): directory next_pow2(2n) (min 16). */
static void jht_new(jht_t* t, uint64_t n) {
  uint64_t d = 16;
  while (d < n * 2) d <<= 1;
  t->dir_size = d;
  t->dir = (uint64_t*)calloc(d, 8);
  t->entries = (jentry_t*)calloc(n + 1, sizeof(jentry_t));
  t->len = 1; /* entry 0 is the null sentinel */
  t->shift = 64 - (unsigned)__builtin_ctzll(d);
}

/* JoinHashTable::insert (:127-151): prepend to the slot's chain, OR the tag. */
static void jht_insert(jht_t* t, uint64_t key, uint32_t row) {
  const uint64_t h = orc_join_hash(key);
  const uint64_t slot = h >> t->shift;
  const uint32_t e = (uint32_t)t->len++;
  const uint64_t old = t->dir[slot];
  t->entries[e].key = key;
  t->entries[e].row_idx = row;
  t->entries[e].next = (uint32_t)(old & 0xffffffffu);
  const uint64_t tag = ((old >> 48) & 0xffff) | bloom_tag(h);
  t->dir[slot] = (tag << 48) | e;
}

/* JoinHashTable::probe_all (:180-198): rows of every chain entry with the key,
 * in chain order (most recently inserted first). */
static uint64_t jht_probe_all(const jht_t* t, uint64_t key, uint32_t* out, uint64_t cap) {
  const uint64_t h = orc_join_hash(key);
  const uint64_t entry = t->dir[h >> t->shift];
  const uint16_t tag = (uint16_t)(entry >> 48), exp = bloom_tag(h);
  if (tag == 0 || (tag & exp) != exp) return 0; /* could_contain, :116-124 */
  uint64_t n = 0;
  for (uint32_t i = (uint32_t)(entry & 0xffffffffu); i; i = t->entries[i].next)
    if (t->entries[i].key == key) {
      if (n < cap) out[n] = t->entries[i].row_idx;
      n++;
    }
  return n;
}

static void jht_free(jht_t* t) {
  free(t->dir);
  free(t->entries);
}

static int cmp_u32(const void* a, const void* b) {
  const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

uint64_t orc_join_pairs(const int64_t* build_keys, uint64_t n_build, const int64_t* probe_keys,
                        uint64_t n_probe, uint64_t* out_probe, uint32_t* out_build,
                        uint64_t capacity) {
  jht_t t;
  jht_new(&t, n_build);
  for (uint64_t r = 0; r < n_build; ++r) jht_insert(&t, (uint64_t)build_keys[r], (uint32_t)r);
  uint64_t np = 0;
  uint32_t buf[4096];
  for (uint64_t i = 0; i < n_probe; ++i) {
    uint64_t m = jht_probe_all(&t, (uint64_t)probe_keys[i], buf, 4096);
    if (m > 4096) m = 4096;
    qsort(buf, m, 4, cmp_u32); /* canonical order: ascending build row */
    for (uint64_t k = 0; k < m; ++k, ++np)
      if (np < capacity) { out_probe[np] = i; out_build[np] = buf[k]; }
  }
  jht_free(&t);
  return np;
}

/* hash_join_with_keys (join.rs:544-939): serial build on the build side,
 * chunked parallel probe (65 536 rows) emitting matches, materialisation of
 * the joined (group key, value) rows, then the accumulator grouping
 * (accum_grouped.rs:198-335: per-chunk maps keyed by the key, merged). */
typedef struct { uint64_t key, count; i128 sum; int used; } acc_t;
typedef struct { acc_t* s; uint64_t cap, n; } amap_t;

static void amap_init(amap_t* m, uint64_t cap) {
  uint64_t c = 16;
  while (c < cap * 2) c <<= 1;
  m->s = (acc_t*)calloc(c, sizeof(acc_t));
  m->cap = c;
  m->n = 0;
}
static acc_t* amap_get(amap_t* m, uint64_t key);
static void amap_grow(amap_t* m) {
  amap_t nm;
  amap_init(&nm, m->cap);
  for (uint64_t i = 0; i < m->cap; ++i)
    if (m->s[i].used) {
      acc_t* a = amap_get(&nm, m->s[i].key);
      a->count = m->s[i].count;
      a->sum = m->s[i].sum;
    }
  free(m->s);
  *m = nm;
}
static acc_t* amap_get(amap_t* m, uint64_t key) {
  if ((m->n + 1) * 2 > m->cap) amap_grow(m);
  uint64_t i = orc_mix64(key) & (m->cap - 1);
  for (;;) {
    acc_t* a = &m->s[i];
    if (!a->used) { a->used = 1; a->key = key; m->n++; return a; }
    if (a->key == key) return a;
    i = (i + 1) & (m->cap - 1);
  }
}

int32_t orc_join_agg(const int64_t* build_keys, const int32_t* build_payload, uint64_t n_build,
                     const int64_t* probe_keys, const int64_t* probe_vals,
                     const int32_t* probe_group, uint64_t row_begin, uint64_t row_end,
                     int32_t group_by, int64_t key_base, uint64_t ngroups, int32_t nthreads,
                     uint64_t* out_count, uint64_t* out_sums, uint64_t* out_oor) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  jht_t t;
  jht_new(&t, n_build);
  for (uint64_t r = 0; r < n_build; ++r) jht_insert(&t, (uint64_t)build_keys[r], (uint32_t)r);
  const uint64_t n = row_end - row_begin;
  const uint64_t nchunks = (n + CHUNK - 1) / CHUNK;
  /* probe: per chunk, the matched (group key, value) rows (materialised) */
  int64_t** jk = (int64_t**)calloc(nchunks ? nchunks : 1, sizeof(int64_t*));
  int64_t** jv = (int64_t**)calloc(nchunks ? nchunks : 1, sizeof(int64_t*));
  uint64_t* jn = (uint64_t*)calloc(nchunks ? nchunks : 1, 8);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
  for (uint64_t c = 0; c < nchunks; ++c) {
    const uint64_t b = row_begin + c * CHUNK, e = b + CHUNK < row_end ? b + CHUNK : row_end;
    uint64_t cap = (e - b) + 16, cnt = 0;
    int64_t* k = (int64_t*)malloc(cap * 8);
    int64_t* v = (int64_t*)malloc(cap * 8);
    uint32_t buf[4096];
    for (uint64_t i = b; i < e; ++i) {
      uint64_t m = jht_probe_all(&t, (uint64_t)probe_keys[i], buf, 4096);
      if (m > 4096) m = 4096;
      for (uint64_t x = 0; x < m; ++x) {
        if (cnt == cap) {
          cap *= 2;
          k = (int64_t*)realloc(k, cap * 8);
          v = (int64_t*)realloc(v, cap * 8);
        }
        k[cnt] = group_by == 0 ? (int64_t)build_payload[buf[x]]
                 : group_by == 1 ? (int64_t)probe_group[i] : key_base;
        v[cnt] = probe_vals[i];
        cnt++;
      }
    }
    jk[c] = k;
    jv[c] = v;
    jn[c] = cnt;
  }
  jht_free(&t);
  /* accumulator grouping over the joined rows */
  amap_t* maps = (amap_t*)calloc(nchunks ? nchunks : 1, sizeof(amap_t));
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
  for (uint64_t c = 0; c < nchunks; ++c) {
    amap_init(&maps[c], 64);
    for (uint64_t i = 0; i < jn[c]; ++i) {
      acc_t* a = amap_get(&maps[c], (uint64_t)jk[c][i]);
      a->count++;
      a->sum += (i128)jv[c][i];
    }
    free(jk[c]);
    free(jv[c]);
  }
  free(jk);
  free(jv);
  free(jn);
  memset(out_count, 0, ngroups * 8);
  memset(out_sums, 0, ngroups * 16);
  uint64_t oor = 0;
  for (uint64_t c = 0; c < nchunks; ++c) {
    for (uint64_t i = 0; i < maps[c].cap; ++i) {
      const acc_t* a = &maps[c].s[i];
      if (!a->used) continue;
      const int64_t g = (int64_t)a->key - key_base;
      if (g < 0 || (uint64_t)g >= ngroups) { oor += a->count; continue; }
      out_count[g] += a->count;
      u128 cur = ((u128)out_sums[g * 2 + 1] << 64) | out_sums[g * 2];
      cur += (u128)a->sum;
      out_sums[g * 2] = (uint64_t)cur;
      out_sums[g * 2 + 1] = (uint64_t)(cur >> 64);
    }
    free(maps[c].s);
  }
  free(maps);
  if (out_oor) *out_oor = oor;
  return 0;
}

/* ---- COUNT(DISTINCT col) ----------------------------------------------------------- */
/* XXH3_64bits of the 8 little-endian bytes of v, seed 0, default secret: the
 * published XXH3 algorithm's 4..8-byte path (xxHash 0.8, which the reference's
 * xxhash-rust 0.8 implements; exec/hll.rs:34 hashes value.to_le_bytes()).  The
 * bit-flip constant is readLE64(secret + 8) ^ readLE64(secret + 16) of the
 * default 192-byte secret; the input64 of that path is the value with its
 * 32-bit halves swapped. */
uint64_t orc_xxh3_u64(uint64_t v) {
  const uint64_t kMx2 = 0x9FB21C651E98DF25ull;
  const uint64_t bitflip = 0x1cad21f72c81017cull ^ 0xdb979083e96dd4deull;
  uint64_t h = ((v << 32) | (v >> 32)) ^ bitflip;
  h ^= ((h << 49) | (h >> 15)) ^ ((h << 24) | (h >> 40));
  h *= kMx2;
  h ^= (h >> 35) + 8;
  h *= kMx2;
  return h ^ (h >> 28);
}

/* HyperLogLog::add_hashed (exec/hll.rs:40-50): register = top 14 bits, rank =
 * leading zeros of the remaining 50 bits (shifted to the top) + 1. */
static void hll_add(uint8_t* regs, uint64_t h) {
  const uint64_t idx = h >> 50;
  const uint64_t w = h << 14;
  const uint8_t lz = (uint8_t)((w ? __builtin_clzll(w) : 64) + 1);
  if (lz > regs[idx]) regs[idx] = lz;
}

/* Registers (16384 bytes) of the selected cells of col[0, n): `sel` is a bitmap
 * (bit i <-> element i) or NULL for all. */
void orc_hll_registers(const void* col, int32_t dtype, uint64_t n, const uint64_t* sel,
                       uint8_t* regs) {
  memset(regs, 0, 1u << 14);
  for (uint64_t i = 0; i < n; ++i) {
    if (sel && !((sel[i >> 6] >> (i & 63)) & 1)) continue;
    hll_add(regs, orc_xxh3_u64((uint64_t)get_native(col, dtype, i)));
  }
}

/* HyperLogLog::estimate (exec/hll.rs:53-75): raw = alpha m^2 / sum 2^-r, the sum
 * taken in register order; linear counting when raw <= 2.5 m and a register is 0;
 * truncated to an integer. */
uint64_t orc_hll_estimate(const uint8_t* regs) {
  const double m = 16384.0;
  const double alpha = 0.7213 / (1.0 + 1.079 / m);
  double sum = 0.0;
  uint64_t zeros = 0;
  for (int i = 0; i < 16384; ++i) {
    sum += pow(2.0, -(double)regs[i]);
    if (regs[i] == 0) zeros++;
  }
  const double raw = alpha * m * m / sum;
  const double est = (raw <= 2.5 * m && zeros > 0) ? m * log(m / (double)zeros) : raw;
  return (uint64_t)est;
}

static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

/* COUNT(DISTINCT col) of the selected cells.  mode 1: exact (the HashSet of
 * count_distinct_masked, vectorized.rs:601-610); 2: HyperLogLog; 0: as the
 * interpreter chooses (aggregate.rs:1379-1400, hll.rs:97-128): exact below
 * 100 000 selected cells, else HyperLogLog. */
uint64_t orc_count_distinct(const void* col, int32_t dtype, uint64_t n, const uint64_t* sel,
                            int32_t mode) {
  uint64_t nsel = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (!sel || ((sel[i >> 6] >> (i & 63)) & 1)) nsel++;
  if (mode == 2 || (mode == 0 && nsel >= 100000)) {
    uint8_t* regs = (uint8_t*)malloc(1u << 14);
    orc_hll_registers(col, dtype, n, sel, regs);
    const uint64_t e = orc_hll_estimate(regs);
    free(regs);
    return e;
  }
  uint64_t* v = (uint64_t*)malloc((nsel ? nsel : 1) * 8);
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (!sel || ((sel[i >> 6] >> (i & 63)) & 1)) v[k++] = (uint64_t)get_native(col, dtype, i);
  qsort(v, k, 8, cmp_u64);
  uint64_t d = 0;
  for (uint64_t i = 0; i < k; ++i)
    if (i == 0 || v[i] != v[i - 1]) d++;
  free(v);
  return d;
}

/* ---- f64 filters (exec/bitmap.rs:479-580) -------------------------------------------
 * op 0 EQ (eps = 0: `==` as the AVX-512 path's _CMP_EQ_OQ, :482-486; eps = 1:
 * filter_eq_f64_epsilon, :505-515), 1 NE (NOT of the epsilon equality, :498-500),
 * 2 LT, 3 LE, 4 GT, 5 GE (:523-580).  Cells are f64 bits; bits[i] for row i. */
void orc_filter_f64(const uint64_t* cells, uint64_t n, int32_t op, int32_t eps, double val,
                    uint8_t* bits) {
  const double tol = 1e-6 * (fabs(val) > 1.0 ? fabs(val) : 1.0);
  memset(bits, 0, (size_t)((n + 7) / 8));
  for (uint64_t i = 0; i < n; ++i) {
    double cv;
    memcpy(&cv, &cells[i], 8);
    int set;
    switch (op) {
      case 0: set = eps ? fabs(cv - val) <= tol : cv == val; break;
      case 1: set = !(fabs(cv - val) <= tol); break;
      case 2: set = cv < val; break;
      case 3: set = cv <= val; break;
      case 4: set = cv > val; break;
      default: set = cv >= val; break;
    }
    if (set) bits[i >> 3] |= (uint8_t)(1u << (i & 7));
  }
}

/* ---- expression predicates: the interpreter's per-row eval -------------------------
 * Value2 (types.rs:8-15 / :351-368) and Interpreter::eval / binop / cmp_* / truthy
 * (query_interpreter/expr.rs:1615-1622, :1626-1800, :2037-2141) over the postfix
 * program of tgp_scan_expr (codes as in tgp_expr_code). */
enum { V_NULL = 0, V_INT = 1, V_FLOAT = 2, V_DATE = 3 };
typedef struct { int tag; int64_t i; double f; int32_t d; } orc_value;

static int v_as_f64(const orc_value* v, double* out) { /* Value2::as_f64 */
  switch (v->tag) {
    case V_INT: *out = (double)v->i; return 1;
    case V_FLOAT: *out = v->f; return 1;
    case V_DATE: *out = (double)v->d; return 1;
    default: return 0;
  }
}
static int v_truthy(const orc_value* v) { /* expr.rs:1615-1622 */
  if (v->tag == V_INT) return v->i != 0;
  if (v->tag == V_FLOAT) return v->f != 0.0;
  return 0;
}
static int v_cmp_eq(const orc_value* a, const orc_value* b) { /* :2106-2124 (no strings) */
  double x, y;
  if (a->tag == V_NULL || b->tag == V_NULL) return 0;
  if (!v_as_f64(a, &x) || !v_as_f64(b, &y)) return 0;
  return x == y;
}
static int v_cmp_lt(const orc_value* a, const orc_value* b) { /* :2126-2138 */
  double x, y;
  if (a->tag == V_NULL || b->tag == V_NULL) return 0;
  if (!v_as_f64(a, &x) || !v_as_f64(b, &y)) return 0;
  return x < y;
}
static int v_cmp_le(const orc_value* a, const orc_value* b) { /* :2139-2141 */
  return v_cmp_lt(a, b) || v_cmp_eq(a, b);
}
static orc_value v_int(int64_t i) { orc_value v = {V_INT, i, 0.0, 0}; return v; }
static orc_value v_float(double f) { orc_value v = {V_FLOAT, 0, f, 0}; return v; }
static orc_value v_null(void) { orc_value v = {V_NULL, 0, 0.0, 0}; return v; }

/* binop (expr.rs:2037-2104); code as tgp_expr_code ADD..OR. */
static orc_value v_binop(int32_t code, const orc_value* lv, const orc_value* rv) {
  if (code >= 5 && code <= 9) { /* Add Sub Mul Div Mod */
    double l, r;
    if (!v_as_f64(lv, &l) || !v_as_f64(rv, &r)) return v_null();
    double res;
    switch (code) {
      case 5: res = l + r; break;
      case 6: res = l - r; break;
      case 7: res = l * r; break;
      case 8:
        if (r == 0.0) return v_null();
        res = l / r;
        break;
      default:
        if (r == 0.0) return v_null();
        res = fmod(l, r);
        break;
    }
    const int ints = lv->tag == V_INT && rv->tag == V_INT;
    if (ints && code != 8 && code != 9) { /* wrapping_add / _sub / _mul */
      const uint64_t li = (uint64_t)lv->i, ri = (uint64_t)rv->i;
      return v_int((int64_t)(code == 5 ? li + ri : code == 6 ? li - ri : li * ri));
    }
    if (ints && code == 9) { /* li % ri; i64::MIN % -1 (a panic in Rust) is taken as 0 */
      if (rv->i == 0) return v_null();
      return v_int(rv->i == -1 ? 0 : lv->i % rv->i);
    }
    return v_float(res);
  }
  switch (code) {
    case 10: return v_int(v_cmp_eq(lv, rv));
    case 11: return v_int(!v_cmp_eq(lv, rv));
    case 12: return v_int(v_cmp_lt(lv, rv));
    case 13: return v_int(v_cmp_le(lv, rv));
    case 14: return v_int(v_cmp_lt(rv, lv));
    case 15: return v_int(v_cmp_le(rv, lv));
    case 16: return v_int(v_truthy(lv) && v_truthy(rv));
    default: return v_int(v_truthy(lv) || v_truthy(rv));
  }
}

/* eval of column `c` at row (expr.rs:1628-1640): Int = the cell as i64, Float = the
 * f64 bits, Date = the cell's low 32 bits as i32. */
static orc_value v_column(const orc_column* c, int64_t kind, uint64_t row) {
  const int64_t cell = get_native(c->data, c->dtype, row);
  orc_value v = v_null();
  if (kind == V_INT) v = v_int(cell);
  else if (kind == V_FLOAT) { v.tag = V_FLOAT; memcpy(&v.f, &cell, 8); }
  else { v.tag = V_DATE; v.d = (int32_t)(uint32_t)(uint64_t)cell; }
  return v;
}

/* The program's value at one row, as truthy (1) or not (0); -1 for a malformed
 * program.  Column c < nleft is read at row_l of cols[c], column c >= nleft at
 * row_r of cols[c] (one combined row of two tables, join.rs:1276-1296). */
static int eval_prog(const orc_column* cols, int32_t ncols, int32_t nleft, uint64_t row_l,
                     uint64_t row_r, const orc_expr_op* prog, int32_t nops) {
  orc_value st[64];
  int sp = 0;
  for (int32_t k = 0; k < nops; ++k) {
    const orc_expr_op* o = &prog[k];
    switch (o->code) {
      case 0:
        if (o->arg < 0 || o->arg >= ncols) return -1;
        st[sp++] = v_column(&cols[o->arg], o->imm, o->arg < nleft ? row_l : row_r);
        break;
      case 1: st[sp++] = v_int(o->imm); break;
      case 2: { orc_value v = v_null(); v.tag = V_FLOAT; memcpy(&v.f, &o->imm, 8); st[sp++] = v; break; }
      case 3: { orc_value v = v_null(); v.tag = V_DATE; v.d = (int32_t)o->imm; st[sp++] = v; break; }
      case 4: st[sp++] = v_null(); break;
      case 18: st[sp - 1] = v_int(!v_truthy(&st[sp - 1])); break;           /* Not :1741-1744 */
      case 19: {                                                             /* Neg :1733-1740 */
        orc_value v = st[sp - 1];
        if (v.tag == V_INT) st[sp - 1] = v_int((int64_t)(0 - (uint64_t)v.i));
        else if (v.tag == V_FLOAT) st[sp - 1] = v_float(-v.f);
        else st[sp - 1] = v_null();
        break;
      }
      case 20: case 21: {                                                    /* :1773-1779 */
        const int in = v_cmp_le(&st[sp - 2], &st[sp - 3]) && v_cmp_le(&st[sp - 3], &st[sp - 1]);
        sp -= 2;
        st[sp - 1] = v_int(o->code == 20 ? in : !in);
        break;
      }
      case 22: case 23: {                                                    /* :1780-1790 */
        int found = 0;
        for (int32_t j = 0; j < o->arg && !found; ++j)
          found = v_cmp_eq(&st[sp - 1 - o->arg], &st[sp - o->arg + j]);
        sp -= o->arg;
        st[sp - 1] = v_int(o->code == 22 ? found : !found);
        break;
      }
      case 24: {                                                             /* Case :1791-1800 */
        const orc_value r = v_truthy(&st[sp - 3]) ? st[sp - 2] : st[sp - 1];
        sp -= 2;
        st[sp - 1] = r;
        break;
      }
      default:
        if (sp < 2 || o->code < 5 || o->code > 17) return -1;
        st[sp - 2] = v_binop(o->code, &st[sp - 2], &st[sp - 1]);
        --sp;
    }
    if (sp < 1 || sp > 60) return -1;
  }
  if (sp != 1) return -1;
  return v_truthy(&st[0]);
}

/* Clears bit (row - row_begin) of bits (LSB-first bytes) for every row of
 * [row_begin, row_end) whose program value is not truthy (build_mask's
 * fallback, expr.rs:1453-1463).  Returns 0, or -1 for a malformed program. */
int32_t orc_scan_expr(const orc_column* cols, int32_t ncols, const orc_expr_op* prog,
                      int32_t nops, uint64_t row_begin, uint64_t row_end, uint8_t* bits) {
  for (uint64_t row = row_begin; row < row_end; ++row) {
    const uint64_t i = row - row_begin;
    if (!((bits[i >> 3] >> (i & 7)) & 1)) continue;
    const int t = eval_prog(cols, ncols, ncols, row, row, prog, nops);
    if (t < 0) return -1;
    if (!t) bits[i >> 3] &= (uint8_t)~(1u << (i & 7));
  }
  return 0;
}

/* The non-equi ON conjuncts of a hash join applied per match
 * (query_interpreter/join.rs:1146-1195 with eval_non_equi_match :1233-1300): the
 * pairs (probe row, build row) of an equi join in probe order, a probe row's
 * matches consecutive (build row no_row: a LEFT row without a match), are kept
 * when the program, over columns cols[0..nleft) of the probe table at the probe
 * row and cols[nleft..ncols) of the build table at the build row, is truthy.  A
 * LEFT join (left != 0) emits (probe row, no_row) for a probe row whose matches
 * all fail.  Returns the number of output pairs, or -1 for a malformed program. */
int64_t orc_join_residual(const orc_column* cols, int32_t ncols, int32_t nleft,
                          const orc_expr_op* prog, int32_t nops, int32_t left,
                          const uint64_t* probe, const uint32_t* build, uint64_t npairs,
                          uint32_t no_row, uint64_t* out_probe, uint32_t* out_build) {
  uint64_t n = 0;
  for (uint64_t i = 0; i < npairs;) {
    uint64_t j = i;
    while (j < npairs && probe[j] == probe[i]) ++j;  /* the probe row's matches [i, j) */
    int passed = 0;
    for (uint64_t k = i; k < j; ++k) {
      if (build[k] == no_row) {  /* no match: LEFT's row as it is */
        out_probe[n] = probe[k];
        out_build[n++] = no_row;
        passed = 1;
        continue;
      }
      const int t = eval_prog(cols, ncols, nleft, probe[k], build[k], prog, nops);
      if (t < 0) return -1;
      if (t) {
        out_probe[n] = probe[k];
        out_build[n++] = build[k];
        passed = 1;
      }
    }
    if (left && !passed) {
      out_probe[n] = probe[i];
      out_build[n++] = no_row;
    }
    i = j;
  }
  return (int64_t)n;
}
