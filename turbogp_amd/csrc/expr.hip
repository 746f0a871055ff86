// f64 column filters and expression predicates.
//
// Reference path replaced:
//   exec/bitmap.rs:479-580, :696-753   filter_{eq,ne,lt,le,gt,ge}_f64,
//                                      filter_eq_f64_epsilon (FLOAT columns)
//   query_interpreter/expr.rs:1453-1463 the per-row eval fallback of build_mask:
//                                      eval(left), eval(right), binop, truthy
//   query_interpreter/expr.rs:1626-1800, :2037-2141  eval / binop / cmp_*
//   query_interpreter/types.rs:351-368  Value2::as_f64 / as_i64
//   query_interpreter/join.rs:1146-1195, :1233-1300  hash_join's non-equi ON conjuncts,
//                                      eval_non_equi_match per match on the combined row
//
// Both narrow or produce selection bitmaps one 64-row word per wave: a lane
// evaluates its row and __ballot gives the word.  The expression arrives as a
// postfix program whose op codes and operand positions are the same for every
// lane, so the interpreter's dispatch is scalar branches; the values (a tag and
// 64 bits) live in a register stack of TGP_MAX_EXPR_DEPTH slots addressed by
// wave-uniform indices.
#include "tgp_internal.h"

namespace tgp {

// ---- f64 filters ---------------------------------------------------------------------------

__device__ __forceinline__ bool f64_keep(int32_t op, int32_t eps, double v, double lit,
                                         double tol) {
  switch (op) {
    case TGP_EQ: return eps ? fabs(v - lit) <= tol : v == lit;
    case TGP_NE: return !(fabs(v - lit) <= tol);
    case TGP_LT: return v < lit;
    case TGP_LE: return v <= lit;
    case TGP_GT: return v > lit;
    default: return v >= lit;
  }
}

// Wave w writes bitmap word w (rows row_begin + 64 w ..) and adds its set bits to *count.
__global__ void __launch_bounds__(256) k_filter_f64(const uint64_t* __restrict__ cells,
                                                    int32_t op, int32_t eps, double lit,
                                                    double tol, uint64_t row_begin, uint64_t n,
                                                    uint64_t* __restrict__ bitmap,
                                                    unsigned long long* count) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwords = (n + 63) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  unsigned long long c = 0;
  for (uint64_t w = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64; w < nwords;
       w += nwaves) {
    const uint64_t i = w * 64 + lane;
    bool keep = false;
    if (i < n) keep = f64_keep(op, eps, __longlong_as_double((long long)cells[row_begin + i]), lit, tol);
    const uint64_t bits = __ballot(keep);
    if (lane == 0) {
      bitmap[w] = bits;
      c += __popcll(bits);
    }
  }
  if (count && lane == 0 && c) atomicAdd(count, c);
}

// ---- expression predicates ---------------------------------------------------------------

constexpr int kExDepth = TGP_MAX_EXPR_DEPTH;
enum : int32_t { kVNull = 0, kVInt = 1, kVFloat = 2, kVDate = 3 };

struct Val {
  uint64_t bits;  // int64, f64 bits, or int32 days (sign-extended)
  int32_t tag;
};

struct ExprProg {
  const void* col[TGP_MAX_COLS];
  int32_t dt[TGP_MAX_COLS];
  tgp_expr_op op[TGP_MAX_EXPR_OPS];
  int32_t nops;
  uint32_t rmask;  // bit c: column c is read at the second row (a join's build row)
};

__device__ __forceinline__ uint64_t cell_bits(const void* base, int32_t dt, uint64_t i) {
  switch (dt) {
    case TGP_I8: return (uint64_t)(int64_t)((const int8_t*)base)[i];
    case TGP_U8: return ((const uint8_t*)base)[i];
    case TGP_I16: return (uint64_t)(int64_t)((const int16_t*)base)[i];
    case TGP_U16: return ((const uint16_t*)base)[i];
    case TGP_I32: return (uint64_t)(int64_t)((const int32_t*)base)[i];
    case TGP_U32: return ((const uint32_t*)base)[i];
    default: return ((const uint64_t*)base)[i];
  }
}

// Value2::as_f64 (None for NULL).
__device__ __forceinline__ bool as_f64(const Val& v, double* out) {
  switch (v.tag) {
    case kVInt: *out = (double)(int64_t)v.bits; return true;
    case kVFloat: *out = __longlong_as_double((long long)v.bits); return true;
    case kVDate: *out = (double)(int32_t)(uint32_t)v.bits; return true;
    default: return false;
  }
}

__device__ __forceinline__ bool cmp_eq(const Val& a, const Val& b) {
  double x, y;
  return as_f64(a, &x) && as_f64(b, &y) && x == y;
}
__device__ __forceinline__ bool cmp_lt(const Val& a, const Val& b) {
  double x, y;
  return as_f64(a, &x) && as_f64(b, &y) && x < y;
}
__device__ __forceinline__ bool cmp_le(const Val& a, const Val& b) {
  return cmp_lt(a, b) || cmp_eq(a, b);
}
__device__ __forceinline__ bool truthy(const Val& v) {
  if (v.tag == kVInt) return v.bits != 0;
  if (v.tag == kVFloat) return __longlong_as_double((long long)v.bits) != 0.0;
  return false;
}
__device__ __forceinline__ Val vint(int64_t x) { return Val{(uint64_t)x, kVInt}; }
__device__ __forceinline__ Val vflt(double x) { return Val{(uint64_t)__double_as_longlong(x), kVFloat}; }
__device__ __forceinline__ Val vbool(bool b) { return vint(b ? 1 : 0); }
__device__ __forceinline__ Val vnull() { return Val{0, kVNull}; }

// binop's arithmetic (expr.rs:2039-2094).
__device__ __forceinline__ Val arith(int32_t code, const Val& l, const Val& r) {
  double lf, rf;
  if (!as_f64(l, &lf) || !as_f64(r, &rf)) return vnull();
  const bool ints = l.tag == kVInt && r.tag == kVInt;
  const int64_t li = (int64_t)l.bits, ri = (int64_t)r.bits;
  switch (code) {
    case TGP_EX_ADD: return ints ? vint((int64_t)((uint64_t)li + (uint64_t)ri)) : vflt(lf + rf);
    case TGP_EX_SUB: return ints ? vint((int64_t)((uint64_t)li - (uint64_t)ri)) : vflt(lf - rf);
    case TGP_EX_MUL: return ints ? vint((int64_t)((uint64_t)li * (uint64_t)ri)) : vflt(lf * rf);
    case TGP_EX_DIV: return rf == 0.0 ? vnull() : vflt(lf / rf);
    default:  // MOD
      if (rf == 0.0) return vnull();
      if (ints) return vint(ri == -1 ? 0 : li % ri);  // MIN % -1 is 0 (Rust panics)
      return vflt(fmod(lf, rf));
  }
}

// Register stack with wave-uniform slot indices: bits and tags in separate
// arrays, read and written through switches (scalar branches, no scratch).
struct Stack {
  uint64_t b[kExDepth];
  int32_t t[kExDepth];
};
static_assert(kExDepth == 8, "the accessors below name 8 slots");
#define TGP_EX_SLOTS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
__device__ __forceinline__ Val sget(const Stack& s, int i) {
  switch (i) {
#define TGP_EX_GET(k) case k: return Val{s.b[k], s.t[k]};
    TGP_EX_SLOTS(TGP_EX_GET)
#undef TGP_EX_GET
  }
  return Val{0, kVNull};
}
__device__ __forceinline__ void sset(Stack& s, int i, const Val& v) {
  switch (i) {
#define TGP_EX_SET(k) case k: s.b[k] = v.bits; s.t[k] = v.tag; break;
    TGP_EX_SLOTS(TGP_EX_SET)
#undef TGP_EX_SET
  }
}
#undef TGP_EX_SLOTS

// The cell of column c (a wave-uniform index) at row `row`.
__device__ __forceinline__ uint64_t col_cell(const ExprProg& P, int c, uint64_t row) {
  switch (c) {
    case 0: return cell_bits(P.col[0], P.dt[0], row);
    case 1: return cell_bits(P.col[1], P.dt[1], row);
    case 2: return cell_bits(P.col[2], P.dt[2], row);
    case 3: return cell_bits(P.col[3], P.dt[3], row);
    case 4: return cell_bits(P.col[4], P.dt[4], row);
    case 5: return cell_bits(P.col[5], P.dt[5], row);
    case 6: return cell_bits(P.col[6], P.dt[6], row);
    default: return cell_bits(P.col[7], P.dt[7], row);
  }
}
static_assert(TGP_MAX_COLS == 8, "col_cell names 8 columns");

// Runs the program (the host checked depths and operands): column c at row_r when
// bit c of P.rmask is set, else at row_l.
__device__ __forceinline__ bool eval_pair(const ExprProg& P, uint64_t row_l, uint64_t row_r) {
  Stack st;
  for (int k = 0; k < kExDepth; ++k) {
    st.b[k] = 0;
    st.t[k] = kVNull;
  }
  int sp = 0;
  for (int k = 0; k < P.nops; ++k) {
    const int32_t code = __builtin_amdgcn_readfirstlane(P.op[k].code);
    const int32_t arg = __builtin_amdgcn_readfirstlane(P.op[k].arg);
    const int64_t imm = P.op[k].imm;
    switch (code) {
      case TGP_EX_COL: {
        const int32_t kind = (int32_t)imm;
        const uint64_t b = col_cell(P, arg, ((P.rmask >> arg) & 1) ? row_r : row_l);
        sset(st, sp++, Val{kind == kVDate ? (uint64_t)(int64_t)(int32_t)(uint32_t)b : b, kind});
        break;
      }
      case TGP_EX_INT: sset(st, sp++, vint(imm)); break;
      case TGP_EX_FLOAT: sset(st, sp++, Val{(uint64_t)imm, kVFloat}); break;
      case TGP_EX_DATE: sset(st, sp++, Val{(uint64_t)(int64_t)(int32_t)imm, kVDate}); break;
      case TGP_EX_NULL: sset(st, sp++, vnull()); break;
      case TGP_EX_NOT: sset(st, sp - 1, vbool(!truthy(sget(st, sp - 1)))); break;
      case TGP_EX_NEG: {
        const Val v = sget(st, sp - 1);
        Val r = vnull();
        if (v.tag == kVInt) r = vint((int64_t)(0ull - v.bits));
        else if (v.tag == kVFloat) r = vflt(-__longlong_as_double((long long)v.bits));
        sset(st, sp - 1, r);
        break;
      }
      case TGP_EX_BETWEEN:
      case TGP_EX_NOT_BETWEEN: {
        const Val v = sget(st, sp - 3), lo = sget(st, sp - 2), hi = sget(st, sp - 1);
        const bool in = cmp_le(lo, v) && cmp_le(v, hi);
        sp -= 2;
        sset(st, sp - 1, vbool(code == TGP_EX_BETWEEN ? in : !in));
        break;
      }
      case TGP_EX_IN:
      case TGP_EX_NOT_IN: {
        const Val v = sget(st, sp - 1 - arg);
        bool found = false;
        for (int j = 0; j < arg; ++j) found = found || cmp_eq(v, sget(st, sp - arg + j));
        sp -= arg;
        sset(st, sp - 1, vbool(code == TGP_EX_IN ? found : !found));
        break;
      }
      case TGP_EX_IF: {
        const Val c = sget(st, sp - 3), a = sget(st, sp - 2), b = sget(st, sp - 1);
        sp -= 2;
        sset(st, sp - 1, truthy(c) ? a : b);
        break;
      }
      default: {  // binary operators
        const Val l = sget(st, sp - 2), r = sget(st, sp - 1);
        Val o;
        switch (code) {
          case TGP_EX_EQ: o = vbool(cmp_eq(l, r)); break;
          case TGP_EX_NE: o = vbool(!cmp_eq(l, r)); break;
          case TGP_EX_LT: o = vbool(cmp_lt(l, r)); break;
          case TGP_EX_LE: o = vbool(cmp_le(l, r)); break;
          case TGP_EX_GT: o = vbool(cmp_lt(r, l)); break;
          case TGP_EX_GE: o = vbool(cmp_le(r, l)); break;
          case TGP_EX_AND: o = vbool(truthy(l) && truthy(r)); break;
          case TGP_EX_OR: o = vbool(truthy(l) || truthy(r)); break;
          default: o = arith(code, l, r); break;
        }
        --sp;
        sset(st, sp - 1, o);
        break;
      }
    }
  }
  return truthy(sget(st, sp - 1));
}

__device__ __forceinline__ bool eval_row(const ExprProg& P, uint64_t row) {
  return eval_pair(P, row, row);
}

// Wave w narrows bitmap word w; a word with no bit set is skipped.
__global__ void __launch_bounds__(256) k_scan_expr(ExprProg P, uint64_t row_begin, uint64_t n,
                                                   uint64_t* __restrict__ bitmap) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwords = (n + 63) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t w = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64; w < nwords;
       w += nwaves) {
    // the word is the same in every lane; readfirstlane takes 32 bits at a time
    const uint64_t raw = bitmap[w];
    const uint64_t word = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(raw >> 32)) << 32) |
                          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)raw);
    if (!word) continue;
    const uint64_t i = w * 64 + lane;
    const bool sel = i < n && ((word >> lane) & 1);
    // every lane runs the program (on a row of the range) so the dispatch stays uniform
    const bool keep = eval_row(P, row_begin + (i < n ? i : n - 1)) && sel;
    const uint64_t bits = __ballot(keep);
    if (lane == 0) bitmap[w] = bits;
  }
}

// ---- residual (non-equi) join conditions ---------------------------------------------
// The pairs (probe row, build row) of an equi join, in probe order with a probe
// row's matches consecutive (build row kNoMatch: a LEFT row without any), are
// filtered by the residual program:
//   k_residual_eval   verdict of each pair (kPass, kKeepAsIs for kNoMatch, kFail);
//                     for LEFT, row_pass marks the probe rows with a passing pair;
//   k_residual_bits   the kept pairs as a bitmap: kPass and kKeepAsIs pairs, and
//                     under LEFT the last pair of a probe row none of whose pairs
//                     passed (it becomes that row's (row, kNoMatch) pair);
//   compaction of the bitmap to pair indices, then k_residual_write.
constexpr uint32_t kNoMatch = 0xffffffffu;
constexpr uint8_t kFail = 0, kKeepAsIs = 1, kPass = 2;

__global__ void __launch_bounds__(256) k_residual_eval(
    ExprProg P, const uint64_t* __restrict__ probe, const uint32_t* __restrict__ build,
    uint64_t n, uint64_t row_begin, uint64_t row_end, uint64_t build_rows, int32_t left,
    uint8_t* __restrict__ verdict, uint8_t* __restrict__ row_pass,
    unsigned long long* __restrict__ bad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t pr = probe[i];
    const uint32_t b = build[i];
    uint8_t v = kFail;
    if (pr < row_begin || pr >= row_end || (b != kNoMatch && b >= build_rows)) {
      atomicAdd(bad, 1ull);  // rows outside the tables: the call fails
    } else if (b == kNoMatch) {
      v = kKeepAsIs;
    } else if (eval_pair(P, pr, b)) {
      v = kPass;
      if (left) row_pass[pr - row_begin] = 1;
    }
    verdict[i] = v;
  }
}

__global__ void __launch_bounds__(256) k_residual_bits(
    const uint64_t* __restrict__ probe, const uint8_t* __restrict__ verdict, uint64_t n,
    uint64_t row_begin, uint64_t row_end, int32_t left, const uint8_t* __restrict__ row_pass,
    uint64_t* __restrict__ bits) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwords = (n + 63) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t w = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64; w < nwords;
       w += nwaves) {
    const uint64_t i = w * 64 + lane;
    bool keep = false;
    if (i < n) {
      const uint8_t v = verdict[i];
      keep = v != kFail;
      if (!keep && left) {
        const uint64_t pr = probe[i];
        const bool last = i + 1 == n || probe[i + 1] != pr;
        // a row outside the range failed the call already; it gets no pair
        keep = last && pr >= row_begin && pr < row_end && !row_pass[pr - row_begin];
      }
    }
    const uint64_t word = __ballot(keep);
    if (lane == 0) bits[w] = word;
  }
}

__global__ void __launch_bounds__(256) k_residual_write(
    const uint64_t* __restrict__ idx, uint64_t m, const uint64_t* __restrict__ probe,
    const uint32_t* __restrict__ build, const uint8_t* __restrict__ verdict,
    uint64_t* __restrict__ out_probe, uint32_t* __restrict__ out_build) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = idx[k];
    out_probe[k] = probe[i];
    out_build[k] = verdict[i] == kPass ? build[i] : kNoMatch;
  }
}

// Checks a postfix program against ncols columns and fills P's ops.
static int32_t load_prog(const char* what, const tgp_column* cols, int32_t ncols,
                         const tgp_expr_op* prog, int32_t nops, ExprProg* P) {
  TGP_REQUIRE(nops >= 1 && nops <= TGP_MAX_EXPR_OPS, "%s: %d ops (1..%d)", what, nops,
              TGP_MAX_EXPR_OPS);
  int depth = 0;
  for (int k = 0; k < nops; ++k) {
    const tgp_expr_op& o = prog[k];
    int pops = 0, pushes = 1;
    switch (o.code) {
      case TGP_EX_COL:
        TGP_REQUIRE(o.arg >= 0 && o.arg < ncols, "%s: op %d reads column %d of %d", what, k,
                    o.arg, ncols);
        TGP_REQUIRE(o.imm == kVInt || o.imm == kVFloat || o.imm == kVDate,
                    "%s: op %d: column kind %lld (TGP_EX_INT, _FLOAT or _DATE)", what, k,
                    (long long)o.imm);
        TGP_REQUIRE(o.imm != kVFloat || dtype_width(cols[o.arg].dtype) == 8,
                    "%s: op %d: a FLOAT column holds f64 bits in 8-byte cells", what, k);
        break;
      case TGP_EX_INT: case TGP_EX_FLOAT: case TGP_EX_DATE: case TGP_EX_NULL: break;
      case TGP_EX_NOT: case TGP_EX_NEG: pops = 1; break;
      case TGP_EX_BETWEEN: case TGP_EX_NOT_BETWEEN: case TGP_EX_IF: pops = 3; break;
      case TGP_EX_IN: case TGP_EX_NOT_IN:
        TGP_REQUIRE(o.arg >= 1, "%s: op %d: IN list of %d values", what, k, o.arg);
        pops = o.arg + 1;
        break;
      default:
        TGP_REQUIRE(o.code >= TGP_EX_ADD && o.code <= TGP_EX_OR, "%s: op %d: code %d", what, k,
                    o.code);
        pops = 2;
    }
    TGP_REQUIRE(depth >= pops, "%s: op %d takes %d values, %d on the stack", what, k, pops,
                depth);
    depth += pushes - pops;
    TGP_REQUIRE(depth <= TGP_MAX_EXPR_DEPTH, "%s: stack deeper than %d at op %d", what,
                TGP_MAX_EXPR_DEPTH, k);
    P->op[k] = o;
  }
  TGP_REQUIRE(depth == 1, "%s: the program leaves %d values (1 expected)", what, depth);
  P->nops = nops;
  return TGP_OK;
}

static unsigned grid_words(uint64_t nwords) {
  uint64_t g = (nwords + 3) / 4;  // 4 waves per workgroup
  if (g > 8192) g = 8192;
  return (unsigned)(g ? g : 1);
}

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_scan_filter_f64(tgp_ctx* ctx, const tgp_column* col, int32_t op, int32_t eq_epsilon,
                            double lit, uint64_t row_begin, uint64_t row_end, uint64_t* d_bitmap,
                            uint64_t* d_count, void* stream) {
  TGP_REQUIRE(ctx, "tgp_scan_filter_f64: null ctx");
  TGP_TRY(check_injected(ctx));
  TGP_TRY(check_column(col, "tgp_scan_filter_f64 column"));
  TGP_REQUIRE(col->dtype == TGP_U64 || col->dtype == TGP_I64,
              "tgp_scan_filter_f64: the column holds f64 bits in 8-byte cells (dtype %d)",
              col->dtype);
  TGP_REQUIRE(op >= TGP_EQ && op <= TGP_GE, "tgp_scan_filter_f64: op %d", op);
  TGP_REQUIRE(row_begin <= row_end && row_end <= col->len,
              "tgp_scan_filter_f64: rows [%llu, %llu) outside the column",
              (unsigned long long)row_begin, (unsigned long long)row_end);
  const uint64_t n = row_end - row_begin;
  if (n == 0) return TGP_OK;
  TGP_REQUIRE(d_bitmap, "tgp_scan_filter_f64: null bitmap");
  // filter_eq_f64_epsilon's tolerance, computed as the reference does
  const double tol = 1e-6 * (std::fabs(lit) > 1.0 ? std::fabs(lit) : 1.0);
  const uint64_t nwords = (n + 63) / 64;
  k_filter_f64<<<grid_words(nwords), 256, 0, pick_stream(ctx, stream)>>>(
      static_cast<const uint64_t*>(col->data), op, eq_epsilon, lit, tol, row_begin, n, d_bitmap,
      reinterpret_cast<unsigned long long*>(d_count));
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_scan_expr(tgp_ctx* ctx, const tgp_column* cols, int32_t ncols,
                      const tgp_expr_op* prog, int32_t nops, uint64_t row_begin,
                      uint64_t row_end, uint64_t* d_bitmap, void* stream) {
  TGP_REQUIRE(ctx && prog, "tgp_scan_expr: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_REQUIRE(ncols >= 0 && ncols <= TGP_MAX_COLS, "tgp_scan_expr: %d columns (0..%d)", ncols,
              TGP_MAX_COLS);
  ExprProg P{};
  for (int c = 0; c < ncols; ++c) {
    TGP_TRY(check_column(&cols[c], "tgp_scan_expr column"));
    TGP_REQUIRE(row_end <= cols[c].len, "tgp_scan_expr: column %d has %llu rows, row_end %llu", c,
                (unsigned long long)cols[c].len, (unsigned long long)row_end);
    P.col[c] = cols[c].data;
    P.dt[c] = cols[c].dtype;
  }
  TGP_TRY(load_prog("tgp_scan_expr", cols, ncols, prog, nops, &P));
  TGP_REQUIRE(row_begin <= row_end, "tgp_scan_expr: row_begin > row_end");
  const uint64_t n = row_end - row_begin;
  if (n == 0) return TGP_OK;
  TGP_REQUIRE(d_bitmap, "tgp_scan_expr: null bitmap");
  k_scan_expr<<<grid_words((n + 63) / 64), 256, 0, pick_stream(ctx, stream)>>>(P, row_begin, n,
                                                                             d_bitmap);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_join_residual_sync(tgp_ctx* ctx, const tgp_column* left_cols, int32_t nleft,
                               const tgp_column* right_cols, int32_t nright,
                               const tgp_expr_op* prog, int32_t nops, int32_t join_type,
                               uint64_t row_begin, uint64_t row_end,
                               const uint64_t* d_probe_idx, const uint32_t* d_build_row,
                               uint64_t npairs, uint64_t* d_out_probe_idx,
                               uint32_t* d_out_build_row, uint64_t* h_nout, void* stream) {
  TGP_REQUIRE(ctx && prog && h_nout, "tgp_join_residual_sync: null argument");
  TGP_TRY(check_injected(ctx));
  *h_nout = 0;
  TGP_REQUIRE(join_type == TGP_JOIN_INNER || join_type == TGP_JOIN_LEFT,
              "tgp_join_residual_sync: join type %d (INNER or LEFT)", join_type);
  TGP_REQUIRE(nleft >= 0 && nright >= 0 && nleft + nright <= TGP_MAX_COLS,
              "tgp_join_residual_sync: %d + %d columns (at most %d)", nleft, nright,
              TGP_MAX_COLS);
  TGP_REQUIRE(row_begin <= row_end, "tgp_join_residual_sync: row_begin > row_end");
  ExprProg P{};
  tgp_column cols[TGP_MAX_COLS];
  uint64_t build_rows = ~0ull;
  for (int c = 0; c < nleft + nright; ++c) {
    const bool right = c >= nleft;
    cols[c] = right ? right_cols[c - nleft] : left_cols[c];
    TGP_TRY(check_column(&cols[c], "tgp_join_residual_sync column"));
    if (right) {
      if (cols[c].len < build_rows) build_rows = cols[c].len;
      P.rmask |= 1u << c;
    } else {
      TGP_REQUIRE(row_end <= cols[c].len,
                  "tgp_join_residual_sync: probe column %d has %llu rows, row_end %llu", c,
                  (unsigned long long)cols[c].len, (unsigned long long)row_end);
    }
    P.col[c] = cols[c].data;
    P.dt[c] = cols[c].dtype;
  }
  TGP_TRY(load_prog("tgp_join_residual_sync", cols, nleft + nright, prog, nops, &P));
  if (npairs == 0) return TGP_OK;
  TGP_REQUIRE(d_probe_idx && d_build_row && d_out_probe_idx && d_out_build_row,
              "tgp_join_residual_sync: null pair array");
  const int32_t left = join_type == TGP_JOIN_LEFT;
  hipStream_t s = pick_stream(ctx, stream);
  const uint64_t nprobe = row_end - row_begin, nwords = (npairs + 63) / 64;
  // slot-1 scratch (the compaction uses slot 0): bad counter | bitmap | indices |
  // row_pass (whole words) | verdicts
  const uint64_t pass_words = left ? (nprobe + 7) / 8 : 0;
  const uint64_t bytes = (1 + nwords + npairs + pass_words) * 8 + npairs;
  void* scr = nullptr;
  TGP_TRY(get_scratch(ctx, s, bytes, &scr, 1));
  auto* bad = static_cast<unsigned long long*>(scr);
  auto* bits = reinterpret_cast<uint64_t*>(bad + 1);
  uint64_t* idx = bits + nwords;
  uint64_t* pass_w = idx + npairs;
  auto* row_pass = reinterpret_cast<uint8_t*>(pass_w);
  auto* verdict = reinterpret_cast<uint8_t*>(pass_w + pass_words);
  TGP_TRY(fill_u64(s, reinterpret_cast<uint64_t*>(bad), 1, 0));
  if (pass_words) TGP_TRY(fill_u64(s, pass_w, pass_words, 0));
  const uint64_t grid = std::min<uint64_t>((npairs + 255) / 256, 8192);
  k_residual_eval<<<(unsigned)grid, 256, 0, s>>>(P, d_probe_idx, d_build_row, npairs, row_begin,
                                                 row_end, build_rows, left, verdict, row_pass,
                                                 bad);
  TGP_HIP(hipGetLastError());
  k_residual_bits<<<grid_words(nwords), 256, 0, s>>>(d_probe_idx, verdict, npairs, row_begin,
                                                     row_end, left, row_pass, bits);
  TGP_HIP(hipGetLastError());
  uint64_t total = 0;
  TGP_TRY(compact_begin(ctx, s, bits, npairs, &total));
  uint64_t nbad = 0;
  TGP_HIP(hipMemcpyAsync(&nbad, bad, 8, hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  TGP_REQUIRE(nbad == 0,
              "tgp_join_residual_sync: %llu pairs name rows outside [row_begin, row_end) or "
              "beyond the build columns", (unsigned long long)nbad);
  if (total) {
    TGP_TRY(compact_finish(ctx, s, bits, npairs, 0, 8, idx));
    k_residual_write<<<(unsigned)std::min<uint64_t>((total + 255) / 256, 8192), 256, 0, s>>>(
        idx, total, d_probe_idx, d_build_row, verdict, d_out_probe_idx, d_out_build_row);
    TGP_HIP(hipGetLastError());
    TGP_HIP(hipStreamSynchronize(s));
  }
  *h_nout = total;
  return TGP_OK;
}

}  // extern "C"
