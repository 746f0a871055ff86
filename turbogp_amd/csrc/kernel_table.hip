// Host-side mirror of the reference's operator interface: the Kernel trait and
// the KernelTable (src/kernel/mod.rs:47-242), with gfx950 kernels registered
// under a device target beside the CPU targets (kernel/cpu.rs:30-41) and the
// memory tier the input lives in (memory/tier.rs:11-29).
//
// Kernel::execute keeps the reference's calling convention: raw `input` /
// `output` pointers laid out as the CPU kernels expect (u64 cells; the HashBuild
// output slot receives a table pointer; the HashProbe input is prefixed by it),
// `KernelParams` and a `KernelResult`.  Inputs that are host pointers are staged
// through HBM; inputs already resident in the HBM tier are used in place.
#include <deque>
#include <memory>
#include <tuple>
#include <vector>

#include "tgp_internal.h"

namespace turbogp {

enum class Operator : int32_t {
  ScanEqU64 = TGP_OP_SCAN_EQ_U64,
  ScanRangeU64 = TGP_OP_SCAN_RANGE_U64,
  ScanMultiPredicate = TGP_OP_SCAN_MULTI_PREDICATE,
  HashBuild = TGP_OP_HASH_BUILD,
  HashProbe = TGP_OP_HASH_PROBE,
  AggregateSumF64 = TGP_OP_AGGREGATE_SUM_F64,
  AggregateCountDistinct = TGP_OP_AGGREGATE_COUNT_DISTINCT,
  SimilarityHamming = TGP_OP_SIMILARITY_HAMMING,
  LeapfrogJoin = TGP_OP_LEAPFROG_JOIN,
};

// A kernel: one operator implementation for one (target, tier).
class Kernel {
 public:
  virtual ~Kernel() = default;
  virtual Operator op() const = 0;
  virtual int32_t target() const { return TGP_TARGET_GFX950; }
  virtual int32_t tier() const = 0;
  virtual const char* name() const = 0;
  virtual int32_t execute(tgp_ctx* ctx, const void* input, void* output,
                          const tgp_kernel_params& p, tgp_kernel_result* r) const = 0;
};

// True when p points into device memory of the HIP runtime.
static bool is_device_ptr(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// Resident or staged u64 cells + a device result word area.
struct Staged {
  const uint64_t* cells = nullptr;
  uint64_t* words = nullptr;  // 2 device words for counts / the first bitmap word
};

// `slot` 1 for operators whose own calls use the stream's slot-0 scratch (the
// f64 sum's limbs, the distinct set), so the staged cells are not overwritten.
static int32_t stage_cells(tgp_ctx* ctx, const void* input, uint64_t n, bool force_copy,
                           Staged* st, int slot = 0) {
  const uint64_t bytes = ((n * 8 + 255) / 256) * 256;
  void* scratch = nullptr;
  TGP_TRY(tgp::get_scratch(ctx, ctx->stream, bytes + 256, &scratch, slot));
  st->words = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(scratch) + bytes);
  const bool dev = input && is_device_ptr(input);
  if (dev && !force_copy && (reinterpret_cast<uintptr_t>(input) & 15) == 0) {
    st->cells = static_cast<const uint64_t*>(input);
  } else if (n) {
    TGP_HIP(hipMemcpyAsync(scratch, input, n * 8,
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    st->cells = static_cast<const uint64_t*>(scratch);
  }
  return TGP_OK;
}

// Count (and first-64 mask) of cells satisfying the AND of preds.
static int32_t scan_cells(tgp_ctx* ctx, const void* input, uint64_t n, const tgp_pred* preds,
                          int npreds, tgp_kernel_result* r) {
  Staged st;
  TGP_TRY(stage_cells(ctx, input, n, false, &st));
  TGP_TRY(tgp::fill_u64(ctx->stream, st.words, 2, 0));
  tgp_column col{st.cells, n, TGP_U64, 0};
  if (n) {
    TGP_TRY(tgp_scan_filter(ctx, &col, 1, preds, npreds, 0, n, nullptr, st.words, ctx->stream));
    TGP_TRY(tgp_scan_filter(ctx, &col, 1, preds, npreds, 0, n < 64 ? n : 64, st.words + 1,
                            nullptr, ctx->stream));
  }
  uint64_t h[2];
  TGP_HIP(hipMemcpyAsync(h, st.words, 16, hipMemcpyDeviceToHost, ctx->stream));
  TGP_HIP(hipStreamSynchronize(ctx->stream));
  r->count = h[0];
  r->sum = 0.0;
  r->mask = h[1];
  return TGP_OK;
}

static int32_t pred_of(int32_t op, uint64_t target, tgp_pred* out) {
  out->col = 0;
  out->is_signed = 0;
  out->_pad = 0;
  out->lit = (int64_t)target;
  switch (op) {  // PredicateOp: Eq, Gt, Lt (kernel/mod.rs:84-92)
    case TGP_PRED_EQ: out->op = TGP_EQ; return TGP_OK;
    case TGP_PRED_GT: out->op = TGP_GT; return TGP_OK;
    case TGP_PRED_LT: out->op = TGP_LT; return TGP_OK;
  }
  tgp::set_error("unknown PredicateOp %d", op);
  return TGP_ERR_INVALID;
}

class ScanEqGfx950 : public Kernel {
 public:
  explicit ScanEqGfx950(int32_t tier) : tier_(tier) {}
  Operator op() const override { return Operator::ScanEqU64; }
  int32_t tier() const override { return tier_; }
  const char* name() const override {
    return tier_ == TGP_TIER_HBM ? "scan_eq_gfx950_hbm" : "scan_eq_gfx950_staged";
  }
  // ScanEqScalar semantics (kernel/scan.rs:26-64): count + mask of the first 64.
  int32_t execute(tgp_ctx* ctx, const void* in, void*, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    tgp_pred pr;
    TGP_TRY(pred_of(TGP_PRED_EQ, p.target_u64, &pr));
    return scan_cells(ctx, in, p.cell_count, &pr, 1, r);
  }

 private:
  int32_t tier_;
};

class ScanRangeGfx950 : public Kernel {
 public:
  explicit ScanRangeGfx950(int32_t tier) : tier_(tier) {}
  Operator op() const override { return Operator::ScanRangeU64; }
  int32_t tier() const override { return tier_; }
  const char* name() const override {
    return tier_ == TGP_TIER_HBM ? "scan_range_gfx950_hbm" : "scan_range_gfx950_staged";
  }
  // ScanRangeScalar (kernel/scan.rs:66-100): low <= c <= high, unsigned.
  int32_t execute(tgp_ctx* ctx, const void* in, void*, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    tgp_pred pr[2] = {{0, TGP_GE, 0, 0, (int64_t)p.low_u64}, {0, TGP_LE, 0, 0, (int64_t)p.high_u64}};
    TGP_TRY(scan_cells(ctx, in, p.cell_count, pr, 2, r));
    r->mask = 0;  // the range kernels report no mask
    return TGP_OK;
  }

 private:
  int32_t tier_;
};

class ScanMultiGfx950 : public Kernel {
 public:
  explicit ScanMultiGfx950(int32_t tier) : tier_(tier) {}
  Operator op() const override { return Operator::ScanMultiPredicate; }
  int32_t tier() const override { return tier_; }
  const char* name() const override {
    return tier_ == TGP_TIER_HBM ? "scan_multi_predicate_gfx950_hbm"
                                 : "scan_multi_predicate_gfx950_staged";
  }
  // ScanMultiPredicateScalar (kernel/scan.rs:126-160): the first
  // min(predicate_count, 3) predicates AND-ed; none => every cell counts.
  int32_t execute(tgp_ctx* ctx, const void* in, void*, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    const int n = p.predicate_count < 3 ? p.predicate_count : 3;
    const uint64_t tg[3] = {p.target_u64, p.target2_u64, p.target3_u64};
    const int32_t ops[3] = {p.pred1_op, p.pred2_op, p.pred3_op};
    tgp_pred pr[3];
    for (int i = 0; i < n; ++i) TGP_TRY(pred_of(ops[i], tg[i], &pr[i]));
    TGP_TRY(scan_cells(ctx, in, p.cell_count, pr, n, r));
    r->mask = 0;
    return TGP_OK;
  }

 private:
  int32_t tier_;
};

class HashBuildGfx950 : public Kernel {
 public:
  Operator op() const override { return Operator::HashBuild; }
  int32_t tier() const override { return TGP_TIER_HBM; }
  const char* name() const override { return "hash_build_gfx950_hbm"; }
  // HashBuildScalar (kernel/hash.rs:118-148): builds over cell_count u64 keys
  // and writes the table handle into the first 8 bytes of `output` (host
  // memory); the caller releases it with tgp_hash_table_destroy.
  int32_t execute(tgp_ctx* ctx, const void* in, void* out, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    TGP_REQUIRE(out != nullptr, "HashBuild: null output slot");
    Staged st;
    TGP_TRY(stage_cells(ctx, in, p.cell_count, false, &st));
    tgp_column keys{st.cells, p.cell_count, TGP_U64, 0};
    tgp_hash_table* t = nullptr;
    TGP_TRY(tgp_hash_build(ctx, &keys, nullptr, &t, ctx->stream));
    *static_cast<tgp_hash_table**>(out) = t;
    r->count = p.cell_count;
    r->sum = 0.0;
    r->mask = 0;
    return TGP_OK;
  }
};

class HashProbeGfx950 : public Kernel {
 public:
  Operator op() const override { return Operator::HashProbe; }
  int32_t tier() const override { return TGP_TIER_HBM; }
  const char* name() const override { return "hash_probe_gfx950_hbm"; }
  // HashProbeScalar (kernel/hash.rs:150-188): input = [table handle][keys];
  // result.count = total number of matches.
  int32_t execute(tgp_ctx* ctx, const void* in, void*, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    TGP_REQUIRE(in != nullptr, "HashProbe: null input");
    const bool dev = is_device_ptr(in);
    tgp_hash_table* t = nullptr;
    if (dev)
      TGP_HIP(hipMemcpy(&t, in, 8, hipMemcpyDeviceToHost));
    else
      std::memcpy(&t, in, 8);
    TGP_REQUIRE(t != nullptr, "HashProbe: null table handle");
    Staged st;
    TGP_TRY(stage_cells(ctx, static_cast<const uint8_t*>(in) + 8, p.cell_count, true, &st));
    TGP_TRY(tgp::fill_u64(ctx->stream, st.words, 1, 0));
    tgp_column keys{st.cells, p.cell_count, TGP_U64, 0};
    TGP_TRY(tgp_hash_probe_count(ctx, t, &keys, 0, p.cell_count, st.words, ctx->stream));
    uint64_t c = 0;
    TGP_HIP(hipMemcpyAsync(&c, st.words, 8, hipMemcpyDeviceToHost, ctx->stream));
    TGP_HIP(hipStreamSynchronize(ctx->stream));
    r->count = c;
    r->sum = 0.0;
    r->mask = 0;
    return TGP_OK;
  }
};

class SumF64Gfx950 : public Kernel {
 public:
  explicit SumF64Gfx950(int32_t tier) : tier_(tier) {}
  Operator op() const override { return Operator::AggregateSumF64; }
  int32_t tier() const override { return tier_; }
  const char* name() const override {
    return tier_ == TGP_TIER_HBM ? "sum_f64_gfx950_hbm" : "sum_f64_gfx950_staged";
  }
  // SumF64Scalar / Avx2 / Avx512 (kernel/aggregate.rs:19-147): input = cell_count
  // f64 bits; result.sum = their sum, result.count = cell_count.  The sum here is
  // exact and rounded once (see tgp_sum_f64_sync); Iterator::sum starts from -0.0.
  int32_t execute(tgp_ctx* ctx, const void* in, void*, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    Staged st;
    TGP_TRY(stage_cells(ctx, in, p.cell_count, false, &st, 1));
    double sum = 0.0;
    TGP_TRY(tgp::sum_f64_exact(ctx, st.cells, p.cell_count, nullptr, true, ctx->stream, &sum));
    r->count = p.cell_count;
    r->sum = sum;
    r->mask = 0;
    return TGP_OK;
  }

 private:
  int32_t tier_;
};

class CountDistinctGfx950 : public Kernel {
 public:
  explicit CountDistinctGfx950(int32_t tier) : tier_(tier) {}
  Operator op() const override { return Operator::AggregateCountDistinct; }
  int32_t tier() const override { return tier_; }
  const char* name() const override {
    return tier_ == TGP_TIER_HBM ? "count_distinct_gfx950_hbm" : "count_distinct_gfx950_staged";
  }
  // CountDistinctScalar (kernel/aggregate.rs:149-185): result.count = the number
  // of distinct u64 cells among the cell_count inputs (an exact device set).
  int32_t execute(tgp_ctx* ctx, const void* in, void*, const tgp_kernel_params& p,
                  tgp_kernel_result* r) const override {
    uint64_t c = 0;
    if (p.cell_count) {
      Staged st;
      TGP_TRY(stage_cells(ctx, in, p.cell_count, false, &st, 1));
      tgp_column col{st.cells, p.cell_count, TGP_U64, 0};
      TGP_TRY(tgp_count_distinct(ctx, &col, 0, p.cell_count, nullptr, TGP_DISTINCT_EXACT, &c,
                                 ctx->stream));
    }
    r->count = c;
    r->sum = 0.0;
    r->mask = 0;
    return TGP_OK;
  }

 private:
  int32_t tier_;
};

// (op, target, tier) -> kernel; select() follows KernelTable::select
// (kernel/mod.rs:209-221) with the device target in the place of the
// detected CPU and the HBM tier as the fallback tier.
class KernelTable {
 public:
  void add(std::shared_ptr<Kernel> k) {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& e : entries_)
      if (e->op() == k->op() && e->target() == k->target() && e->tier() == k->tier()) {
        e = std::move(k);  // register() replaces by key (mod.rs:229-232)
        return;
      }
    entries_.push_back(std::move(k));
  }
  const Kernel* select(Operator op, int32_t tier) const {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& e : entries_)
      if (e->op() == op && e->target() == TGP_TARGET_GFX950 && e->tier() == tier) return e.get();
    for (auto& e : entries_)
      if (e->op() == op && e->target() == TGP_TARGET_GFX950 && e->tier() == TGP_TIER_HBM)
        return e.get();
    for (auto& e : entries_)
      if (e->op() == op) return e.get();
    return nullptr;
  }
  size_t size() const {
    std::lock_guard<std::mutex> lk(mu_);
    return entries_.size();
  }
  const Kernel* at(size_t i) const {
    std::lock_guard<std::mutex> lk(mu_);
    return i < entries_.size() ? entries_[i].get() : nullptr;
  }

 private:
  mutable std::mutex mu_;
  std::deque<std::shared_ptr<Kernel>> entries_;
};

}  // namespace turbogp

struct tgp_kernel_table {
  tgp_ctx* ctx = nullptr;
  turbogp::KernelTable table;
};

// A kernel handle is the table-owned Kernel object itself.
struct tgp_kernel : turbogp::Kernel {};

using namespace turbogp;

static const tgp_kernel* as_handle(const Kernel* k) {
  return reinterpret_cast<const tgp_kernel*>(k);
}
static const Kernel* as_kernel(const tgp_kernel* k) { return reinterpret_cast<const Kernel*>(k); }

// Context of the table a kernel was selected from: kernels are shared across
// tables of one process, so execute() receives the context through a registry.
static std::mutex g_owner_mu;
static std::map<const Kernel*, tgp_ctx*> g_owner;

extern "C" {

int32_t tgp_kernel_table_new(tgp_ctx* ctx, tgp_kernel_table** out) {
  TGP_REQUIRE(ctx && out, "tgp_kernel_table_new: null argument");
  auto* t = new tgp_kernel_table();
  t->ctx = ctx;
  // register_scan_kernels / register_hash_kernels / register_aggregate_kernels
  // (kernel/mod.rs:250-346)
  for (int32_t tier : {TGP_TIER_HBM, TGP_TIER_L3, TGP_TIER_DDR5}) {
    t->table.add(std::make_shared<ScanEqGfx950>(tier));
    t->table.add(std::make_shared<ScanRangeGfx950>(tier));
    t->table.add(std::make_shared<ScanMultiGfx950>(tier));
  }
  t->table.add(std::make_shared<HashBuildGfx950>());
  t->table.add(std::make_shared<HashProbeGfx950>());
  // the reference registers SumF64 under L3 and CountDistinct under Ddr5
  for (int32_t tier : {TGP_TIER_HBM, TGP_TIER_L3}) t->table.add(std::make_shared<SumF64Gfx950>(tier));
  for (int32_t tier : {TGP_TIER_HBM, TGP_TIER_DDR5})
    t->table.add(std::make_shared<CountDistinctGfx950>(tier));
  {
    std::lock_guard<std::mutex> lk(g_owner_mu);
    for (size_t i = 0; i < t->table.size(); ++i) g_owner[t->table.at(i)] = ctx;
  }
  *out = t;
  return TGP_OK;
}

int32_t tgp_kernel_table_destroy(tgp_kernel_table* t) {
  if (!t) return TGP_OK;
  {
    std::lock_guard<std::mutex> lk(g_owner_mu);
    for (size_t i = 0; i < t->table.size(); ++i) g_owner.erase(t->table.at(i));
  }
  delete t;
  return TGP_OK;
}

int32_t tgp_kernel_table_select(tgp_kernel_table* t, int32_t op, int32_t tier, tgp_kernel** out) {
  TGP_REQUIRE(t && out, "tgp_kernel_table_select: null argument");
  *out = const_cast<tgp_kernel*>(as_handle(t->table.select(static_cast<Operator>(op), tier)));
  return TGP_OK;
}

int32_t tgp_kernel_table_len(tgp_kernel_table* t, int32_t* out_n) {
  TGP_REQUIRE(t && out_n, "tgp_kernel_table_len: null argument");
  *out_n = (int32_t)t->table.size();
  return TGP_OK;
}

int32_t tgp_kernel_table_entry(tgp_kernel_table* t, int32_t i, int32_t* op, int32_t* target,
                               int32_t* tier, const char** name) {
  TGP_REQUIRE(t, "tgp_kernel_table_entry: null table");
  const Kernel* k = i >= 0 ? t->table.at((size_t)i) : nullptr;
  TGP_REQUIRE(k != nullptr, "tgp_kernel_table_entry: index %d out of range", i);
  return tgp_kernel_info(as_handle(k), op, target, tier, name);
}

int32_t tgp_kernel_info(const tgp_kernel* kh, int32_t* op, int32_t* target, int32_t* tier,
                        const char** name) {
  TGP_REQUIRE(kh, "tgp_kernel_info: null kernel");
  const Kernel* k = as_kernel(kh);
  if (op) *op = (int32_t)k->op();
  if (target) *target = k->target();
  if (tier) *tier = k->tier();
  if (name) *name = k->name();
  return TGP_OK;
}

int32_t tgp_kernel_execute(const tgp_kernel* kh, const void* input, void* output,
                           const tgp_kernel_params* params, tgp_kernel_result* result) {
  TGP_REQUIRE(kh && params && result, "tgp_kernel_execute: null argument");
  *result = tgp_kernel_result{0, 0.0, 0};
  const Kernel* k = as_kernel(kh);
  tgp_ctx* ctx = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_owner_mu);
    auto it = g_owner.find(k);
    if (it != g_owner.end()) ctx = it->second;
  }
  TGP_REQUIRE(ctx != nullptr, "tgp_kernel_execute: kernel's table was destroyed");
  TGP_TRY(tgp::check_injected(ctx));
  TGP_REQUIRE(params->cell_count == 0 || input != nullptr, "tgp_kernel_execute: null input");
  int32_t rc = k->execute(ctx, input, output, *params, result);
  if (rc != TGP_OK) *result = tgp_kernel_result{0, 0.0, 0};
  return rc;
}

}  // extern "C"
