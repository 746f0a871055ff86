// Run-time compiled shape-specialised kernels.
//
// spec.hip compiles the single-query kernels (spec_body / wave_body of
// spec_kernels.h) ahead of time for the benchmark shapes only: each Shape fixes
// column dtypes, which column each predicate, key part and factor reads, and the
// factor forms, and there are far too many combinations to compile them all.
// Every other query that extract_shape can express (any column dtypes and
// order, up to four range predicates, divided key parts, product sums, an input
// selection bitmap) is compiled here on first use with hipRTC from the same
// device headers (embedded in the library at build time), loaded as a module
// on each device that runs it and cached for the life of the process, keyed by
// the Shape's bytes.  Compiling takes a second or two per shape, so a caller
// may instead start it on a background thread and keep using the generic
// engine until it is done (the policy lives in spec.hip).
#include <hip/hiprtc.h>

#include <chrono>
#include <condition_variable>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "jit.h"

namespace tgp {
namespace {

#include "_build/jit_headers.inc"

static_assert(sizeof(Shape) % sizeof(int) == 0, "a Shape is a sequence of ints");
constexpr int kShapeInts = (int)(sizeof(Shape) / sizeof(int));

// The program of one shape: the Shape rebuilt from its ints with a constexpr bit
// cast (no field-by-field spelling to keep in step with the struct), and two
// extern "C" kernels with the ahead-of-time kernels' signatures.
std::string program_source(const Shape& s) {
  int w[kShapeInts];
  std::memcpy(w, &s, sizeof(Shape));
  std::string src = "#include \"spec_kernels.h\"\nnamespace tgp {\nstruct JitShapeInts { int w[";
  src += std::to_string(kShapeInts) + "]; };\nconstexpr JitShapeInts kJitInts = {{";
  for (int i = 0; i < kShapeInts; ++i) {
    if (i) src += ",";
    src += std::to_string(w[i]);
  }
  src +=
      "}};\nconstexpr Shape kJitShape = __builtin_bit_cast(Shape, kJitInts);\n}  // namespace tgp\n"
      "extern \"C\" __global__ void __launch_bounds__(256) tgp_jit_spec(tgp::SpecArgs A) {\n"
      "  tgp::spec_body<tgp::kJitShape>(A);\n}\n"
      "extern \"C\" __global__ void __launch_bounds__(256) tgp_jit_wave(tgp::SpecArgs A, "
      "uint32_t rshift, uint32_t chunk, unsigned long long* counter) {\n"
      "  tgp::wave_body<tgp::kJitShape>(A, rshift, chunk, counter);\n}\n";
  return src;
}

// One shape's compile (shared by the devices) and its module on each device.
struct Entry {
  enum State { kPending, kCompiled, kFailed } state = kPending;
  std::vector<char> code;  // the code object, once compiled
  std::string err;
  struct Loaded {
    hipModule_t mod = nullptr;
    JitShapeKernels k;
  };
  std::map<int, Loaded> dev;
};

// The cache and its lock live for the whole process and are never destroyed: a
// background compile still running while the process exits must not find them
// gone.  Waiting for it at exit instead is not safe: the exit handlers of the
// libraries the compile loads (run before this library's) can leave it stuck.
// jit_drain, called when a context is destroyed, waits for the compiles while
// everything is still alive.
struct JitState {
  std::mutex mu;
  std::condition_variable done;
  std::map<std::string, std::unique_ptr<Entry>> cache;  // key: the Shape's bytes
  JitStats stats;
  int pending = 0;  // background compiles not finished
};
JitState& g_jit = *new JitState;

// Compiles s for `arch` into a code object (no device calls: safe off the
// calling thread).  Returns false with the message in *err on failure.
bool compile(const Shape& s, const std::string& arch, std::vector<char>* code, std::string* err) {
  const std::string src = program_source(s);
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "tgp_jit_shape.hip", kJitNumHeaders,
                          kJitHeaderText, kJitHeaderNames) != HIPRTC_SUCCESS) {
    *err = "hiprtcCreateProgram failed";
    return false;
  }
  const std::string opt_arch = "--gpu-architecture=" + arch;
  // the flags of the ahead-of-time build (Makefile HIPFLAGS) that apply here
  const char* opts[] = {opt_arch.c_str(), "-O3", "-std=c++20", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    *err = "run-time compile of a query shape failed: " + log.substr(0, 1500);
    return false;
  }
  size_t code_n = 0;
  hiprtcGetCodeSize(prog, &code_n);
  code->resize(code_n);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return true;
}

// Runs the compile of entry e (registered as pending) and publishes its result.
void compile_entry(Entry* e, Shape s, std::string arch, bool background) {
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<char> code;
  std::string err;
  const bool ok = compile(s, arch, &code, &err);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::lock_guard<std::mutex> lk(g_jit.mu);
  if (ok) {
    e->code = std::move(code);
    e->state = Entry::kCompiled;
    ++g_jit.stats.compiled;
  } else {
    e->err = std::move(err);
    e->state = Entry::kFailed;
    ++g_jit.stats.failed;
  }
  g_jit.stats.compile_ms += ms;
  if (background) --g_jit.pending;
  g_jit.done.notify_all();
}

}  // namespace

int32_t jit_shape_kernels(const Shape& s, bool wait, JitShapeKernels* out, bool* ready) {
  *ready = false;
  int device = 0;
  TGP_HIP(hipGetDevice(&device));
  const std::string key(reinterpret_cast<const char*>(&s), sizeof(Shape));
  std::unique_lock<std::mutex> lk(g_jit.mu);
  auto it = g_jit.cache.find(key);
  if (it == g_jit.cache.end()) {
    hipDeviceProp_t prop;
    TGP_HIP(hipGetDeviceProperties(&prop, device));
    it = g_jit.cache.emplace(key, std::make_unique<Entry>()).first;
    Entry* e = it->second.get();
    if (wait) {
      lk.unlock();
      compile_entry(e, s, prop.gcnArchName, false);
      lk.lock();
    } else {
      ++g_jit.pending;
      std::thread(compile_entry, e, s, std::string(prop.gcnArchName), true).detach();
    }
  }
  Entry* e = it->second.get();
  if (e->state == Entry::kPending) {
    if (!wait) return TGP_OK;
    g_jit.done.wait(lk, [&] { return e->state != Entry::kPending; });
  }
  if (e->state == Entry::kFailed) {
    set_error("%s", e->err.c_str());
    return TGP_ERR_HIP;
  }
  auto d = e->dev.find(device);
  if (d == e->dev.end()) {
    Entry::Loaded L;
    TGP_HIP(hipModuleLoadData(&L.mod, e->code.data()));
    TGP_HIP(hipModuleGetFunction(&L.k.spec, L.mod, "tgp_jit_spec"));
    TGP_HIP(hipModuleGetFunction(&L.k.wave, L.mod, "tgp_jit_wave"));
    d = e->dev.emplace(device, L).first;
  }
  ++g_jit.stats.launches;
  *out = d->second.k;
  *ready = true;
  return TGP_OK;
}

void jit_drain() {
  std::unique_lock<std::mutex> lk(g_jit.mu);
  g_jit.done.wait_for(lk, std::chrono::minutes(2), [] { return g_jit.pending == 0; });
}

JitStats jit_stats() {
  std::lock_guard<std::mutex> lk(g_jit.mu);
  return g_jit.stats;
}

}  // namespace tgp
