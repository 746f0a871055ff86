// Run-time compiled shape-specialised kernels (jit.hip).
#pragma once

#include "spec_kernels.h"

namespace tgp {

// The two kernels of one shape: tgp_jit_spec(SpecArgs) and
// tgp_jit_wave(SpecArgs, uint32_t rshift, uint32_t chunk, unsigned long long*
// counter), the signatures of k_spec<S> and k_wave<S>.
struct JitShapeKernels {
  hipFunction_t spec = nullptr;
  hipFunction_t wave = nullptr;
};

struct JitStats {
  uint64_t compiled = 0, failed = 0, launches = 0;
  double compile_ms = 0;
};

// The kernels of shape s on the current device (*ready), compiled and loaded on
// first use.  wait = false: a shape not compiled yet starts compiling on a
// background thread and the call returns TGP_OK with *ready false.  Thread
// safe; a failed compile is remembered and its error returned again.
int32_t jit_shape_kernels(const Shape& s, bool wait, JitShapeKernels* out, bool* ready);
JitStats jit_stats();
// Waits (at most two minutes) for the background compiles started so far.
void jit_drain();

}  // namespace tgp
