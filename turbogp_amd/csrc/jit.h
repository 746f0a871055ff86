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

// The kernels of one shape: compiled ahead of time (spec, wave: k_spec<S>, k_wave<S>)
// or at run time (jit, used when spec is null).
struct ShapeKernels {
  const void* spec = nullptr;
  const void* wave = nullptr;
  JitShapeKernels jit;
  int32_t occupancy(bool use_wave, uint32_t lds, int* per_cu) const {
    if (spec) {
      TGP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, use_wave ? wave : spec, kBlock,
                                                           lds));
    } else {
      TGP_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
          per_cu, use_wave ? jit.wave : jit.spec, kBlock, lds));
    }
    return TGP_OK;
  }
  int32_t launch(bool use_wave, uint64_t grid, uint32_t lds, hipStream_t st, void** args) const {
    if (spec) {
      TGP_HIP(hipLaunchKernel(use_wave ? wave : spec, dim3((unsigned)grid), dim3(kBlock), args,
                              lds, st));
    } else {
      TGP_HIP(hipModuleLaunchKernel(use_wave ? jit.wave : jit.spec, (unsigned)grid, 1, 1, kBlock,
                                    1, 1, lds, st, args, nullptr));
    }
    return TGP_OK;
  }
};

// The kernels of shape s: its ahead-of-time ones, else the run-time compiled ones
// (wait: compiled now; else a background compile starts and *ok stays false until
// it is done).  *ok is false for a failed compile too (its error stays cached).
int32_t find_shape_kernels(const Shape& s, bool wait, ShapeKernels* K, bool* ok);
// Launches shape S over A on st, zeroing A's result first unless `zeroed` (the
// caller cleared it).  counter: one device word for the wave form's dynamic chunks
// (null: the stream's scratch, which may allocate).
int32_t launch_shape(tgp_ctx* ctx, const Shape& S, const ShapeKernels& K, SpecArgs A,
                     hipStream_t st, unsigned long long* counter, bool zeroed);

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
