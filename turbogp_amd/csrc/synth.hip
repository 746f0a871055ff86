// Seeded, counter-based synthetic columns generated on the device.
//
// Row i of column c is a pure function of (seed, c, i), so any row range of a
// table can be produced independently and bit-identically by the CPU oracle.
#include "tgp_internal.h"

namespace tgp {

// SplitMix64 finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014); a bijection on 64-bit words.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ void store_native(uint8_t* base, int32_t dtype, uint64_t i, int64_t v) {
  switch (dtype) {
    case TGP_I8: case TGP_U8: base[i] = (uint8_t)v; break;
    case TGP_I16: case TGP_U16: reinterpret_cast<uint16_t*>(base)[i] = (uint16_t)v; break;
    case TGP_I32: case TGP_U32: reinterpret_cast<uint32_t*>(base)[i] = (uint32_t)v; break;
    default: reinterpret_cast<uint64_t*>(base)[i] = (uint64_t)v; break;
  }
}

__global__ void __launch_bounds__(256) k_synth_uniform(uint8_t* out, int32_t dtype, uint64_t key,
                                                       int64_t lo, uint64_t range, uint64_t b,
                                                       uint64_t e) {
  for (uint64_t i = b + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < e;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = mix64(key ^ i);
    // range == 0 encodes the full 2^64 span
    const int64_t v = range ? lo + (int64_t)__umul64hi(u, range) : (int64_t)u;
    store_native(out, dtype, i, v);
  }
}

__global__ void __launch_bounds__(256) k_synth_cdf(uint8_t* out, int32_t dtype, uint64_t key,
                                                   const double* __restrict__ cdf, uint64_t n,
                                                   const int64_t* __restrict__ table, uint64_t b,
                                                   uint64_t e) {
  for (uint64_t i = b + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < e;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = mix64(key ^ i);
    const double x = (double)(u >> 11) * 0x1.0p-53;
    uint64_t lo = 0, hi = n;  // count of cdf entries <= x
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (cdf[mid] <= x) lo = mid + 1; else hi = mid;
    }
    if (lo >= n) lo = n - 1;
    store_native(out, dtype, i, table[lo]);
  }
}

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_synth_uniform(tgp_ctx* ctx, tgp_column* col, uint64_t seed, uint64_t column_id,
                          int64_t lo, int64_t hi, uint64_t row_begin, uint64_t row_end,
                          void* stream) {
  TGP_REQUIRE(ctx && col && col->data, "tgp_synth_uniform: null argument");
  TGP_REQUIRE(dtype_width(col->dtype) > 0, "tgp_synth_uniform: bad dtype");
  TGP_REQUIRE(lo <= hi, "tgp_synth_uniform: lo > hi");
  TGP_REQUIRE(row_begin <= row_end && row_end <= col->len, "tgp_synth_uniform: bad row range");
  if (row_begin == row_end) return TGP_OK;
  const uint64_t range = (uint64_t)hi - (uint64_t)lo + 1;  // wraps to 0 for the full span
  const uint64_t key = seed ^ (column_id << 40);
  uint64_t g = (row_end - row_begin + 255) / 256;
  if (g > 16384) g = 16384;
  k_synth_uniform<<<(unsigned)g, 256, 0, pick_stream(ctx, stream)>>>(
      (uint8_t*)col->data, col->dtype, key, lo, range, row_begin, row_end);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

int32_t tgp_synth_lookup_cdf(tgp_ctx* ctx, tgp_column* out, uint64_t seed, uint64_t column_id,
                             const double* d_cdf, uint64_t n_cdf, const int64_t* d_table,
                             uint64_t row_begin, uint64_t row_end, void* stream) {
  TGP_REQUIRE(ctx && out && out->data && d_cdf && d_table && n_cdf > 0,
              "tgp_synth_lookup_cdf: null argument");
  TGP_REQUIRE(dtype_width(out->dtype) > 0, "tgp_synth_lookup_cdf: bad dtype");
  TGP_REQUIRE(row_begin <= row_end && row_end <= out->len, "tgp_synth_lookup_cdf: bad row range");
  if (row_begin == row_end) return TGP_OK;
  const uint64_t key = seed ^ (column_id << 40);
  uint64_t g = (row_end - row_begin + 255) / 256;
  if (g > 16384) g = 16384;
  k_synth_cdf<<<(unsigned)g, 256, 0, pick_stream(ctx, stream)>>>(
      (uint8_t*)out->data, out->dtype, key, d_cdf, n_cdf, d_table, row_begin, row_end);
  TGP_HIP(hipGetLastError());
  return TGP_OK;
}

}  // extern "C"
